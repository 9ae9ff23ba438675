// Kernel 3: BAM record chain across block seams, and kernel 4: record-to-column scatter.
//
// The reference tests every inflated byte offset with is_record (one thread per byte,
// reading the name and CIGAR at each), radix-sorts the candidates and chains them
// with pointer jumping (src/cubam/bam.py:403-418, src/cubam/_kernels.py:250-289,
// 676-702). Here the inflated data is cut into segments of REC_SEG bytes; a wavefront
// stages its segment in LDS, finds the first offset of it that passes the structural test
// (confirmed by the record after it), walks block_size from there over the records that
// start in the segment, and records where it left. The walks chain when every exit is the
// start found in the segment it lands in (segment 0 starts at the known record
// boundary), which a second pass checks; a walk that does not chain is repaired by
// re-walking it from the true exit of the walk before it.
#pragma once
#include "common.h"

namespace cbm {

// 7 KiB: four wavefronts' stages (segment, overlap) fit five workgroups a CU; 8 KiB fit four
// and walked 4% slower, 6 KiB walked 5% slower
#ifndef CBM_REC_SEG
#define CBM_REC_SEG 7168
#endif
constexpr int64_t REC_SEG = CBM_REC_SEG;
constexpr int REC_NT = 256;

enum { SEG_END_DATA = 1, SEG_END_INVALID = 2 };

// per-segment state, each an int64 array of nseg entries
struct RecWs {
    int64_t* f;     // walk entry, -1 = no record starts in the segment
    int64_t* e;     // walk exit
    int64_t* cnt;   // complete records
    int64_t* cig;   // sum n_cigar_op
    int64_t* seqb;  // sum ceil(l_seq/2)
    int64_t* qual;  // sum l_seq
    int64_t* flag;  // SEG_END_*
    int64_t* base;  // exclusive scan of cnt [nseg + 1]
    int64_t* pred;  // walks that land on each segment
    int64_t* hdr;   // [0] bad links [1..7] spare
    int64_t* nrel;  // record starts the walk recorded in rel, or -1 (rec_emit walks the segment again)
    uint16_t* rel;  // [nseg][REC_MAXR]: each record start of the walk, less its first
    uint32_t* pre;  // [nseg][REC_MAXR]: the sums over the walk's records before each (pre_pack)
    int64_t* cig_base;   // exclusive scans over the segments (walks off the chain count 0) [nseg + 1]
    int64_t* seq_base;   //   of seqb
    int64_t* qual_base;  //   of qual
    void* scan_ws;
    int64_t nseg;
};

// record starts a segment's walk records: a record takes at least 4 + 32 + 1 bytes, so at most
// REC_SEG / 37 + 1 of them start in a segment
constexpr int REC_MAXR = (int)(REC_SEG / 37 + 2);

// The records before record r of a walk lie inside its segment (each ends where the next starts,
// at or before r), and a record holds its 4 * n_cigar_op CIGAR bytes and l_seq quality bytes
// (the structural test checks both fit in block_size): over them n_cigar_op sums to less than
// REC_SEG / 4 and l_seq to less than REC_SEG, and at most REC_MAXR have an odd l_seq. Packed:
// bits 0-10 the n_cigar_op sum, 11-23 the l_seq sum, 24-31 the odd l_seq count.
static_assert(REC_SEG <= 8192 && REC_MAXR < 256, "the packed in-segment sums need REC_SEG <= 8 KiB");
__host__ __device__ constexpr uint32_t pre_pack(uint32_t cig, uint32_t qual, uint32_t odd) {
    return cig | (qual << 11) | (odd << 24);
}
__device__ __forceinline__ uint32_t pre_cig(uint32_t p) { return p & 0x7ffu; }
__device__ __forceinline__ uint32_t pre_qual(uint32_t p) { return (p >> 11) & 0x1fffu; }
// bases before the record: sum of ceil(l_seq / 2) = (sum of l_seq + odd ones) / 2
__device__ __forceinline__ uint32_t pre_seqb(uint32_t p) { return (pre_qual(p) + (p >> 24)) >> 1; }

inline int64_t rec_nseg(int64_t len) { return (len + REC_SEG - 1) / REC_SEG; }

inline RecWs rec_ws_layout_seg(void* ws, int64_t nseg) {
    RecWs w;
    int64_t* p = (int64_t*)ws;
    w.nseg = nseg;
    w.hdr = p;
    p += 8;
    w.f = p; p += nseg;
    w.e = p; p += nseg;
    w.cnt = p; p += nseg;
    w.cig = p; p += nseg;
    w.seqb = p; p += nseg;
    w.qual = p; p += nseg;
    w.flag = p; p += nseg;
    w.pred = p; p += nseg;
    w.base = p; p += nseg + 1;
    w.nrel = p; p += nseg;
    w.rel = (uint16_t*)p; p += (nseg * REC_MAXR + 3) / 4;
    w.pre = (uint32_t*)p; p += (nseg * REC_MAXR + 1) / 2;
    w.cig_base = p; p += nseg + 1;
    w.seq_base = p; p += nseg + 1;
    w.qual_base = p; p += nseg + 1;
    w.scan_ws = p;
    return w;
}

inline RecWs rec_ws_layout(void* ws, int64_t len) { return rec_ws_layout_seg(ws, rec_nseg(len)); }

inline int64_t rec_ws_bytes_seg(int64_t nseg) {
    return 8 * (8 + 10 * nseg + 1) + 8 * ((nseg * REC_MAXR + 3) / 4) + 8 * ((nseg * REC_MAXR + 1) / 2) +
           8 * 3 * (nseg + 1) + scan_workspace(nseg) + 64;
}

inline int64_t rec_ws_bytes(int64_t len) { return rec_ws_bytes_seg(rec_nseg(len)); }

// a record start confirmed by the record after it (or by the end of the data)
template <class A>
__device__ __forceinline__ bool record_start_at(const A& a, int64_t n, int64_t o, int n_ref) {
    if (!is_record_at(a, n, o, n_ref)) return false;
    int64_t nx = o + 4 + (int64_t)a.u32(o);
    if (nx + 36 > n) return true;
    if (nx + 4 + (int64_t)a.u32(nx) > n) return true;
    return is_record_at(a, n, nx, n_ref);
}

struct WalkOut {
    int64_t e, cnt, cig, seqb, qual, flag, odd;
};

// records that start in [o, hi) from o, a record start
__device__ __forceinline__ WalkOut walk(const uint8_t* b, int64_t n, int64_t o, int64_t hi, int n_ref) {
    WalkOut w = {o, 0, 0, 0, 0, 0, 0};
    while (o < hi) {
        if (o + 4 > n) { w.flag = SEG_END_DATA; break; }
        int64_t bs = (int64_t)ld_u32(b, o);
        if (o + 4 + bs > n) { w.flag = SEG_END_DATA; break; }
        if (!is_record(b, n, o, n_ref)) { w.flag = SEG_END_INVALID; break; }
        uint32_t lseq = ld_u32(b, o + 20);
        w.cnt++;
        w.cig += ld_u16(b, o + 16);
        w.seqb += (lseq + 1) / 2;
        w.qual += lseq;
        o += 4 + bs;
    }
    if (o >= n && !w.flag) w.flag = SEG_END_DATA;
    w.e = o;
    return w;
}

// Each walk that does not end the data must land on the start found in that segment,
// with no walk in the segments it skips; count the walks landing on each segment.
__global__ void __launch_bounds__(REC_NT) rec_links(int64_t start, RecWs w) {
    int64_t k = (int64_t)blockIdx.x * REC_NT + threadIdx.x;
    int64_t bad = 0;
    if (k < w.nseg && w.f[k] >= 0 && !w.flag[k]) {
        int64_t j = (w.e[k] - start) / REC_SEG;
        if (j >= w.nseg || w.f[j] != w.e[k]) bad++;
        else {
            for (int64_t i = k + 1; i < j; i++)
                if (w.f[i] >= 0) bad++;
            atomicAdd((unsigned long long*)&w.pred[j], 1ull);
        }
    }
    bad = wave_sum(bad);
    if (lane_id() == 0 && bad) atomicAdd((unsigned long long*)&w.hdr[0], (unsigned long long)bad);
}

// Every walk but the first must be landed on exactly once; walks that stop early. (The record
// and CSR totals come from the scans over the segments.)
__global__ void __launch_bounds__(REC_NT) rec_sum(RecWs w, int64_t* result) {
    int64_t k = (int64_t)blockIdx.x * REC_NT + threadIdx.x;
    int64_t bad = 0, inv = 0, tail = -1;
    if (k < w.nseg && w.f[k] >= 0) {
        if (k > 0 && w.pred[k] != 1) bad++;
        if (w.flag[k]) tail = w.e[k];
        if (w.flag[k] == SEG_END_INVALID) inv++;
    }
    bad = wave_sum(bad);
    inv = wave_sum(inv);
    for (int d = 32; d; d >>= 1) {
        int64_t t = __shfl_xor(tail, d, CBM_WAVE);
        tail = t > tail ? t : tail;
    }
    if (lane_id() == 0) {
        if (bad) atomicAdd((unsigned long long*)&w.hdr[0], (unsigned long long)bad);
        if (inv) atomicAdd((unsigned long long*)&result[CBM_REC_INVALID], (unsigned long long)inv);
        if (tail >= 0) atomicMax((long long*)&result[CBM_REC_TAIL], (long long)tail);
    }
}

// a segment's value, 0 for a walk off the chain (f < 0)
struct SegVal {
    const int64_t* f;
    const int64_t* v;
    __device__ int64_t operator()(int64_t k) const { return f[k] >= 0 ? v[k] : 0; }
};

__global__ void rec_result_init(int64_t* result, int64_t start, RecWs w) {
    for (int i = threadIdx.x; i < CBM_REC_RESULT_LEN; i += blockDim.x) result[i] = 0;
    if (threadIdx.x == 0) {
        result[CBM_REC_TAIL] = start;
        result[CBM_REC_NSEG] = w.nseg;
    }
    for (int i = threadIdx.x; i < 8; i += blockDim.x) w.hdr[i] = 0;
}

__global__ void rec_result_links(int64_t* result, RecWs w) { result[CBM_REC_BAD_LINKS] = w.hdr[0]; }

// CSR offset columns written with the record offsets (null pointers: none), [n_rec + 1] each
struct CsrOut {
    int64_t* cig;
    int64_t* seq;
    int64_t* qual;
};

// offsets of the records of each walk, at the walk's base, and with them the CSR offsets of their
// CIGAR words, bases and qualities (the segment's scanned base plus the walk's in-segment sums)
// a wavefront per segment: lanes copy its recorded starts, coalesced; a segment whose starts were
// not recorded is walked by lane 0. Entry n_rec of a CSR column is written with the record
// n_rec when there is one (a caller taking fewer records than the chain holds).
constexpr int EMIT_SEGS = REC_NT / CBM_WAVE;  // segments per workgroup
__global__ void __launch_bounds__(REC_NT) rec_emit(const uint8_t* b, int64_t start, int64_t n, RecWs w,
                                                   int64_t* off, int64_t n_rec, CsrOut csr) {
    const int lane = (int)threadIdx.x % CBM_WAVE;
    const int64_t k = (int64_t)blockIdx.x * EMIT_SEGS + (int64_t)threadIdx.x / CBM_WAVE;
    if (k >= w.nseg) return;
    int64_t o = w.f[k];
    if (o < 0) return;
    const int64_t c = w.cnt[k], idx = w.base[k];
    const int64_t cb = csr.cig ? w.cig_base[k] : 0, sb = csr.seq ? w.seq_base[k] : 0,
                  qb = csr.qual ? w.qual_base[k] : 0;
    if (w.nrel[k] == c) {  // the walk recorded every start: no chain of dependent loads
        const uint16_t* rel = w.rel + k * REC_MAXR;
        const uint32_t* pre = w.pre + k * REC_MAXR;
        for (int64_t i = lane; i < c && idx + i <= n_rec; i += CBM_WAVE) {
            if (idx + i < n_rec) off[idx + i] = o + rel[i];
            const uint32_t p = pre[i];
            if (csr.cig) csr.cig[idx + i] = cb + pre_cig(p);
            if (csr.seq) csr.seq[idx + i] = sb + pre_seqb(p);
            if (csr.qual) csr.qual[idx + i] = qb + pre_qual(p);
        }
        return;
    }
    if (lane) return;
    int64_t sc = 0, ss = 0, sq = 0;
    for (int64_t i = 0; i < c && idx + i <= n_rec; i++) {
        if (idx + i < n_rec) off[idx + i] = o;
        if (csr.cig) csr.cig[idx + i] = cb + sc;
        if (csr.seq) csr.seq[idx + i] = sb + ss;
        if (csr.qual) csr.qual[idx + i] = qb + sq;
        const uint32_t lseq = ld_u32(b, o + 20);
        sc += ld_u16(b, o + 16);
        ss += (lseq + 1) / 2;
        sq += lseq;
        o += 4 + (int64_t)ld_u32(b, o);
    }
}

// entry n_rec of the CSR columns when n_rec is the whole chain: the totals
__global__ void rec_csr_totals(RecWs w, int64_t n_rec, CsrOut csr) {
    if (w.base[w.nseg] != n_rec) return;
    if (csr.cig) csr.cig[n_rec] = w.cig_base[w.nseg];
    if (csr.seq) csr.seq[n_rec] = w.seq_base[w.nseg];
    if (csr.qual) csr.qual[n_rec] = w.qual_base[w.nseg];
}

// ------------------------------------------------------------------ kernel 3 over many spans
// A batch of requests inflates into one buffer; the records of request s are the bytes
// [start[s], end[s]) of it, both record boundaries. Each span is cut into its own
// segments (segbase[s] is the first), so one launch sequence chains every span. A span
// must end exactly at its end: a record that runs past it, or a tail of fewer bytes than
// a record, marks the span invalid.
struct SpanTab {
    const int64_t* start;
    const int64_t* end;
    const int64_t* segbase;  // n + 1 entries, nondecreasing
    const int32_t* n_ref;    // references of the file each span comes from
    int64_t n;
};

// the span holding segment k: the largest s with segbase[s] <= k (empty spans are skipped)
__device__ __forceinline__ int64_t span_of(const SpanTab& t, int64_t k) {
    int64_t lo = 0, hi = t.n - 1;
    while (lo < hi) {
        int64_t mid = (lo + hi + 1) >> 1;
        if (t.segbase[mid] <= k) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ void __launch_bounds__(REC_NT) recs_links(SpanTab t, RecWs w) {
    int64_t k = (int64_t)blockIdx.x * REC_NT + threadIdx.x;
    int64_t bad = 0;
    if (k < w.nseg && w.f[k] >= 0 && !w.flag[k]) {
        int64_t s = span_of(t, k);
        int64_t j = t.segbase[s] + (w.e[k] - t.start[s]) / REC_SEG;
        if (j >= t.segbase[s + 1] || w.f[j] != w.e[k]) bad++;
        else {
            for (int64_t i = k + 1; i < j; i++)
                if (w.f[i] >= 0) bad++;
            atomicAdd((unsigned long long*)&w.pred[j], 1ull);
        }
    }
    bad = wave_sum(bad);
    if (lane_id() == 0 && bad) atomicAdd((unsigned long long*)&w.hdr[0], (unsigned long long)bad);
}

__global__ void __launch_bounds__(REC_NT) recs_sum(SpanTab t, RecWs w, int64_t* result) {
    int64_t k = (int64_t)blockIdx.x * REC_NT + threadIdx.x;
    int64_t bad = 0, inv = 0;
    if (k < w.nseg && w.f[k] >= 0) {
        int64_t s = span_of(t, k);
        if (k > t.segbase[s] && w.pred[k] != 1) bad++;
        if (w.flag[k] == SEG_END_INVALID || (w.flag[k] && w.e[k] != t.end[s])) inv++;
    }
    bad = wave_sum(bad);
    inv = wave_sum(inv);
    if (lane_id() == 0) {
        if (bad) atomicAdd((unsigned long long*)&w.hdr[0], (unsigned long long)bad);
        if (inv) atomicAdd((unsigned long long*)&result[CBM_REC_INVALID], (unsigned long long)inv);
    }
}

// first record of each span (span_rec[s] = base of its first segment, span_rec[n] = total)
__global__ void recs_bases(SpanTab t, RecWs w, int64_t* span_rec) {
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s <= t.n; s += (int64_t)gridDim.x * blockDim.x)
        span_rec[s] = w.base[t.segbase[s]];
}

// ------------------------------------------------------------------ the walk, a wavefront a segment
// Each wavefront stages its segment and the first RW_OVER bytes after it in LDS (16-byte loads),
// then finds the segment's first record start with a lane per candidate offset, and walks the
// records with every lane: the fixed fields are read alike by all lanes, the read name and the
// CIGAR ops a byte or an op a lane. Bytes past the staged window are read from global memory.
constexpr int RW_WAVES = 4;
constexpr int RW_OVER = 512;
constexpr int RW_STAGE = (int)REC_SEG + RW_OVER + 16;

struct StagedBytes {
    const CBM_GLOBAL uint8_t* g;  // the data
    const CBM_LDS uint8_t* l;     // LDS copy of g[base, base + lim), 4-byte aligned, readable 4 bytes past lim
    int64_t base, lim;
    __device__ __forceinline__ uint32_t u8(int64_t o) const {
        const int64_t r = o - base;
        if ((uint64_t)r < (uint64_t)lim) return l[r];
        return g[o];
    }
    // wider reads inside the stage: the two words that hold them, realigned
    __device__ __forceinline__ uint32_t u32(int64_t o) const {
        const int64_t r = o - base;
        if ((uint64_t)r + 4 <= (uint64_t)lim) {
            const CBM_LDS uint32_t* w = (const CBM_LDS uint32_t*)l + (r >> 2);
            return __builtin_amdgcn_alignbyte(w[1], w[0], (uint32_t)(r & 3));
        }
        return u8(o) | (u8(o + 1) << 8) | (u8(o + 2) << 16) | (u8(o + 3) << 24);
    }
    __device__ __forceinline__ uint32_t u16(int64_t o) const {
        const int64_t r = o - base;
        if ((uint64_t)r + 2 <= (uint64_t)lim) {
            const CBM_LDS uint32_t* w = (const CBM_LDS uint32_t*)l + (r >> 2);
            return __builtin_amdgcn_alignbyte(w[1], w[0], (uint32_t)(r & 3)) & 0xffffu;
        }
        return u8(o) | (u8(o + 1) << 8);
    }
};

// walk() with the wavefront: records that start in [o, hi) from o, a record start. The chain
// through block_size is followed 64 records at a time by the wavefront as one (scalar registers;
// lane j keeps the j-th offset), then each lane tests its record; the first that fails ends the
// walk there. The sums over the records (n_cigar_op, l_seq, odd l_seq) are DPP scans and ballots.

template <class A>
__device__ WalkOut walk_wave(const A& a, int64_t n, int64_t o, int64_t hi, int n_ref, int lane, uint16_t* rel,
                            uint32_t* pre) {
    WalkOut w = {o, 0, 0, 0, 0, 0, 0};
    o = uniform64(o);
    n = uniform64(n);
    hi = uniform64(hi);
    const int64_t o0 = o;
    while (o < hi && !w.flag) {
        int64_t mine = -1;
        int nrec = 0;
        while (nrec < CBM_WAVE && o < hi) {
            if (o + 4 > n) { w.flag = SEG_END_DATA; break; }
            const int64_t bs = (int64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)a.u32(o));
            if (o + 4 + bs > n) { w.flag = SEG_END_DATA; break; }
            if (lane == nrec) mine = o;
            nrec++;
            o += 4 + bs;
        }
        const bool bad = mine >= 0 && !is_record_at(a, n, mine, n_ref);
        const uint64_t badm = __ballot(bad);
        const int good = badm ? __ffsll((unsigned long long)badm) - 1 : nrec;
        uint32_t ncig = 0, lseq = 0;
        if (lane < good) {
            lseq = a.u32(mine + 20);
            ncig = a.u16(mine + 16);
        }
        // sums over the walk's records before each (pre_pack): n_cigar_op and l_seq by inclusive
        // scans (a record's l_seq is at most 2^22, so 64 of them fit 32 bits), odd l_seq by ballot
        const uint32_t ic = (uint32_t)wave_scan_sum((int)ncig), iq = (uint32_t)wave_scan_sum((int)lseq);
        const uint64_t oddm = __ballot(lseq & 1u);
        const uint32_t xodd = __builtin_amdgcn_mbcnt_hi((uint32_t)(oddm >> 32),
                                                        __builtin_amdgcn_mbcnt_lo((uint32_t)oddm, 0u));
        if (lane < good && w.cnt + lane < REC_MAXR) {
            rel[w.cnt + lane] = (uint16_t)(mine - o0);
            pre[w.cnt + lane] = pre_pack((uint32_t)w.cig + ic - ncig, (uint32_t)w.qual + iq - lseq,
                                         (uint32_t)w.odd + xodd);
        }
        const int64_t tc = (uint32_t)__builtin_amdgcn_readlane((int)ic, CBM_WAVE - 1);
        const int64_t tq = (uint32_t)__builtin_amdgcn_readlane((int)iq, CBM_WAVE - 1);
        const int64_t todd = __popcll(oddm);
        w.cnt += good;
        w.cig += tc;
        w.qual += tq;
        w.seqb += (tq + todd) / 2;  // sum of ceil(l_seq / 2)
        w.odd += todd;
        if (badm) {
            o = uniform64(__shfl(mine, good, CBM_WAVE));
            w.flag = SEG_END_INVALID;
        }
    }
    if (o >= n && !w.flag) w.flag = SEG_END_DATA;
    w.e = o;
    return w;
}

// Segment k of one chunk (SPANS false: the data [start, n), n_ref references) or of a batch of
// spans (SPANS true: span table t): its first record start f (the span or chunk start for its
// first segment) and the walk from f over the records that start in the segment.
template <bool SPANS>
__global__ void __launch_bounds__(RW_WAVES * CBM_WAVE)
    rec_walk_wave(const uint8_t* b, int64_t start, int64_t n_all, int n_ref_all, SpanTab t, RecWs w) {
    __shared__ __align__(16) uint8_t stg[RW_WAVES][RW_STAGE + 16];  // 16 spare: word reads past the end
    const int wv = threadIdx.x / CBM_WAVE, lane = threadIdx.x % CBM_WAVE;
    const int64_t k = (int64_t)blockIdx.x * RW_WAVES + wv;
    const bool have = k < w.nseg;  // wave-uniform
    int64_t lo = 0, hi = 0, n = 0, first = 0;
    int n_ref = n_ref_all;
    if (have) {
        if constexpr (SPANS) {
            const int64_t s = span_of(t, k);
            const int64_t k0 = t.segbase[s];
            n = t.end[s];
            n_ref = t.n_ref[s];
            lo = t.start[s] + (k - k0) * REC_SEG;
            first = k == k0;
        } else {
            n = n_all;
            lo = start + k * REC_SEG;
            first = k == 0;
        }
        hi = lo + REC_SEG < n ? lo + REC_SEG : n;
    }
    // the window starts at the 16-byte boundary at or before lo (at lo itself, byte by byte, if that
    // boundary lies before the data)
    int64_t abase = lo - (int64_t)((uintptr_t)(b + lo) & 15);
    const bool vec = abase >= 0;
    if (!vec) abase = lo;
    const int64_t lim = have ? (n - abase < RW_STAGE ? n - abase : RW_STAGE) : 0;
    uint8_t* l = stg[wv];
    const int nvec = vec ? (int)(lim / 16) : 0;
    // the vectors a lane stages, loaded four at a time before any is written (a vector past the
    // window loads the lane's first one again), so that the loads are in flight together
    for (int i0 = lane; i0 < nvec; i0 += 4 * CBM_WAVE) {
        uint4 t[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int i = i0 + u * CBM_WAVE;
            t[u] = *(const uint4*)(b + abase + 16LL * (i < nvec ? i : i0));
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int i = i0 + u * CBM_WAVE;
            if (i < nvec) ((uint4*)l)[i] = t[u];
        }
    }
    for (int i = nvec * 16 + lane; i < lim; i += CBM_WAVE) l[i] = b[abase + i];
    __syncthreads();
    if (!have) return;
    const StagedBytes a{(const CBM_GLOBAL uint8_t*)b, (const CBM_LDS uint8_t*)l, abase, lim};
    int64_t f = -1;
    if (first) {
        f = lo;
    } else {
        for (int64_t o0 = lo; o0 < hi; o0 += CBM_WAVE) {
            const int64_t o = o0 + lane;
            const uint64_t m = __ballot(o < hi && record_start_at(a, n, o, n_ref));
            if (m) {
                f = o0 + __ffsll((unsigned long long)m) - 1;
                break;
            }
        }
    }
    WalkOut r = {-1, 0, 0, 0, 0, 0, 0};
    if (f >= 0) r = walk_wave(a, n, f, hi, n_ref, lane, w.rel + k * REC_MAXR, w.pre + k * REC_MAXR);
    if (lane == 0) {
        w.f[k] = f;
        w.nrel[k] = r.cnt <= REC_MAXR ? r.cnt : -1;
        w.pred[k] = 0;
        w.e[k] = r.e;
        w.cnt[k] = r.cnt;
        w.cig[k] = r.cig;
        w.seqb[k] = r.seqb;
        w.qual[k] = r.qual;
        w.flag[k] = r.flag;
    }
}

// ------------------------------------------------------------------ kernel 4: scatter
constexpr int SC_NT = 256;

// decode_fixed (_kernels.py:292-319) and find_tag (_kernels.py:334-388) of the record at p + o;
// gbase + (a position in p) is that position in the data (for the tag value offsets). Also
// writes n_cigar_op and l_seq of the record to ncig[i] and lseq[i] for the CSR offsets.
// 2- and 4-byte reads at any offset: from LDS as two aligned words realigned (LDS true: p is
// a 4-byte aligned stage readable 4 bytes past the record), else byte by byte
template <bool LDS>
__device__ __forceinline__ uint32_t rd32(const uint8_t* p, int64_t o) {
    if constexpr (LDS) {
        const uint32_t* w = (const uint32_t*)p + (o >> 2);
        return __builtin_amdgcn_alignbyte(w[1], w[0], (uint32_t)(o & 3));
    } else {
        return ld_u32(p, o);
    }
}
template <bool LDS>
__device__ __forceinline__ uint32_t rd16(const uint8_t* p, int64_t o) {
    if constexpr (LDS) return rd32<true>(p, o) & 0xffffu;
    else return ld_u16(p, o);
}
// the first zero byte at or after v, or end (a Z or H value's NUL)
template <bool LDS>
__device__ __forceinline__ int64_t zero_after(const uint8_t* p, int64_t v, int64_t end) {
    if constexpr (LDS) {
        for (int64_t x = v; x < end; x += 4) {
            const uint32_t w = rd32<true>(p, x);
            const uint32_t z = (w - 0x01010101u) & ~w & 0x80808080u;  // high bit of each zero byte
            if (z) {
                const int64_t at = x + (__ffs(z) - 1) / 8;
                return at < end ? at : end;
            }
        }
        return end;
    } else {
        int64_t x = v;
        while (x < end && p[x] != 0) x++;
        return x;
    }
}

template <bool LDS>
__device__ __forceinline__ void parse_record(const uint8_t* p, int64_t o, int64_t gbase, int64_t i,
                                             const cbm_columns& c, int32_t* ncig_out, int32_t* lseq_out) {
    int32_t pos = (int32_t)rd32<LDS>(p, o + 8);
    uint32_t lseq = rd32<LDS>(p, o + 20);
    if (c.ref) c.ref[i] = (int32_t)rd32<LDS>(p, o + 4);
    if (c.pos) c.pos[i] = pos;
    if (c.mapq) c.mapq[i] = p[o + 13];
    if (c.flag) c.flag[i] = (uint16_t)rd16<LDS>(p, o + 18);
    if (c.nref) c.nref[i] = (int32_t)rd32<LDS>(p, o + 24);
    if (c.npos) c.npos[i] = (int32_t)rd32<LDS>(p, o + 28);
    if (c.tlen) c.tlen[i] = (int32_t)rd32<LDS>(p, o + 32);
    if (c.l_seq) c.l_seq[i] = (int32_t)lseq;
    uint32_t lrn = p[o + 12], ncig = rd16<LDS>(p, o + 16);
    if (ncig_out) ncig_out[i] = (int32_t)ncig;
    if (lseq_out) lseq_out[i] = (int32_t)lseq;
    int64_t cg = o + 36 + lrn, rlen = 0;
    uint32_t cl = 0, cr = 0;
    for (uint32_t k = 0; k < ncig; k++) {
        uint32_t v = rd32<LDS>(p, cg + 4LL * k), op = v & 15, len = v >> 4;
        if (ref_consuming(op)) rlen += len;
        if (op == 4) {
            if (k == 0) cl = len;
            else cr = len;
        }
    }
    if (c.end) c.end[i] = pos + (int32_t)rlen;
    if (c.clip_left) c.clip_left[i] = (uint16_t)(cl > 65535 ? 65535 : cl);
    if (c.clip_right) c.clip_right[i] = (uint16_t)(cr > 65535 ? 65535 : cr);
    int nt = c.n_tags;
    if (nt <= 0) return;
    // aux fields, once for every tag asked for
    int64_t rec_end = o + 4 + (int64_t)rd32<LDS>(p, o);
    int64_t a = cg + 4LL * ncig + (lseq + 1) / 2 + lseq;
    uint32_t found = 0;
    bool bad = false;
    uint8_t ty_out[4] = {0, 0, 0, 0};
    int64_t vo[4] = {0, 0, 0, 0}, iv[4] = {0, 0, 0, 0};
    int32_t vl[4] = {0, 0, 0, 0};
    uint32_t want = (1u << nt) - 1;
    while (a + 3 <= rec_end && found != want) {
        uint32_t ty = p[a + 2];
        int64_t v = a + 3, nx;
        switch (ty) {
            case 'A': case 'c': case 'C': nx = v + 1; break;
            case 's': case 'S': nx = v + 2; break;
            case 'i': case 'I': case 'f': nx = v + 4; break;
            case 'Z': case 'H':
                nx = zero_after<LDS>(p, v, rec_end) + 1;
                break;
            case 'B': {
                if (v + 5 > rec_end) { nx = -1; break; }
                uint32_t sub = p[v];
                int64_t cnt = (int64_t)rd32<LDS>(p, v + 1);
                int sz = (sub == 'c' || sub == 'C') ? 1 : (sub == 's' || sub == 'S') ? 2 : 4;
                nx = v + 5 + cnt * sz;
                break;
            }
            default: nx = -1;
        }
        if (nx < 0 || nx > rec_end) { bad = true; break; }
        uint32_t t0 = p[a], t1 = p[a + 1];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            if (t < nt && !(found & (1u << t)) && t0 == (uint8_t)c.tag_names[2 * t] &&
                t1 == (uint8_t)c.tag_names[2 * t + 1]) {
                found |= 1u << t;
                ty_out[t] = (uint8_t)ty;
                vo[t] = gbase + v;
                switch (ty) {
                    case 'c': iv[t] = (int8_t)p[v]; break;
                    case 'C': case 'A': iv[t] = p[v]; break;
                    case 's': iv[t] = (int16_t)rd16<LDS>(p, v); break;
                    case 'S': iv[t] = rd16<LDS>(p, v); break;
                    case 'i': iv[t] = (int32_t)rd32<LDS>(p, v); break;
                    case 'I': iv[t] = rd32<LDS>(p, v); break;
                    case 'Z': case 'H': vl[t] = (int32_t)(nx - 1 - v); break;
                    case 'B': vl[t] = (int32_t)(nx - v); break;
                    default: break;
                }
            }
        }
        a = nx;
    }
#pragma unroll
    for (int t = 0; t < 4; t++) {
        if (t >= nt) break;
        // a malformed aux area before the tag makes it 255 (the reference stops there)
        uint8_t ty = (bad && !(found & (1u << t))) ? 255 : ty_out[t];
        if (c.tag_type[t]) c.tag_type[t][i] = ty;
        if (c.tag_offset[t]) c.tag_offset[t][i] = (found & (1u << t)) ? vo[t] : 0;
        if (c.tag_length[t]) c.tag_length[t][i] = (found & (1u << t)) ? vl[t] : 0;
        if (c.tag_value[t]) c.tag_value[t][i] = (found & (1u << t)) ? iv[t] : 0;
    }
}

// One workgroup per FIELDS_NT consecutive records. Records found by the chain are back to back,
// so the workgroup's records are one byte range: it is staged in LDS with coalesced 16-byte
// loads, and each thread decodes its record from there (one thread per record reading the
// data in place would touch 64 cache lines per load instruction of a wavefront). A record
// outside the staged range (offsets that are not consecutive records, or past the stage's end
// when the range is longer than the stage) is decoded from global memory.
// 24 KiB: 128 records of up to 190 bytes fit, and six workgroups share a CU's LDS (a 40 KiB stage,
// three workgroups, made the columns 1.45x slower); a longer range is staged in part
#ifndef CBM_SC_STAGE
#define CBM_SC_STAGE 24576
#endif
constexpr int SC_STAGE = CBM_SC_STAGE;
#ifndef CBM_FIELDS_NT
#define CBM_FIELDS_NT 128
#endif
constexpr int FIELDS_NT = CBM_FIELDS_NT;  // records per workgroup
constexpr int FIELDS_THREADS = 256;  // threads per workgroup (all stage and copy; FIELDS_NT decode)

// Stage data bytes [a0, a0 + nb) (a0 16-byte aligned in memory, nb <= SC_STAGE) into st with
// 16-byte loads, four in flight per thread.
__device__ __forceinline__ void stage_records(uint8_t* st, const uint8_t* b, int64_t a0, int64_t nb) {
    constexpr int SB = 4;
    for (int64_t k0 = (int64_t)threadIdx.x * 16; k0 < nb; k0 += (int64_t)FIELDS_THREADS * 16 * SB) {
        uint4 t[SB];
        // every load issued, unconditionally (a vector past the range loads the thread's first one
        // again), so that the SB loads are in flight together and t stays in registers
#pragma unroll
        for (int u = 0; u < SB; u++) {
            const int64_t k = k0 + (int64_t)u * FIELDS_THREADS * 16;
            t[u] = *(const uint4*)(b + a0 + (k < nb ? k : k0));
        }
#pragma unroll
        for (int u = 0; u < SB; u++) {
            const int64_t k = k0 + (int64_t)u * FIELDS_THREADS * 16;
            if (k < nb) *(uint4*)(st + k) = t[u];
        }
    }
}

// n_cigar_op and l_seq of each record, for the CSR offsets (a thread a record)
__global__ void __launch_bounds__(256) rec_sizes(const uint8_t* b, const int64_t* off, int64_t n, int32_t* ncig,
                                                 int32_t* lseq) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int64_t o = off[i];
    ncig[i] = (int32_t)ld_u16(b, o + 16);
    lseq[i] = (int32_t)ld_u32(b, o + 20);
}

// The columns. A workgroup stages its FIELDS_NT consecutive records in LDS; each of its first
// FIELDS_NT threads decodes the fixed fields and tags of one record from there, then the CSR
// columns are written: each column's output range is contiguous across the records, so threads
// take consecutive pieces of it, find the record of a piece's first element by a binary
// search over the records' CSR offsets (in LDS) and read the elements from the stage. Writes
// are coalesced whatever the records' lengths. Records outside the stage are read in place.
__device__ __forceinline__ int csr_record(const int64_t* sh_off, int m, int64_t j) {
    // the record r with sh_off[r] <= j < sh_off[r + 1]
    int lo = 0, hi = m - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (sh_off[mid] <= j) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// CIGAR words: one thread per word
__device__ __forceinline__ void csr_fill_words(uint32_t* out, int64_t base, int64_t end, int m,
                                               const int64_t* sh_off, const int64_t* sh_src, const uint8_t* st,
                                               int64_t a0, int64_t hi, const uint8_t* b) {
    const int64_t total = end - base;
    const CBM_LDS uint32_t* sw = (const CBM_LDS uint32_t*)st;
    const CBM_GLOBAL uint8_t* bg = (const CBM_GLOBAL uint8_t*)b;
    for (int64_t j = threadIdx.x; j < total; j += FIELDS_THREADS) {
        const int r = csr_record(sh_off, m, j);
        const int64_t src = sh_src[r] + 4 * (j - sh_off[r]);
        uint32_t v;
        if (src >= a0 && src + 4 <= hi) {
            const int x = (int)(src - a0);
            v = __builtin_amdgcn_alignbyte(sw[(x >> 2) + 1], sw[x >> 2], (uint32_t)(x & 3));
        } else {
            v = (uint32_t)bg[src] | ((uint32_t)bg[src + 1] << 8) | ((uint32_t)bg[src + 2] << 16) |
                ((uint32_t)bg[src + 3] << 24);
        }
        out[base + j] = v;
    }
}

// Bases or qualities: each thread fills one 16-byte aligned piece of the output (one search,
// then a walk over the records it spans), stored with one 16-byte store when the piece lies
// wholly in this workgroup's range.
__device__ __forceinline__ void csr_fill_bytes(uint8_t* out, int64_t base, int64_t end, int m,
                                               const int64_t* sh_off, const int64_t* sh_src, const uint8_t* st,
                                               int64_t a0, int64_t hi, const uint8_t* b) {
    if (end <= base) return;
    const uintptr_t ob = (uintptr_t)out;
    // pieces: [p, p + 16) with out + p 16-byte aligned
    const int64_t p0 = base - (int64_t)((ob + base) & 15);
    for (int64_t p = p0 + 16 * (int64_t)threadIdx.x; p < end; p += 16 * (int64_t)FIELDS_THREADS) {
        const int64_t lo = p < base ? base : p, up = p + 16 < end ? p + 16 : end;
        int r = csr_record(sh_off, m, lo - base);
        int64_t nxt = r + 1 < m ? sh_off[r + 1] + base : end;
        int64_t src = sh_src[r] + (lo - base - sh_off[r]);
        uint8_t v[16];
#pragma unroll
        for (int k = 0; k < 16; k++) {
            const int64_t j = p + k;
            uint8_t x = 0;
            if (j >= lo && j < up) {
                while (j >= nxt) {  // next record (empty ones skipped)
                    r++;
                    src = sh_src[r];
                    nxt = r + 1 < m ? sh_off[r + 1] + base : end;
                }
                if (src >= a0 && src < hi) x = ((const CBM_LDS uint8_t*)st)[src - a0];
                else x = ((const CBM_GLOBAL uint8_t*)b)[src];
                src++;
            }
            v[k] = x;
        }
        if (lo == p && up == p + 16) {
            uint4 w;
            w.x = v[0] | (v[1] << 8) | (v[2] << 16) | ((uint32_t)v[3] << 24);
            w.y = v[4] | (v[5] << 8) | (v[6] << 16) | ((uint32_t)v[7] << 24);
            w.z = v[8] | (v[9] << 8) | (v[10] << 16) | ((uint32_t)v[11] << 24);
            w.w = v[12] | (v[13] << 8) | (v[14] << 16) | ((uint32_t)v[15] << 24);
            *(uint4*)(out + p) = w;
        } else {
            for (int64_t j = lo; j < up; j++) out[j] = v[j - p];
        }
    }
}

__global__ void __launch_bounds__(FIELDS_THREADS) rec_columns(const uint8_t* b, const int64_t* off, int64_t n,
                                                         cbm_columns c) {
    __shared__ __align__(16) uint8_t st[SC_STAGE + 16];
    __shared__ int64_t sh_off[FIELDS_NT + 1], sh_src[FIELDS_NT];
    // [0] stage start, [1] stage end, [2] the range's first offset, [3] the offset after the
    // workgroup's records (the next record's, or the last record's end), [4..6] / [7..9] the
    // CIGAR, base and quality CSR offsets of the first record / after the last
    __shared__ int64_t range[10];
    const int64_t i0 = (int64_t)blockIdx.x * FIELDS_NT;
    if (i0 >= n) return;
    const int m = (int)(n - i0 < FIELDS_NT ? n - i0 : FIELDS_NT);
    const int t = threadIdx.x;
    // every load of the workgroup's offsets at once: the record offsets and the three CSR columns'
    // offsets of records i0 .. i0 + m (CSR columns hold n + 1 entries)
    const bool cig_on = c.cigar && c.cigar_off, seq_on = c.seq && c.seq_off, qual_on = c.qual && c.qual_off;
    int64_t o = 0, co = 0, so = 0, qo = 0;
    if (t <= m) {
        const int64_t i = i0 + t;
        if (i < n) o = off[i];
        if (cig_on) co = c.cigar_off[i];
        if (seq_on) so = c.seq_off[i];
        if (qual_on) qo = c.qual_off[i];
        if (t == 0) range[2] = o, range[4] = co, range[5] = so, range[6] = qo;
        if (t == m) range[7] = co, range[8] = so, range[9] = qo;
        // the records of a chain are back to back: the next record's offset ends the range; the
        // last workgroup reads the last record's size
        if (t == m && i < n) range[3] = o;
        if (t == m - 1 && i0 + m == n) range[3] = o + 4 + (int64_t)ld_u32(b, o);
    }
    __syncthreads();
    if (t == 0) {
        const int64_t lo = range[2], end = range[3];
        const int64_t a0 = lo - (int64_t)(((uintptr_t)(b + lo)) & 15);
        range[0] = a0;
        // as much of the range as the stage holds; records past it (or outside it, for offsets
        // that are not in order) are read in place
        range[1] = end >= lo ? (end - a0 <= SC_STAGE ? end : a0 + SC_STAGE) : a0;
    }
    __syncthreads();
    const int64_t a0 = range[0], hi = range[1];
    stage_records(st, b, a0, hi - a0);
    __syncthreads();
    // fixed fields and tags (decode_fixed and find_tag)
    if (t < m) {
        const int64_t i = i0 + t;
        bool staged = o >= a0 && o + 36 <= hi;
        if (staged) staged = o + 4 + (int64_t)ld_u32(st, o - a0) <= hi;
        if (staged) parse_record<true>(st, o - a0, a0, i, c, nullptr, nullptr);
        else parse_record<false>(b, o, 0, i, c, nullptr, nullptr);
    }
    if (!cig_on && !seq_on && !qual_on) return;
    // where each record's CIGAR, bases and qualities start (data offsets)
    int64_t cig = 0, sq = 0, ql = 0;
    if (t < m) {
        const bool staged = o >= a0 && o + 36 <= hi;
        const uint8_t* p = staged ? st : b;
        const int64_t r = staged ? o - a0 : o;
        const uint32_t lrn = p[r + 12], ncig = ld_u16(p, r + 16), lseq = ld_u32(p, r + 20);
        cig = o + 36 + lrn;
        sq = cig + 4LL * ncig;
        ql = sq + (lseq + 1) / 2;
    }
    if (cig_on) {
        if (t < m) {
            sh_off[t] = co - range[4];
            sh_src[t] = cig;
        }
        __syncthreads();
        csr_fill_words(c.cigar, range[4], range[7], m, sh_off, sh_src, st, a0, hi, b);
        __syncthreads();
    }
    if (seq_on) {
        if (t < m) {
            sh_off[t] = so - range[5];
            sh_src[t] = sq;
        }
        __syncthreads();
        csr_fill_bytes(c.seq, range[5], range[8], m, sh_off, sh_src, st, a0, hi, b);
        __syncthreads();
    }
    if (qual_on) {
        if (t < m) {
            sh_off[t] = qo - range[6];
            sh_src[t] = ql;
        }
        __syncthreads();
        csr_fill_bytes(c.qual, range[6], range[9], m, sh_off, sh_src, st, a0, hi, b);
    }
}

// per-record sizes for the CSR offsets, from the arrays rec_sizes writes
struct RecNcig {
    const int32_t* ncig;
    __device__ int64_t operator()(int64_t i) const { return ncig[i]; }
};
struct RecSeqB {
    const int32_t* lseq;
    __device__ int64_t operator()(int64_t i) const { return ((int64_t)(uint32_t)lseq[i] + 1) / 2; }
};
struct RecQual {
    const int32_t* lseq;
    __device__ int64_t operator()(int64_t i) const { return (int64_t)(uint32_t)lseq[i]; }
};

}  // namespace cbm
