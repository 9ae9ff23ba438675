// Kernel 3: BAM record chain across block seams, and kernel 4: record-to-column scatter.
//
// The reference tests every inflated byte offset with is_record (one thread per byte,
// reading the name and CIGAR at each), radix-sorts the candidates and chains them
// with pointer jumping (src/cubam/bam.py:403-418, src/cubam/_kernels.py:250-289,
// 676-702). Here the inflated data is cut into segments of SEG bytes; each thread
// finds the first offset of its segment that passes the structural test (confirmed by
// the record after it), walks block_size from there over the records that start in
// its segment, and records where it left. The walks chain when every exit is the
// start found in the segment it lands in (segment 0 starts at the known record
// boundary), which a second pass checks; a walk that does not chain is repaired by
// re-walking it from the true exit of the walk before it.
#pragma once
#include "common.h"

namespace cbm {

constexpr int64_t REC_SEG = 4096;
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
    void* scan_ws;
    int64_t nseg;
};

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
    w.scan_ws = p;
    return w;
}

inline RecWs rec_ws_layout(void* ws, int64_t len) { return rec_ws_layout_seg(ws, rec_nseg(len)); }

inline int64_t rec_ws_bytes_seg(int64_t nseg) { return 8 * (8 + 9 * nseg + 1) + scan_workspace(nseg) + 64; }

inline int64_t rec_ws_bytes(int64_t len) { return rec_ws_bytes_seg(rec_nseg(len)); }

// a record start confirmed by the record after it (or by the end of the data)
__device__ __forceinline__ bool record_start(const uint8_t* b, int64_t n, int64_t o, int n_ref) {
    if (!is_record(b, n, o, n_ref)) return false;
    int64_t nx = o + 4 + (int64_t)ld_u32(b, o);
    if (nx + 36 > n) return true;
    if (nx + 4 + (int64_t)ld_u32(b, nx) > n) return true;
    return is_record(b, n, nx, n_ref);
}

struct WalkOut {
    int64_t e, cnt, cig, seqb, qual, flag;
};

// records that start in [o, hi) from o, a record start
__device__ __forceinline__ WalkOut walk(const uint8_t* b, int64_t n, int64_t o, int64_t hi, int n_ref) {
    WalkOut w = {o, 0, 0, 0, 0, 0};
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

__global__ void __launch_bounds__(REC_NT) rec_walk(const uint8_t* b, int64_t start, int64_t n, int n_ref, RecWs w) {
    int64_t k = (int64_t)blockIdx.x * REC_NT + threadIdx.x;
    if (k >= w.nseg) return;
    int64_t lo = start + k * REC_SEG, hi = lo + REC_SEG < n ? lo + REC_SEG : n;
    int64_t f = -1;
    if (k == 0) f = start;
    else
        for (int64_t o = lo; o < hi; o++)
            if (record_start(b, n, o, n_ref)) { f = o; break; }
    w.f[k] = f;
    w.pred[k] = 0;
    WalkOut r = {-1, 0, 0, 0, 0, 0};
    if (f >= 0) r = walk(b, n, f, hi, n_ref);
    w.e[k] = r.e;
    w.cnt[k] = r.cnt;
    w.cig[k] = r.cig;
    w.seqb[k] = r.seqb;
    w.qual[k] = r.qual;
    w.flag[k] = r.flag;
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

// Every walk but the first must be landed on exactly once; then sum the walks.
__global__ void __launch_bounds__(REC_NT) rec_sum(RecWs w, int64_t* result) {
    int64_t k = (int64_t)blockIdx.x * REC_NT + threadIdx.x;
    int64_t bad = 0, cnt = 0, cig = 0, seqb = 0, qual = 0, inv = 0, tail = -1;
    if (k < w.nseg && w.f[k] >= 0) {
        if (k > 0 && w.pred[k] != 1) bad++;
        cnt = w.cnt[k];
        cig = w.cig[k];
        seqb = w.seqb[k];
        qual = w.qual[k];
        if (w.flag[k]) tail = w.e[k];
        if (w.flag[k] == SEG_END_INVALID) inv++;
    }
    bad = wave_sum(bad);
    cnt = wave_sum(cnt);
    cig = wave_sum(cig);
    seqb = wave_sum(seqb);
    qual = wave_sum(qual);
    inv = wave_sum(inv);
    for (int d = 32; d; d >>= 1) {
        int64_t t = __shfl_xor(tail, d, CBM_WAVE);
        tail = t > tail ? t : tail;
    }
    if (lane_id() == 0) {
        if (bad) atomicAdd((unsigned long long*)&w.hdr[0], (unsigned long long)bad);
        if (cnt) atomicAdd((unsigned long long*)&result[CBM_REC_N], (unsigned long long)cnt);
        if (cig) atomicAdd((unsigned long long*)&result[CBM_REC_CIGAR], (unsigned long long)cig);
        if (seqb) atomicAdd((unsigned long long*)&result[CBM_REC_SEQ], (unsigned long long)seqb);
        if (qual) atomicAdd((unsigned long long*)&result[CBM_REC_QUAL], (unsigned long long)qual);
        if (inv) atomicAdd((unsigned long long*)&result[CBM_REC_INVALID], (unsigned long long)inv);
        if (tail >= 0) atomicMax((long long*)&result[CBM_REC_TAIL], (long long)tail);
    }
}

__global__ void rec_result_init(int64_t* result, int64_t start, RecWs w) {
    for (int i = threadIdx.x; i < CBM_REC_RESULT_LEN; i += blockDim.x) result[i] = 0;
    if (threadIdx.x == 0) {
        result[CBM_REC_TAIL] = start;
        result[CBM_REC_NSEG] = w.nseg;
    }
    for (int i = threadIdx.x; i < 8; i += blockDim.x) w.hdr[i] = 0;
}

__global__ void rec_result_links(int64_t* result, RecWs w) { result[CBM_REC_BAD_LINKS] = w.hdr[0]; }

// offsets of the records of each walk, at the walk's base
__global__ void __launch_bounds__(REC_NT) rec_emit(const uint8_t* b, int64_t start, int64_t n, RecWs w,
                                                   int64_t* off, int64_t n_rec) {
    int64_t k = (int64_t)blockIdx.x * REC_NT + threadIdx.x;
    if (k >= w.nseg) return;
    int64_t o = w.f[k];
    if (o < 0) return;
    int64_t c = w.cnt[k], idx = w.base[k];
    for (int64_t i = 0; i < c && idx + i < n_rec; i++) {
        off[idx + i] = o;
        o += 4 + (int64_t)ld_u32(b, o);
    }
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

__global__ void __launch_bounds__(REC_NT) recs_walk(const uint8_t* b, SpanTab t, RecWs w) {
    int64_t k = (int64_t)blockIdx.x * REC_NT + threadIdx.x;
    if (k >= w.nseg) return;
    int64_t s = span_of(t, k);
    int64_t s0 = t.start[s], n = t.end[s], k0 = t.segbase[s];
    int n_ref = t.n_ref[s];
    int64_t lo = s0 + (k - k0) * REC_SEG, hi = lo + REC_SEG < n ? lo + REC_SEG : n;
    int64_t f = -1;
    if (k == k0) f = s0;
    else
        for (int64_t o = lo; o < hi; o++)
            if (record_start(b, n, o, n_ref)) { f = o; break; }
    w.f[k] = f;
    w.pred[k] = 0;
    WalkOut r = {-1, 0, 0, 0, 0, 0};
    if (f >= 0) r = walk(b, n, f, hi, n_ref);
    w.e[k] = r.e;
    w.cnt[k] = r.cnt;
    w.cig[k] = r.cig;
    w.seqb[k] = r.seqb;
    w.qual[k] = r.qual;
    w.flag[k] = r.flag;
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
    int64_t bad = 0, cnt = 0, cig = 0, seqb = 0, qual = 0, inv = 0;
    if (k < w.nseg && w.f[k] >= 0) {
        int64_t s = span_of(t, k);
        if (k > t.segbase[s] && w.pred[k] != 1) bad++;
        cnt = w.cnt[k];
        cig = w.cig[k];
        seqb = w.seqb[k];
        qual = w.qual[k];
        if (w.flag[k] == SEG_END_INVALID || (w.flag[k] && w.e[k] != t.end[s])) inv++;
    }
    bad = wave_sum(bad);
    cnt = wave_sum(cnt);
    cig = wave_sum(cig);
    seqb = wave_sum(seqb);
    qual = wave_sum(qual);
    inv = wave_sum(inv);
    if (lane_id() == 0) {
        if (bad) atomicAdd((unsigned long long*)&w.hdr[0], (unsigned long long)bad);
        if (cnt) atomicAdd((unsigned long long*)&result[CBM_REC_N], (unsigned long long)cnt);
        if (cig) atomicAdd((unsigned long long*)&result[CBM_REC_CIGAR], (unsigned long long)cig);
        if (seqb) atomicAdd((unsigned long long*)&result[CBM_REC_SEQ], (unsigned long long)seqb);
        if (qual) atomicAdd((unsigned long long*)&result[CBM_REC_QUAL], (unsigned long long)qual);
        if (inv) atomicAdd((unsigned long long*)&result[CBM_REC_INVALID], (unsigned long long)inv);
    }
}

// first record of each span (span_rec[s] = base of its first segment, span_rec[n] = total)
__global__ void recs_bases(SpanTab t, RecWs w, int64_t* span_rec) {
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s <= t.n; s += (int64_t)gridDim.x * blockDim.x)
        span_rec[s] = w.base[t.segbase[s]];
}

// ------------------------------------------------------------------ kernel 4: scatter
constexpr int SC_NT = 256;

// decode_fixed (_kernels.py:292-319) and find_tag (_kernels.py:334-388), one thread per record
__global__ void __launch_bounds__(SC_NT) rec_fields(const uint8_t* b, const int64_t* off, int64_t n,
                                                    cbm_columns c) {
    int64_t i = (int64_t)blockIdx.x * SC_NT + threadIdx.x;
    if (i >= n) return;
    int64_t o = off[i];
    int32_t p = (int32_t)ld_u32(b, o + 8);
    uint32_t lseq = ld_u32(b, o + 20);
    if (c.ref) c.ref[i] = (int32_t)ld_u32(b, o + 4);
    if (c.pos) c.pos[i] = p;
    if (c.mapq) c.mapq[i] = b[o + 13];
    if (c.flag) c.flag[i] = (uint16_t)ld_u16(b, o + 18);
    if (c.nref) c.nref[i] = (int32_t)ld_u32(b, o + 24);
    if (c.npos) c.npos[i] = (int32_t)ld_u32(b, o + 28);
    if (c.tlen) c.tlen[i] = (int32_t)ld_u32(b, o + 32);
    if (c.l_seq) c.l_seq[i] = (int32_t)lseq;
    uint32_t lrn = b[o + 12], ncig = ld_u16(b, o + 16);
    int64_t cg = o + 36 + lrn, rlen = 0;
    uint32_t cl = 0, cr = 0;
    for (uint32_t k = 0; k < ncig; k++) {
        uint32_t v = ld_u32(b, cg + 4LL * k), op = v & 15, len = v >> 4;
        if (ref_consuming(op)) rlen += len;
        if (op == 4) {
            if (k == 0) cl = len;
            else cr = len;
        }
    }
    if (c.end) c.end[i] = p + (int32_t)rlen;
    if (c.clip_left) c.clip_left[i] = (uint16_t)(cl > 65535 ? 65535 : cl);
    if (c.clip_right) c.clip_right[i] = (uint16_t)(cr > 65535 ? 65535 : cr);
    int nt = c.n_tags;
    if (nt <= 0) return;
    // aux fields, once for every tag asked for
    int64_t rec_end = o + 4 + (int64_t)ld_u32(b, o);
    int64_t a = cg + 4LL * ncig + (lseq + 1) / 2 + lseq;
    uint32_t found = 0;
    bool bad = false;
    uint8_t ty_out[4] = {0, 0, 0, 0};
    int64_t vo[4] = {0, 0, 0, 0}, iv[4] = {0, 0, 0, 0};
    int32_t vl[4] = {0, 0, 0, 0};
    uint32_t want = (1u << nt) - 1;
    while (a + 3 <= rec_end && found != want) {
        uint32_t ty = b[a + 2];
        int64_t v = a + 3, nx;
        switch (ty) {
            case 'A': case 'c': case 'C': nx = v + 1; break;
            case 's': case 'S': nx = v + 2; break;
            case 'i': case 'I': case 'f': nx = v + 4; break;
            case 'Z': case 'H':
                nx = v;
                while (nx < rec_end && b[nx] != 0) nx++;
                nx++;
                break;
            case 'B': {
                if (v + 5 > rec_end) { nx = -1; break; }
                uint32_t sub = b[v];
                int64_t cnt = (int64_t)ld_u32(b, v + 1);
                int sz = (sub == 'c' || sub == 'C') ? 1 : (sub == 's' || sub == 'S') ? 2 : 4;
                nx = v + 5 + cnt * sz;
                break;
            }
            default: nx = -1;
        }
        if (nx < 0 || nx > rec_end) { bad = true; break; }
        uint32_t t0 = b[a], t1 = b[a + 1];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            if (t < nt && !(found & (1u << t)) && t0 == (uint8_t)c.tag_names[2 * t] &&
                t1 == (uint8_t)c.tag_names[2 * t + 1]) {
                found |= 1u << t;
                ty_out[t] = (uint8_t)ty;
                vo[t] = v;
                switch (ty) {
                    case 'c': iv[t] = (int8_t)b[v]; break;
                    case 'C': case 'A': iv[t] = b[v]; break;
                    case 's': iv[t] = (int16_t)ld_u16(b, v); break;
                    case 'S': iv[t] = ld_u16(b, v); break;
                    case 'i': iv[t] = (int32_t)ld_u32(b, v); break;
                    case 'I': iv[t] = ld_u32(b, v); break;
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

// CSR columns: a wavefront copies the CIGAR words, packed bases and qualities of 64
// consecutive records, one record at a time with all lanes, so writes are contiguous.
__global__ void __launch_bounds__(SC_NT) rec_copy_var(const uint8_t* b, const int64_t* off, int64_t n,
                                                      cbm_columns c) {
    int lane = threadIdx.x % CBM_WAVE;
    int64_t g = ((int64_t)blockIdx.x * SC_NT + threadIdx.x) / CBM_WAVE;
    int64_t r0 = g * CBM_WAVE;
    if (r0 >= n) return;
    int64_t my = r0 + lane;
    int64_t src = 0, co = 0, so = 0, qo = 0;
    int32_t ncig = 0, lseq = 0;
    if (my < n) {
        int64_t o = off[my];
        uint32_t lrn = b[o + 12];
        ncig = (int32_t)ld_u16(b, o + 16);
        lseq = (int32_t)ld_u32(b, o + 20);
        src = o + 36 + lrn;
        co = c.cigar_off ? c.cigar_off[my] : 0;
        so = c.seq_off ? c.seq_off[my] : 0;
        qo = c.qual_off ? c.qual_off[my] : 0;
    }
    int m = (int)(n - r0 < CBM_WAVE ? n - r0 : CBM_WAVE);
    for (int r = 0; r < m; r++) {
        int64_t s = __shfl(src, r, CBM_WAVE);
        int32_t nc = __shfl(ncig, r, CBM_WAVE), ls = __shfl(lseq, r, CBM_WAVE);
        if (c.cigar) {
            int64_t d = __shfl(co, r, CBM_WAVE);
            for (int k = lane; k < nc; k += CBM_WAVE) c.cigar[d + k] = ld_u32(b, s + 4LL * k);
        }
        int64_t sq = s + 4LL * nc;
        int32_t nb = (ls + 1) / 2;
        if (c.seq) {
            int64_t d = __shfl(so, r, CBM_WAVE);
            for (int k = lane; k < nb; k += CBM_WAVE) c.seq[d + k] = b[sq + k];
        }
        if (c.qual) {
            int64_t d = __shfl(qo, r, CBM_WAVE);
            for (int k = lane; k < ls; k += CBM_WAVE) c.qual[d + k] = b[sq + nb + k];
        }
    }
}

// per-record sizes for the CSR offsets
struct RecNcig {
    const uint8_t* b;
    const int64_t* off;
    __device__ int64_t operator()(int64_t i) const { return ld_u16(b, off[i] + 16); }
};
struct RecSeqB {
    const uint8_t* b;
    const int64_t* off;
    __device__ int64_t operator()(int64_t i) const { return (ld_u32(b, off[i] + 20) + 1) / 2; }
};
struct RecQual {
    const uint8_t* b;
    const int64_t* off;
    __device__ int64_t operator()(int64_t i) const { return ld_u32(b, off[i] + 20); }
};

}  // namespace cbm
