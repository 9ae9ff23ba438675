// Kernel 2, workgroup variant: one BGZF block per workgroup, decoded by all its
// threads at once, output staged in LDS, CRC32 fused.
//
// DEFLATE is serial within a Huffman block, but Huffman codes resynchronise: a
// decoder started at an arbitrary bit soon falls onto the true symbol boundaries
// (the speculative self-synchronising decode of Weissenberger and Schmidt,
// "Massively Parallel Huffman Decoding on GPUs", ICPP 2018). Each round of a
// Huffman block gives thread t the SEG bits that start at base + t * SEG (fewer
// in the last round of the final Huffman block, where the payload's end bounds it).
// Pass A decodes speculatively, from 1.5 * SEG bits before the segment (so that the
// walk is most likely on the true boundaries when it enters it) to the first symbol
// that starts at or after the next segment, and keeps checkpoints inside its
// segment: the symbol start before every few symbols, with the output bytes and
// matches counted up to there. The walks are then linked: thread t's true entry
// is thread t-1's exit; a thread whose entry changed decodes again from it until
// it lands on a checkpoint (from there its first walk was right, and its count
// from there is known) - repeated until nothing changes, within each wavefront
// by lane shifts and across wavefronts through LDS. A workgroup prefix sum of the
// walks' counts places them, and pass C writes the literals into the LDS output
// window and the matches, in output order, into a per-workgroup list in global
// scratch. Pass D applies the matches. Variants whose staging area holds a bit per
// output byte keep the round's pending bytes in a bitmap there and apply, step by
// step, every match whose source bytes are no longer pending; the others go
// frontier by frontier (with F the destination of the first match not yet
// applied, every byte before F is final).
// A dynamic block's code lengths are decoded the same way by the 64 lanes of one
// wavefront, CL_SEG bits a lane.
// The block's output then leaves LDS in one coalesced store, and its CRC32 is
// computed from LDS by all threads (slices combined with x^(8n) mod P, the
// method of zlib's crc32_combine).
//
// Global memory is touched by the coalesced staging of each round's compressed
// bits into LDS, the match list, and the final store; no load waits behind the
// decoder's byte stores, which all go to LDS.
#pragma once
#include "inflate.h"

namespace cbm {

enum { SV_POS = 0, SV_FINAL, SV_TYPE, SV_ERR, SV_LEN, SV_SRC, SV_OP, SV_CLEAR, SV_TEND, SV_CHANGED, SV_DONE, SV_OK,
       SV_BLOCK, SV_MAX, SV_SITE, SV_F0, SV_F1, SV_F2, SV_MCNT, SV_N };

// matches each thread holds in pass D (fewer in the widest workgroups, to save registers)
template <int NT>
constexpr int md_k() { return NT >= 512 ? 4 : 8; }

// The distance table of these kernels: an 8-bit root takes nearly every distance code in one
// lookup, and its worst case (30 symbols, codes of at most 15 bits, in canonical order) is 400
// entries, fewer than the 592 a 6-bit root can need (zlib's bound for its 6-bit root)
constexpr int WDBITS = 8, WDTAB = 400;

// TOK: the decode phase of the two-phase inflate (tokens to global memory, no output window in LDS)
template <int NT, int SEG, bool TOK = false>
struct WaveSmem {
    // staged compressed bits of a round; at least 1 << LBITS words, the table build's scratch
    static constexpr int IN_WORDS = NT * SEG / 32 + 8 > (1 << LBITS) + 8 ? NT * SEG / 32 + 8 : (1 << LBITS) + 8;
    static_assert(IN_WORDS >= (1 << LBITS), "the staging area is the table build's scratch");
    // the tables and the staging area first: their offsets fit the 16-bit offset field of the
    // LDS instructions, so a lookup needs no address add
    uint32_t lit[LTAB];
    uint16_t dist[WDTAB];
    uint32_t in[IN_WORDS];
    union {
        struct {  // header and table build
            uint16_t code[320];
            uint8_t lens[320];
        };
        struct {  // rounds: each walk's exit and how it stopped
            int32_t exit_[NT];
            uint8_t stop_[NT];
        };
    };
    int32_t red[NT / CBM_WAVE];
    int64_t red64[NT / CBM_WAVE];
    int32_t lcount[16];
    int32_t lnext[16];
    int32_t chunkcnt[5][16];
    int32_t sv[SV_N];  // shared scalars (SV_*)
    alignas(16) uint8_t out[TOK ? 16 : 65536 + 16];  // the block's output window (not with TOK)
};


// the check that failed, reported in bits 8-15 of a block's status
enum { SITE_HEADER = 1, SITE_STORED = 2, SITE_TYPE = 3, SITE_CLCODE = 4, SITE_LENGTHS = 5, SITE_TABLES = 6,
       SITE_ROUND = 7, SITE_DIST = 8, SITE_OVERRUN = 9, SITE_SIZE = 10, SITE_ROUND_SIZE = 11 };

enum { STOP_NONE = 0, STOP_EOB = 1, STOP_BAD = 2 };

// decoder statistics (cycles of thread 0 per phase, and rounds), summed over blocks, in builds with
// CBM_INFLATE_STATS=1 (they cost registers)
#ifndef CBM_INFLATE_STATS
#define CBM_INFLATE_STATS 0
#endif
// passes of match forwarding in pass D, and the bits (in half segments) a speculative walk
// starts before its segment
#ifndef CBM_INFLATE_FWD_PASSES
#define CBM_INFLATE_FWD_PASSES 3
#endif
// symbols between two checkpoints of a speculative walk
#ifndef CBM_INFLATE_CP_STEPS
#define CBM_INFLATE_CP_STEPS 3
#endif
#ifndef CBM_INFLATE_OVL_HALVES
#define CBM_INFLATE_OVL_HALVES 3
#endif
enum { STAT_HEADER = 0, STAT_PASS_A, STAT_LINK, STAT_PASS_B, STAT_PASS_C, STAT_PASS_D, STAT_STORE, STAT_ROUNDS,
       STAT_DSTEPS, STAT_LINKIT, STAT_H_STAGE, STAT_H_PARSE, STAT_H_CL, STAT_H_CLDEC, STAT_BLOCKEND, STAT_H_STAGE2,
       STAT_CLWIN, STAT_DYN, STAT_D_MARK, STAT_D_LOOP, STAT_MBIG, STAT_MSUM, STAT_CLLINK, STAT_D_IDLE, STAT_D_LANEIT, STAT_D_WAVEM, STAT_N };
// the inflate workspace: the block counter, 7 spare words, the statistics, then the match lists
constexpr int INFLATE_WS_HDR = 512;
static_assert(8 * (8 + STAT_N) <= INFLATE_WS_HDR, "statistics overrun the workspace header");

// matches of a round, per workgroup, in global scratch: a block of at most 65,536 bytes has at
// most 21,846 matches of 3 bytes or more
constexpr int MLIST_CAP = 21888;

// length bases and extra bits of RFC 1951 3.2.5, computed
__device__ __forceinline__ void len_code(int c, int& base, int& extra) {
    if (c < 8) {
        base = c + 3;
        extra = 0;
    } else if (c < 28) {
        extra = (c >> 2) - 1;
        base = ((4 + (c & 3)) << extra) + 3;
    } else {
        base = 258;
        extra = 0;
    }
}

// 64 bits of the staged stream from bit p (p relative to the payload start)
template <class S>
__device__ __forceinline__ uint64_t peek64(const S& s, int p, int wbase) {
    const int w = (p >> 5) - wbase;
    const uint32_t lo = s.in[w], mid = s.in[w + 1], hi = s.in[w + 2];
    const uint32_t sh = (uint32_t)p & 31;
    return ((uint64_t)__builtin_amdgcn_alignbit(hi, mid, sh) << 32) | __builtin_amdgcn_alignbit(mid, lo, sh);
}

// The LDS byte address of a round's staging area less 4 * wbase (wave-uniform), so that the staged
// word holding bit p sits at sb + 4 * (p >> 5): a walk's step then finds its words with one shift
// and one add of a scalar
template <class S>
__device__ __forceinline__ uint32_t stage_base(const S& s, int wbase) {
    return (uint32_t)(uintptr_t)(const CBM_LDS uint32_t*)s.in - 4u * (uint32_t)wbase;
}
// peek64 through stage_base's address
__device__ __forceinline__ uint64_t peek64_sb(uint32_t sb, int p) {
    // the word index by a bit-field extract, which the compiler keeps (a shift it would fold with the
    // scaling into a shift, a mask and an add), then one shift-and-add
    const CBM_LDS uint32_t* q = (const CBM_LDS uint32_t*)(uintptr_t)(sb + __builtin_amdgcn_ubfe((uint32_t)p, 5, 27) * 4u);
    const uint32_t lo = q[0], mid = q[1], hi = q[2];
    const uint32_t sh = (uint32_t)p & 31;
    return ((uint64_t)__builtin_amdgcn_alignbit(hi, mid, sh) << 32) | __builtin_amdgcn_alignbit(mid, lo, sh);
}

// 32 bits of the staged stream from bit p
template <class S>
__device__ __forceinline__ uint32_t peek32(const S& s, int p, int wbase) {
    const int w = (p >> 5) - wbase;
    return __builtin_amdgcn_alignbit(s.in[w + 1], s.in[w], p & 31);
}


// one code-length symbol at bit p, looked up in the code-length table (7-bit, no subtables) held
// in the distance table's space: its bits (code and extra bits; 0 for an invalid code), the
// symbol (0..18) and how many lengths it gives
struct ClSym {
    int bits, sym, rep;
};
#ifndef CBM_CL_SEG
#define CBM_CL_SEG 24
#endif
constexpr int CL_SEG = CBM_CL_SEG;  // bits a lane takes: more than the 14 of the longest symbol
static_assert(CL_SEG >= 15 && CL_SEG <= 32, "a lane's walk records its starts in 32 bits and covers a symbol");

// the symbol at bit r of a lane's window (win holds 64 staged bits from the lane's first bit; a
// symbol starts at r < CL_SEG and takes at most 14 bits)
template <class S>
__device__ __forceinline__ ClSym cl_sym(const S& s, uint64_t win, int r) {
    const uint32_t v = (uint32_t)(win >> r);
    const uint32_t e = s.dist[v & 127];
    const int l = e & 15, sym = (e >> 4) & 0x1ff;
    const uint32_t x = v >> l;
    ClSym c{l, sym, 1};
    if (!l) return c;
    if (sym == 16) c.bits = l + 2, c.rep = 3 + (int)(x & 3);
    else if (sym == 17) c.bits = l + 3, c.rep = 3 + (int)(x & 7);
    else if (sym == 18) c.bits = l + 7, c.rep = 11 + (int)(x & 127);
    return c;
}

struct Sym {
    int kind;  // 0 literal, 1 match, 2 end of block, 3 invalid
    int bits;  // bits used
    int val;   // literal byte or match length
    int dist;
};

// Decode tables of this variant. Literal/length entries are 32 bits:
//   leaf:    bit 31 = 0, bits 0-3 = code length (0: no code), bits 4-6 = extra bits of a length,
//            bits 8-9 = kind (0 literal, 1 length, 2 end of block, 3 symbol 286 or 287, or no
//            code), bits 16-24 = the literal byte or the length's base (RFC 1951 3.2.5)
//   pointer: bit 31 = 1, bits 0-3 = index bits of the subtable, bits 4-14 = its offset
// Distance entries are 16 bits:
//   leaf:    bit 15 = 0, bits 0-3 = code length (0: no code, and symbols 30 and 31, which only a
//            fixed block can name), bits 4-7 = extra bits e, bits 8-9 = m, the distance's base
//            being (m << e) + 1 (RFC 1951 3.2.5: m = the symbol below 4, else 2 + its low bit)
//   pointer: bit 15 = 1, bits 0-3 = index bits, bits 4-14 = offset
// A subtable leaf holds the whole code length, root bits included. The code-length code's table
// keeps bits 0-3 = length and bits 4-12 = symbol.
constexpr uint32_t KIND_LIT = 0, KIND_LEN = 1, KIND_EOB = 2, KIND_BAD = 3;

__device__ __forceinline__ uint32_t lit_entry(int sym, int len) {
    if (sym < 256) return ((uint32_t)sym << 16) | (uint32_t)len;
    if (sym == 256) return (KIND_EOB << 8) | (uint32_t)len;
    if (sym > 285) return (KIND_BAD << 8) | (uint32_t)len;
    int base, extra;
    len_code(sym - 257, base, extra);
    return ((uint32_t)base << 16) | (KIND_LEN << 8) | ((uint32_t)extra << 4) | (uint32_t)len;
}
__device__ __forceinline__ uint32_t dist_entry(int sym, int len) {
    if (sym >= 30) return 0;  // invalid distance symbols decode as no code
    const int extra = sym < 4 ? 0 : (sym >> 1) - 1, m = sym < 4 ? sym : 2 | (sym & 1);
    return ((uint32_t)m << 8) | ((uint32_t)extra << 4) | (uint32_t)len;
}

// decode one symbol (a literal/length code and, for a length, its distance) from the next
// 64 bits of the stream; a symbol uses at most 15 + 5 + 15 + 13 = 48 bits. A code the table
// does not hold (distance symbols 30 and 31 included) gives kind 3; its bit count is then
// meaningless (0 for a literal/length code), and every walk stops at such a symbol.
template <class S>
__device__ __forceinline__ Sym decode_v(const S& s, uint64_t v) {
    const uint32_t v0 = (uint32_t)v;
    uint32_t e = s.lit[v0 & ((1u << LBITS) - 1)];
    if ((int32_t)e < 0) e = s.lit[((e >> 4) & 0x7ff) + __builtin_amdgcn_ubfe(v0, LBITS, e & 15)];
    const int cl = (int)(e & 15);
    Sym r;
    r.kind = (int)((e >> 8) & 3);  // entries without a code hold KIND_BAD
    r.bits = cl;
    r.val = (int)(e >> 16);
    r.dist = 0;
    if (r.kind == (int)KIND_LEN) {
        const int ex = (int)((e >> 4) & 7), u = cl + ex;
        r.val += (int)__builtin_amdgcn_ubfe(v0, cl, ex);
        const uint32_t x = (uint32_t)(v >> u);
        uint32_t d = s.dist[x & ((1u << WDBITS) - 1)];
        if (d & 0x8000) d = s.dist[((d >> 4) & 0x7ff) + __builtin_amdgcn_ubfe(x, WDBITS, d & 15)];
        const int dl = (int)(d & 15), de = (int)((d >> 4) & 15), dm = (int)((d >> 8) & 3);
        r.kind = dl ? (int)KIND_LEN : (int)KIND_BAD;
        r.bits = u + dl + de;
        r.dist = (dm << de) + 1 + (int)__builtin_amdgcn_ubfe(x, dl, de);
    }
    return r;
}

// how a walk stopped, from its last symbol's kind and its position after it: STOP_EOB at the
// end of block, STOP_BAD for a bad code or a symbol that ends past the payload, else STOP_NONE
__device__ __forceinline__ int walk_stop(int kind, int p, int in_bits) {
    return kind >= 2 ? kind - 1 : p > in_bits ? STOP_BAD : STOP_NONE;
}
// a symbol's output as a walk counts it: matches << 20 | bytes. The symbol that stops a walk (an end
// of block or a bad code) counts as a byte here; the walk takes it back once it has stopped
// (stop_count), which saves a select every step
__device__ __forceinline__ uint32_t sym_count(const Sym& y) {
    return y.kind == 1 ? (1u << 20) + (uint32_t)y.val : 1u;
}
__device__ __forceinline__ uint32_t stop_count(int last_kind) { return last_kind >= 2 ? 1u : 0u; }

// A walk's position on the staged stream. Each symbol is decoded from 64 bits read afresh from
// LDS at the position (three words, two funnel shifts): cheaper than keeping a register window
// and testing its fill every step.
struct Win {
    int p;
};
template <class S>
__device__ __forceinline__ void win_at(Win& w, const S&, int p, int) {
    w.p = p;
}
template <class S>
__device__ __forceinline__ Sym win_next(Win& w, const S& s, uint32_t sb) {
    const Sym y = decode_v(s, peek64_sb(sb, w.p));
    w.p += y.bits;
    return y;
}

// ------------------------------------------------------------------ cooperative table build
// Canonical Huffman decode table of lens[0:nsym) into tab (root bits, cap entries), with the
// checks of build_table; scratch32 holds 1 << root words. Returns (in all threads) success.
template <int NT, class S, class T>
__device__ bool coop_build(S& s, const uint8_t* lens, int nsym, int root, T* tab, int cap, int type,
                           uint32_t* scratch32) {
    constexpr uint32_t PTR = sizeof(T) == 4 ? 0x80000000u : 0x8000u;
    const int tid = threadIdx.x, lane = tid % CBM_WAVE, wave = tid / CBM_WAVE;
    const int rsize = 1 << root;
    if (tid < 16) s.lcount[tid] = 0;
    __syncthreads();
    for (int i = tid; i < nsym; i += NT) atomicAdd(&s.lcount[lens[i]], 1);
    __syncthreads();
    if (tid == 0) {
        int max = 15;
        while (max >= 1 && s.lcount[max] == 0) max--;
        int ok = 1, left = 1;
        for (int l = 1; l < 16; l++) {
            left = (left << 1) - s.lcount[l];
            if (left < 0) ok = 0;
        }
        if (ok && max > 0 && left > 0 && (type == TAB_CODES || max != 1)) ok = 0;
        int code = 0;
        s.lcount[0] = 0;
        for (int l = 1; l < 16; l++) {
            code = (code + s.lcount[l - 1]) << 1;
            s.lnext[l] = code;
        }
        s.sv[SV_OK] = ok;
        s.sv[SV_MAX] = max;
        s.sv[SV_CLEAR] = (left > 0 || max > root || max == 0) ? 1 : 0;  // root needs clearing
    }
    __syncthreads();
    if (!s.sv[SV_OK]) return false;
    const int max = s.sv[SV_MAX];
    if (s.sv[SV_CLEAR]) {  // entries without a code: kind 3 in a literal/length table, length 0
        const T none = (T)(type == TAB_LENS ? KIND_BAD << 8 : 0u);
        for (int i = tid; i < rsize; i += NT) tab[i] = none;
    }
    // canonical codes: rank of each symbol among symbols of its length, by 64-symbol chunks
    const int nchunk = (nsym + CBM_WAVE - 1) / CBM_WAVE;
    for (int c = wave; c < nchunk; c += NT / CBM_WAVE) {
        int sidx = c * CBM_WAVE + lane;
        int L = sidx < nsym ? lens[sidx] : 0;
        int rank = 0;
        for (int l = 1; l <= 15; l++) {
            uint64_t m = __ballot(L == l);
            if (L == l) rank = __popcll(m & ((1ull << lane) - 1));
            if (lane == 0) s.chunkcnt[c][l] = __popcll(m);
        }
        if (sidx < nsym) s.code[sidx] = (uint16_t)rank;
    }
    __syncthreads();
    for (int i = tid; i < nsym; i += NT) {
        int L = lens[i];
        if (!L) continue;
        int c = i / CBM_WAVE, base = s.lnext[L];
        for (int k = 0; k < c; k++) base += s.chunkcnt[k][L];
        s.code[i] = (uint16_t)(base + s.code[i]);
    }
    __syncthreads();
    if (max > root) {
        // deepest code under each root prefix (atomicMax on 32-bit scratch), then place subtables
        for (int i = tid; i < rsize; i += NT) scratch32[i] = 0;
        __syncthreads();
        for (int i = tid; i < nsym; i += NT) {
            int L = lens[i];
            if (L > root) atomicMax(&scratch32[bitrev(s.code[i], L) & (rsize - 1)], (uint32_t)(L - root));
        }
        __syncthreads();
        // exclusive prefix of subtable sizes over prefixes (each thread owns consecutive prefixes)
        const int per = (rsize + NT - 1) / NT;
        int lo = tid * per, hi = lo + per < rsize ? lo + per : rsize;
        int sz = 0;
        for (int i = lo; i < hi; i++) sz += scratch32[i] ? (1 << scratch32[i]) : 0;
        int tot;
        int off = rsize + block_exclusive_sum<NT, int>(sz, s.red, &tot);
        if (rsize + tot > cap) return false;  // uniform: every thread sees the same total
        for (int i = lo; i < hi; i++) {
            int m = scratch32[i];
            if (m) {
                tab[i] = (T)(PTR | ((uint32_t)off << 4) | (uint32_t)m);
                off += 1 << m;
            }
        }
        __syncthreads();
    }
    for (int i = tid; i < nsym; i += NT) {
        int L = lens[i];
        if (!L) continue;
        uint32_t rev = bitrev(s.code[i], L);
        const T e = (T)(type == TAB_LENS ? lit_entry(i, L) : type == TAB_DISTS ? dist_entry(i, L) : (uint32_t)((i << 4) | L));
        if (L <= root) {
            for (uint32_t j = rev; j < (uint32_t)rsize; j += 1u << L) tab[j] = e;
        } else {
            const uint32_t p = tab[rev & (rsize - 1)];
            const int m = p & 15, off = (p >> 4) & 0x7ff, sl = L - root;
            for (uint32_t j = rev >> root; j < (1u << m); j += 1u << sl) tab[off + j] = e;
        }
    }
    __syncthreads();
    return true;
}

// ------------------------------------------------------------------ match copies in LDS
// Copy n bytes from s to d (s + n <= d: no overlap) in the output window: bytes up to a 4-byte
// boundary of d, then whole aligned words (the source read as aligned words and realigned
// with v_alignbyte), then the last bytes. Source words may hold bytes outside the range that
// other threads are writing: those bytes are shifted out, never used.
__device__ __forceinline__ void copy_bytes(uint8_t* out, int d, int s, int n) {
    while (n > 0 && (d & 3)) {
        out[d++] = out[s++];
        n--;
    }
    if (n >= 4) {
        const uint32_t* w = (const uint32_t*)out;
        uint32_t* wd = (uint32_t*)out;
        const int sh = s & 3;
        // 16 bytes a step: five loads, then four stores, so one LDS round trip per step (the
        // source never overlaps the destination here)
        for (; n >= 16; n -= 16, d += 16, s += 16) {
            const int sw = s >> 2;
            const uint32_t a0 = w[sw], a1 = w[sw + 1], a2 = w[sw + 2], a3 = w[sw + 3], a4 = w[sw + 4];
            const int dw = d >> 2;
            if (sh == 0) {
                wd[dw] = a0, wd[dw + 1] = a1, wd[dw + 2] = a2, wd[dw + 3] = a3;
            } else {
                wd[dw] = __builtin_amdgcn_alignbyte(a1, a0, (uint32_t)sh);
                wd[dw + 1] = __builtin_amdgcn_alignbyte(a2, a1, (uint32_t)sh);
                wd[dw + 2] = __builtin_amdgcn_alignbyte(a3, a2, (uint32_t)sh);
                wd[dw + 3] = __builtin_amdgcn_alignbyte(a4, a3, (uint32_t)sh);
            }
        }
        if (sh == 0) {
            for (; n >= 4; n -= 4, d += 4, s += 4) wd[d >> 2] = w[s >> 2];
        } else if (n >= 4) {
            uint32_t lo = w[s >> 2];
            for (; n >= 4; n -= 4, d += 4, s += 4) {
                const uint32_t hi = w[(s >> 2) + 1];
                wd[d >> 2] = __builtin_amdgcn_alignbyte(hi, lo, (uint32_t)sh);
                lo = hi;
            }
        }
    }
    while (n > 0) {
        out[d++] = out[s++];
        n--;
    }
}

// Apply a match: len bytes at dest copied from dist bytes back (RFC 1951 3.2.3 semantics: an
// overlapping copy repeats the dist bytes before dest)
__device__ __forceinline__ void lds_copy_match(uint8_t* out, int dest, int dist, int len) {
    const int src = dest - dist;
    if (dist >= len) {
        copy_bytes(out, dest, src, len);
    } else if (dist >= 8) {
        // dest + k * dist .. : each piece repeats [src, src + dist), which is final
        for (int k = 0; k < len; k += dist) copy_bytes(out, dest + k, src, len - k < dist ? len - k : dist);
    } else {
        // the dist bytes before dest repeat: read them once, then only write
        uint64_t pat = 0;
        for (int q = 0; q < dist; q++) pat |= (uint64_t)out[src + q] << (8 * q);
        int r = 0;
        for (int q = 0; q < len; q++) {
            out[dest + q] = (uint8_t)(pat >> (8 * r));
            r = r + 1 == dist ? 0 : r + 1;
        }
    }
}

// Set (or clear) the bits [d, d + n) of a bitmap in LDS, a word at a time.
__device__ __forceinline__ void mark_bits(uint32_t* bm, int d, int n, bool set) {
    const int e = d + n;
    for (int w = d >> 5; w <= (e - 1) >> 5; w++) {
        const int a = w * 32 > d ? w * 32 : d, b = w * 32 + 32 < e ? w * 32 + 32 : e;
        const uint32_t m = (b - a == 32 ? 0xffffffffu : ((1u << (b - a)) - 1u)) << (a & 31);
        if (set) atomicOr(&bm[w], m);
        else atomicAnd(&bm[w], ~m);
    }
}

// Any bit of [a, b) set?
__device__ __forceinline__ bool any_bits(const uint32_t* bm, int a, int b) {
    for (int w = a >> 5; w <= (b - 1) >> 5; w++) {
        const int x = w * 32 > a ? w * 32 : a, y = w * 32 + 32 < b ? w * 32 + 32 : b;
        const uint32_t m = (y - x == 32 ? 0xffffffffu : ((1u << (y - x)) - 1u)) << (x & 31);
        if (__atomic_load_n(&bm[w], __ATOMIC_RELAXED) & m) return true;
    }
    return false;
}

// pass D, per wavefront: more lanes with ready matches than this and each lane copies its own;
// idle polls (with a short sleep each) before a round is given up as bad
#ifndef CBM_D_LANE_COPY
#define CBM_D_LANE_COPY 8
#endif
constexpr int D_LANE_COPY = CBM_D_LANE_COPY;
constexpr int D_IDLE_MAX = 1 << 16;

// Are the bytes match m (dest | dist << 17 | len << 33) reads final: is none of them (at or after
// op, the round's output start) pending in the bitmap?
__device__ __forceinline__ bool match_ready(const uint32_t* pend, uint64_t m, int op) {
    const int dest = (int)(m & 0x1ffff), dist = (int)((m >> 17) & 0xffff);
    const int len = (int)((m >> 33) & 0x1ff), src = dest - dist;
    const int lo = src < op ? op : src, hi = src + (dist < len ? dist : len);
    return !(lo < hi && any_bits(pend, lo, hi));
}

// lane `from`'s x (from is wavefront-uniform): two readlanes, no LDS round trip
__device__ __forceinline__ uint64_t lane_u64(uint64_t x, int from) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, from);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), from);
    return ((uint64_t)hi << 32) | lo;
}

// Match m copied by the lanes of a wavefront, a byte a lane: the bytes read, [src, src + len)
// or the dist bytes before dest, never overlap the bytes written
__device__ __forceinline__ void wave_copy_match(uint8_t* out, uint64_t m, int lane) {
    const int dest = (int)(m & 0x1ffff), dist = (int)((m >> 17) & 0xffff);
    const int len = (int)((m >> 33) & 0x1ff), src = dest - dist;
    if (dist >= len) {
        for (int i = lane; i < len; i += CBM_WAVE) out[dest + i] = out[src + i];
    } else {
        for (int i = lane; i < len; i += CBM_WAVE) out[dest + i] = out[src + (int)((uint32_t)i % (uint32_t)dist)];
    }
}

// ------------------------------------------------------------------ staging
// Stage payload words [w0, w0 + IN_WORDS) (word = 4 payload bytes; payload byte 0 = in[0])
// into LDS; bytes at or past `limit` (payload bytes + trailer) read as zero. A thread issues all
// its loads before it writes any word, so a stage costs one global-memory latency, not one per
// word; the aligned source pointer is derived from `in` so the loads stay global loads.
template <int NT, class S>
__device__ void stage(S& s, const uint8_t* in, int w0, int limit) {
    const int sh = (int)((uintptr_t)in & 3);
    const uint32_t* al = (const uint32_t*)(in - sh);
    constexpr int PER = (S::IN_WORDS + NT - 1) / NT;
    uint32_t x[PER], y[PER];
#pragma unroll
    for (int u = 0; u < PER; u++) {
        const int i = (int)threadIdx.x + u * NT, byte0 = 4 * (w0 + i);
        x[u] = 0, y[u] = 0;
        if (i < S::IN_WORDS && byte0 + 4 <= limit) {
            const int k = (byte0 + sh) >> 2;
            x[u] = al[k];
            if (sh) y[u] = al[k + 1];
        }
    }
#pragma unroll
    for (int u = 0; u < PER; u++) {
        const int i = (int)threadIdx.x + u * NT, byte0 = 4 * (w0 + i);
        if (i >= S::IN_WORDS) break;
        uint32_t w = 0;
        if (byte0 + 4 <= limit) {
            w = sh ? __builtin_amdgcn_alignbit(y[u], x[u], (uint32_t)(8 * sh)) : x[u];
        } else if (byte0 < limit) {
            for (int j = 0; j < 4 && byte0 + j < limit; j++) w |= (uint32_t)in[byte0 + j] << (8 * j);
        }
        s.in[i] = w;
    }
}

// n bytes from s to d in global memory (the two do not overlap), by the nt threads of a workgroup
// (this one is t): bytes up to a 16-byte boundary of d, then 16-byte stores of d, each from five
// aligned words of s realigned (four pieces a thread loaded before any is stored, so that the
// loads are in flight together), then the last bytes. Reads up to 3 bytes before s and past
// s + n, in the same buffer.
__device__ __forceinline__ void copy_global(uint8_t* __restrict__ d, const uint8_t* __restrict__ s, int n, int t,
                                            int nt) {
    int head = (int)((16 - ((uintptr_t)d & 15)) & 15);
    if (head > n) head = n;
    for (int i = t; i < head; i += nt) d[i] = s[i];
    d += head, s += head, n -= head;
    const int nv = n >> 4;
    const uint32_t sh = (uint32_t)((uintptr_t)s & 3);
    const uint32_t* sw = (const uint32_t*)(s - sh);
    for (int v0 = t; v0 < nv; v0 += 4 * nt) {
        uint32_t a[4][5];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int v = v0 + j * nt < nv ? v0 + j * nt : v0;  // (past the last piece: this thread's first again)
#pragma unroll
            for (int k = 0; k < 5; k++) a[j][k] = sw[4 * v + k];
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int v = v0 + j * nt;
            if (v < nv) {
                uint4 o;
                o.x = __builtin_amdgcn_alignbyte(a[j][1], a[j][0], sh);
                o.y = __builtin_amdgcn_alignbyte(a[j][2], a[j][1], sh);
                o.z = __builtin_amdgcn_alignbyte(a[j][3], a[j][2], sh);
                o.w = __builtin_amdgcn_alignbyte(a[j][4], a[j][3], sh);
                *(uint4*)(d + 16 * v) = o;
            }
        }
    }
    for (int i = 16 * nv + t; i < n; i += nt) d[i] = s[i];
}

// ------------------------------------------------------------------ CRC32 tables
// Slicing-by-4 tables (Kounavis and Berry, 2008): table k maps a byte followed by k zero bytes,
// for the reflected polynomial 0xEDB88320. Built at compile time; read through the vector memory
// path (they stay in the L1 cache), which leaves the LDS, shared by every phase, to the decoders.
struct CrcTables {
    uint32_t t[4][256];
};
constexpr CrcTables make_crc_tables() {
    CrcTables c{};
    for (int i = 0; i < 256; i++) {
        uint32_t x = (uint32_t)i;
        for (int k = 0; k < 8; k++) x = (x & 1) ? 0xEDB88320u ^ (x >> 1) : x >> 1;
        c.t[0][i] = x;
    }
    for (int k = 1; k < 4; k++)
        for (int i = 0; i < 256; i++) c.t[k][i] = (c.t[k - 1][i] >> 8) ^ c.t[0][c.t[k - 1][i] & 255];
    return c;
}
__constant__ CrcTables kCrcTab = make_crc_tables();

// ------------------------------------------------------------------ the kernel
// two workgroups share a CU (the LDS holds two output windows): ask for the registers that let
// both run, NT / 128 wavefronts per SIMD
//
// TOK (the decode phase of the two-phase inflate): no output window. Literals go straight to their
// place in dst and matches to the block's list in `pool` (MCAP_TOK entries a block, in output
// order; mcount[b] of them); status[b] is CBM_BLOCK_OK, or BLOCK_DEFER for a block this phase
// does not finish (any check that fails, or more matches than its list holds), which the
// one-phase kernel decodes again. No CRC32 here. Speculative walks start at most TOK_OVL bits
// before their segment.
// idx / n_dev (may be null): decode blocks idx[0 .. *n_dev) instead of 0 .. n_blocks.
constexpr int MCAP_TOK = 6144;
template <int MPL, int SPAN>
__device__ __forceinline__ void resolve_block(uint8_t* out, int L, uint32_t expect, uint32_t* crc_out_b,
                                              int32_t* status_b, int64_t* bad, const uint64_t* ml, int M,
                                              uint8_t* sp, uint32_t* pend, const uint32_t (*T)[256], uint32_t xk,
                                              int lane);
#ifndef CBM_TOK_OVL
#define CBM_TOK_OVL 256
#endif
// FUSE (with TOK): the wavefront resolves each block it decoded at once (resolve_block, in the
// staging area, which is free until the next block stages), instead of a resolve launch after
template <int NT, int SEG, bool TOK = false, bool FUSE = false>
__global__ void __launch_bounds__(NT, (TOK ? 4 : NT >= 256 ? NT / 128 : 1)) inflate_wave(const uint8_t* src, const int64_t* blk_off,
                                                   const int32_t* bsize, const int32_t* isize,
                                                   const uint32_t* expect, const int64_t* out_off,
                                                   int64_t n_blocks, uint8_t* dst, uint32_t* crc_out,
                                                   int32_t* status, int64_t* bad_count,
                                                   unsigned long long* next_block, uint64_t* mlists,
                                                   unsigned long long* stats, const int64_t* idx,
                                                   const int64_t* n_dev, uint64_t* pool, int32_t* mcount,
                                                   int64_t mcap, int64_t kmax) {
    using S = WaveSmem<NT, SEG, TOK>;
    static_assert(!FUSE || (TOK && NT == CBM_WAVE), "a fused resolve is one wavefront's");
    static_assert(!FUSE || S::IN_WORDS * 4 >= 2048 + 48 + ((2048 + 16) / 32 + 2) * 4,
                  "the staging area holds the resolve's window and bitmap");
    extern __shared__ __align__(16) uint8_t smem_raw[];
    S& s = *(S*)smem_raw;
    const int tid = threadIdx.x;
    if (n_dev) {  // a list of blocks: workgroups past its length have nothing to take (uniform)
        const int64_t nd = *n_dev;
        if ((int64_t)blockIdx.x >= (nd < kmax ? nd : kmax)) return;
    }
    uint64_t* mlist = TOK ? nullptr : mlists + (size_t)blockIdx.x * MLIST_CAP;
#if CBM_INFLATE_STATS
    // cycles per phase, kept by thread 0 and added to stats[] once per block
    long long acc[STAT_N] = {};
    long long tlast = tid == 0 ? clock64() : 0;
#define PHASE(k)                          \
    do {                                  \
        if (tid == 0) {                   \
            long long t_ = clock64();     \
            acc[k] += t_ - tlast;         \
            tlast = t_;                   \
        }                                 \
    } while (0)
#define COUNT(k)                \
    do {                        \
        if (tid == 0) acc[k]++; \
    } while (0)
#else
#define PHASE(k) \
    do {         \
    } while (0)
#define COUNT(k) \
    do {         \
    } while (0)
    (void)stats;
#endif
    // this thread's CRC32 slice factor x^(8 * SLICE * (NT - 1 - tid)) mod P, computed with the
    // workgroup's first block (a launch over the blocks another kernel deferred often has none)
    constexpr int SLICE = (65536 + NT - 1) / NT;
    uint32_t xfac = 0;
    bool have_xfac = false;
    __syncthreads();

    for (;;) {
        if (tid == 0) s.sv[SV_BLOCK] = (int)atomicAdd(next_block, 1ull);
        __syncthreads();
        const int64_t kb = s.sv[SV_BLOCK];
        if (kb >= (n_dev ? *n_dev : n_blocks) || kb >= kmax) break;
        if (!have_xfac) {
            uint32_t r = 0x80000000u, sq = 0x40000000u;  // x^0, x^1
            for (int j = 0; j < 3; j++) sq = crc_mulmod(sq, sq);  // x^8
            uint32_t k = (uint32_t)SLICE * (uint32_t)(NT - 1 - tid);
            while (k) {
                if (k & 1) r = crc_mulmod(r, sq);
                sq = crc_mulmod(sq, sq);
                k >>= 1;
            }
            xfac = r;
            have_xfac = true;
        }
        const int64_t b = idx ? idx[kb] : kb;
        // this block's match list (TOK) and output. The output offset is read once (a store through
        // dst may alias out_off, so a read inside a store loop would be repeated at every store),
        // and both bases are made wave-uniform, so that pass C's stores take them from scalar
        // registers with a 32-bit offset
        uint64_t* blist = TOK ? pool + (size_t)uniform64(kb) * mcap : nullptr;
        uint8_t* const bdst = dst + uniform64(out_off[b]);
        const uint8_t* in = src + blk_off[b] + 18;
        const int in_bytes = bsize[b] - 26;
        const int in_bits = 8 * in_bytes;
        const int limit = in_bytes + 8;
        const int cap = isize[b];
        if (tid == 0) {
            s.sv[SV_POS] = 0;
            s.sv[SV_OP] = 0;
            s.sv[SV_ERR] = CBM_BLOCK_OK;
            s.sv[SV_SITE] = 0;
            s.sv[SV_FINAL] = 0;
            s.sv[SV_MCNT] = 0;
        }
        __syncthreads();
        while (!s.sv[SV_FINAL] && s.sv[SV_ERR] == CBM_BLOCK_OK) {
            // ---- block header, read by thread 0 from a fresh staging at pos
            const int pos0 = s.sv[SV_POS];
            const int wb0 = pos0 >> 5;
            stage<NT>(s, in, wb0, limit);
            __syncthreads();
            PHASE(STAT_H_STAGE);
            if (tid == 0) {
                __builtin_amdgcn_s_setprio(3);  // the rest of the workgroup waits
                int p = pos0;
                if (p + 3 > in_bits) s.sv[SV_ERR] = CBM_BLOCK_DEFLATE, s.sv[SV_SITE] = SITE_HEADER;
                uint64_t v = peek64(s, p, wb0);
                s.sv[SV_FINAL] = (int)(v & 1);
                int type = (int)((v >> 1) & 3);
                s.sv[SV_TYPE] = type;
                p += 3;
                if (type == 0) {
                    p = (p + 7) & ~7;
                    uint64_t w = peek64(s, p, wb0);
                    uint32_t len = (uint32_t)(w & 0xffff), nlen = (uint32_t)((w >> 16) & 0xffff);
                    p += 32;
                    if (len != (~nlen & 0xffffu) || (p >> 3) + (int)len > in_bytes)
                        s.sv[SV_ERR] = CBM_BLOCK_DEFLATE, s.sv[SV_SITE] = SITE_STORED;
                    else if (s.sv[SV_OP] + (int)len > cap)
                        s.sv[SV_ERR] = CBM_BLOCK_SIZE, s.sv[SV_SITE] = SITE_STORED;
                    s.sv[SV_LEN] = (int)len;
                    s.sv[SV_SRC] = p >> 3;
                    p += 8 * (int)len;
                } else if (type == 2) {
                    uint32_t hlit = (uint32_t)((v >> 3) & 31) + 257, hdist = (uint32_t)((v >> 8) & 31) + 1;
                    uint32_t hclen = (uint32_t)((v >> 13) & 15) + 4;
                    p += 14;
                    if (hlit > 286 || hdist > 30) s.sv[SV_ERR] = CBM_BLOCK_DEFLATE, s.sv[SV_SITE] = SITE_HEADER;
                    // the code-length code's lengths, 3 bits each, are read by 19 threads below
                    s.sv[SV_F0] = (int)hclen;
                    s.sv[SV_F1] = p;
                    p += 3 * (int)hclen;
                    s.sv[SV_LEN] = (int)hlit;
                    s.sv[SV_SRC] = (int)hdist;
                } else if (type == 3) {
                    s.sv[SV_ERR] = CBM_BLOCK_DEFLATE, s.sv[SV_SITE] = SITE_TYPE;
                }
                s.sv[SV_POS] = p;
                __builtin_amdgcn_s_setprio(0);
            }
            __syncthreads();
            PHASE(STAT_H_PARSE);
            if (s.sv[SV_ERR] != CBM_BLOCK_OK) break;
            const int type = s.sv[SV_TYPE];
            if (type == 0) {
                const int len = s.sv[SV_LEN], sp = s.sv[SV_SRC], op = s.sv[SV_OP];
                if constexpr (TOK) {
                    copy_global(bdst + op, in + sp, len, tid, NT);
                } else {
                    for (int i = tid; i < len; i += NT) s.out[op + i] = in[sp + i];
                }
                __syncthreads();
                if (tid == 0) s.sv[SV_OP] = op + len;
                __syncthreads();
                continue;
            }
            if (type == 1) {
                for (int i = tid; i < 288; i += NT) s.lens[i] = i < 144 ? 8 : i < 256 ? 9 : i < 280 ? 7 : 8;
                __syncthreads();
                coop_build<NT>(s, s.lens, 288, LBITS, s.lit, LTAB, TAB_LENS, s.in);
                for (int i = tid; i < 32; i += NT) s.lens[i] = 5;
                __syncthreads();
                coop_build<NT>(s, s.lens, 32, WDBITS, s.dist, WDTAB, TAB_DISTS, s.in);
            } else {
                const int hlit = s.sv[SV_LEN], hdist = s.sv[SV_SRC];
                if (tid < 19) {
                    const int h = s.sv[SV_F0], q = s.sv[SV_F1] + 3 * tid;
                    s.lens[kClOrder[tid]] = tid < h ? (uint8_t)(peek32(s, q, wb0) & 7) : 0;
                }
                __syncthreads();
                // the code-length code, then the code lengths (wavefront 0); the code-length
                // table goes in the distance table's space
                bool ok = coop_build<NT>(s, s.lens, 19, CBITS, s.dist, WDTAB, TAB_CODES, s.in);
                if (!ok) {
                    if (tid == 0) s.sv[SV_ERR] = CBM_BLOCK_DEFLATE, s.sv[SV_SITE] = SITE_CLCODE;
                    __syncthreads();
                    break;
                }
                PHASE(STAT_H_CL);
                COUNT(STAT_DYN);
                const int p1 = s.sv[SV_POS];
                const int wb1 = p1 >> 5;
                stage<NT>(s, in, wb1, limit);
                __syncthreads();
                PHASE(STAT_H_STAGE2);
                if (tid < CBM_WAVE) {
                    // wavefront 0 decodes the code lengths, CL_SEG bits a lane: each lane walks its
                    // bits from their first bit, the walks are linked as the rounds link theirs (a
                    // lane whose entry moves walks again until it meets a symbol start it saw), then
                    // the lanes count their lengths and write them at offsets a scan gives. Up to
                    // 64 * CL_SEG bits a pass; the lengths of a block take at most 316 * 14 bits.
                    __builtin_amdgcn_s_setprio(3);  // the rest of the workgroup waits
                    const int lane = tid;
                    const int total = hlit + hdist;
                    int wbase = p1, n0 = 0, prev0 = -1, err = 0, endp = -1;
                    while (true) {
                        COUNT(STAT_CLWIN);
                        // every symbol start a pass decodes lies below wbase + 64 * CL_SEG
                        if (((wbase + CBM_WAVE * CL_SEG) >> 5) - wb1 + 2 >= S::IN_WORDS) {
                            err = 2;
                            break;
                        }
                        const int a = wbase + lane * CL_SEG, aend = a + CL_SEG;
                        const uint64_t win = peek64(s, a, wb1);  // every symbol this lane decodes
                        uint32_t seen = 0;
                        int p = a, bad = 0;
                        while (p < aend) {
                            seen |= 1u << (p - a);
                            const ClSym c = cl_sym(s, win, p - a);
                            if (!c.bits || p + c.bits > in_bits) {
                                bad = 1;
                                break;
                            }
                            p += c.bits;
                        }
                        // p: the first symbol start at or past aend, or the start of a symbol that
                        // is invalid or runs past the input (bad)
                        const int a_exit = p, a_bad = bad;
                        int entry = a;
                        while (true) {
                            const int want = wave_up1(p, wbase), wbad = wave_up1(bad, 0);
                            bool changed = false;
                            if (!wbad && want != entry) {
                                int q = want, nb = 0;
                                while (q < aend) {
                                    if (q >= a && ((seen >> (q - a)) & 1u)) {
                                        q = a_exit;
                                        nb = a_bad;
                                        break;
                                    }
                                    const ClSym c = cl_sym(s, win, q - a);
                                    if (!c.bits || q + c.bits > in_bits) {
                                        nb = 1;
                                        break;
                                    }
                                    q += c.bits;
                                }
                                changed = q != p || nb != bad;
                                p = q;
                                bad = nb;
                                entry = want;
                            }
                            COUNT(STAT_CLLINK);
                            if (!__ballot(changed)) break;
                        }
                        // lanes up to the first that stopped on a bad symbol hold the true walk
                        const uint64_t badm = __ballot(bad);
                        const int fb = badm ? __ffsll((unsigned long long)badm) - 1 : CBM_WAVE;
                        const bool alive = lane <= fb;
                        int cnt = 0, last = -1;
                        if (alive) {
                            for (int q = entry; q < p;) {
                                const ClSym c = cl_sym(s, win, q - a);
                                if (!c.bits) break;  // not met: the symbols before p are valid
                                cnt += c.rep;
                                last = c.sym < 16 ? c.sym : c.sym == 16 ? last : 0;
                                q += c.bits;
                            }
                        }
                        // offsets (a sum over the lanes before) and the length a leading 16 repeats
                        const int inc = wave_scan_sum(cnt);
                        const int key = wave_scan_max(last >= 0 ? ((lane + 1) << 5) | last : 0);
                        const int kex = wave_up1(key, 0);
                        int n = n0 + inc - cnt, prev = kex ? (kex & 31) : prev0, myerr = 0, myend = -1;
                        if (alive && n < total) {
                            int q = entry;
                            while (q < p && n < total) {
                                const ClSym c = cl_sym(s, win, q - a);
                                int val = 0;
                                if (!c.bits) {
                                    myerr = 1;
                                    break;
                                }
                                if (c.sym < 16) {
                                    val = c.sym;
                                } else if (c.sym == 16) {
                                    if (prev < 0) {
                                        myerr = 1;  // a repeat with no length before it
                                        break;
                                    }
                                    val = prev;
                                }
                                if (n + c.rep > total) {
                                    myerr = 1;  // a repeat past the last length
                                    break;
                                }
                                for (int k = 0; k < c.rep; k++) s.lens[n + k] = (uint8_t)val;
                                n += c.rep;
                                prev = val;
                                q += c.bits;
                            }
                            if (!myerr) {
                                if (n == total) myend = q;
                                else if (bad) myerr = 1;  // the lengths end on a bad symbol
                            }
                        }
                        if (__ballot(myerr)) {
                            err = 1;
                            break;
                        }
                        const uint64_t em = __ballot(myend >= 0);
                        if (em) {
                            endp = __builtin_amdgcn_readlane(myend, __ffsll((unsigned long long)em) - 1);
                            break;
                        }
                        // every lane's walk is true and the lengths go on past this pass
                        n0 = __builtin_amdgcn_readlane(n0 + inc, CBM_WAVE - 1);
                        const int k63 = __builtin_amdgcn_readlane(key, CBM_WAVE - 1);
                        prev0 = k63 ? (k63 & 31) : prev0;
                        wbase = __builtin_amdgcn_readlane(p, CBM_WAVE - 1);
                    }
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_s_setprio(0);
                    if (lane == 0) {
                        if (!err && s.lens[256] == 0) err = 1;
                        s.sv[SV_ERR] = err == 1 ? CBM_BLOCK_DEFLATE : CBM_BLOCK_OK;
                        if (err) s.sv[SV_SITE] = SITE_LENGTHS;
                        s.sv[SV_CHANGED] = err == 2;
                        s.sv[SV_POS] = endp;
                    }
                }
                __syncthreads();
                PHASE(STAT_H_CLDEC);
                if (s.sv[SV_CHANGED]) {
                    // code lengths longer than a staging window: cannot happen for 320 lengths of
                    // at most 7 + 7 bits (4,480 bits) with windows of NT * SEG >= 8,192 bits
                    if (tid == 0) s.sv[SV_ERR] = CBM_BLOCK_DEFLATE, s.sv[SV_SITE] = SITE_LENGTHS;
                    __syncthreads();
                }
                if (s.sv[SV_ERR] != CBM_BLOCK_OK) break;
                ok = coop_build<NT>(s, s.lens, hlit, LBITS, s.lit, LTAB, TAB_LENS, s.in);
                if (ok) {
                    // distance lengths follow the literal lengths
                    for (int i = tid; i < hdist; i += NT) s.code[i] = s.lens[hlit + i];
                    __syncthreads();
                    for (int i = tid; i < hdist; i += NT) s.lens[i] = (uint8_t)s.code[i];
                    __syncthreads();
                    ok = coop_build<NT>(s, s.lens, hdist, WDBITS, s.dist, WDTAB, TAB_DISTS, s.in);
                }
                if (!ok) {
                    if (tid == 0) s.sv[SV_ERR] = CBM_BLOCK_DEFLATE, s.sv[SV_SITE] = SITE_TABLES;
                    __syncthreads();
                    break;
                }
            }
            PHASE(STAT_HEADER);
            // ---- Huffman-coded data, round by round
            for (;;) {
                const int base = s.sv[SV_POS];
                const int wb = base >> 5;
                const uint32_t sb = stage_base(s, wb);
                stage<NT>(s, in, wb, limit);
                __syncthreads();
                // segment bits this round: SEG, or fewer in the last round of the final Huffman
                // block, where the payload's end bounds the bits left (a multiple of 32 bits that
                // covers them, so every thread has a share)
                const int bits_left = in_bits - base;
                const int seg = s.sv[SV_FINAL] && bits_left < NT * SEG
                                    ? max(32, ((bits_left + NT - 1) / NT + 31) & ~31)
                                    : SEG;
                const int a = base + tid * seg;         // speculative start
                const int aend = a + seg;
                const bool active = a < in_bits;        // thread 0 is always active
                // pass A. Threads after the first start OVL bits before their segment, so that the
                // walk has most likely fallen onto the true symbol boundaries when it enters the
                // segment. A walk that stops (end of block or a bad code) before its segment starts
                // again from the segment start. Inside the segment the walk keeps NCP checkpoints:
                // checkpoint 0 is its first symbol start at or after a, and checkpoint j the start
                // j * CP_STEPS symbols later, each with the output bytes and matches counted from
                // checkpoint 0 up to there; and the count to its exit.
                const int OVL = TOK ? min(seg * CBM_INFLATE_OVL_HALVES / 2, CBM_TOK_OVL)
                                    : seg * CBM_INFLATE_OVL_HALVES / 2;
#ifdef CBM_INFLATE_NCP
                constexpr int NCP = CBM_INFLATE_NCP, CP_STEPS = CBM_INFLATE_CP_STEPS;
#else
                constexpr int NCP = SEG / 32 < 8 ? SEG / 32 : 8, CP_STEPS = CBM_INFLATE_CP_STEPS;
#endif
                int cp[NCP];
                uint32_t cum[NCP];
#pragma unroll
                for (int j = 0; j < NCP; j++) cp[j] = 0x7fffffff, cum[j] = 0;
                uint32_t run = 0;  // matches << 20 | bytes, from checkpoint 0
                int p = a, stop = STOP_NONE;
                if (active) {
                    Win w;
                    win_at(w, s, a - OVL > base ? a - OVL : base, wb);
                    // the overlap: symbol lengths only, up to the first symbol start at or after a
                    while (w.p < a) {
                        const Sym y = win_next(w, s, sb);
                        if (y.kind >= 2 || w.p > in_bits) {
                            win_at(w, s, a, wb);
                            break;
                        }
                    }
                    // the segment, with a checkpoint every CP_STEPS symbols (the same step for all
                    // lanes of a wavefront, so recording one costs no test per symbol)
                    // a walk goes on while its symbols are literals or matches and it is inside
                    // both its segment and the payload; how it stopped is worked out once, after
                    const int lim = min(aend, in_bits + 1);
                    int lk = 0;  // the last symbol's kind
                    bool walking = w.p < aend;
                    auto step = [&]() {
                        const Sym y = win_next(w, s, sb);
                        run += sym_count(y);
                        lk = y.kind;
                        walking = y.kind < 2 && w.p < lim;
                    };
#pragma unroll
                    for (int j = 0; j < NCP; j++) {
                        if (walking) cp[j] = w.p, cum[j] = run;
                        for (int k = 0; k < CP_STEPS && walking; k++) step();
                    }
                    while (walking) step();
                    p = w.p;
                    stop = walk_stop(lk, p, in_bits);
                    run -= stop_count(lk);
                }
                s.exit_[tid] = p;
                s.stop_[tid] = active ? stop : STOP_BAD;
                // the walk's exit is the exit from a only if it passed through a; else the first
                // link step walks again from the predecessor's exit
                int entry = cp[0] == a ? a : -1;
                // a later walk that lands on a checkpoint of this first walk takes this walk's end
                // and counts from there
                const int a_exit = p, a_stop = stop;
                const uint32_t a_run = run;
                __syncthreads();
                PHASE(STAT_PASS_A);
                // link the walks. Each wavefront runs its own links to a fixed point, lane t taking
                // lane t-1's exit by a DPP shift, with no workgroup barrier; only exits that cross a
                // wavefront boundary go through LDS, and a pass repeats while one of them moved.
                const int lane = tid % CBM_WAVE;
                int pub = active ? stop : STOP_BAD;  // the stop a successor sees
                if (tid == 0) s.sv[SV_F0] = s.sv[SV_F1] = 0;
                __syncthreads();
                for (int it = 0;; it++) {
                    const int w0 = tid == 0 ? base : s.exit_[tid - 1];
                    const int ps0 = tid == 0 ? STOP_NONE : s.stop_[tid - 1];
                    for (;;) {
                        const int want = wave_up1(p, w0);
                        const int pstop = wave_up1(pub, ps0);
                        const bool redo = active && pstop == STOP_NONE && want != entry;
                        if (!__ballot(redo)) break;
                        if (redo) {
                            // walk from the new entry until a checkpoint of the first walk
                            int lk = 0, sj = -1;
                            uint32_t r2 = 0;
                            Win w;
                            win_at(w, s, want, wb);
                            while (w.p < aend) {
#pragma unroll
                                for (int j = 0; j < NCP; j++)
                                    if (w.p == cp[j]) sj = j;
                                if (sj >= 0) break;
                                Sym y = win_next(w, s, sb);
                                r2 += sym_count(y);
                                lk = y.kind;
                                if (y.kind >= 2 || w.p > in_bits) break;
                            }
                            int np = w.p;
                            int nstop = walk_stop(lk, np, in_bits);
                            r2 -= stop_count(lk);
                            if (sj >= 0) {
                                uint32_t c = 0;
#pragma unroll
                                for (int j = 0; j < NCP; j++)
                                    if (j == sj) c = cum[j];
                                np = a_exit;
                                nstop = a_stop;
                                r2 += a_run - c;
                            }
                            entry = want;
                            run = r2;
                            p = np;
                            stop = nstop;
                            pub = stop;
                        }
                        COUNT(STAT_LINKIT);
                    }
                    __syncthreads();  // every read of the exits of this pass is done
                    s.exit_[tid] = p;
                    s.stop_[tid] = pub;
                    if (tid == 0) s.sv[SV_F0 + (it + 1) % 2] = 0;
                    __syncthreads();
                    if (lane == 0 && tid > 0 && (s.exit_[tid - 1] != w0 || s.stop_[tid - 1] != ps0))
                        s.sv[SV_F0 + it % 2] = 1;
                    __syncthreads();
                    if (!s.sv[SV_F0 + it % 2]) break;
                }
                PHASE(STAT_LINK);
                // the first walk that stops ends the round (and the Huffman block, on EOB)
                if (tid == 0) s.sv[SV_TEND] = NT;
                __syncthreads();
                if (stop != STOP_NONE) atomicMin(&s.sv[SV_TEND], tid);
                __syncthreads();
                const int tend = s.sv[SV_TEND];
                const bool live = tid <= tend && tid < NT;
                // output bytes and matches of each live walk, entry .. exit, as the walks counted them
                const int cnt = live ? (int)(run & 0xfffff) : 0, nmatch = live ? (int)(run >> 20) : 0;
                int64_t tot2;
                const int64_t pre = block_exclusive_sum<NT, int64_t>(((int64_t)nmatch << 32) | (uint32_t)cnt,
                                                                      (int64_t*)s.red64, &tot2);
                const int ob = (int)(uint32_t)pre, mbase = (int)(pre >> 32);
                const int total = (int)(uint32_t)tot2, mtotal = (int)(tot2 >> 32);
                const int op = s.sv[SV_OP];
                const int tstop = tend < NT ? s.stop_[tend] : STOP_NONE;
                if (tstop == STOP_BAD || op + total > cap || (TOK && s.sv[SV_MCNT] + mtotal > mcap)) {
                    __syncthreads();
                    if (tid == 0) {
                        s.sv[SV_ERR] = tstop == STOP_BAD ? CBM_BLOCK_DEFLATE : CBM_BLOCK_SIZE;
                        s.sv[SV_SITE] = tstop == STOP_BAD ? SITE_ROUND : SITE_ROUND_SIZE;
                    }
                    __syncthreads();
                    break;
                }
                PHASE(STAT_PASS_B);
                // pass C: literals go to their place; matches go to the list, in output order
                if (live && cnt) {
                    const int e = s.exit_[tid];
                    int o = op + ob, k = mbase + (TOK ? s.sv[SV_MCNT] : 0);
                    bool bad = false;
                    Win w;
                    win_at(w, s, tid == 0 ? base : s.exit_[tid - 1], wb);
                    // the walk repeats a walk of pass A or of the links, which met no bad code and
                    // stops after an end of block; a distance reaching before the block's first
                    // byte marks the block bad (the stores stay inside its output and list)
                    int kind = 0;
                    while (w.p < e && kind < 2) {
                        const Sym y = win_next(w, s, sb);
                        kind = y.kind;
                        bad |= kind == 1 && y.dist > o;
                        // o | dist << 17 | len << 33 (a distance of 32,768 reaches bit 32)
                        const uint64_t mrec = ((uint64_t)(((uint32_t)y.dist >> 15) | ((uint32_t)y.val << 1)) << 32) |
                                              ((uint32_t)o | ((uint32_t)y.dist << 17));
                        if constexpr (TOK) {
                            if (kind == 0) bdst[(uint32_t)o] = (uint8_t)y.val;
                            else if (kind == 1) blist[(uint32_t)k++] = mrec;
                        } else {
                            if (kind == 0) s.out[o] = (uint8_t)y.val;
                            else if (kind == 1) mlist[k++] = mrec;
                        }
                        o += kind == 1 ? y.val : 1;  // (past an end of block, o is not used)
                    }
                    if (bad) {
                        atomicOr(&s.sv[SV_ERR], CBM_BLOCK_DEFLATE);
                        s.sv[SV_SITE] = SITE_DIST;
                    }
                }
                // the pass's LDS writes before pass D reads them (TOK: its global writes are read
                // only by later launches)
                if constexpr (!TOK) __threadfence_block();
                __syncthreads();
                if (s.sv[SV_ERR] != CBM_BLOCK_OK) break;
                PHASE(STAT_PASS_C);
                // pass D: apply the round's matches. A match can be copied once the bytes it reads
                // are final. Variants whose staging area holds a bit per output byte (64 Ki bits)
                // keep the round's pending output bytes there: each step, every thread applies each
                // of its matches (up to MD_K, in list order) whose source bytes are no longer
                // pending, then clears its destination bits. The first unapplied match always has
                // its source final (sources precede destinations), so every step applies one at
                // least. Other variants use the frontier F, the destination of the first unapplied
                // match: every byte before F is final. Rounds with more matches than the threads
                // hold go through one wavefront in list order.
                if constexpr (!TOK) {  // TOK: the resolve phase applies the block's matches
                constexpr int MD_K = md_k<NT>();
#if CBM_INFLATE_STATS
                if (tid == 0) acc[STAT_MSUM] += mtotal, acc[STAT_MBIG] += mtotal > MD_K * NT;
#endif
                if (mtotal <= MD_K * NT) {
                    uint64_t mm[MD_K];
                    uint32_t left = 0;
#pragma unroll
                    for (int u = 0; u < MD_K; u++) {
                        const int k = tid + u * NT;
                        mm[u] = k < mtotal ? __atomic_load_n(&mlist[k], __ATOMIC_RELAXED) : 0;
                        if (k < mtotal) left |= 1u << u;
                    }
                    if constexpr (S::IN_WORDS >= 2048 + 8) {
                        uint32_t* pend = s.in;
                        const int w0 = op >> 5, w1 = (op + total + 31) >> 5;
                        // match forwarding: a match that reads only bytes another match of the
                        // round writes, both without overlap (dist >= len), reads that match's
                        // source instead (its dist grows by the other's). Each pass looks the
                        // covering match up (binary search over the round's destinations, kept as
                        // 16-bit offsets in a free part of the staging area) in the list as the
                        // last pass left it, so chains shorten like pointer jumping; pass D then
                        // needs fewer steps.
                        const int need = (mtotal + 1) / 2;
                        const int tabw = w0 >= need ? 0 : S::IN_WORDS - w1 >= need ? w1 : -1;
                        if (tabw >= 0 && mtotal > 1) {
                            uint16_t* dst16 = (uint16_t*)(s.in + tabw);
#pragma unroll
                            for (int u = 0; u < MD_K; u++)
                                if (left & (1u << u)) dst16[tid + u * NT] = (uint16_t)((int)(mm[u] & 0x1ffff) - op);
                            __syncthreads();
                            for (int pass = 0; pass < CBM_INFLATE_FWD_PASSES; pass++) {
                                uint64_t nm[MD_K];
#pragma unroll
                                for (int u = 0; u < MD_K; u++) {
                                    nm[u] = mm[u];
                                    if (!(left & (1u << u))) continue;
                                    const int d = (int)(mm[u] & 0x1ffff), dist = (int)((mm[u] >> 17) & 0xffff);
                                    const int len = (int)((mm[u] >> 33) & 0x1ff), src = d - dist;
                                    if (dist < len || src < op) continue;
                                    // the last match (before this one) whose destination starts at or before src
                                    int lo = 0, hi = tid + u * NT;
                                    while (lo < hi) {
                                        const int mid = (lo + hi) >> 1;
                                        if ((int)dst16[mid] <= src - op) lo = mid + 1;
                                        else hi = mid;
                                    }
                                    if (lo == 0) continue;
                                    const uint64_t m2 = __atomic_load_n(&mlist[lo - 1], __ATOMIC_RELAXED);
                                    const int d2 = (int)(m2 & 0x1ffff), dist2 = (int)((m2 >> 17) & 0xffff);
                                    const int len2 = (int)((m2 >> 33) & 0x1ff);
                                    if (dist2 >= len2 && src + len <= d2 + len2)
                                        nm[u] = (uint64_t)d | ((uint64_t)(dist + dist2) << 17) | ((uint64_t)len << 33);
                                }
                                __syncthreads();  // every read of the list by this pass is done
#pragma unroll
                                for (int u = 0; u < MD_K; u++)
                                    if (nm[u] != mm[u]) {
                                        mm[u] = nm[u];
                                        mlist[tid + u * NT] = nm[u];
                                    }
                                __threadfence_block();
                                __syncthreads();
                            }
                        }
                        for (int i = w0 + tid; i < w1; i += NT) pend[i] = 0;
                        if (tid == 0) s.sv[SV_F0] = 0;
                        __syncthreads();
#pragma unroll
                        for (int u = 0; u < MD_K; u++)
                            if (left & (1u << u)) mark_bits(pend, (int)(mm[u] & 0x1ffff), (int)((mm[u] >> 33) & 0x1ff), true);
                        __syncthreads();
                        PHASE(STAT_D_MARK);
                        // every wavefront applies its lanes' matches as their sources become final,
                        // with no workgroup barrier between steps: the unapplied match with the
                        // lowest destination always has a final source (every source lies before its
                        // destination), so some wavefront always progresses. A step with many ready
                        // lanes lets each lane copy its own matches; otherwise the wavefront copies
                        // the ready matches one after another, a byte a lane. A wavefront with nothing
                        // ready sleeps briefly; one that waits past a bound reports the block bad
                        // instead of spinning on (which no valid round reaches).
                        int idle = 0;
                        while (__ballot(left != 0)) {
                            uint32_t rdy = 0;
#pragma unroll
                            for (int u = 0; u < MD_K; u++)
                                if ((left & (1u << u)) && match_ready(pend, mm[u], op)) rdy |= 1u << u;
                            const uint64_t lanes = __ballot(rdy != 0);
                            if (!lanes) {
                                COUNT(STAT_D_IDLE);
                                if (++idle > D_IDLE_MAX) {
                                    if ((tid % CBM_WAVE) == 0) {
                                        atomicOr(&s.sv[SV_ERR], CBM_BLOCK_DEFLATE);
                                        s.sv[SV_SITE] = SITE_ROUND;
                                    }
                                    break;
                                }
                                __builtin_amdgcn_s_sleep(1);
                                continue;
                            }
                            left &= ~rdy;
                            __threadfence_block();  // read the sources after their bits were seen clear
                            if (__popcll(lanes) > D_LANE_COPY) {
                                COUNT(STAT_D_LANEIT);
#pragma unroll
                                for (int u = 0; u < MD_K; u++)
                                    if (rdy & (1u << u))
                                        lds_copy_match(s.out, (int)(mm[u] & 0x1ffff), (int)((mm[u] >> 17) & 0xffff),
                                                       (int)((mm[u] >> 33) & 0x1ff));
                            } else {
#pragma unroll
                                for (int u = 0; u < MD_K; u++) {
                                    uint64_t bal = __ballot(rdy & (1u << u));
                                    while (bal) {
                                        const int from = __ffsll((unsigned long long)bal) - 1;
                                        bal &= bal - 1;
                                        COUNT(STAT_D_WAVEM);
                                        wave_copy_match(s.out, lane_u64(mm[u], from), tid % CBM_WAVE);
                                    }
                                }
                            }
                            __threadfence_block();  // the bytes land before their bits clear
#pragma unroll
                            for (int u = 0; u < MD_K; u++)
                                if (rdy & (1u << u))
                                    mark_bits(pend, (int)(mm[u] & 0x1ffff), (int)((mm[u] >> 33) & 0x1ff), false);
                            COUNT(STAT_DSTEPS);
                        }
                        __syncthreads();
                        PHASE(STAT_D_LOOP);
                    } else {
                        if (tid == 0) {
                            s.sv[SV_F0] = mtotal ? (int)(mm[0] & 0x1ffff) : 0x7fffffff;
                            s.sv[SV_F0 + 1] = 0x7fffffff;
                            s.sv[SV_F0 + 2] = 0x7fffffff;
                        }
                        __syncthreads();
                        for (int step = 0;; step++) {
                            const int F = s.sv[SV_F0 + step % 3];
                            if (F == 0x7fffffff) break;
#pragma unroll
                            for (int u = 0; u < MD_K; u++) {
                                if (!(left & (1u << u))) continue;
                                const uint64_t m = mm[u];
                                const int dest = (int)(m & 0x1ffff), dist = (int)((m >> 17) & 0xffff);
                                const int len = (int)((m >> 33) & 0x1ff), src = dest - dist;
                                if (src + (dist < len ? dist : len) > F) continue;
                                left &= ~(1u << u);
                                lds_copy_match(s.out, dest, dist, len);
                            }
                            // the next frontier: a minimum within the wavefront, then one atomic per wavefront
                            int mine = left ? (int)(mm[__ffs(left) - 1] & 0x1ffff) : 0x7fffffff;
#pragma unroll
                            for (int d = 32; d; d >>= 1) {
                                int o = __shfl_xor(mine, d, CBM_WAVE);
                                mine = o < mine ? o : mine;
                            }
                            if ((tid % CBM_WAVE) == 0 && mine != 0x7fffffff)
                                atomicMin(&s.sv[SV_F0 + (step + 1) % 3], mine);
                            if (tid == 0) s.sv[SV_F0 + (step + 2) % 3] = 0x7fffffff;
                            __syncthreads();
                            COUNT(STAT_DSTEPS);
                        }
                    }
                } else if (tid < CBM_WAVE) {
                    const int lane = tid;
                    for (int k0 = 0; k0 < mtotal; k0 += CBM_WAVE) {
                        const uint64_t mine =
                            k0 + lane < mtotal ? __atomic_load_n(&mlist[k0 + lane], __ATOMIC_RELAXED) : 0;
                        const int nb = mtotal - k0 < CBM_WAVE ? mtotal - k0 : CBM_WAVE;
                        for (int j = 0; j < nb; j++) {
                            const uint64_t m = __shfl(mine, j, CBM_WAVE);
                            const int dest = (int)(m & 0x1ffff), dist = (int)((m >> 17) & 0xffff);
                            const int len = (int)((m >> 33) & 0x1ff), src = dest - dist;
                            if (dist >= len) {
                                for (int i = lane; i < len; i += CBM_WAVE) s.out[dest + i] = s.out[src + i];
                            } else {
                                for (int i = lane; i < len; i += CBM_WAVE) s.out[dest + i] = s.out[src + i % dist];
                            }
                        }
                    }
                }
                }  // !TOK
                __syncthreads();
                PHASE(STAT_PASS_D);
                COUNT(STAT_ROUNDS);
                if (s.sv[SV_ERR] != CBM_BLOCK_OK) break;
                if (tid == 0) {
                    s.sv[SV_MCNT] += mtotal;
                    s.sv[SV_OP] = op + total;
                    s.sv[SV_POS] = tend < NT ? s.exit_[tend] : s.exit_[NT - 1];
                }
                __syncthreads();
                if (tend < NT) break;  // end of the Huffman block
            }
        }
        // ---- block end: size, store, CRC32
        __syncthreads();
        PHASE(STAT_BLOCKEND);
        int err = s.sv[SV_ERR], site = s.sv[SV_SITE];
        const int op = s.sv[SV_OP];
        if (err == CBM_BLOCK_OK && s.sv[SV_POS] > in_bits) err = CBM_BLOCK_DEFLATE, site = SITE_OVERRUN;
        if (err == CBM_BLOCK_OK && op != cap) err = CBM_BLOCK_SIZE, site = SITE_SIZE;
        if constexpr (TOK) {
            // matches and CRC32 come later (resolve and CRC kernels); a block with any failed check
            // is decoded again by the one-phase kernel, which reports it exactly
            if (tid == 0) {
                status[b] = err == CBM_BLOCK_OK ? CBM_BLOCK_OK : BLOCK_DEFER;
                mcount[b] = s.sv[SV_MCNT];
#if CBM_INFLATE_STATS
                if (stats)
                    for (int k = 0; k < STAT_N; k++) atomicAdd(&stats[k], (unsigned long long)acc[k]), acc[k] = 0;
#endif
            }
            __syncthreads();  // also: the block's literals and match list, written above, are visible
            if constexpr (FUSE) {
                if (err == CBM_BLOCK_OK) {
                    uint8_t* const win = (uint8_t*)s.in;  // 16-byte aligned (the tables before it are)
                    resolve_block<1, 2048>(bdst, cap, expect[b], crc_out + b, status + b, bad_count, blist,
                                           s.sv[SV_MCNT], win, (uint32_t*)(win + 2048 + 48), kCrcTab.t, xfac, tid);
                }
                __syncthreads();  // the staging area is free again for the next block
            }
            continue;
        }
        const int nout = err == CBM_BLOCK_OK ? cap : (op < cap ? op : cap);
        uint8_t* d = dst + out_off[b];
        {
            // the output in 16-byte stores where d + i is 16-byte aligned (a head and a tail of
            // single bytes), each read from LDS as five words realigned
            const int head = (int)((16 - ((uintptr_t)d & 15)) & 15) < nout ? (int)((16 - ((uintptr_t)d & 15)) & 15) : nout;
            if (tid < head) d[tid] = s.out[tid];
            const int nvec = (nout - head) / 16;
            const uint32_t* ow = (const uint32_t*)s.out;
            for (int v = tid; v < nvec; v += NT) {
                const int i = head + 16 * v, w0 = i >> 2;
                const uint32_t sh = (uint32_t)(i & 3);
                const uint32_t x0 = ow[w0], x1 = ow[w0 + 1], x2 = ow[w0 + 2], x3 = ow[w0 + 3], x4 = ow[w0 + 4];
                uint4 q;
                q.x = __builtin_amdgcn_alignbyte(x1, x0, sh);
                q.y = __builtin_amdgcn_alignbyte(x2, x1, sh);
                q.z = __builtin_amdgcn_alignbyte(x3, x2, sh);
                q.w = __builtin_amdgcn_alignbyte(x4, x3, sh);
                *(uint4*)(d + i) = q;
            }
            for (int i = head + 16 * nvec + tid; i < nout; i += NT) d[i] = s.out[i];
        }
        // CRC32 of the output: slice tid covers [nout - SLICE*(NT - tid), nout - SLICE*(NT-1-tid)),
        // slicing-by-4 with the tables in global memory (kCrcTab, L1-resident): four bytes a step,
        // read from LDS as aligned words realigned by the slice's offset
        const uint32_t(*crcG)[256] = kCrcTab.t;
        uint32_t c = 0;
        {
            int e2 = nout - SLICE * (NT - 1 - tid), s2 = e2 - SLICE;
            if (s2 < 0) s2 = 0;
            if (e2 > s2) {
                uint32_t x = 0xFFFFFFFFu;
                int k = s2;
                const uint32_t* ow = (const uint32_t*)s.out;
                const uint32_t sh = (uint32_t)s2 & 3;
                uint32_t lo = ow[s2 >> 2];
                for (; k + 4 <= e2; k += 4) {
                    const uint32_t hi = ow[(k >> 2) + 1];
                    x ^= __builtin_amdgcn_alignbyte(hi, lo, sh);
                    lo = hi;
                    x = crcG[3][x & 255] ^ crcG[2][(x >> 8) & 255] ^ crcG[1][(x >> 16) & 255] ^ crcG[0][x >> 24];
                }
                for (; k < e2; k++) x = crcG[0][(x ^ s.out[k]) & 255] ^ (x >> 8);
                c = crc_mulmod(xfac, x ^ 0xFFFFFFFFu);
            }
        }
        for (int dd = 32; dd; dd >>= 1) c ^= __shfl_xor(c, dd, CBM_WAVE);
        if ((tid % CBM_WAVE) == 0) s.red[tid / CBM_WAVE] = (int)c;
        __syncthreads();
        if (tid == 0) {
            uint32_t x = 0;
            for (int k = 0; k < NT / CBM_WAVE; k++) x ^= (uint32_t)s.red[k];
            crc_out[b] = x;
            if (err == CBM_BLOCK_OK && x != expect[b]) err = CBM_BLOCK_CRC;
            status[b] = err | (err != CBM_BLOCK_OK ? site << 8 : 0);
            PHASE(STAT_STORE);
#if CBM_INFLATE_STATS
            if (stats)
                for (int k = 0; k < STAT_N; k++) atomicAdd(&stats[k], (unsigned long long)acc[k]), acc[k] = 0;
#endif
            if (err != CBM_BLOCK_OK) atomicAdd((unsigned long long*)bad_count, 1ull);
        }
        __syncthreads();
    }
#undef PHASE
#undef COUNT
}

// ------------------------------------------------------------------ resolve phase (two-phase inflate)
// A wavefront per block that the decode phase finished (status OK), then the block's CRC32;
// WPG blocks a workgroup. The block's matches (in output order) go in windows of up to 64 * MPL,
// MPL a lane, whose output span [lo, hi) fits SPAN bytes: the span is staged in LDS from dst (its
// literals are in place; LDS byte shift + p - lo holds byte p, so 16-byte global accesses are
// 16-byte LDS accesses). Far matches (their source wholly before the span, so final in dst, and
// not overlapping) are copied at once, each by its lane. The others are marked pending in a
// bitmap, and each step applies every one whose source holds no pending byte: the lowest pending
// match always qualifies. Then the span goes back to dst. A match reads only final bytes,
// [src, src + len) or the dist bytes before dest that an overlapping copy repeats.
constexpr int RES_MAX_DIST = 32768;

template <int SPAN, int WPG>
struct ResSmem {
    uint32_t crc_t[4][256];
    uint32_t crc_x[64];
    alignas(16) uint8_t span[WPG][SPAN + 48];
    uint32_t pend[WPG][(SPAN + 16) / 32 + 2];
};

// LDS order among the lanes of one wavefront: the wavefront's LDS operations run in program
// order, so a compiler barrier is all it takes (no wait on global memory)
__device__ __forceinline__ void res_wave_fence() { __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront"); }

__device__ __forceinline__ void res_fence() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// zlib CRC32 (pre- and post-inverted) of p[0:n) in global memory: bytes up to a 16-byte boundary,
// then 16 bytes a load (loads of a 64-byte stretch issued together), slicing-by-4 from LDS tables
__device__ __forceinline__ uint32_t crc_bytes16(const uint8_t* p, int n, const uint32_t (*T)[256]) {
    uint32_t c = 0xFFFFFFFFu;
    int k = 0;
    while (k < n && (((uintptr_t)(p + k)) & 15)) c = T[0][(c ^ p[k++]) & 255] ^ (c >> 8);
    for (; k + 64 <= n; k += 64) {
        uint4 q[4];
#pragma unroll
        for (int u = 0; u < 4; u++) q[u] = *(const uint4*)(p + k + 16 * u);
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const uint32_t w[4] = {q[u].x, q[u].y, q[u].z, q[u].w};
#pragma unroll
            for (int v = 0; v < 4; v++) {
                c ^= w[v];
                c = T[3][c & 255] ^ T[2][(c >> 8) & 255] ^ T[1][(c >> 16) & 255] ^ T[0][c >> 24];
            }
        }
    }
    for (; k + 4 <= n; k += 4) {
        c ^= *(const uint32_t*)(p + k);
        c = T[3][c & 255] ^ T[2][(c >> 8) & 255] ^ T[1][(c >> 16) & 255] ^ T[0][c >> 24];
    }
    for (; k < n; k++) c = T[0][(c ^ p[k]) & 255] ^ (c >> 8);
    return c ^ 0xFFFFFFFFu;
}

// len bytes from g (dst) to d (LDS), 32 a pass: the aligned words that hold the pass's bytes are
// loaded together (at most 3 bytes outside the range, inside the same buffer), then realigned
// with v_alignbyte and written a byte at a time
__device__ __forceinline__ void res_copy_far(const uint8_t* g, uint8_t* d, int len) {
    for (int i0 = 0; i0 < len; i0 += 32) {
        const int sh = (int)((uintptr_t)(g + i0) & 3);
        const uint32_t* w = (const uint32_t*)(g + i0 - sh);
        const int nb = len - i0 < 32 ? len - i0 : 32;
        const int nw = (nb + sh + 3) >> 2;
        uint32_t x[9];
#pragma unroll
        for (int j = 0; j < 9; j++) x[j] = j < nw ? w[j] : 0u;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const uint32_t v = __builtin_amdgcn_alignbyte(x[j + 1], x[j], (uint32_t)sh);
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (4 * j + k < nb) d[i0 + 4 * j + k] = (uint8_t)(v >> (8 * k));
        }
    }
}

// One block's resolve: its M matches (ml, in output order) applied to its L bytes at out, in
// windows of up to 64 * MPL matches whose output fits SPAN bytes, then its CRC32 checked
// against expect (into *crc_out_b, *status_b and the bad count). sp: SPAN + 48 bytes of LDS,
// 16-byte aligned; pend: (SPAN + 16) / 32 + 2 words of LDS; T the slicing-by-4 tables; xk this
// lane's x^(8 * 1024 * (63 - lane)) mod P. A window that cannot progress leaves the block to the
// one-phase kernel (status BLOCK_DEFER).
template <int MPL, int SPAN>
__device__ __forceinline__ void resolve_block(uint8_t* out, int L, uint32_t expect, uint32_t* crc_out_b, int32_t* status_b,
                              int64_t* bad, const uint64_t* ml, int M, uint8_t* sp, uint32_t* pend,
                              const uint32_t (*T)[256], uint32_t xk, int lane) {
    bool stuck = false;
    for (int k0 = 0; k0 < M && !stuck;) {
        int dest[MPL], dist[MPL], len[MPL];
#pragma unroll
        for (int u = 0; u < MPL; u++) {
            const int k = k0 + u * CBM_WAVE + lane;
            const uint64_t m = k < M ? ml[k] : 0;
            dest[u] = (int)(m & 0x1ffff), dist[u] = (int)((m >> 17) & 0xffff), len[u] = (int)((m >> 33) & 0x1ff);
        }
        const int lo = __builtin_amdgcn_readlane(dest[0], 0);
        // dest + len grows with the match index (outputs are disjoint and in order): the
        // matches that fit are a prefix, c of them (match k0's at most 258 bytes always fits)
        int c = 0;
#pragma unroll
        for (int u = 0; u < MPL; u++)
            c += __popcll(__ballot(k0 + u * CBM_WAVE + lane < M && dest[u] + len[u] - lo <= SPAN));
        int hi = 0;
#pragma unroll
        for (int u = 0; u < MPL; u++)
            if ((c - 1) / CBM_WAVE == u) hi = __builtin_amdgcn_readlane(dest[u] + len[u], (c - 1) % CBM_WAVE);
        const int nspan = hi - lo;
        const int shift = (int)((uintptr_t)(out + lo) & 15);
        // stage the span: the 16-byte aligned pieces that cover it
        const uint8_t* g0 = out + lo - shift;
        const int npiece = (shift + nspan + 15) >> 4;
        for (int q = lane; q < npiece; q += CBM_WAVE) *(uint4*)(sp + 16 * q) = *(const uint4*)(g0 + 16 * q);
        for (int w = lane; w <= ((shift + nspan) >> 5); w += CBM_WAVE) pend[w] = 0;
        bool near[MPL];
#pragma unroll
        for (int u = 0; u < MPL; u++) {
            const bool mine = u * CBM_WAVE + lane < c;
            const int src = dest[u] - dist[u];
            // far: needs nothing of the window, so its lane copies it at once
            const bool far = mine && dist[u] >= len[u] && src + len[u] <= lo;
            if (far) res_copy_far(out + src, sp + shift + dest[u] - lo, len[u]);
            near[u] = mine && !far;
        }
        res_wave_fence();
#pragma unroll
        for (int u = 0; u < MPL; u++)
            if (near[u]) mark_bits(pend, shift + dest[u] - lo, len[u], true);
        res_wave_fence();
        uint32_t left = 0;
#pragma unroll
        for (int u = 0; u < MPL; u++) left |= near[u] ? 1u << u : 0u;
        int idle = 0;
        while (__ballot(left != 0)) {
            uint32_t rdy = 0;
#pragma unroll
            for (int u = 0; u < MPL; u++) {
                if (!(left & (1u << u))) continue;
                const int src = dest[u] - dist[u], e = src + (dist[u] < len[u] ? dist[u] : len[u]);
                const int a0 = src < lo ? lo : src;
                if (!(a0 < e && any_bits(pend, shift + a0 - lo, shift + e - lo))) rdy |= 1u << u;
            }
            if (!__ballot(rdy != 0) && ++idle > (1 << 12)) {  // cannot happen (the lowest is ready):
                stuck = true;                                  // leave the block to the one-phase kernel
                break;
            }
            // the ready matches, one at a time, each copied by the whole wavefront (a byte a
            // lane): their destinations are disjoint and pending, their sources final, so they
            // go in any order. Byte i of a match reads src + i, or src + i mod dist when it
            // overlaps its own output (a repeat of the dist final bytes before dest).
#pragma unroll
            for (int u = 0; u < MPL; u++) {
                for (uint64_t rm = __ballot((rdy >> u) & 1u); rm; rm &= rm - 1) {
                    const int l = __builtin_ctzll(rm);
                    const int md = __builtin_amdgcn_readlane(dest[u], l);
                    const int dl = __builtin_amdgcn_readlane(dist[u], l);
                    const int ln = __builtin_amdgcn_readlane(len[u], l);
                    const int src = md - dl;
                    uint8_t* d = sp + shift + md - lo;
                    if (src >= lo) {  // the source is in the span: LDS only
                        const uint8_t* sl = sp + shift + src - lo;
                        if (dl >= ln) {
                            for (int i = lane; i < ln; i += CBM_WAVE) d[i] = sl[i];
                        } else {
                            for (int i = lane; i < ln; i += CBM_WAVE) d[i] = sl[i % dl];
                        }
                    } else {  // partly before the span (in dst), partly in it
                        for (int i = lane; i < ln; i += CBM_WAVE) {
                            const int q = src + (dl >= ln ? i : i % dl);
                            d[i] = q < lo ? out[q] : sp[shift + q - lo];
                        }
                    }
                }
            }
            res_wave_fence();  // the bytes land before their bits clear
#pragma unroll
            for (int u = 0; u < MPL; u++)
                if (rdy & (1u << u)) mark_bits(pend, shift + dest[u] - lo, len[u], false);
            res_wave_fence();
            left &= ~rdy;
        }
        // the span back to dst: single bytes up to a 16-byte boundary, 16-byte pieces, single bytes
        {
            const int head = ((16 - shift) & 15) < nspan ? ((16 - shift) & 15) : nspan;
            if (lane < head) out[lo + lane] = sp[shift + lane];
            const int body = (nspan - head) >> 4;
            for (int q = lane; q < body; q += CBM_WAVE)
                *(uint4*)(out + lo + head + 16 * q) = *(const uint4*)(sp + shift + head + 16 * q);
            for (int i = head + 16 * body + lane; i < nspan; i += CBM_WAVE) out[lo + i] = sp[shift + i];
        }
        res_fence();  // later windows read these bytes from dst
        k0 += c;
    }
    if (stuck) {
        if (lane == 0) *status_b = BLOCK_DEFER;
        return;
    }
    // CRC32: lane j takes the j-th of 64 stretches of 1,024 bytes counted back from the end
    const int e = L - (CBM_WAVE - 1 - lane) * 1024, st = e - 1024 > 0 ? e - 1024 : 0;
    uint32_t cr = e > st ? crc_mulmod(xk, crc_bytes16(out + st, e - st, T)) : 0;
    for (int d = 32; d; d >>= 1) cr ^= __shfl_xor(cr, d, CBM_WAVE);
    if (lane == 0) {
        *crc_out_b = cr;
        if (cr != expect) {
            *status_b = CBM_BLOCK_CRC;
            atomicAdd((unsigned long long*)bad, 1ull);
        }
    }
}

template <int MPL, int SPAN, int WPG>
__global__ void __launch_bounds__(WPG * CBM_WAVE) inflate_resolve(uint8_t* dst, const int64_t* out_off,
                                                                  const int32_t* isize, const uint32_t* expect,
                                                                  uint32_t* crc_out, int32_t* status, int64_t* bad,
                                                                  const uint64_t* pool, const int32_t* mcount,
                                                                  int64_t n_blocks, const int64_t* idx,
                                                                  const int64_t* n_dev, int64_t mcap, int64_t kmax) {
    static_assert(SPAN >= 258 && SPAN % 32 == 0, "a window holds one match at least");
    __shared__ ResSmem<SPAN, WPG> s;
    // blocks idx[0 .. min(*n_dev, kmax)) with idx, else 0 .. min(n_blocks, kmax); list kb of the pool
    const int64_t n = n_dev ? *n_dev : n_blocks, lim = n < kmax ? n : kmax;
    if ((int64_t)blockIdx.x * WPG >= lim) return;  // (uniform) no block: skip the tables
    crc_tables(s.crc_t, s.crc_x);
    const int wv = (int)threadIdx.x / CBM_WAVE, lane = (int)threadIdx.x % CBM_WAVE;
    uint8_t* sp = s.span[wv];
    uint32_t* pend = s.pend[wv];
    for (int64_t kb = (int64_t)blockIdx.x * WPG + wv; kb < lim; kb += (int64_t)gridDim.x * WPG) {
        const int64_t b = idx ? idx[kb] : kb;
        if (status[b] != CBM_BLOCK_OK) continue;  // left to the one-phase kernel
        resolve_block<MPL, SPAN>(dst + out_off[b], isize[b], expect[b], crc_out + b, status + b, bad,
                                 pool + (size_t)kb * mcap, mcount[b], sp, pend, s.crc_t,
                                 s.crc_x[CBM_WAVE - 1 - lane], lane);
    }
}

// ------------------------------------------------------------------ resolve phase with a history ring
// The block's output passes once through a ring of RING bytes of LDS (a wavefront's), in stretches
// of RES_FL bytes aligned to 16 bytes of dst (u = p + a0, a0 = the block's misalignment): each
// stretch is staged from dst once (its literals are in place), receives the matches whose output
// lies in it, and once final goes back to dst once (if a match wrote into it) while its CRC32 is
// taken from LDS. Windows of up to 64 matches whose output fits SPAN bytes are applied as in
// inflate_resolve; a match source is read from the ring while the ring still holds it (at least
// RING - SPAN - RES_FL bytes back from the window) and from dst before that (final bytes written
// back earlier). Staging stretch s reuses the slot of stretch s - RING, which is final by then:
// every byte before the window is, and RING >= SPAN + RES_FL keeps s - RING + RES_FL <= the
// window's start.
// CRC32: lane j takes bytes [32j, 32j + 32) of each stretch and keeps A_j = A_j * x^(8 RES_FL) xor
// crc(piece) (the multiplication by table, kCrcAdv); the last, partial stretch is cut into 32-byte
// pieces counted back from the block's end, and the block's CRC is xor_j x^(8 * 32 (63 - j)) *
// (A_j * x^(8 r) xor c_j), r being the bytes of the partial stretch (zlib's crc32_combine rule).
constexpr int RES_FL = 2048;
// 16 bytes as a plain vector (loads and stores in an explicit address space)
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr uint32_t crc_mulmod_c(uint32_t a, uint32_t b) {
    uint32_t p = 0;
    for (int k = 0; k < 32; k++) {
        if (a & 0x80000000u) p ^= b;
        a <<= 1;
        b = (b & 1) ? (b >> 1) ^ 0xEDB88320u : b >> 1;
    }
    return p;
}
// kCrcAdv.t[k][v] = (v << 8k) * x^(8 * RES_FL) mod P: a CRC value times x^(8 RES_FL) is the xor of
// four lookups, one per byte
struct CrcAdv {
    uint32_t t[4][256];
};
constexpr CrcAdv make_crc_adv() {
    uint32_t x = 0x40000000u;                                // x^1
    for (int j = 0; j < 14; j++) x = crc_mulmod_c(x, x);     // x^(2^14) = x^(8 * 2048)
    CrcAdv c{};
    for (int k = 0; k < 4; k++)
        for (int v = 0; v < 256; v++) c.t[k][v] = crc_mulmod_c((uint32_t)v << (8 * k), x);
    return c;
}
static_assert(RES_FL == 2048, "kCrcAdv is built for 2,048-byte stretches");
__constant__ CrcAdv kCrcAdv = make_crc_adv();

// zlib CRC32 of r[s, e) (LDS, e - s <= 32 or any length, s <= e), 4 bytes a step where aligned
__device__ __forceinline__ uint32_t crc_lds(const CBM_LDS uint8_t* r, int s, int e, const uint32_t (*T)[256]) {
    if (e <= s) return 0;
    uint32_t x = 0xFFFFFFFFu;
    int k = s;
    for (; k < e && (k & 3); k++) x = T[0][(x ^ r[k]) & 255] ^ (x >> 8);
    for (; k + 4 <= e; k += 4) {
        x ^= *(const CBM_LDS uint32_t*)(r + k);
        x = T[3][x & 255] ^ T[2][(x >> 8) & 255] ^ T[1][(x >> 16) & 255] ^ T[0][x >> 24];
    }
    for (; k < e; k++) x = T[0][(x ^ r[k]) & 255] ^ (x >> 8);
    return x ^ 0xFFFFFFFFu;
}

template <int SPAN, int WPG, int RING>
struct RingSmem {
    alignas(16) uint8_t ring[WPG][RING];
    uint32_t pend[WPG][SPAN / 32 + 2];
};

template <int SPAN, int WPG, int RING>
__global__ void __launch_bounds__(WPG * CBM_WAVE) inflate_resolve_ring(uint8_t* dst, const int64_t* out_off,
                                                                       const int32_t* isize, const uint32_t* expect,
                                                                       uint32_t* crc_out, int32_t* status,
                                                                       int64_t* bad, const uint64_t* pool,
                                                                       const int32_t* mcount, int64_t n_blocks) {
    static_assert(SPAN >= 258 && SPAN % 32 == 0, "a window holds one match at least");
    static_assert((RING & (RING - 1)) == 0 && RING % RES_FL == 0 && RING >= SPAN + RES_FL,
                  "the ring holds the window and the stretch staged past it");
    static_assert(RING / RES_FL <= 32, "dirty stretches are bits of a word");
    constexpr int MASK = RING - 1, NSLOT = RING / RES_FL;
    __shared__ RingSmem<SPAN, WPG, RING> s;
    const int wv = (int)threadIdx.x / CBM_WAVE, lane = (int)threadIdx.x % CBM_WAVE;
    CBM_LDS uint8_t* const ring = (CBM_LDS uint8_t*)s.ring[wv];
    uint32_t* pend = s.pend[wv];
    const uint32_t(*T)[256] = kCrcTab.t;
    const uint32_t(*TX)[256] = kCrcAdv.t;
    // this lane's factor x^(8 * 32 * (63 - lane)), and x^(8 * 2^lane) for the partial stretch
    uint32_t P, E;
    {
        uint32_t sq = 0x40000000u;
        for (int j = 0; j < 8; j++) sq = crc_mulmod(sq, sq);  // x^256
        uint32_t r = 0x80000000u;
        for (int k = CBM_WAVE - 1 - lane; k; k >>= 1) {
            if (k & 1) r = crc_mulmod(r, sq);
            sq = crc_mulmod(sq, sq);
        }
        P = r;
        uint32_t e = 0x40000000u;
        for (int j = 0; j < 3; j++) e = crc_mulmod(e, e);  // x^8
        for (int j = 0; j < lane && j < 12; j++) e = crc_mulmod(e, e);
        E = e;
    }
    for (int64_t b = (int64_t)blockIdx.x * WPG + wv; b < n_blocks; b += (int64_t)gridDim.x * WPG) {
        if (status[b] != CBM_BLOCK_OK) continue;  // left to the one-phase kernel
        const int M = mcount[b], L = isize[b];
        uint8_t* out = dst + out_off[b];
        const int a0 = (int)((uintptr_t)out & 15);
        CBM_GLOBAL uint8_t* const oal = (CBM_GLOBAL uint8_t*)(out - a0);  // 16-byte aligned: byte p is oal[a0 + p]
        const int uend = a0 + L;
        const uint64_t* ml = pool + (size_t)b * MCAP_TOK;
        int st = 0, fin = 0;  // staged and written-back ends (u, multiples of RES_FL)
        uint32_t dirty = 0;   // ring slots holding match output not yet written back
        bool wb = false;      // written back since the last fence
        uint32_t acc = 0;
        // stretch [fin, fin + RES_FL) (final, wholly inside the block's end): CRC piece, write-back
        auto flush = [&]() {
            const CBM_LDS uint8_t* r = ring + (fin & MASK);
            const int p0 = 32 * lane, ps = fin + p0 < a0 ? a0 - fin : p0;
            const uint32_t c = crc_lds(r, ps, p0 + 32, T);
            acc = TX[0][acc & 255] ^ TX[1][(acc >> 8) & 255] ^ TX[2][(acc >> 16) & 255] ^ TX[3][acc >> 24] ^ c;
            const int slot = (fin / RES_FL) & (NSLOT - 1);
            if ((dirty >> slot) & 1u) {
#pragma unroll
                for (int h = 0; h < RES_FL / 16 / CBM_WAVE; h++) {
                    const int q = lane + h * CBM_WAVE, u = fin + 16 * q;
                    if (u >= a0 && u + 16 <= uend) {
                        *(CBM_GLOBAL u32x4*)(oal + u) = *(const CBM_LDS u32x4*)(r + 16 * q);
                    } else {
                        for (int i = 0; i < 16; i++)
                            if (u + i >= a0 && u + i < uend) oal[u + i] = r[16 * q + i];
                    }
                }
                dirty &= ~(1u << slot);
                wb = true;
            }
            fin += RES_FL;
        };
        // stage whole stretches until [.., target) is in the ring. The stretch after the staged
        // ones is loaded into registers ahead (nx, for nx_at), so that its latency passes while
        // the windows before it are applied; its bytes are literals, final since the decode phase.
        constexpr int PPL = RES_FL / 16 / CBM_WAVE;  // 16-byte pieces a lane
        u32x4 nx[PPL];
        int nx_at = -1;
        auto prefetch = [&]() {
            nx_at = st;
#pragma unroll
            for (int h = 0; h < PPL; h++) {
                const int u = st + 16 * (lane + h * CBM_WAVE);
                nx[h] = u < uend ? *(const CBM_GLOBAL u32x4*)(oal + u) : u32x4{0u, 0u, 0u, 0u};
            }
        };
        auto stage_to = [&](int target) {
            if (st >= target) return;
            while (st < target) {
                while (fin <= st - RING) flush();  // the slot's previous stretch goes back first
                CBM_LDS uint8_t* r = ring + (st & MASK);
                if (nx_at != st) prefetch();
#pragma unroll
                for (int h = 0; h < PPL; h++) *(CBM_LDS u32x4*)(r + 16 * (lane + h * CBM_WAVE)) = nx[h];
                st += RES_FL;
            }
            if (st < uend) prefetch();
            res_wave_fence();
        };
        bool stuck = false;
        uint64_t m = lane < M ? ml[lane] : 0;  // this window's matches (the next window's load ahead)
        for (int k0 = 0; k0 < M && !stuck;) {
            const int k = k0 + lane;
            const int dest = (int)(m & 0x1ffff), dist = (int)((m >> 17) & 0xffff), len = (int)((m >> 33) & 0x1ff);
            const int lo = __builtin_amdgcn_readlane(dest, 0);
            // dest + len grows with the match index: the matches that fit are a prefix
            const int c = __popcll(__ballot(k < M && dest + len - lo <= SPAN));
            const int hi = __builtin_amdgcn_readlane(dest + len, c - 1);
            const int ulo = a0 + lo, uhi = a0 + hi;
            const uint64_t mn = k0 + c + lane < M ? ml[k0 + c + lane] : 0;
            stage_to(uhi);
            const int rbase = st - RING;  // the ring holds [rbase, st); below that, dst (final)
            for (int w = lane; w <= ((uhi - ulo) >> 5); w += CBM_WAVE) pend[w] = 0;
            dirty |= ((2u << ((uhi - 1) / RES_FL - ulo / RES_FL)) - 1u) << ((ulo / RES_FL) & (NSLOT - 1));
            dirty |= dirty >> NSLOT;  // slots wrap
            dirty &= (NSLOT == 32 ? 0xffffffffu : (1u << NSLOT) - 1u);
            const bool mine = lane < c;
            const int usrc = a0 + dest - dist, udst = a0 + dest;
            // a source wholly before the window (and not overlapping its output) is final: copy it now
            const bool now = mine && dist >= len && dest - dist + len <= lo;
            if (wb && __ballot(now && usrc < rbase)) {
                res_fence();  // write-backs land before they are read
                wb = false;
            }
            if (now) {
                for (int i0 = 0; i0 < len; i0 += 32) {
                    const int us = usrc + i0, sh = us & 3, base = us - sh;
                    const int nb = len - i0 < 32 ? len - i0 : 32;
                    const int nw = (nb + sh + 3) >> 2;
                    uint32_t x[9];
#pragma unroll
                    for (int j = 0; j < 9; j++) {
                        const int wpos = base + 4 * j;
                        uint32_t v = 0u;
                        if (j < nw) {
                            if (wpos < rbase) v = *(const CBM_GLOBAL uint32_t*)(oal + wpos);
                            else v = *(const CBM_LDS uint32_t*)(ring + (wpos & MASK));
                        }
                        x[j] = v;
                    }
#pragma unroll
                    for (int j = 0; j < 8; j++) {
                        const uint32_t v = __builtin_amdgcn_alignbyte(x[j + 1], x[j], (uint32_t)sh);
#pragma unroll
                        for (int t = 0; t < 4; t++)
                            if (4 * j + t < nb) ring[(udst + i0 + 4 * j + t) & MASK] = (uint8_t)(v >> (8 * t));
                    }
                }
            }
            const bool near = mine && !now;
            res_wave_fence();
            if (near) mark_bits(pend, udst - ulo, len, true);
            res_wave_fence();
            bool left = near;
            int idle = 0;
            while (__ballot(left)) {
                bool rdy = false;
                if (left) {
                    const int e = usrc + (dist < len ? dist : len), a = usrc < ulo ? ulo : usrc;
                    rdy = !(a < e && any_bits(pend, a - ulo, e - ulo));
                }
                const uint64_t rm0 = __ballot(rdy);
                if (!rm0 && ++idle > (1 << 12)) {  // cannot happen (the lowest is ready):
                    stuck = true;                   // leave the block to the one-phase kernel
                    break;
                }
                if (wb && rm0) {
                    res_fence();
                    wb = false;
                }
                // each ready match copied by the whole wavefront, a byte a lane (sources final,
                // destinations disjoint and pending); an overlapping match repeats its dist bytes
                for (uint64_t rm = rm0; rm; rm &= rm - 1) {
                    const int l = __builtin_ctzll(rm);
                    const int md = __builtin_amdgcn_readlane(udst, l);
                    const int dl = __builtin_amdgcn_readlane(dist, l);
                    const int ln = __builtin_amdgcn_readlane(len, l);
                    const int sr = md - dl;
                    for (int i = lane; i < ln; i += CBM_WAVE) {
                        const int q = sr + (dl >= ln ? i : (int)((uint32_t)i % (uint32_t)dl));
                        uint8_t v;
                        if (q < rbase) v = oal[q];
                        else v = ring[q & MASK];
                        ring[(md + i) & MASK] = v;
                    }
                }
                res_wave_fence();  // the bytes land before their bits clear
                if (rdy) mark_bits(pend, udst - ulo, len, false);
                res_wave_fence();
                left = left && !rdy;
            }
            k0 += c;
            m = mn;
        }
        if (stuck) {
            if (lane == 0) status[b] = BLOCK_DEFER;
            continue;
        }
        // the rest of the block: stage, then every whole stretch, then the partial one
        stage_to(uend);
        while (fin + RES_FL <= uend) flush();
        const int f0 = fin > a0 ? fin : a0;
        const int r = uend - f0 > 0 ? uend - f0 : 0;
        uint32_t cj = 0;
        if (r) {
            const CBM_LDS uint8_t* rr = ring + (fin & MASK);
            const int e = uend - 32 * (CBM_WAVE - 1 - lane), sb = e - 32 > f0 ? e - 32 : f0;
            cj = crc_lds(rr, sb - fin, e - fin, T);
            const int slot = (fin / RES_FL) & (NSLOT - 1);
            if ((dirty >> slot) & 1u)
                for (int i = lane; i < uend - fin; i += CBM_WAVE)
                    if (fin + i >= a0) oal[fin + i] = rr[i];
        }
        uint32_t xr = 0x80000000u;  // x^(8 r)
        for (int j = 0; j < 12; j++)
            if ((r >> j) & 1) xr = crc_mulmod(xr, (uint32_t)__builtin_amdgcn_readlane((int)E, j));
        uint32_t cr = crc_mulmod(P, crc_mulmod(acc, xr) ^ cj);
        for (int d = 32; d; d >>= 1) cr ^= __shfl_xor(cr, d, CBM_WAVE);
        if (lane == 0) {
            crc_out[b] = cr;
            if (cr != expect[b]) {
                status[b] = CBM_BLOCK_CRC;
                atomicAdd((unsigned long long*)bad, 1ull);
            }
        }
    }
}

// the blocks left to the one-phase kernel, listed (in any order) with their count
__global__ void inflate_defer_list(const int32_t* status, int64_t n_blocks, int64_t* idx, int64_t* count) {
    for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b < n_blocks; b += (int64_t)gridDim.x * blockDim.x)
        if (status[b] == BLOCK_DEFER) idx[atomicAdd((unsigned long long*)count, 1ull)] = b;
}

}  // namespace cbm
