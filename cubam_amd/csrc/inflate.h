// Kernel 2: per-block DEFLATE inflate (RFC 1951: stored, fixed and dynamic Huffman
// blocks) with per-lane decode tables in LDS, then a wavefront-cooperative CRC32
// check of every output block.
//
// The reference hands this to nvCOMP (src/cubam/bgzf.py:587-611) and never checks
// the CRC (docs/reading.md:48-51). Here every decoding lane of a wavefront owns
// one BGZF block at a time: its Huffman tables live in its own slice of LDS, so a
// wavefront inflates `blocks_per_wave` blocks side by side and the workgroup
// holds `waves_per_group` such wavefronts; both are tunables (the LDS budget per
// CU bounds their product). The CRC pass then gives one wavefront to each block.
#pragma once
#include "common.h"

namespace cbm {

// ------------------------------------------------------------------ RFC 1951 constants
__constant__ uint16_t kLenBase[29] = {3,  4,  5,  6,  7,  8,  9,  10, 11,  13,  15,  17,  19,  23, 27,
                                      31, 35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258};
__constant__ uint8_t kLenExtra[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2,
                                      2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
__constant__ uint16_t kDistBase[30] = {1,    2,    3,    4,    5,    7,     9,     13,    17,  25,
                                       33,   49,   65,   97,   129,  193,   257,   385,   513, 769,
                                       1025, 1537, 2049, 3073, 4097, 6145, 8193, 12289, 16385, 24577};
__constant__ uint8_t kDistExtra[30] = {0, 0, 0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 6,
                                       6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13};
// order of the code-length code lengths in a dynamic block header
__constant__ uint8_t kClOrder[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};

// ------------------------------------------------------------------ decode tables
// Two-level tables of 16-bit entries, indexed by the next bits of the stream (LSB first).
//   leaf:    bit 15 = 0, bits 0-3 = bits to consume at this level (0 = invalid code),
//            bits 4-12 = symbol
//   pointer: bit 15 = 1, bits 0-3 = index bits of the subtable, bits 4-14 = its offset
// Root sizes 9 (literal/length) and 6 (distance) bits; with subtables of the smallest
// power of two that holds every code under their prefix, complete codes of at most 15
// bits need at most 852 and 592 entries (the bounds zlib's enough.c derives).
constexpr int LBITS = 9, DBITS = 6, CBITS = 7;
constexpr int LTAB = 852, DTAB = 592;

struct LaneTables {
    uint16_t lit[LTAB];
    uint16_t dist[DTAB];
    uint16_t cnt[16];
    uint16_t nxt[16];
    uint16_t nx2[16];
    uint8_t lens[320];
};  // 3,304 bytes
constexpr int LANE_BYTES = (sizeof(LaneTables) + 15) / 16 * 16;

struct FixedTables {
    uint16_t lit[1 << LBITS];  // fixed literal/length code: 7..9 bits, no subtables
    uint16_t dist[1 << DBITS]; // fixed distance code: 5 bits
};

enum { TAB_CODES = 0, TAB_LENS = 1, TAB_DISTS = 2 };

// status of a block the decode phase of the two-phase inflate leaves to the one-phase kernel
// (never left in a status array the caller sees)
constexpr int BLOCK_DEFER = 0xff;

__device__ __forceinline__ uint32_t bitrev(uint32_t c, int len) { return __builtin_bitreverse32(c) >> (32 - len); }

// Canonical Huffman decode table for code lengths lens[0:nsym) (RFC 1951 section 3.2.2).
// Over-subscribed lengths are an error; incomplete ones are an error unless the longest
// code has one bit (and never for the code-length code): the rules zlib applies. An
// empty code gives a table of invalid entries. Returns false on error.
__device__ bool build_table(const uint8_t* lens, int nsym, int root, uint16_t* tab, int cap, int type,
                            LaneTables* s) {
    uint16_t* count = s->cnt;
    uint16_t* next = s->nxt;
    uint16_t* next2 = s->nx2;
    for (int l = 0; l < 16; l++) count[l] = 0;
    for (int i = 0; i < nsym; i++) count[lens[i]]++;
    int max = 15;
    while (max >= 1 && count[max] == 0) max--;
    int rsize = 1 << root;
    if (max == 0) {
        for (int i = 0; i < rsize; i++) tab[i] = 0;
        return true;
    }
    int left = 1;
    for (int l = 1; l < 16; l++) {
        left = (left << 1) - count[l];
        if (left < 0) return false;
    }
    if (left > 0 && (type == TAB_CODES || max != 1)) return false;
    count[0] = 0;
    uint32_t code = 0;
    for (int l = 1; l < 16; l++) {
        code = (code + count[l - 1]) << 1;
        next[l] = (uint16_t)code;
        next2[l] = (uint16_t)code;
    }
    if (left > 0 || max > root)
        for (int i = 0; i < rsize; i++) tab[i] = 0;
    uint32_t rmask = rsize - 1;
    if (max > root) {
        // pass A: the deepest code under each root prefix
        for (int s2 = 0; s2 < nsym; s2++) {
            int L = lens[s2];
            if (!L) continue;
            uint32_t c = next2[L]++;
            if (L > root) {
                uint32_t pre = bitrev(c, L) & rmask;
                if ((uint32_t)(L - root) > tab[pre]) tab[pre] = (uint16_t)(L - root);
            }
        }
        // pass B: place the subtables
        int off = rsize;
        for (int p = 0; p < rsize; p++) {
            int m = tab[p];
            if (m) {
                if (off + (1 << m) > cap) return false;
                tab[p] = (uint16_t)(0x8000 | (off << 4) | m);
                off += 1 << m;
            }
        }
    }
    // pass C: fill leaves
    for (int s2 = 0; s2 < nsym; s2++) {
        int L = lens[s2];
        if (!L) continue;
        uint32_t rev = bitrev(next[L]++, L);
        if (L <= root) {
            uint16_t e = (uint16_t)((s2 << 4) | L);
            for (uint32_t i = rev; i < (uint32_t)rsize; i += 1u << L) tab[i] = e;
        } else {
            uint16_t p = tab[rev & rmask];
            int m = p & 15, off = (p >> 4) & 0x7ff, sl = L - root;
            uint16_t e = (uint16_t)((s2 << 4) | sl);
            for (uint32_t i = rev >> root; i < (1u << m); i += 1u << sl) tab[off + i] = e;
        }
    }
    return true;
}

// ------------------------------------------------------------------ bit reader
// Reads the payload LSB first through a 64-bit buffer, refilled 32 bits at a time from
// 8-byte aligned words loaded two words ahead.
struct Bits {
    const uint8_t* base;  // payload start
    const uint64_t* p;    // next aligned word to prefetch
    uint64_t pf0, pf1;
    int pfb;  // bits of pf0 not yet moved: 64 or 32
    uint64_t buf;
    int cnt;
    int64_t loaded;  // bits moved into buf, counted from the payload start

    __device__ __forceinline__ void init(const uint8_t* b, int64_t pos) {
        base = b;
        uintptr_t a = (uintptr_t)(b + pos);
        const uint64_t* q = (const uint64_t*)(a & ~(uintptr_t)7);
        int sh = (int)(a & 7);
        uint64_t w0 = q[0];
        pf0 = q[1];
        pf1 = q[2];
        p = q + 3;
        pfb = 64;
        buf = w0 >> (8 * sh);
        cnt = 64 - 8 * sh;
        loaded = 8 * pos + cnt;
    }
    __device__ __forceinline__ void refill() {
        if (cnt <= 32) {
            uint32_t w;
            if (pfb == 64) {
                w = (uint32_t)pf0;
                pfb = 32;
            } else {
                w = (uint32_t)(pf0 >> 32);
                pf0 = pf1;
                pf1 = *p++;
                pfb = 64;
            }
            buf |= (uint64_t)w << cnt;
            cnt += 32;
            loaded += 32;
        }
    }
    __device__ __forceinline__ uint32_t peek(int n) const { return (uint32_t)buf & ((1u << n) - 1); }
    __device__ __forceinline__ void drop(int n) {
        buf >>= n;
        cnt -= n;
    }
    __device__ __forceinline__ uint32_t take(int n) {
        uint32_t v = peek(n);
        drop(n);
        return v;
    }
    __device__ __forceinline__ int64_t consumed() const { return loaded - cnt; }
};

// decode one symbol with a two-level table; returns (symbol << 4 | 0) or -1 when invalid
__device__ __forceinline__ int decode_sym(Bits& br, const uint16_t* tab, int root) {
    uint32_t e = tab[br.peek(root)];
    if (e & 0x8000) {
        int m = e & 15, off = (e >> 4) & 0x7ff;
        uint32_t idx = (uint32_t)(br.buf >> root) & ((1u << m) - 1);
        uint32_t e2 = tab[off + idx];
        int l = e2 & 15;
        if (!l) return -1;
        br.drop(root + l);
        return (e2 >> 4) & 0x1ff;
    }
    int l = e & 15;
    if (!l) return -1;
    br.drop(l);
    return (e >> 4) & 0x1ff;
}

// ------------------------------------------------------------------ one block
__device__ __forceinline__ void copy_match(uint8_t* out, int32_t op, int32_t dist, int32_t len) {
    uint8_t* d = out + op;
    const uint8_t* s = d - dist;
    if (dist >= 8) {
        for (int32_t k = 0; k < len; k += 8) {
            uint8_t t[8];
#pragma unroll
            for (int j = 0; j < 8; j++) t[j] = (k + j < len) ? s[k + j] : 0;
#pragma unroll
            for (int j = 0; j < 8; j++)
                if (k + j < len) d[k + j] = t[j];
        }
    } else {
        uint8_t t[8];
#pragma unroll
        for (int j = 0; j < 8; j++) t[j] = (j < dist) ? s[j] : 0;
        int r = 0;
        for (int32_t k = 0; k < len; k++) {
            uint8_t v = t[0];
#pragma unroll
            for (int j = 1; j < 8; j++)
                if (j == r) v = t[j];
            d[k] = v;
            r = (r + 1 == dist) ? 0 : r + 1;
        }
    }
}

// Inflate one raw DEFLATE stream of `in_len` bytes at `in` into out[0:cap). Returns a
// CBM_BLOCK_* status (never CRC).
__device__ int inflate_one(const uint8_t* in, int64_t in_len, uint8_t* out, int32_t cap, LaneTables* T,
                           const FixedTables* F) {
    Bits br;
    br.init(in, 0);
    const int64_t limit = 8 * in_len + 64;  // bits that may be read before giving up
    int32_t op = 0;
    bool final = false;
    while (!final) {
        br.refill();
        if (br.consumed() + 3 > 8 * in_len) return CBM_BLOCK_DEFLATE;
        final = br.take(1);
        uint32_t type = br.take(2);
        const uint16_t *lt, *dt;
        if (type == 0) {
            // stored: skip to a byte boundary, LEN and its complement, then LEN bytes
            br.drop(br.cnt & 7);
            br.refill();
            uint32_t len = br.take(16), nlen = br.take(16);
            if (len != (~nlen & 0xffffu)) return CBM_BLOCK_DEFLATE;
            int64_t pos = br.consumed() >> 3;
            if (pos + len > in_len) return CBM_BLOCK_DEFLATE;
            if (op + (int32_t)len > cap) return CBM_BLOCK_SIZE;
            const uint8_t* s = in + pos;
            uint8_t* d = out + op;
            for (uint32_t k = 0; k < len; k += 8) {
                uint8_t t[8];
#pragma unroll
                for (int j = 0; j < 8; j++) t[j] = (k + j < len) ? s[k + j] : 0;
#pragma unroll
                for (int j = 0; j < 8; j++)
                    if (k + j < len) d[k + j] = t[j];
            }
            op += (int32_t)len;
            br.init(in, pos + len);
            continue;
        } else if (type == 1) {
            lt = F->lit;
            dt = F->dist;
        } else if (type == 2) {
            br.refill();
            uint32_t hlit = br.take(5) + 257, hdist = br.take(5) + 1, hclen = br.take(4) + 4;
            if (hlit > 286 || hdist > 30) return CBM_BLOCK_DEFLATE;
            uint8_t* lens = T->lens;
            for (int i = 0; i < 19; i++) lens[i] = 0;
            for (uint32_t i = 0; i < hclen; i++) {
                br.refill();
                lens[kClOrder[i]] = (uint8_t)br.take(3);
            }
            // the code-length code: its table goes in the literal table's space, and the
            // lengths it decodes overwrite its own lengths once the table is built
            if (!build_table(lens, 19, CBITS, T->lit, LTAB, TAB_CODES, T)) return CBM_BLOCK_DEFLATE;
            uint32_t total = hlit + hdist, n = 0;
            while (n < total) {
                br.refill();
                int sym = decode_sym(br, T->lit, CBITS);
                if (sym < 0) return CBM_BLOCK_DEFLATE;
                if (sym < 16) {
                    lens[n++] = (uint8_t)sym;
                } else {
                    uint32_t rep;
                    uint8_t val = 0;
                    if (sym == 16) {
                        if (n == 0) return CBM_BLOCK_DEFLATE;
                        val = lens[n - 1];
                        rep = 3 + br.take(2);
                    } else if (sym == 17) {
                        rep = 3 + br.take(3);
                    } else {
                        rep = 11 + br.take(7);
                    }
                    if (n + rep > total) return CBM_BLOCK_DEFLATE;
                    while (rep--) lens[n++] = val;
                }
                if (br.loaded > limit) return CBM_BLOCK_DEFLATE;
            }
            if (lens[256] == 0) return CBM_BLOCK_DEFLATE;
            if (!build_table(lens, (int)hlit, LBITS, T->lit, LTAB, TAB_LENS, T)) return CBM_BLOCK_DEFLATE;
            if (!build_table(lens + hlit, (int)hdist, DBITS, T->dist, DTAB, TAB_DISTS, T)) return CBM_BLOCK_DEFLATE;
            lt = T->lit;
            dt = T->dist;
        } else {
            return CBM_BLOCK_DEFLATE;
        }
        // Huffman-coded data up to the end-of-block symbol
        for (;;) {
            br.refill();
            if (br.loaded > limit) return CBM_BLOCK_DEFLATE;
            int sym = decode_sym(br, lt, LBITS);
            if (sym < 0) return CBM_BLOCK_DEFLATE;
            if (sym < 256) {
                if (op >= cap) return CBM_BLOCK_SIZE;
                out[op++] = (uint8_t)sym;
                continue;
            }
            if (sym == 256) break;
            sym -= 257;
            if (sym >= 29) return CBM_BLOCK_DEFLATE;
            int32_t len = kLenBase[sym] + (int32_t)br.take(kLenExtra[sym]);
            br.refill();
            int ds = decode_sym(br, dt, DBITS);
            if (ds < 0 || ds >= 30) return CBM_BLOCK_DEFLATE;
            int32_t dist = kDistBase[ds] + (int32_t)br.take(kDistExtra[ds]);
            if (dist > op) return CBM_BLOCK_DEFLATE;
            if (op + len > cap) return CBM_BLOCK_SIZE;
            copy_match(out, op, dist, len);
            op += len;
        }
    }
    if (br.consumed() > 8 * in_len) return CBM_BLOCK_DEFLATE;
    return op == cap ? CBM_BLOCK_OK : CBM_BLOCK_SIZE;
}

__device__ void build_fixed(FixedTables* F, LaneTables* T) {
    uint8_t* lens = T->lens;
    for (int i = 0; i < 144; i++) lens[i] = 8;
    for (int i = 144; i < 256; i++) lens[i] = 9;
    for (int i = 256; i < 280; i++) lens[i] = 7;
    for (int i = 280; i < 288; i++) lens[i] = 8;
    build_table(lens, 288, LBITS, F->lit, 1 << LBITS, TAB_LENS, T);
    for (int i = 0; i < 32; i++) lens[i] = 5;
    build_table(lens, 32, DBITS, F->dist, 1 << DBITS, TAB_DISTS, T);
}

// One workgroup: waves_per_group wavefronts; lanes < bpw of each inflate blocks
// b = decoder, decoder + n_decoders, ... (the blocks of a BAM file are nearly all full).
__global__ void inflate_lanes(const uint8_t* src, const int64_t* blk_off, const int32_t* bsize,
                              const int32_t* isize, const int64_t* out_off, int64_t n_blocks, uint8_t* dst,
                              int32_t* status, int bpw) {
    extern __shared__ __align__(16) uint8_t smem[];
    FixedTables* F = (FixedTables*)smem;
    uint8_t* lanes = smem + (sizeof(FixedTables) + 15) / 16 * 16;
    int wave = threadIdx.x / CBM_WAVE, lane = threadIdx.x % CBM_WAVE;
    int wpg = blockDim.x / CBM_WAVE;
    if (threadIdx.x == 0) build_fixed(F, (LaneTables*)lanes);
    __syncthreads();
    if (lane >= bpw) return;
    LaneTables* T = (LaneTables*)(lanes + (size_t)(wave * bpw + lane) * LANE_BYTES);
    int64_t dec = ((int64_t)blockIdx.x * wpg + wave) * bpw + lane;
    int64_t ndec = (int64_t)gridDim.x * wpg * bpw;
    for (int64_t b = dec; b < n_blocks; b += ndec) {
        int64_t o = blk_off[b];
        int32_t bs = bsize[b];
        int st = inflate_one(src + o + 18, bs - 26, dst + out_off[b], isize[b], T, F);
        status[b] = st;
    }
}

// ------------------------------------------------------------------ CRC32
// zlib's CRC-32 (reflected polynomial 0xEDB88320). Each wavefront takes one block; lane j
// takes the j-th of 64 segments of 1,024 bytes counted back from the block end, so the
// factor that moves a segment's CRC into place, x^(8 * 1024 * (63 - j)) mod P, depends on
// the lane only. crc(A B) = crc(A) * x^(8 |B|) + crc(B) (zlib's crc32_combine), summed
// over lanes with XOR.
constexpr int CRC_SEG = 1024;
constexpr int CRC_NT = 256;

__device__ __forceinline__ uint32_t crc_mulmod(uint32_t a, uint32_t b) {
    // a * b modulo P in the reflected representation (bit 31 = x^0)
    uint32_t p = 0;
    for (int k = 0; k < 32; k++) {
        if (a & 0x80000000u) p ^= b;
        a <<= 1;
        b = (b & 1) ? (b >> 1) ^ 0xEDB88320u : b >> 1;
    }
    return p;
}

__device__ void crc_tables(uint32_t (*T)[256], uint32_t* X) {
    for (int i = threadIdx.x; i < 256; i += blockDim.x) {
        uint32_t c = i;
        for (int k = 0; k < 8; k++) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
        T[0][i] = c;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 256; i += blockDim.x) {
        uint32_t c = T[0][i];
        for (int t = 1; t < 4; t++) {
            c = (c >> 8) ^ T[0][c & 255];
            T[t][i] = c;
        }
    }
    if (threadIdx.x < 64) {
        // X[k] = x^(8 * CRC_SEG * k): powers x^(2^j) by squaring from x^1 (bit 30)
        uint32_t k = threadIdx.x, r = 0x80000000u, sq = 0x40000000u;
        for (int j = 0; j < 13; j++) sq = crc_mulmod(sq, sq);  // x^(2^13) = x^(8 * 1024)
        while (k) {
            if (k & 1) r = crc_mulmod(r, sq);
            sq = crc_mulmod(sq, sq);
            k >>= 1;
        }
        X[threadIdx.x] = r;
    }
    __syncthreads();
}

// zlib CRC32 (pre- and post-inverted) of p[0:n), byte-table with 4-byte steps
__device__ __forceinline__ uint32_t crc_bytes(const uint8_t* p, int64_t n, const uint32_t (*T)[256]) {
    uint32_t c = 0xFFFFFFFFu;
    int64_t k = 0;
    while (k < n && (((uintptr_t)(p + k)) & 3)) c = T[0][(c ^ p[k++]) & 255] ^ (c >> 8);
    for (; k + 4 <= n; k += 4) {
        c ^= *(const uint32_t*)(p + k);
        c = T[3][c & 255] ^ T[2][(c >> 8) & 255] ^ T[1][(c >> 16) & 255] ^ T[0][c >> 24];
    }
    for (; k < n; k++) c = T[0][(c ^ p[k]) & 255] ^ (c >> 8);
    return c ^ 0xFFFFFFFFu;
}

// CRC32 of each block output; status CRC on mismatch; result[0] += blocks not OK.
__global__ void __launch_bounds__(CRC_NT) crc_blocks(const uint8_t* dst, const int64_t* out_off,
                                                     const int32_t* isize, const uint32_t* expect, int64_t n,
                                                     uint32_t* crc, int32_t* status, int64_t* bad) {
    __shared__ uint32_t T[4][256];
    __shared__ uint32_t X[64];
    crc_tables(T, X);
    int wave = threadIdx.x / CBM_WAVE, lane = threadIdx.x % CBM_WAVE;
    int64_t nw = (int64_t)gridDim.x * (CRC_NT / CBM_WAVE);
    for (int64_t b = (int64_t)blockIdx.x * (CRC_NT / CBM_WAVE) + wave; b < n; b += nw) {
        if (status[b] == BLOCK_DEFER) continue;  // decoded again, CRC32 included, by the one-phase kernel
        int64_t len = isize[b];
        int64_t e = len - (int64_t)(63 - lane) * CRC_SEG, s = e - CRC_SEG;
        if (s < 0) s = 0;
        uint32_t c = 0;
        if (e > s) c = crc_mulmod(X[63 - lane], crc_bytes(dst + out_off[b] + s, e - s, T));
        for (int d = 32; d; d >>= 1) c ^= __shfl_xor(c, d, CBM_WAVE);
        if (lane == 0) {
            crc[b] = c;
            int st = status[b];
            if (st == CBM_BLOCK_OK && c != expect[b]) status[b] = st = CBM_BLOCK_CRC;
            if (st != CBM_BLOCK_OK) atomicAdd((unsigned long long*)bad, 1ull);
        }
    }
}

// CRC32 of arbitrary spans (one wavefront each): segments of ceil(len / 64) bytes per lane,
// moved into place with x^(8 * bytes after the segment) from the x^(2^j) powers.
__global__ void __launch_bounds__(CRC_NT) crc_spans(const uint8_t* data, const int64_t* off, const int64_t* len,
                                                    int64_t n, uint32_t* out) {
    __shared__ uint32_t T[4][256];
    __shared__ uint32_t X[64];
    crc_tables(T, X);
    int wave = threadIdx.x / CBM_WAVE, lane = threadIdx.x % CBM_WAVE;
    int64_t nw = (int64_t)gridDim.x * (CRC_NT / CBM_WAVE);
    for (int64_t w = (int64_t)blockIdx.x * (CRC_NT / CBM_WAVE) + wave; w < n; w += nw) {
        int64_t l = len[w], seg = (l + 63) / 64;
        int64_t s = lane * seg < l ? lane * seg : l, e = s + seg < l ? s + seg : l;
        uint32_t c = e > s ? crc_bytes(data + off[w] + s, e - s, T) : 0;
        uint64_t after = (uint64_t)(l - e);
        uint32_t m = 0x80000000u, sq = 0x40000000u;  // x^0, x^1
        for (int j = 0; j < 3; j++) sq = crc_mulmod(sq, sq);  // x^8
        while (after) {
            if (after & 1) m = crc_mulmod(m, sq);
            sq = crc_mulmod(sq, sq);
            after >>= 1;
        }
        c = e > s ? crc_mulmod(m, c) : 0;
        for (int d = 32; d; d >>= 1) c ^= __shfl_xor(c, d, CBM_WAVE);
        if (lane == 0) out[w] = c;
    }
}

}  // namespace cbm
