// Kernel 5: BGZF compression for the write-back path. One wavefront compresses one BGZF
// block (at most 65,280 payload bytes, as htslib cuts them) into its own output slot:
//
//   1. Match search, 64 positions at a time: each lane hashes the 4 bytes at its position,
//      looks the hash up in the wavefront's LDS table (the last position seen with it, from
//      earlier windows), extends the match by comparing bytes, then every lane inserts its
//      position. A greedy parse (a match of 4 or more bytes is taken, else a literal) walks
//      the window with uniform readlanes; the lanes on the parse write their symbols to a
//      scratch list in order and count code frequencies in LDS. This is zlib's level-1
//      strategy (greedy, one candidate per hash), laid out for a 64-wide wavefront.
//   2. Dynamic Huffman codes (RFC 1951 3.2.7) for literals/lengths, distances and the code
//      lengths: Huffman trees by the two-queue method over frequency-sorted leaves, depths
//      limited to 15 (7 for code lengths) by zlib's overflow rebalancing (trees.c,
//      gen_bitlen), canonical codes. A block whose dynamic encoding would not be shorter
//      than a stored block is stored.
//   3. Bit packing, 64 symbols at a time: a wavefront prefix sum of the symbols' bit
//      lengths places each one; lanes OR their bits into an LDS staging window and the
//      complete words go out with one coalesced store.
//   4. BGZF framing: the gzip header with the BC extra field (BSIZE), the trailer with the
//      CRC32 (computed by all lanes, slices combined with x^(8n) mod P) and ISIZE.
//
// Compressed bytes are not meant to equal zlib's; the payload and CRC32 of every block are.
#pragma once
#include "inflate.h"

namespace cbm {

constexpr int DF_WAVES = 4;                  // blocks (wavefronts) per workgroup
constexpr int DF_NT = DF_WAVES * CBM_WAVE;
constexpr int DF_HBITS = 12;                 // hash table entries per wavefront: 4096
constexpr int DF_MAX_IN = 65280;             // payload bytes per block at most
constexpr int DF_SLOT = 65536 + 256;         // output bytes per block slot (a BGZF block is <= 65,536)
constexpr int DF_MIN_MATCH = 4;
constexpr int DF_STAGE = 128;                // bit-packing staging words

struct DfWave {
    uint16_t head[1 << DF_HBITS];  // last position with each hash, 0xFFFF = none
    uint32_t lfreq[288];           // literal/length symbol frequencies (286 used)
    uint32_t dfreq[32];            // distance symbol frequencies (30 used)
    uint32_t cfreq[20];            // code-length symbol frequencies (19 used)
    uint8_t len[320];              // code lengths: literal/length [0, 286), distance [288, 318)
    uint8_t clen[20];              // code-length code lengths
    uint16_t code[320];            // bit-reversed canonical codes, same layout as len
    uint16_t ccode[20];
    uint16_t rle[320];             // run-length coded code lengths: symbol | extra << 5
    // Huffman tree scratch: leaves sorted by frequency, internal nodes after them
    uint32_t nf[2 * 288];
    uint16_t leaf[288];
    int16_t parent[2 * 288];
    uint8_t depth[2 * 288];
    uint32_t stage[DF_STAGE];      // bit-packing window
    int32_t sv[16];                // shared scalars
};

enum { DV_NSYM = 0, DV_NRLE, DV_HLIT, DV_HDIST, DV_HCLEN, DV_BITS };

// length symbol (257..285) and extra bits of a match length 3..258 (RFC 1951 3.2.5)
__device__ __forceinline__ void df_len_sym(int len, int& sym, int& ebits, int& eval) {
    if (len <= 10) {
        sym = 254 + len;
        ebits = 0;
        eval = 0;
    } else if (len == 258) {
        sym = 285;
        ebits = 0;
        eval = 0;
    } else {
        const int v = len - 3;                 // 8 .. 254
        const int e = 29 - __clz(v);           // extra bits: floor(log2 v) - 2
        sym = 257 + 4 * (e + 1) + ((v >> e) & 3);
        ebits = e;
        eval = v & ((1 << e) - 1);
    }
}

// distance symbol (0..29) and extra bits of a distance 1..32768
__device__ __forceinline__ void df_dist_sym(int dist, int& sym, int& ebits, int& eval) {
    const int v = dist - 1;
    if (v < 4) {
        sym = v;
        ebits = 0;
        eval = 0;
    } else {
        const int e = 31 - __clz(v) - 1;  // extra bits: floor(log2 v) - 1
        sym = 2 * (e + 1) + ((v >> e) & 1);
        ebits = e;
        eval = v & ((1 << e) - 1);
    }
}

__device__ __forceinline__ int df_len_extra(int sym) {  // extra bits of length symbol 257..285
    const int c = sym - 257;
    return (c < 8 || c == 28) ? 0 : (c >> 2) - 1;
}
__device__ __forceinline__ int df_dist_extra(int sym) { return sym < 4 ? 0 : (sym >> 1) - 1; }

__device__ __forceinline__ uint32_t df_hash(uint32_t x) { return (x * 2654435761u) >> (32 - DF_HBITS); }

// Code lengths (at most maxbits) of a Huffman code for freq[0:n) into len[0:n), by one wavefront.
// Symbols of zero frequency get length 0. One symbol alone gets length 1; with `complete`, a
// second symbol gets length 1 too (a code-length code must be complete, RFC 1951 3.2.7).
__device__ void df_lengths(DfWave& w, const uint32_t* freq, int n, int maxbits, uint8_t* len,
                           bool complete = false) {
    const int lane = threadIdx.x % CBM_WAVE;
    // leaves: symbols of nonzero frequency, sorted by (frequency, symbol)
    int m = 0;
    for (int s0 = 0; s0 < n; s0 += CBM_WAVE) {
        const int s = s0 + lane;
        const bool nz = s < n && freq[s] != 0;
        const uint64_t b = __ballot(nz);
        if (nz) w.leaf[m + __popcll(b & ((1ull << lane) - 1))] = (uint16_t)s;
        m += __popcll(b);
    }
    for (int s = lane; s < n; s += CBM_WAVE) len[s] = 0;
    __builtin_amdgcn_wave_barrier();
    if (m == 0) return;
    if (m == 1) {
        if (lane == 0) {
            len[w.leaf[0]] = 1;
            if (complete) len[w.leaf[0] == 0 ? 1 : 0] = 1;
        }
        __builtin_amdgcn_wave_barrier();
        return;
    }
    // rank sort of the leaves (each lane ranks a few; ties by symbol)
    uint16_t mine[5];
    int rank[5];
#pragma unroll
    for (int k = 0; k < 5; k++) {
        const int i = lane + k * CBM_WAVE;
        mine[k] = i < m ? w.leaf[i] : 0;
        rank[k] = 0;
    }
    for (int j = 0; j < m; j++) {
        const uint16_t sj = w.leaf[j];
        const uint32_t fj = freq[sj];
#pragma unroll
        for (int k = 0; k < 5; k++) {
            const int i = lane + k * CBM_WAVE;
            if (i < m) {
                const uint32_t fi = freq[mine[k]];
                rank[k] += (fj < fi || (fj == fi && sj < mine[k])) ? 1 : 0;
            }
        }
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int k = 0; k < 5; k++) {
        const int i = lane + k * CBM_WAVE;
        if (i < m) {
            w.leaf[rank[k]] = mine[k];
            w.nf[rank[k]] = freq[mine[k]];
        }
    }
    __builtin_amdgcn_wave_barrier();
    if (lane == 0) {
        // two queues: leaves [0, m) in order, internal nodes [m, 2m - 1) in creation order
        int qa = 0, qb = m, nb = m;
        for (int k = 0; k < m - 1; k++) {
            int pick[2];
            for (int t = 0; t < 2; t++) {
                if (qa < m && (qb >= nb || w.nf[qa] <= w.nf[qb])) pick[t] = qa++;
                else pick[t] = qb++;
            }
            w.nf[nb] = w.nf[pick[0]] + w.nf[pick[1]];
            w.parent[pick[0]] = (int16_t)nb;
            w.parent[pick[1]] = (int16_t)nb;
            nb++;
        }
        // depths from the root (the last node) down, clamped at maxbits; every node (leaf or
        // internal) whose depth had to be clamped counts in overflow (zlib's gen_bitlen)
        const int root = nb - 1;
        int blc[16] = {0};
        int overflow = 0;
        w.depth[root] = 0;
        for (int k = root - 1; k >= 0; k--) {
            int d = w.depth[w.parent[k]] + 1;
            if (d > maxbits) {
                d = maxbits;
                overflow++;
            }
            w.depth[k] = (uint8_t)d;
            if (k < m) blc[d]++;
        }
        while (overflow > 0) {
            int bits = maxbits - 1;
            while (bits > 1 && blc[bits] == 0) bits--;
            blc[bits]--;       // a leaf moves down one level
            blc[bits + 1] += 2;  // it and an overflowed leaf become brothers there
            blc[maxbits]--;
            overflow -= 2;
        }
        // longest codes to the least frequent leaves
        int k = 0;
        for (int bits = maxbits; bits >= 1; bits--)
            for (int c = blc[bits]; c > 0; c--) len[w.leaf[k++]] = (uint8_t)bits;
    }
    __builtin_amdgcn_wave_barrier();
}

// Canonical codes of len[0:n) (bit-reversed, to go out least significant bit first)
__device__ void df_codes(const uint8_t* len, int n, uint16_t* code) {
    const int lane = threadIdx.x % CBM_WAVE;
    if (lane == 0) {
        int blc[16] = {0}, next[16];
        for (int s = 0; s < n; s++) blc[len[s]]++;
        blc[0] = 0;
        int c = 0;
        for (int b = 1; b < 16; b++) {
            c = (c + blc[b - 1]) << 1;
            next[b] = c;
        }
        for (int s = 0; s < n; s++) {
            const int l = len[s];
            code[s] = l ? (uint16_t)bitrev((uint32_t)next[l]++, l) : 0;
        }
    }
    __builtin_amdgcn_wave_barrier();
}

// Pack one item per lane (bits v, nb <= 58 of them, nb = 0 for none) at the running bit
// position pos (uniform) of the wavefront's output words; complete words are stored.
__device__ __forceinline__ void df_put(DfWave& w, uint32_t* out, int64_t& pos, uint64_t v, int nb) {
    const int lane = threadIdx.x % CBM_WAVE;
    const int incl = wave_inclusive_sum(nb);
    const int excl = incl - nb;
    const int total = __shfl(incl, CBM_WAVE - 1, CBM_WAVE);
    const int64_t w0 = pos >> 5;
    if (nb) {
        const int64_t q = pos + excl;
        int idx = (int)((q >> 5) - w0);
        const int sh = (int)(q & 31);
        // up to three words: bits [sh, sh + nb)
        atomicOr(&w.stage[idx], (uint32_t)(v << sh));
        if (sh + nb > 32) atomicOr(&w.stage[idx + 1], (uint32_t)(v >> (32 - sh)));
        if (sh + nb > 64) atomicOr(&w.stage[idx + 2], (uint32_t)(v >> (64 - sh)));
    }
    __builtin_amdgcn_wave_barrier();
    const int64_t npos = pos + total;
    const int full = (int)((npos >> 5) - w0);  // complete words now
    for (int k = lane; k < full; k += CBM_WAVE) out[w0 + k] = w.stage[k];
    __builtin_amdgcn_wave_barrier();
    const uint32_t carry = w.stage[full];
    __builtin_amdgcn_wave_barrier();
    for (int k = lane; k < DF_STAGE; k += CBM_WAVE) w.stage[k] = k == 0 ? carry : 0;
    __builtin_amdgcn_wave_barrier();
    pos = npos;
}

// One item from lane 0 only
__device__ __forceinline__ void df_put1(DfWave& w, uint32_t* out, int64_t& pos, uint64_t v, int nb) {
    const int lane = threadIdx.x % CBM_WAVE;
    df_put(w, out, pos, lane == 0 ? v : 0, lane == 0 ? nb : 0);
}

// kernel: blocks [0, n) of src (in_off, in_len) into slots of DF_SLOT bytes; bsize[b] = BGZF block size
__global__ void __launch_bounds__(DF_NT) deflate_blocks(const uint8_t* src, const int64_t* in_off,
                                                        const int32_t* in_len, int64_t n, uint8_t* slots,
                                                        int32_t* bsize, uint32_t* crc_out, uint32_t* symbols,
                                                        int64_t sym_stride) {
    __shared__ DfWave waves[DF_WAVES];
    __shared__ uint32_t T[4][256];
    __shared__ uint32_t X[64];
    crc_tables(T, X);  // X[k] = x^(8 * 1024 * k)
    const int wid = threadIdx.x / CBM_WAVE, lane = threadIdx.x % CBM_WAVE;
    DfWave& w = waves[wid];
    uint32_t* syms = symbols + (size_t)(blockIdx.x * DF_WAVES + wid) * (size_t)sym_stride;
    const int64_t nw = (int64_t)gridDim.x * DF_WAVES;
    for (int64_t b = (int64_t)blockIdx.x * DF_WAVES + wid; b < n; b += nw) {
        const uint8_t* in = src + in_off[b];
        const int L = in_len[b];
        uint8_t* slot = slots + (size_t)b * DF_SLOT;
        // ---- reset
        for (int i = lane; i < (1 << DF_HBITS); i += CBM_WAVE) w.head[i] = 0xFFFF;
        for (int i = lane; i < 288; i += CBM_WAVE) w.lfreq[i] = 0;
        if (lane < 32) w.dfreq[lane] = 0;
        if (lane < 20) w.cfreq[lane] = 0;
        __builtin_amdgcn_wave_barrier();
        // ---- 1. matches and the greedy parse
        int cur = 0, nsym = 0;
        for (int w0 = 0; w0 < L; w0 += CBM_WAVE) {
            const int i = w0 + lane;
            int mlen = 0, mdist = 0, byte = 0;
            uint32_t h = 0;
            const bool hashable = i + DF_MIN_MATCH <= L;
            if (i < L) byte = in[i];
            if (hashable) {
                const uint32_t x = ld_u32(in, i);
                h = df_hash(x);
                const int c = w.head[h];
                if (c != 0xFFFF && i - c <= 32768) {
                    const int lim = L - i < 258 ? L - i : 258;
                    int k = 0;
                    while (k + 4 <= lim && ld_u32(in, c + k) == ld_u32(in, i + k)) k += 4;
                    while (k < lim && in[c + k] == in[i + k]) k++;
                    if (k >= DF_MIN_MATCH) {
                        mlen = k;
                        mdist = i - c;
                    }
                }
            }
            __builtin_amdgcn_wave_barrier();
            if (hashable) w.head[h] = (uint16_t)i;
            // the parse over this window (uniform)
            uint64_t on = 0;
            const int wend = w0 + CBM_WAVE < L ? w0 + CBM_WAVE : L;
            while (cur < wend) {
                const int l = __builtin_amdgcn_readlane(mlen, cur - w0);
                on |= 1ull << (cur - w0);
                cur += l ? l : 1;
            }
            if (on >> lane & 1) {
                const int k = nsym + __popcll(on & ((1ull << lane) - 1));
                if (mlen) {
                    syms[k] = 0x80000000u | ((uint32_t)(mlen - 3) << 16) | (uint32_t)(mdist - 1);
                    int ls, le, lv, ds, de, dv;
                    df_len_sym(mlen, ls, le, lv);
                    df_dist_sym(mdist, ds, de, dv);
                    atomicAdd(&w.lfreq[ls], 1u);
                    atomicAdd(&w.dfreq[ds], 1u);
                } else {
                    syms[k] = (uint32_t)byte;
                    atomicAdd(&w.lfreq[byte], 1u);
                }
            }
            nsym += __popcll(on);
        }
        if (lane == 0) w.lfreq[256] = 1;  // end of block
        __builtin_amdgcn_wave_barrier();
        // ---- 2. codes
        df_lengths(w, w.lfreq, 286, 15, w.len);
        df_lengths(w, w.dfreq, 30, 15, w.len + 288);
        // HLIT / HDIST, then the run-length coding of the lengths and its code
        if (lane == 0) {
            int hlit = 286;
            while (hlit > 257 && w.len[hlit - 1] == 0) hlit--;
            int hdist = 30;
            while (hdist > 1 && w.len[288 + hdist - 1] == 0) hdist--;
            if (w.len[288] == 0 && hdist == 1) w.len[288] = 1;  // no match: one distance code of length 1
            w.sv[DV_HLIT] = hlit;
            w.sv[DV_HDIST] = hdist;
            const int tot = hlit + hdist;
            int nr = 0, k = 0;
            while (k < tot) {
                const int v = k < hlit ? w.len[k] : w.len[288 + k - hlit];
                int run = 1;
                while (k + run < tot && (k + run < hlit ? w.len[k + run] : w.len[288 + k + run - hlit]) == v) run++;
                int left = run;
                if (v == 0) {
                    while (left >= 11) {
                        const int r = left < 138 ? left : 138;
                        w.rle[nr++] = (uint16_t)(18 | ((r - 11) << 5));
                        w.cfreq[18]++;
                        left -= r;
                    }
                    if (left >= 3) {
                        w.rle[nr++] = (uint16_t)(17 | ((left - 3) << 5));
                        w.cfreq[17]++;
                        left = 0;
                    }
                } else {
                    w.rle[nr++] = (uint16_t)v;
                    w.cfreq[v]++;
                    left--;
                    while (left >= 3) {
                        const int r = left < 6 ? left : 6;
                        w.rle[nr++] = (uint16_t)(16 | ((r - 3) << 5));
                        w.cfreq[16]++;
                        left -= r;
                    }
                }
                while (left-- > 0) {
                    w.rle[nr++] = (uint16_t)v;
                    w.cfreq[v]++;
                }
                k += run;
            }
            w.sv[DV_NRLE] = nr;
        }
        __builtin_amdgcn_wave_barrier();
        df_lengths(w, w.cfreq, 19, 7, w.clen, true);
        df_codes(w.len, 286, w.code);
        df_codes(w.len + 288, 30, w.code + 288);
        df_codes(w.clen, 19, w.ccode);
        // sizes: dynamic block bits against a stored block
        if (lane == 0) {
            int hclen = 19;
            while (hclen > 4 && w.clen[kClOrder[hclen - 1]] == 0) hclen--;
            w.sv[DV_HCLEN] = hclen;
            int64_t bits = 3 + 14 + 3 * hclen;
            for (int k = 0; k < w.sv[DV_NRLE]; k++) {
                const int s = w.rle[k] & 31;
                bits += w.clen[s] + (s == 16 ? 2 : s == 17 ? 3 : s == 18 ? 7 : 0);
            }
            for (int s = 0; s < 286; s++)
                if (w.lfreq[s]) bits += (int64_t)w.lfreq[s] * (w.len[s] + (s > 256 ? df_len_extra(s) : 0));
            for (int s = 0; s < 30; s++)
                if (w.dfreq[s]) bits += (int64_t)w.dfreq[s] * (w.len[288 + s] + df_dist_extra(s));
            w.sv[DV_BITS] = (int)bits;
        }
        __builtin_amdgcn_wave_barrier();
        const int64_t dyn_bytes = ((int64_t)w.sv[DV_BITS] + 7) / 8;
        const bool stored = dyn_bytes >= (int64_t)L + 5;
        const int pay = stored ? L + 5 : (int)dyn_bytes;
        const int bs = 18 + pay + 8;
        bool bad_block = false;
        // ---- 3. the Deflate stream from byte 18 of the slot
        uint32_t* ow = (uint32_t*)slot;
        if (stored) {
            for (int k = lane; k < L; k += CBM_WAVE) slot[23 + k] = in[k];
            if (lane == 0) {
                slot[18] = 1;  // final stored block
                slot[19] = (uint8_t)L;
                slot[20] = (uint8_t)(L >> 8);
                slot[21] = (uint8_t)~L;
                slot[22] = (uint8_t)(~L >> 8);
            }
        } else {
            for (int k = lane; k < DF_STAGE; k += CBM_WAVE) w.stage[k] = 0;
            __builtin_amdgcn_wave_barrier();
            int64_t pos = 18 * 8;
            const int hlit = w.sv[DV_HLIT], hdist = w.sv[DV_HDIST], hclen = w.sv[DV_HCLEN];
            df_put1(w, ow, pos, 1 | (2 << 1) | ((uint64_t)(hlit - 257) << 3) | ((uint64_t)(hdist - 1) << 8) |
                                    ((uint64_t)(hclen - 4) << 13),
                    17);
            // code-length code lengths, 3 bits each in kClOrder, one lane each
            df_put(w, ow, pos, lane < hclen ? w.clen[kClOrder[lane < 19 ? lane : 0]] : 0, lane < hclen ? 3 : 0);
            const int nr = w.sv[DV_NRLE];
            for (int k0 = 0; k0 < nr; k0 += CBM_WAVE) {
                uint64_t v = 0;
                int nb = 0;
                if (k0 + lane < nr) {
                    const int r = w.rle[k0 + lane], s = r & 31, x = r >> 5;
                    const int xb = s == 16 ? 2 : s == 17 ? 3 : s == 18 ? 7 : 0;
                    v = (uint64_t)w.ccode[s] | ((uint64_t)x << w.clen[s]);
                    nb = w.clen[s] + xb;
                }
                df_put(w, ow, pos, v, nb);
            }
            __threadfence();  // this wavefront's symbol stores are read back below
            for (int k0 = 0; k0 < nsym; k0 += CBM_WAVE) {
                uint64_t v = 0;
                int nb = 0;
                if (k0 + lane < nsym) {
                    const uint32_t y = syms[k0 + lane];
                    if (y & 0x80000000u) {
                        const int mlen = (int)((y >> 16) & 0x1ff) + 3, mdist = (int)(y & 0xffff) + 1;
                        int ls, le, lv, ds, de, dv;
                        df_len_sym(mlen, ls, le, lv);
                        df_dist_sym(mdist, ds, de, dv);
                        v = w.code[ls];
                        nb = w.len[ls];
                        v |= (uint64_t)lv << nb;
                        nb += le;
                        v |= (uint64_t)w.code[288 + ds] << nb;
                        nb += w.len[288 + ds];
                        v |= (uint64_t)dv << nb;
                        nb += de;
                    } else {
                        v = w.code[y];
                        nb = w.len[y];
                    }
                }
                df_put(w, ow, pos, v, nb);
            }
            df_put1(w, ow, pos, w.code[256], w.len[256]);
            // the last partial word (bits past the end are zero: byte padding)
            if (lane == 0 && (pos & 31)) ow[pos >> 5] = w.stage[0];
            if (pos - 18 * 8 != (int64_t)w.sv[DV_BITS]) bad_block = true;  // cannot happen: sizes agree
        }
        // ---- 4. CRC32 of the payload (lane slices of 1,024 bytes from the end, as crc_blocks)
        uint32_t c = 0;
        {
            const int64_t e = (int64_t)L - (int64_t)(63 - lane) * CRC_SEG;
            int64_t s = e - CRC_SEG;
            if (s < 0) s = 0;
            if (e > s) c = crc_mulmod(X[63 - lane], crc_bytes(in + s, e - s, T));
            for (int d = 32; d; d >>= 1) c ^= __shfl_xor(c, d, CBM_WAVE);
        }
        __threadfence();  // the stream's words are written before the header bytes beside them
        if (lane == 0) {
            const uint8_t hdr[16] = {31, 139, 8, 4, 0, 0, 0, 0, 0, 255, 6, 0, 66, 67, 2, 0};
            for (int k = 0; k < 16; k++) slot[k] = hdr[k];
            slot[16] = (uint8_t)((bs - 1) & 255);
            slot[17] = (uint8_t)((bs - 1) >> 8);
            uint8_t* tr = slot + 18 + pay;
            for (int k = 0; k < 4; k++) {
                tr[k] = (uint8_t)(c >> (8 * k));
                tr[4 + k] = (uint8_t)((uint32_t)L >> (8 * k));
            }
            bsize[b] = bad_block ? -1 : bs;
            if (crc_out) crc_out[b] = c;
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// Copy n spans src[src_off[i] : + len[i]) to dst[dst_off[i] :), one wavefront per span.
__global__ void __launch_bounds__(256) copy_spans(const uint8_t* src, const int64_t* src_off, const int64_t* len,
                                                  const int64_t* dst_off, int64_t n, uint8_t* dst) {
    const int lane = threadIdx.x % CBM_WAVE;
    const int64_t nw = (int64_t)gridDim.x * (256 / CBM_WAVE);
    for (int64_t i = (int64_t)blockIdx.x * (256 / CBM_WAVE) + threadIdx.x / CBM_WAVE; i < n; i += nw) {
        const uint8_t* s = src + src_off[i];
        uint8_t* d = dst + dst_off[i];
        const int64_t l = len[i];
        for (int64_t k = lane; k < l; k += CBM_WAVE) d[k] = s[k];
    }
}

}  // namespace cbm
