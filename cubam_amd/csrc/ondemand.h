// On-demand record kernels: what applications compute from a RecordBatch after read_bam
// (the reference's RecordBatch.name_hash / aligned_length / blocks / base_at and Tag.hash /
// Tag.packed, src/cubam/bam.py:221-248, 322-389, kernels in src/cubam/_kernels.py:322-330,
// 391-437, 498-571). One thread per record; each reads the record in place from the batch's
// inflated data.
#pragma once
#include "common.h"

namespace cbm {

constexpr uint64_t FNV_BASIS = 1469598103934665603ull;  // the reference's basis (bam.py:169)
constexpr uint64_t FNV_PRIME = 1099511628211ull;
constexpr int OD_NT = 256;

__device__ __forceinline__ bool aligned_op(uint32_t op) { return op == 0 || op == 7 || op == 8; }  // M = X

// FNV-1a 64 of the read name (without its NUL)
__global__ void __launch_bounds__(OD_NT) od_name_hash(const uint8_t* b, const int64_t* off, int64_t n, uint64_t* h) {
    const int64_t i = (int64_t)blockIdx.x * OD_NT + threadIdx.x;
    if (i >= n) return;
    const int64_t o = off[i];
    const uint32_t lrn = b[o + 12];
    uint64_t x = FNV_BASIS;
    for (uint32_t k = 0; k + 1 < lrn; k++) x = (x ^ b[o + 36 + k]) * FNV_PRIME;
    h[i] = x;
}

// FNV-1a 64 of spans b[voff[i] : + vlen[i]); 0 where vlen[i] < 0
__global__ void __launch_bounds__(OD_NT) od_hash_spans(const uint8_t* b, const int64_t* voff, const int32_t* vlen,
                                                       int64_t n, uint64_t* out) {
    const int64_t i = (int64_t)blockIdx.x * OD_NT + threadIdx.x;
    if (i >= n) return;
    const int32_t l = vlen[i];
    if (l < 0) {
        out[i] = 0;
        return;
    }
    uint64_t x = FNV_BASIS;
    const uint8_t* p = b + voff[i];
    for (int32_t k = 0; k < l; k++) x = (x ^ p[k]) * FNV_PRIME;
    out[i] = x;
}

// [ACGT]{1,24}(-[0-9]{1,3})? packed: 2 bits a base from bit 0, the base count at bit 48, the
// suffix + 1 at bit 53 (0: no suffix). Code 0 and ok = 0 for a value that does not fit.
__global__ void __launch_bounds__(OD_NT) od_pack_acgt(const uint8_t* b, const int64_t* voff, const int32_t* vlen,
                                                      int64_t n, uint64_t* out, uint8_t* ok) {
    const int64_t i = (int64_t)blockIdx.x * OD_NT + threadIdx.x;
    if (i >= n) return;
    const int32_t l = vlen[i];
    const uint8_t* p = b + voff[i];
    uint64_t code = 0, suffix = 0;
    int nb = 0, k = 0;
    bool good = l > 0;
    for (; good && k < l && p[k] != '-'; k++) {
        const uint8_t c = p[k];
        const int x = c == 'A' ? 0 : c == 'C' ? 1 : c == 'G' ? 2 : c == 'T' ? 3 : 4;
        if (x == 4 || nb == 24) good = false;
        else code |= (uint64_t)x << (2 * nb++);
    }
    if (good && k < l) {  // a '-': one to three digits follow
        k++;
        int nd = 0;
        for (; k < l; k++, nd++) {
            if (p[k] < '0' || p[k] > '9' || nd == 3) {
                good = false;
                break;
            }
            suffix = suffix * 10 + (p[k] - '0');
        }
        if (nd == 0) good = false;
        suffix += 1;
    }
    if (nb == 0) good = false;
    out[i] = good ? (code | ((uint64_t)nb << 48) | (suffix << 53)) : 0;
    ok[i] = good ? 1 : 0;
}

// Aligned bases (CIGAR M, = and X) of each read
__global__ void __launch_bounds__(OD_NT) od_aligned_len(const uint8_t* b, const int64_t* off, int64_t n, int32_t* out) {
    const int64_t i = (int64_t)blockIdx.x * OD_NT + threadIdx.x;
    if (i >= n) return;
    const int64_t o = off[i];
    const uint32_t lrn = b[o + 12], ncig = ld_u16(b, o + 16);
    const int64_t cg = o + 36 + lrn;
    int32_t s = 0;
    for (uint32_t k = 0; k < ncig; k++) {
        const uint32_t v = ld_u32(b, cg + 4LL * k);
        if (aligned_op(v & 15)) s += (int32_t)(v >> 4);
    }
    out[i] = s;
}

// Aligned blocks: runs of M, = and X (and D unless split_del) between reference skips (N), and
// between deletions with split_del; insertions and clips do not break a block. fill = 0 counts
// each read's blocks into cnt; fill = 1 writes them from first[i].
__global__ void __launch_bounds__(OD_NT) od_blocks(const uint8_t* b, const int64_t* off, int64_t n, int split_del,
                                                   int fill, const int64_t* first, int32_t* cnt, int64_t* rec,
                                                   int32_t* start, int32_t* end) {
    const int64_t i = (int64_t)blockIdx.x * OD_NT + threadIdx.x;
    if (i >= n) return;
    const int64_t o = off[i];
    const uint32_t lrn = b[o + 12], ncig = ld_u16(b, o + 16);
    const int64_t cg = o + 36 + lrn;
    int32_t x = (int32_t)ld_u32(b, o + 8), s = x, k = 0;
    const int64_t w = fill ? first[i] : 0;
    for (uint32_t c = 0; c < ncig; c++) {
        const uint32_t v = ld_u32(b, cg + 4LL * c), op = v & 15, l = v >> 4;
        const bool adv = aligned_op(op) || (op == 2 && !split_del);
        const bool gap = op == 3 || (op == 2 && split_del);
        if (adv) {
            x += (int32_t)l;
        } else if (gap) {
            if (x > s) {
                if (fill) {
                    rec[w + k] = i;
                    start[w + k] = s;
                    end[w + k] = x;
                }
                k++;
            }
            x += (int32_t)l;
            s = x;
        }
    }
    if (x > s) {
        if (fill) {
            rec[w + k] = i;
            start[w + k] = s;
            end[w + k] = x;
        }
        k++;
    }
    if (!fill) cnt[i] = k;
}

// The base of read rec[j] at reference position refpos[j]: 0..3 for A C G T, 4 for another
// base, -1 in a deletion, -2 in a reference skip (N), -3 where the read does not cover it.
__global__ void __launch_bounds__(OD_NT) od_base_at(const uint8_t* b, const int64_t* off, const int64_t* rec,
                                                    const int32_t* refpos, int64_t m, int8_t* base) {
    const int64_t j = (int64_t)blockIdx.x * OD_NT + threadIdx.x;
    if (j >= m) return;
    const int64_t o = off[rec[j]];
    const int32_t target = refpos[j];
    const uint32_t lrn = b[o + 12], ncig = ld_u16(b, o + 16), lseq = ld_u32(b, o + 20);
    const int64_t cg = o + 36 + lrn, sq = cg + 4LL * ncig;
    int64_t x = (int32_t)ld_u32(b, o + 8), y = 0;
    int8_t r = -3;
    for (uint32_t k = 0; k < ncig; k++) {
        const uint32_t v = ld_u32(b, cg + 4LL * k), op = v & 15, l = v >> 4;
        if (aligned_op(op)) {
            if (target < x + l) {
                const int64_t q = y + (target - x);
                if (target >= x && q < (int64_t)lseq) {
                    const uint32_t c = (b[sq + q / 2] >> ((q & 1) ? 0 : 4)) & 15;
                    r = c == 1 ? 0 : c == 2 ? 1 : c == 4 ? 2 : c == 8 ? 3 : 4;
                }
                break;
            }
            x += l;
            y += l;
        } else if (op == 2 || op == 3) {
            if (target < x + l) {
                r = target >= x ? (op == 2 ? -1 : -2) : -3;
                break;
            }
            x += l;
        } else if (op == 1 || op == 4) {
            y += l;
        }
    }
    base[j] = r;
}

}  // namespace cbm
