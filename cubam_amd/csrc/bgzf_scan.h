// Kernel 1: BGZF block-boundary scan.
//
// The reference tests every byte with one thread, appends candidates with
// warp atomics, radix-sorts them, then syncs the host to check the chain
// (src/cubam/_kernels.py:69-89, 662-702; src/cubam/bgzf.py:549-584). Here a
// workgroup scans a 16 KiB tile with 16-byte vector loads and a cheap 0x1f
// byte filter, tiles are counted, prefix-summed and re-scanned so that the
// candidates come out already in order (no sort), and the chain check, the
// block fields and the output offsets stay on the device. The host reads the
// result slots once.
#pragma once
#include "common.h"

namespace cbm {

constexpr int BGZF_TILE = 16384;
constexpr int BGZF_NT = 256;  // 64 bytes per thread: 4 vectors of 16

// The fixed BGZF header fields at o, BSIZE + 1 >= 26, and the block inside [0, n), read through an
// accessor a (a.u8 at a data offset)
template <class A>
__device__ __forceinline__ bool bgzf_candidate_at(const A& a, int64_t n, int64_t o) {
    if (o + 18 > n) return false;
    if (!(a.u8(o) == 31 && a.u8(o + 1) == 139 && a.u8(o + 2) == 8 && a.u8(o + 3) == 4 && a.u8(o + 10) == 6 &&
          a.u8(o + 11) == 0 && a.u8(o + 12) == 66 && a.u8(o + 13) == 67 && a.u8(o + 14) == 2 && a.u8(o + 15) == 0))
        return false;
    int64_t bs = (int64_t)(a.u8(o + 16) | (a.u8(o + 17) << 8)) + 1;
    return bs >= 26 && o + bs <= n;
}
__device__ __forceinline__ bool bgzf_candidate(const uint8_t* b, int64_t n, int64_t o) {
    return bgzf_candidate_at(GlobalBytes{b}, n, o);
}

// A tile's bytes [base, base + len) staged in LDS, the rest read in place
struct TileBytes {
    const uint8_t* g;
    const uint8_t* l;
    int64_t base, len;
    __device__ __forceinline__ uint32_t u8(int64_t o) const {
        const int64_t r = o - base;
        return (uint64_t)r < (uint64_t)len ? l[r] : g[o];
    }
};

// A 16-byte mask of the bytes of a 4-word vector that equal 0x1f
__device__ __forceinline__ uint32_t mask_1f(uint4 v) {
    uint32_t m = 0;
    uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        uint32_t x = w[k] ^ 0x1f1f1f1fu;
        // zero-byte detection, exact per byte: a byte is zero iff its high bit is set here
        uint32_t t = ~(((x & 0x7f7f7f7fu) + 0x7f7f7f7fu) | x | 0x7f7f7f7fu);
#pragma unroll
        for (int j = 0; j < 4; j++)
            if (t & (0x80u << (8 * j))) m |= 1u << (4 * k + j);
    }
    return m;
}

// Candidates among the 64 offsets [o0, o0 + 64) of one thread: returns a 64-bit mask.
__device__ __forceinline__ uint64_t thread_candidates(const uint8_t* b, int64_t n, int64_t o0) {
    uint64_t mask = 0;
#pragma unroll
    for (int v = 0; v < 4; v++) {
        int64_t o = o0 + 16 * v;
        if (o >= n) break;
        uint32_t m;
        if (o + 16 <= n) {
            uint4 q = *(const uint4*)(b + o);  // buffers are 16-byte aligned at tile starts
            m = mask_1f(q);
        } else {
            m = 0;
            for (int j = 0; o + j < n; j++)
                if (b[o + j] == 31) m |= 1u << j;
        }
        while (m) {
            int j = __builtin_ctz(m);
            m &= m - 1;
            if (bgzf_candidate(b, n, o + j)) mask |= 1ull << (16 * v + j);
        }
    }
    return mask;
}

// Block starts a tile keeps for bgzf_emit (a 16 KiB tile of blocks of average size holds one or
// two); a tile with more is scanned again there
constexpr int BGZF_SLOTS = 8;

// Candidates of each tile: their count to tile_cnt, and themselves, in order, to the tile's slots
// when they fit there
// A whole tile's candidates: its bytes are loaded four vectors a thread at once and kept in LDS,
// where the header test of each 0x1f byte reads them (bytes past the tile are read in place)
__device__ __forceinline__ uint64_t tile_candidates(const uint8_t* b, int64_t n, int64_t t0, uint8_t* tile) {
    const int64_t o0 = t0 + (int64_t)threadIdx.x * 64;
    uint4 q[4];
#pragma unroll
    for (int v = 0; v < 4; v++) q[v] = *(const uint4*)(b + o0 + 16 * v);  // buffers are 16-byte aligned at tile starts
#pragma unroll
    for (int v = 0; v < 4; v++) ((uint4*)tile)[threadIdx.x * 4 + v] = q[v];
    __syncthreads();
    const TileBytes a{b, tile, t0, BGZF_TILE};
    uint64_t mask = 0;
#pragma unroll
    for (int v = 0; v < 4; v++) {
        uint32_t m = mask_1f(q[v]);
        while (m) {
            const int j = __builtin_ctz(m);
            m &= m - 1;
            if (bgzf_candidate_at(a, n, o0 + 16 * v + j)) mask |= 1ull << (16 * v + j);
        }
    }
    return mask;
}

__global__ void __launch_bounds__(BGZF_NT) bgzf_count(const uint8_t* b, int64_t n, int64_t* tile_cnt,
                                                      int64_t* slots) {
    __shared__ int64_t sh[BGZF_NT / CBM_WAVE];
    __shared__ __align__(16) uint8_t tile[BGZF_TILE];
    const int64_t t0 = (int64_t)blockIdx.x * BGZF_TILE;
    int64_t o0 = t0 + (int64_t)threadIdx.x * 64;
    // a whole tile (all but the last, usually) through LDS; the last, partial one as before
    uint64_t m = t0 + BGZF_TILE <= n ? tile_candidates(b, n, t0, tile) : thread_candidates(b, n, o0);
    int64_t tot;
    int64_t k = block_exclusive_sum<BGZF_NT, int64_t>((int64_t)__popcll(m), sh, &tot);
    if (tot <= BGZF_SLOTS) {
        int64_t* slot = slots + (int64_t)blockIdx.x * BGZF_SLOTS;
        while (m) {
            const int j = __builtin_ctzll(m);
            m &= m - 1;
            slot[k++] = o0 + j;
        }
    }
    if (threadIdx.x == 0) tile_cnt[blockIdx.x] = tot;
}

// The candidates in order at cand[tile_base[t] ..], a thread per tile: copied from the tile's slots,
// or (a tile with more than its slots hold) scanned again by the whole workgroup. *total is the
// number of candidates (the base after the last tile).
__global__ void __launch_bounds__(BGZF_NT) bgzf_emit(const uint8_t* b, int64_t n, const int64_t* tile_base,
                                                     const int64_t* total, int64_t tiles, const int64_t* slots,
                                                     int64_t cap, int64_t* cand) {
    __shared__ int64_t sh[BGZF_NT / CBM_WAVE];
    __shared__ int64_t over[BGZF_NT];
    __shared__ int n_over;
    if (threadIdx.x == 0) n_over = 0;
    __syncthreads();
    const int64_t t = (int64_t)blockIdx.x * BGZF_NT + threadIdx.x;
    if (t < tiles) {
        const int64_t k0 = tile_base[t], c = (t + 1 < tiles ? tile_base[t + 1] : *total) - k0;
        if (c > BGZF_SLOTS) {
            over[atomicAdd(&n_over, 1)] = t;
        } else {
            for (int64_t j = 0; j < c && k0 + j < cap; j++) cand[k0 + j] = slots[t * BGZF_SLOTS + j];
        }
    }
    __syncthreads();
    const int no = n_over;
    for (int i = 0; i < no; i++) {
        const int64_t u = over[i];
        const int64_t o0 = u * BGZF_TILE + (int64_t)threadIdx.x * 64;
        uint64_t m = thread_candidates(b, n, o0);
        int64_t tot;
        int64_t k = tile_base[u] + block_exclusive_sum<BGZF_NT, int64_t>((int64_t)__popcll(m), sh, &tot);
        while (m) {
            const int j = __builtin_ctzll(m);
            m &= m - 1;
            if (k < cap) cand[k] = o0 + j;
            k++;
        }
        __syncthreads();  // sh is read by every thread before the next tile's sum writes it
    }
}

// ws layout (int64): [0] ncand (scan total) [1] bad links [2] corrupt flag [3..7] spare,
// then tile counts/bases [ntiles + 1], then scan scratch.
struct BgzfWs {
    int64_t* hdr;
    int64_t* tiles;
    void* scan_ws;
};

// Each chained candidate must end where the next begins (the usual case: no spurious headers).
__global__ void bgzf_links(const uint8_t* b, const int64_t* cand, const int64_t* ncand_d, int64_t cap,
                           int64_t* bad) {
    int64_t m = *ncand_d;
    if (m > cap) return;  // handled by the serial walk
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t local = 0;
    for (; i + 1 < m; i += stride) {
        int64_t c = cand[i];
        int64_t nxt = c + (int64_t)ld_u16(b, c + 16) + 1;
        if (nxt != cand[i + 1]) local++;
    }
    if (local) atomicAdd((unsigned long long*)bad, (unsigned long long)local);
}

// Decide the chain: the candidates themselves (fast path) or a serial walk from offset 0.
__global__ void bgzf_finish(const uint8_t* b, int64_t n, int eof, int64_t cap, int64_t* blk, const int64_t* hdr,
                            int64_t* result) {
    int64_t m = hdr[0], bad = hdr[1];
    int64_t status = CBM_SCAN_OK, nb = 0, used = 0;
    if (m == 0) {
        status = (eof && n) ? CBM_SCAN_TRUNCATED : CBM_SCAN_OK;
    } else if (m <= cap && bad == 0) {
        if (blk[0] != 0) status = CBM_SCAN_NO_BLOCK;
        else {
            nb = m;
            int64_t c = blk[m - 1];
            used = c + (int64_t)ld_u16(b, c + 16) + 1;
        }
    } else {
        // some header inside a block's data looks like a block: walk the true chain from 0
        if (!bgzf_candidate(b, n, 0)) status = CBM_SCAN_NO_BLOCK;
        else {
            int64_t o = 0;
            while (o < n && bgzf_candidate(b, n, o)) {
                if (nb < cap) blk[nb] = o;
                nb++;
                o += (int64_t)ld_u16(b, o + 16) + 1;
            }
            used = o;
            if (nb > cap) status = CBM_SCAN_CAP;
        }
    }
    if (status == CBM_SCAN_OK && eof && used != n) status = CBM_SCAN_TRUNCATED;
    if (status != CBM_SCAN_OK) nb = 0, used = 0;
    result[CBM_SCAN_NBLOCKS] = nb;
    result[CBM_SCAN_USED] = used;
    result[CBM_SCAN_STATUS] = status;
    result[CBM_SCAN_NCAND] = m;
    result[CBM_SCAN_TOTAL_OUT] = 0;
}

// Per block: sizes, trailer CRC and ISIZE, and the checks of bgzf.py:574-584.
__global__ void bgzf_fields(const uint8_t* b, const int64_t* blk, const int64_t* result, int32_t* bsize,
                            int32_t* isize, uint32_t* crc, int64_t* corrupt) {
    int64_t m = result[CBM_SCAN_NBLOCKS];
    int64_t local = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        int64_t o = blk[i];
        int32_t bs = (int32_t)ld_u16(b, o + 16) + 1;
        uint32_t is = ld_u32(b, o + bs - 4);
        bsize[i] = bs;
        crc[i] = ld_u32(b, o + bs - 8);
        if (is == 0) {
            // the empty block must hold the empty final fixed-Huffman stream (03 00)
            if (bs != 28 || b[o + 18] != 3 || b[o + 19] != 0) local++;
        } else if (is > 65536) {
            local++;
            is = 0;
        }
        isize[i] = (int32_t)is;
    }
    if (local) atomicAdd((unsigned long long*)corrupt, (unsigned long long)local);
}

__global__ void bgzf_total(const int64_t* corrupt, int64_t* result) {
    if (result[CBM_SCAN_STATUS] == CBM_SCAN_OK && *corrupt) {
        result[CBM_SCAN_STATUS] = CBM_SCAN_CORRUPT;
        result[CBM_SCAN_NBLOCKS] = 0;
        result[CBM_SCAN_USED] = 0;
        result[CBM_SCAN_TOTAL_OUT] = 0;
    }
}

struct Bgzf16Count {  // tile counts for the candidate scan
    const int64_t* p;
    __device__ int64_t operator()(int64_t i) const { return p[i]; }
};

}  // namespace cbm
