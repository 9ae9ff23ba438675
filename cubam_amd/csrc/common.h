// Shared device helpers of libcubam_amd (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/cubam_amd.h"

#define CBM_WAVE 64
// pointers whose address space the compiler keeps: a read through one is an LDS read or a global
// load, never a flat load (which waits on both counters)
#define CBM_LDS __attribute__((address_space(3)))
#define CBM_GLOBAL __attribute__((address_space(1)))

namespace cbm {

// ------------------------------------------------------------------ errors (host)
void set_error(const char* fmt, ...);
int check_launch(const char* what);

// ------------------------------------------------------------------ byte access
__device__ __forceinline__ uint32_t ld_u8(const uint8_t* b, int64_t o) { return b[o]; }
__device__ __forceinline__ uint32_t ld_u16(const uint8_t* b, int64_t o) {
    return (uint32_t)b[o] | ((uint32_t)b[o + 1] << 8);
}
__device__ __forceinline__ uint32_t ld_u32(const uint8_t* b, int64_t o) {
    return (uint32_t)b[o] | ((uint32_t)b[o + 1] << 8) | ((uint32_t)b[o + 2] << 16) | ((uint32_t)b[o + 3] << 24);
}

// ------------------------------------------------------------------ SAMv1 helpers
// reg2bin of SAMv1 section 5.3 (the reference's _kernels.py:235-243)
__device__ __forceinline__ int reg2bin(int beg, int end) {
    --end;
    if (beg >> 14 == end >> 14) return ((1 << 15) - 1) / 7 + (beg >> 14);
    if (beg >> 17 == end >> 17) return ((1 << 12) - 1) / 7 + (beg >> 17);
    if (beg >> 20 == end >> 20) return ((1 << 9) - 1) / 7 + (beg >> 20);
    if (beg >> 23 == end >> 23) return ((1 << 6) - 1) / 7 + (beg >> 23);
    if (beg >> 26 == end >> 26) return ((1 << 3) - 1) / 7 + (beg >> 26);
    return 0;
}

// CIGAR ops M D N = X consume the reference (_kernels.py:245-247)
__device__ __forceinline__ bool ref_consuming(uint32_t op) {
    return op == 0 || op == 2 || op == 3 || op == 7 || op == 8;
}

// A complete, structurally valid record starts at o: the test of the reference's
// is_record (_kernels.py:250-278): block_size in [33, 4 MiB] and inside the data,
// reference ids in [-1, n_ref), positions >= -1, l_read_name >= 2, flag < 4096,
// l_seq <= 4 Mi, the fields fit block_size, a NUL-terminated printable name without
// '@', CIGAR ops <= 8 and bin == reg2bin(pos, pos + max(rlen, 1)) for mapped positions.
// The test reads bytes through an accessor A (u8, u16, u32 at a data offset): GlobalBytes reads
// the data in place; the record walk reads a copy staged in LDS.
struct GlobalBytes {
    const uint8_t* b;
    __device__ __forceinline__ uint32_t u8(int64_t o) const { return b[o]; }
    __device__ __forceinline__ uint32_t u16(int64_t o) const { return ld_u16(b, o); }
    __device__ __forceinline__ uint32_t u32(int64_t o) const { return ld_u32(b, o); }
};

// the fixed-field part of the test; on success the fields the rest needs
struct RecHead {
    int32_t pos;
    uint32_t lrn, ncig;
};
template <class A>
__device__ __forceinline__ bool record_head(const A& a, int64_t n, int64_t o, int n_ref, RecHead& h) {
    if (o < 0 || o + 36 > n) return false;
    int32_t bs = (int32_t)a.u32(o);
    if (bs < 33 || bs > (1 << 22) || o + 4 + bs > n) return false;
    int32_t ref = (int32_t)a.u32(o + 4), pos = (int32_t)a.u32(o + 8);
    int32_t nref = (int32_t)a.u32(o + 24), npos = (int32_t)a.u32(o + 28);
    if (ref < -1 || ref >= n_ref || pos < -1 || nref < -1 || nref >= n_ref || npos < -1) return false;
    uint32_t lrn = a.u8(o + 12), ncig = a.u16(o + 16), flag = a.u16(o + 18);
    uint32_t lseq = a.u32(o + 20);
    if (lrn < 2 || flag >= 4096 || lseq > (1u << 22)) return false;
    if (32LL + lrn + 4LL * ncig + (lseq + 1) / 2 + lseq > bs) return false;
    if (a.u8(o + 36 + lrn - 1) != 0) return false;
    h.pos = pos;
    h.lrn = lrn;
    h.ncig = ncig;
    return true;
}

// bin == reg2bin(pos, pos + max(rlen, 1)) for a mapped position
template <class A>
__device__ __forceinline__ bool record_bin_ok(const A& a, int64_t o, int32_t pos, int64_t rlen) {
    if (pos < 0) return true;
    int end = pos + (int)(rlen > 0 ? rlen : 1);
    return a.u16(o + 14) == (uint32_t)reg2bin(pos, end);
}

template <class A>
__device__ inline bool is_record_at(const A& a, int64_t n, int64_t o, int n_ref) {
    RecHead h;
    if (!record_head(a, n, o, n_ref, h)) return false;
    int64_t nm = o + 36;
    // the name, four characters a read (a read may pass the record's end by up to 3 bytes, which
    // CBM_PAD keeps readable)
    const uint32_t nn = h.lrn - 1;
    for (uint32_t i = 0; i < nn; i += 4) {
        const uint32_t w = a.u32(nm + i), v = w ^ 0x40404040u;
        // bit 7 of each byte: below 33, above 126, or '@' (per-byte subtraction and addition
        // across the word: a borrow or carry reaches a byte only above a byte that fails the same
        // test, so some name byte fails if and only if a name byte's bit is set)
        const uint32_t bad = (((w - 0x21212121u) & ~w) | ((w + 0x01010101u) | w) | ((v - 0x01010101u) & ~v)) &
                             0x80808080u;
        const uint32_t keep = nn - i >= 4 ? 0xffffffffu : (1u << (8 * (nn - i))) - 1u;
        if (bad & keep) return false;
    }
    int64_t cg = nm + h.lrn, rlen = 0;
    for (uint32_t i = 0; i < h.ncig; i++) {
        uint32_t v = a.u32(cg + 4LL * i), op = v & 15;
        if (op > 8) return false;
        if (ref_consuming(op)) rlen += v >> 4;
    }
    return record_bin_ok(a, o, h.pos, rlen);
}

__device__ inline bool is_record(const uint8_t* b, int64_t n, int64_t o, int n_ref) {
    return is_record_at(GlobalBytes{b}, n, o, n_ref);
}

// ------------------------------------------------------------------ wave64 / workgroup scans
__device__ __forceinline__ int lane_id() { return __lane_id(); }

// a value all lanes hold alike, in scalar registers (the first active lane's)
__device__ __forceinline__ int64_t uniform64(int64_t x) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)x);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)x >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

template <typename T>
__device__ __forceinline__ T wave_inclusive_sum(T v) {
#pragma unroll
    for (int d = 1; d < CBM_WAVE; d <<= 1) {
        T u = __shfl_up(v, d, CBM_WAVE);
        if (lane_id() >= d) v += u;
    }
    return v;
}

// Wavefront shifts and scans on the data-parallel primitives (DPP: register moves between
// lanes, no LDS): lane i gets x of lane i - 1 (lane 0: fill)
__device__ __forceinline__ int wave_up1(int x, int fill) {
    return __builtin_amdgcn_update_dpp(fill, x, 0x138 /* wave_shr:1 */, 0xf, 0xf, false);
}
// sum (max) of x over lanes 0..i: steps of 1, 2, 4, 8 lanes within rows of 16 (a lane with no
// source adds 0), then row 15 into rows 1 and 3, then lane 31 into rows 2 and 3
__device__ __forceinline__ int wave_scan_sum(int v) {
    v += __builtin_amdgcn_update_dpp(0, v, 0x111 /* row_shr:1 */, 0xf, 0xf, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x112 /* row_shr:2 */, 0xf, 0xf, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x114 /* row_shr:4 */, 0xf, 0xf, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x118 /* row_shr:8 */, 0xf, 0xf, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x142 /* row_bcast:15 */, 0xa, 0xf, true);
    v += __builtin_amdgcn_update_dpp(0, v, 0x143 /* row_bcast:31 */, 0xc, 0xf, true);
    return v;
}
__device__ __forceinline__ int wave_scan_max(int v) {  // v >= 0
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, true));
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, true));
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, true));
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, true));
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, true));
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, true));
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int d = CBM_WAVE / 2; d; d >>= 1) v += __shfl_xor(v, d, CBM_WAVE);
    return v;
}

// Exclusive prefix sum over a workgroup of NT threads; *total gets the sum. `sh` holds NT/64 values.
template <int NT, typename T>
__device__ __forceinline__ T block_exclusive_sum(T v, T* sh, T* total) {
    constexpr int NW = NT / CBM_WAVE;
    T inc = wave_inclusive_sum(v);
    int w = threadIdx.x / CBM_WAVE;
    if (lane_id() == CBM_WAVE - 1) sh[w] = inc;
    __syncthreads();
    T base = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NW; i++) {
        T s = sh[i];
        if (i < w) base += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return base + inc - v;
}

// ------------------------------------------------------------------ device-wide exclusive scan
// out[i] = sum(in[0..i)), out[n] = total, for n given in device memory (n_d) with a host upper
// bound n_max that sizes the grid. Three launches: tile sums, a scan of the tile sums in one
// workgroup, and the tile scans with their bases. `In` maps an index to an int64 value.
constexpr int SCAN_NT = 256;
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_TILE = SCAN_NT * SCAN_ITEMS;

inline int64_t scan_tiles(int64_t n_max) { return (n_max + SCAN_TILE - 1) / SCAN_TILE; }
inline int64_t scan_workspace(int64_t n_max) { return (scan_tiles(n_max) + 2) * 8; }

template <typename In>
__global__ void __launch_bounds__(SCAN_NT) scan_tile_sums(In in, const int64_t* n_d, int64_t n_max,
                                                           int64_t* tile_sum) {
    __shared__ int64_t sh[SCAN_NT / CBM_WAVE];
    int64_t n = n_d ? *n_d : n_max;
    int64_t t = blockIdx.x;
    int64_t base = t * SCAN_TILE;
    if (base >= n && t > 0) return;
    int64_t s = 0;
#pragma unroll 4
    for (int k = 0; k < SCAN_ITEMS; k++) {
        int64_t i = base + (int64_t)k * SCAN_NT + threadIdx.x;
        if (i < n) s += in(i);
    }
    int64_t tot;
    block_exclusive_sum<SCAN_NT>(s, sh, &tot);
    if (threadIdx.x == 0) tile_sum[t] = tot;
}

__global__ void __launch_bounds__(1024) scan_tile_bases(int64_t* tile_sum, const int64_t* n_d, int64_t n_max,
                                                         int64_t* total_out);

template <typename In, typename Out>
__global__ void __launch_bounds__(SCAN_NT) scan_apply(In in, Out out, const int64_t* n_d, int64_t n_max,
                                                       const int64_t* tile_base) {
    __shared__ int64_t sh[SCAN_NT / CBM_WAVE];
    int64_t n = n_d ? *n_d : n_max;
    int64_t t = blockIdx.x;
    int64_t base = t * SCAN_TILE;
    if (base >= n && !(n == 0 && t == 0)) return;
    // each thread owns SCAN_ITEMS consecutive items
    int64_t v[SCAN_ITEMS];
    int64_t s = 0;
    int64_t first = base + (int64_t)threadIdx.x * SCAN_ITEMS;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        int64_t i = first + k;
        v[k] = i < n ? in(i) : 0;
        s += v[k];
    }
    int64_t tot;
    int64_t run = block_exclusive_sum<SCAN_NT>(s, sh, &tot) + tile_base[t];
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        int64_t i = first + k;
        if (i < n) out(i, run);
        if (i == n) out(i, run);  // out[n] = total
        run += v[k];
    }
    if (base + SCAN_TILE == n && threadIdx.x == SCAN_NT - 1) out(n, run);
}

// Host driver. ws must hold scan_workspace(n_max) bytes; total_d (may be null) gets the sum.
template <typename In, typename Out>
int exclusive_scan(In in, Out out, const int64_t* n_d, int64_t n_max, void* ws, int64_t* total_d,
                   hipStream_t st) {
    int64_t tiles = scan_tiles(n_max);
    if (tiles == 0) tiles = 1;
    int64_t* tile_sum = (int64_t*)ws;
    hipLaunchKernelGGL(scan_tile_sums<In>, dim3((unsigned)tiles), dim3(SCAN_NT), 0, st, in, n_d, n_max, tile_sum);
    hipLaunchKernelGGL(scan_tile_bases, dim3(1), dim3(1024), 0, st, tile_sum, n_d, n_max, total_d);
    hipLaunchKernelGGL((scan_apply<In, Out>), dim3((unsigned)tiles), dim3(SCAN_NT), 0, st, in, out, n_d, n_max,
                       tile_sum);
    return check_launch("exclusive_scan");
}

// Common index maps
struct LoadI32 {
    const int32_t* p;
    __device__ int64_t operator()(int64_t i) const { return p[i]; }
};
struct LoadI64 {
    const int64_t* p;
    __device__ int64_t operator()(int64_t i) const { return p[i]; }
};
struct StoreI64 {
    int64_t* p;
    __device__ void operator()(int64_t i, int64_t v) const { p[i] = v; }
};

// Fill kernel (graph-safe replacement for hipMemsetAsync)
__global__ void fill_i64(int64_t* p, int64_t n, int64_t v);

}  // namespace cbm
