// libcubam_amd: the C ABI declared in include/cubam_amd.h, over the gfx950 kernels of
// bgzf_scan.h (kernel 1), inflate.h (kernel 2), records.h (kernels 3 and 4).
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <map>
#include <mutex>
#include <vector>

#include "bgzf_scan.h"
#include "common.h"
#include "deflate.h"
#include "inflate.h"
#include "inflate_wave.h"
#include "ondemand.h"
#include "records.h"

namespace cbm {

static thread_local char g_err[512];

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

int check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("%s: %s", what, hipGetErrorString(e));
        return CBM_ERR_LAUNCH;
    }
    return CBM_OK;
}

__global__ void __launch_bounds__(1024) scan_tile_bases(int64_t* tile_sum, const int64_t* n_d, int64_t n_max,
                                                         int64_t* total_out) {
    __shared__ int64_t sh[1024 / CBM_WAVE];
    int64_t n = n_d ? *n_d : n_max;
    int64_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (tiles == 0) tiles = 1;
    // each thread owns a run of consecutive tiles
    int64_t per = (tiles + 1023) / 1024;
    int64_t lo = threadIdx.x * per, hi = lo + per < tiles ? lo + per : tiles;
    int64_t s = 0;
    for (int64_t t = lo; t < hi; t++) s += tile_sum[t];
    int64_t tot;
    int64_t run = block_exclusive_sum<1024, int64_t>(s, sh, &tot);
    for (int64_t t = lo; t < hi; t++) {
        int64_t v = tile_sum[t];
        tile_sum[t] = run;
        run += v;
    }
    if (threadIdx.x == 0 && total_out) *total_out = tot;
}

__global__ void fill_i64(int64_t* p, int64_t n, int64_t v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

static int fill(int64_t* p, int64_t n, int64_t v, hipStream_t st) {
    if (n <= 0) return CBM_OK;
    int64_t g = (n + 255) / 256;
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(fill_i64, dim3((unsigned)g), dim3(256), 0, st, p, n, v);
    return check_launch("fill");
}

static int device_cus() {
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        (void)hipGetDevice(&dev);
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess) cus = prop.multiProcessorCount;
        if (cus <= 0) cus = 256;
    }
    return cus;
}

// ------------------------------------------------------------------ record chain walk of one segment
__global__ void rec_walk_one(const uint8_t* b, int64_t start, int64_t n, int n_ref, RecWs w, int64_t k,
                             int64_t entry) {
    int64_t lo = start + k * REC_SEG, hi = lo + REC_SEG < n ? lo + REC_SEG : n;
    (void)lo;
    WalkOut r = walk(b, n, entry, hi, n_ref);
    w.f[k] = entry;
    w.nrel[k] = -1;  // its starts are not recorded: rec_emit walks it
    w.e[k] = r.e;
    w.cnt[k] = r.cnt;
    w.cig[k] = r.cig;
    w.seqb[k] = r.seqb;
    w.qual[k] = r.qual;
    w.flag[k] = r.flag;
    w.pred[k] = 1;
}

__global__ void rec_zero_dropped(RecWs w) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < w.nseg && w.f[k] < 0) w.cnt[k] = 0;
}

__global__ void rec_pred_fill(RecWs w) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < w.nseg; k += (int64_t)gridDim.x * blockDim.x)
        w.pred[k] = w.f[k] >= 0 ? 1 : 0;
}

// Offsets of data[start, n) that pass is_record, counted (a thread an offset, grid-stride). The
// structural test reads at most 3 bytes past a record that fits, which CBM_PAD keeps readable.
__global__ void __launch_bounds__(256) rec_candidates_count(const uint8_t* b, int64_t start, int64_t n, int n_ref,
                                                            unsigned long long* count) {
    unsigned long long c = 0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t o = start + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; o < n; o += stride)
        c += is_record(b, n, o, n_ref) ? 1ull : 0ull;
    c = wave_sum(c);
    if (lane_id() == 0 && c) atomicAdd(count, c);
}

// Host side of the chain repair: walk each span's chain from its start, re-walking every
// segment whose speculative entry is not the true exit of the walk before it.
static int repair_load(const RecWs& w, std::vector<int64_t>& f, std::vector<int64_t>& e, std::vector<int64_t>& flag,
                       std::vector<char>& keep, hipStream_t st) {
    int64_t ns = w.nseg;
    f.assign(ns, 0);
    e.assign(ns, 0);
    flag.assign(ns, 0);
    keep.assign(ns, 0);
    if (hipMemcpyAsync(f.data(), w.f, ns * 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(e.data(), w.e, ns * 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(flag.data(), w.flag, ns * 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        set_error("record chain repair: copy failed");
        return CBM_ERR_LAUNCH;
    }
    return CBM_OK;
}

static int repair_span(const uint8_t* data, int64_t s0, int64_t end, int64_t k0, int32_t n_ref, const RecWs& w,
                       std::vector<int64_t>& f, std::vector<int64_t>& e, std::vector<int64_t>& flag,
                       std::vector<char>& keep, hipStream_t st) {
    int64_t p = s0;
    while (p < end) {
        int64_t k = k0 + (p - s0) / REC_SEG;
        if (f[k] != p) {
            // rec_walk_one places segment k at start + k * REC_SEG
            hipLaunchKernelGGL(rec_walk_one, dim3(1), dim3(1), 0, st, data, s0 - k0 * REC_SEG, end, n_ref, w, k, p);
            int rc = check_launch("rec_walk_one");
            if (rc) return rc;
            if (hipMemcpyAsync(&e[k], w.e + k, 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipMemcpyAsync(&flag[k], w.flag + k, 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess) {
                set_error("record chain repair: copy failed");
                return CBM_ERR_LAUNCH;
            }
            f[k] = p;
        }
        keep[k] = 1;
        if (flag[k]) break;
        p = e[k];
    }
    return CBM_OK;
}

// walks off the chains no longer count
static int repair_store(const RecWs& w, std::vector<int64_t>& f, const std::vector<char>& keep, hipStream_t st) {
    for (int64_t k = 0; k < w.nseg; k++)
        if (!keep[k] && f[k] >= 0) f[k] = -1;
    if (hipMemcpyAsync(w.f, f.data(), w.nseg * 8, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        set_error("record chain repair: copy failed");
        return CBM_ERR_LAUNCH;
    }
    return CBM_OK;
}

// workgroups of the persistent inflate launch: the LDS of a CU holds two block buffers
static int64_t inflate_max_groups() { return (int64_t)device_cus() * 2; }

// one launch of inflate_wave<NT, SEG, TOK, FUSE>: a persistent grid taking blocks from next_block
template <int NT, int SEG, bool TOK, bool FUSE = false>
static int launch_inflate_wave(const uint8_t* src, const int64_t* blk_off, const int32_t* bsize, const int32_t* isize,
                               const uint32_t* crc_expect, const int64_t* out_off, int64_t n_blocks, uint8_t* dst,
                               uint32_t* crc, int32_t* status, int64_t* result, int64_t groups, int64_t max_groups,
                               int64_t* next_block, uint64_t* mlists, int64_t* stats, const int64_t* idx,
                               const int64_t* n_dev, uint64_t* pool, int32_t* mcount, hipStream_t st,
                               int64_t mcap = MCAP_TOK, int64_t kmax = INT64_MAX) {
    const size_t lds = sizeof(WaveSmem<NT, SEG, TOK>);
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute((const void*)inflate_wave<NT, SEG, TOK, FUSE>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        attr = true;
    }
    // workgroups that fit a CU's 160 KiB of LDS (allocated in 512-byte granules); the decode phase
    // asks for four wavefronts a SIMD, so at most 16 of its one-wavefront workgroups run on a CU
    size_t per_cu = (160 * 1024) / ((lds + 511) / 512 * 512);
    if (TOK && per_cu > 16) per_cu = 16;
    int64_t g = groups > 0 ? groups : (int64_t)device_cus() * (int64_t)(per_cu ? per_cu : 1);
    if (g > max_groups) g = max_groups;  // the workspace holds that many match lists (one-phase)
    if (g > n_blocks) g = n_blocks;
    if (g < 1) g = 1;
    hipLaunchKernelGGL((inflate_wave<NT, SEG, TOK, FUSE>), dim3((unsigned)g), dim3(NT), lds, st, src, blk_off, bsize, isize,
                       crc_expect, out_off, n_blocks, dst, crc, status, result, (unsigned long long*)next_block,
                       mlists, (unsigned long long*)stats, idx, n_dev, pool, mcount, mcap, kmax);
    return check_launch("inflate_wave");
}

template <int NT, int SEG>
static int launch_wave(const uint8_t* src, const int64_t* blk_off, const int32_t* bsize, const int32_t* isize,
                       const uint32_t* crc_expect, const int64_t* out_off, int64_t n_blocks, uint8_t* dst,
                       uint32_t* crc, int32_t* status, int64_t* result, int64_t groups, int64_t max_groups,
                       int64_t* counter, hipStream_t st) {
    uint64_t* mlists = (uint64_t*)((uint8_t*)counter + INFLATE_WS_HDR);
    return launch_inflate_wave<NT, SEG, false>(src, blk_off, bsize, isize, crc_expect, out_off, n_blocks, dst, crc,
                                               status, result, groups, max_groups, counter, mlists, counter + 8,
                                               nullptr, nullptr, nullptr, nullptr, st);
}

// bytes the two-phase inflate adds to the workspace for n blocks: the deferred-block list, the
// match counts and a match list of MCAP_TOK entries a block
static int64_t tok_ws_bytes(int64_t n) { return ((12 * n + 255) / 256) * 256 + n * (int64_t)MCAP_TOK * 8; }

// The two-phase inflate: the decode phase (a wavefront a block, SEG-bit segments, literals to
// dst, matches to the block's list), the resolve phase (the matches, in dst), the CRC32 of every
// block the decode phase finished, then the one-phase kernel over the blocks it left.
// ws: the one-phase workspace (counter, statistics, max_groups match lists), then tok_ws_bytes(n).
// the resolve phase: MPL matches a lane, SPAN-byte windows, WPG blocks a workgroup
template <int MPL, int SPAN, int WPG>
static int launch_resolve(uint8_t* dst, const int64_t* out_off, const int32_t* isize, const uint32_t* crc_expect,
                          uint32_t* crc, int32_t* status, int64_t* result, const uint64_t* pool, const int32_t* mcount,
                          int64_t n_blocks, hipStream_t st, const int64_t* idx = nullptr,
                          const int64_t* n_dev = nullptr, int64_t mcap = MCAP_TOK, int64_t kmax = INT64_MAX) {
    // the workgroups that fit a CU's LDS
    const int64_t per_cu = (160 * 1024) / (int64_t)((sizeof(ResSmem<SPAN, WPG>) + 511) / 512 * 512);
    int64_t g = (n_blocks + WPG - 1) / WPG;
    if (g > (int64_t)device_cus() * per_cu) g = (int64_t)device_cus() * per_cu;
    hipLaunchKernelGGL((inflate_resolve<MPL, SPAN, WPG>), dim3((unsigned)g), dim3(WPG * CBM_WAVE), 0, st, dst, out_off,
                       isize, crc_expect, crc, status, result, pool, mcount, n_blocks, idx, n_dev, mcap, kmax);
    return check_launch("inflate_resolve");
}

// the resolve phase with a history ring: RING bytes of LDS a wavefront, WPG blocks a workgroup
template <int RING, int WPG>
static int launch_resolve_ring(uint8_t* dst, const int64_t* out_off, const int32_t* isize, const uint32_t* crc_expect,
                               uint32_t* crc, int32_t* status, int64_t* result, const uint64_t* pool,
                               const int32_t* mcount, int64_t n_blocks, hipStream_t st) {
    using K = RingSmem<2048, WPG, RING>;
    const int64_t per_cu = (160 * 1024) / (int64_t)((sizeof(K) + 511) / 512 * 512);
    int64_t g = (n_blocks + WPG - 1) / WPG;
    if (g > (int64_t)device_cus() * per_cu) g = (int64_t)device_cus() * per_cu;
    hipLaunchKernelGGL((inflate_resolve_ring<2048, WPG, RING>), dim3((unsigned)g), dim3(WPG * CBM_WAVE), 0, st, dst,
                       out_off, isize, crc_expect, crc, status, result, pool, mcount, n_blocks);
    return check_launch("inflate_resolve_ring");
}

// The decode and resolve phases of blocks [0, n_blocks) of the arrays passed (a part of a call's
// blocks: the pointers start at the part's first block), on st; next_block counts the decode
// phase's blocks.
template <int SEG, bool FUSE>
static int tok_part(const uint8_t* src, const int64_t* blk_off, const int32_t* bsize, const int32_t* isize,
                    const uint32_t* crc_expect, const int64_t* out_off, int64_t n_blocks, uint8_t* dst,
                    uint32_t* crc, int32_t* status, int64_t* result, int64_t groups, int res_mpl, int res_wpg,
                    int64_t* next_block, int64_t* stats, uint64_t* pool, int32_t* mcount, hipStream_t st) {
    int rc = launch_inflate_wave<64, SEG, true, FUSE>(src, blk_off, bsize, isize, crc_expect, out_off, n_blocks, dst,
                                                      crc, status, result, groups, (int64_t)1 << 40, next_block,
                                                      nullptr, stats, nullptr, nullptr, pool, mcount, st);
    if (rc) return rc;
    switch (FUSE ? 0 : res_mpl * 100 + res_wpg) {
        case 0:  // resolved by the decode launch
            break;
#define RES_CASE(MPL_, WPG_)                                                                                     \
    case MPL_ * 100 + WPG_:                                                                                      \
        rc = launch_resolve<MPL_, 2048 * MPL_, WPG_>(dst, out_off, isize, crc_expect, crc, status, result, pool, \
                                                     mcount, n_blocks, st);                                      \
        break;
        RES_CASE(1, 4)
        RES_CASE(1, 2)
        RES_CASE(1, 8)
        RES_CASE(2, 4)
        RES_CASE(2, 2)
#undef RES_CASE
#define RING_CASE(KIB_, WPG_)                                                                                   \
    case KIB_ * 100 + WPG_:                                                                                     \
        rc = launch_resolve_ring<KIB_ * 1024, WPG_>(dst, out_off, isize, crc_expect, crc, status, result, pool, \
                                                    mcount, n_blocks, st);                                      \
        break;
        RING_CASE(4, 1)
        RING_CASE(4, 2)
        RING_CASE(4, 4)
        RING_CASE(8, 1)
        RING_CASE(8, 2)
        RING_CASE(8, 4)
        RING_CASE(16, 1)
        RING_CASE(16, 2)
        RING_CASE(32, 1)
#undef RING_CASE
        default:
            set_error("cbm_inflate: two-phase resolve shape (blocks_per_wave, waves_per_group) not built");
            return CBM_ERR_ARG;
    }
    return rc;
}

// A side stream for each (device, calling stream), with the events that fork work onto it and
// join it back; created on first use and kept for the life of the process. Streams at the
// default priority.
struct SideStream {
    hipStream_t s = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    std::mutex mu;  // held from the fork to the join, so that calls on one stream pair their events
};
static std::mutex g_side_mu;
static std::map<std::pair<int, hipStream_t>, SideStream> g_side;

// *out = null when there is none yet and `create` is false (no stream is created while st captures
// a graph: work forked under capture goes to a side stream made by an earlier call on st)
static int side_stream(hipStream_t st, bool create, SideStream** out) {
    *out = nullptr;
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) {
        set_error("cbm_inflate: no current device");
        return CBM_ERR_LAUNCH;
    }
    std::lock_guard<std::mutex> lock(g_side_mu);
    const auto key = std::make_pair(dev, st);
    auto it = g_side.find(key);
    if (it != g_side.end() && it->second.s) {
        *out = &it->second;
        return CBM_OK;
    }
    if (!create) return CBM_OK;
    hipStream_t ns = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    if (hipStreamCreateWithFlags(&ns, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&join, hipEventDisableTiming) != hipSuccess) {
        if (join) (void)hipEventDestroy(join);
        if (fork) (void)hipEventDestroy(fork);
        if (ns) (void)hipStreamDestroy(ns);
        set_error("cbm_inflate: side stream or event creation failed");
        return CBM_ERR_LAUNCH;
    }
    SideStream& e = g_side[key];
    e.fork = fork;
    e.join = join;
    e.s = ns;
    *out = &e;
    return CBM_OK;
}

// Is st capturing a graph (or is its state unknown)?
static bool capturing(hipStream_t st) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone;
}

template <int SEG, bool FUSE = false>
static int launch_tok(const uint8_t* src, const int64_t* blk_off, const int32_t* bsize, const int32_t* isize,
                      const uint32_t* crc_expect, const int64_t* out_off, int64_t n_blocks, uint8_t* dst,
                      uint32_t* crc, int32_t* status, int64_t* result, int64_t groups, int res_mpl, int res_wpg,
                      void* ws, int64_t ws_groups, int64_t max_groups, bool split, hipStream_t st) {
    // [0] decode-phase next block (first part), [1] one-phase next block, [2] blocks the decode
    // phase left, [3] decode-phase next block of the second part, [4] next block of the second
    // pass over the blocks left, [5] blocks left after it
    int64_t* counter = (int64_t*)ws;
    uint64_t* mlists = (uint64_t*)((uint8_t*)ws + INFLATE_WS_HDR);
    uint8_t* tok = (uint8_t*)mlists + ws_groups * (int64_t)MLIST_CAP * 8;
    int64_t* idx = (int64_t*)tok;
    int32_t* mcount = (int32_t*)(idx + n_blocks);
    uint64_t* pool = (uint64_t*)(tok + ((12 * n_blocks + 255) / 256) * 256);
    int rc;
    SideStream* side = nullptr;
    if (split && !FUSE && n_blocks >= 2 && (rc = side_stream(st, !capturing(st), &side))) return rc;
    if (!side) {
        rc = tok_part<SEG, FUSE>(src, blk_off, bsize, isize, crc_expect, out_off, n_blocks, dst, crc, status, result,
                                 groups, res_mpl, res_wpg, counter, counter + 8, pool, mcount, st);
        if (rc) return rc;
    } else {
        const int64_t cut = n_blocks * 3 / 5 > 0 ? n_blocks * 3 / 5 : 1;
        std::lock_guard<std::mutex> lock(side->mu);
        if (hipEventRecord(side->fork, st) != hipSuccess || hipStreamWaitEvent(side->s, side->fork, 0) != hipSuccess) {
            set_error("cbm_inflate: fork to the side stream failed");
            return CBM_ERR_LAUNCH;
        }
        rc = tok_part<SEG, FUSE>(src, blk_off, bsize, isize, crc_expect, out_off, cut, dst, crc, status, result,
                                 groups, res_mpl, res_wpg, counter, counter + 8, pool, mcount, st);
        // the second part's launches go to the side stream even when the first part's failed, so
        // that the join below always follows them
        const int rc2 = tok_part<SEG, FUSE>(src, blk_off + cut, bsize + cut, isize + cut, crc_expect + cut,
                                            out_off + cut, n_blocks - cut, dst, crc + cut, status + cut, result,
                                            groups, res_mpl, res_wpg, counter + 3, counter + 8,
                                            pool + cut * (int64_t)MCAP_TOK, mcount + cut, side->s);
        if (hipEventRecord(side->join, side->s) != hipSuccess || hipStreamWaitEvent(st, side->join, 0) != hipSuccess) {
            set_error("cbm_inflate: join from the side stream failed");
            return CBM_ERR_LAUNCH;
        }
        if (rc || rc2) return rc ? rc : rc2;
    }
    int64_t lg = (n_blocks + 255) / 256;
    if (lg > 1024) lg = 1024;
    hipLaunchKernelGGL(inflate_defer_list, dim3((unsigned)lg), dim3(256), 0, st, status, n_blocks, idx, counter + 2);
    if ((rc = check_launch("inflate_defer_list"))) return rc;
    // the deferred blocks once more through both phases, each with a list of MLIST_CAP entries (as
    // many matches as a block can hold) in the pool the first pass is done with: list kb for the
    // kb-th deferred block, as many as the pool holds
    const int64_t kmax2 = n_blocks * (int64_t)MCAP_TOK / MLIST_CAP;
    if ((rc = launch_inflate_wave<64, SEG, true>(src, blk_off, bsize, isize, crc_expect, out_off, n_blocks, dst, crc,
                                                 status, result, groups, (int64_t)1 << 40, counter + 4, nullptr,
                                                 counter + 8, idx, counter + 2, pool, mcount, st, MLIST_CAP, kmax2)))
        return rc;
    if ((rc = launch_resolve<1, 2048, 4>(dst, out_off, isize, crc_expect, crc, status, result, pool, mcount,
                                         n_blocks, st, idx, counter + 2, MLIST_CAP, kmax2)))
        return rc;
    // what is left (blocks that fail a check, and any past the pool) to the one-phase kernel
    hipLaunchKernelGGL(inflate_defer_list, dim3((unsigned)lg), dim3(256), 0, st, status, n_blocks, idx, counter + 5);
    if ((rc = check_launch("inflate_defer_list"))) return rc;
    return launch_inflate_wave<512, 128, false>(src, blk_off, bsize, isize, crc_expect, out_off, n_blocks, dst, crc,
                                                status, result, 0, max_groups, counter + 1, mlists, counter + 8, idx,
                                                counter + 5, nullptr, nullptr, st);
}

// The chain's totals and per-segment bases: exclusive scans over the segments of the record
// counts (w.base), n_cigar_op sums, base-byte sums and l_seq sums (walks off the chain count 0),
// their totals into the result slots.
static int chain_totals(const RecWs& w, int64_t* result, hipStream_t st) {
    const int64_t* none = nullptr;
    int rc;
    if ((rc = exclusive_scan(SegVal{w.f, w.cnt}, StoreI64{w.base}, none, w.nseg, w.scan_ws, result + CBM_REC_N, st)))
        return rc;
    if ((rc = exclusive_scan(SegVal{w.f, w.cig}, StoreI64{w.cig_base}, none, w.nseg, w.scan_ws,
                             result + CBM_REC_CIGAR, st)))
        return rc;
    if ((rc = exclusive_scan(SegVal{w.f, w.seqb}, StoreI64{w.seq_base}, none, w.nseg, w.scan_ws,
                             result + CBM_REC_SEQ, st)))
        return rc;
    return exclusive_scan(SegVal{w.f, w.qual}, StoreI64{w.qual_base}, none, w.nseg, w.scan_ws,
                          result + CBM_REC_QUAL, st);
}

// the offsets of the first n_rec records of a chain and, for non-null columns, their CSR offsets
static int chain_emit(const uint8_t* data, int64_t start, int64_t n, const RecWs& w, int64_t* offsets, int64_t n_rec,
                      CsrOut csr, hipStream_t st) {
    unsigned g = (unsigned)((w.nseg + EMIT_SEGS - 1) / EMIT_SEGS);
    hipLaunchKernelGGL(rec_emit, dim3(g), dim3(REC_NT), 0, st, data, start, n, w, offsets, n_rec, csr);
    int rc = check_launch("rec_emit");
    if (rc || !(csr.cig || csr.seq || csr.qual)) return rc;
    hipLaunchKernelGGL(rec_csr_totals, dim3(1), dim3(1), 0, st, w, n_rec, csr);
    return check_launch("rec_csr_totals");
}

}  // namespace cbm

using namespace cbm;

#define ARG(cond, msg)              \
    do {                            \
        if (!(cond)) {              \
            set_error("%s", msg);   \
            return CBM_ERR_ARG;     \
        }                           \
    } while (0)

extern "C" {

const char* cbm_version(void) { return "cubam_amd 0.1 (gfx950)"; }

int cbm_last_error(char* buf, int64_t len) {
    int n = (int)strlen(g_err);
    if (buf && len > 0) {
        int k = n < len - 1 ? n : (int)len - 1;
        memcpy(buf, g_err, k);
        buf[k] = 0;
    }
    return n;
}

// ------------------------------------------------------------------ kernel 1
static int64_t bgzf_max_blocks(int64_t n) { return n / 26 + 2; }

int64_t cbm_bgzf_scan_workspace(int64_t n) {
    int64_t tiles = (n + BGZF_TILE - 1) / BGZF_TILE + 1;
    int64_t m = bgzf_max_blocks(n);
    return 8 * (8 + tiles + 2) + 8 * BGZF_SLOTS * tiles + scan_workspace(m > tiles ? m : tiles) + 64;
}

int cbm_bgzf_scan(const uint8_t* buf, int64_t n, int eof, int64_t cap, int64_t* blk_off, int32_t* bsize,
                  int32_t* isize, uint32_t* crc, int64_t* out_off, int64_t* result, void* ws, int64_t ws_bytes,
                  void* stream) {
    hipStream_t st = (hipStream_t)stream;
    ARG(n >= 0 && cap >= 0 && result && ws, "cbm_bgzf_scan: bad arguments");
    ARG(ws_bytes >= cbm_bgzf_scan_workspace(n), "cbm_bgzf_scan: workspace too small");
    ARG(n == 0 || (buf && ((uintptr_t)buf & 15) == 0), "cbm_bgzf_scan: buf must be 16-byte aligned");
    ARG(cap == 0 || (blk_off && bsize && isize && crc && out_off), "cbm_bgzf_scan: null output");
    if (cap > bgzf_max_blocks(n)) cap = bgzf_max_blocks(n);  // a block has at least 26 bytes
    int64_t* hdr = (int64_t*)ws;
    int64_t tiles = (n + BGZF_TILE - 1) / BGZF_TILE;
    int64_t* tcnt = hdr + 8;
    int64_t* slots = tcnt + tiles + 2;
    void* sws = slots + BGZF_SLOTS * tiles;
    int rc = fill(hdr, 8, 0, st);
    if (rc) return rc;
    if (tiles > 0) {
        hipLaunchKernelGGL(bgzf_count, dim3((unsigned)tiles), dim3(BGZF_NT), 0, st, buf, n, tcnt, slots);
        if ((rc = check_launch("bgzf_count"))) return rc;
        if ((rc = exclusive_scan(LoadI64{tcnt}, StoreI64{tcnt}, (const int64_t*)nullptr, tiles, sws, hdr, st)))
            return rc;
        hipLaunchKernelGGL(bgzf_emit, dim3((unsigned)((tiles + BGZF_NT - 1) / BGZF_NT)), dim3(BGZF_NT), 0, st, buf, n,
                           tcnt, hdr, tiles, slots, cap, blk_off);
        if ((rc = check_launch("bgzf_emit"))) return rc;
        int64_t g = (cap + 255) / 256;
        if (g > 1024) g = 1024;
        if (g < 1) g = 1;
        hipLaunchKernelGGL(bgzf_links, dim3((unsigned)g), dim3(256), 0, st, buf, blk_off, hdr, cap, hdr + 1);
        if ((rc = check_launch("bgzf_links"))) return rc;
    }
    hipLaunchKernelGGL(bgzf_finish, dim3(1), dim3(1), 0, st, buf, n, eof, cap, blk_off, hdr, result);
    if ((rc = check_launch("bgzf_finish"))) return rc;
    if (cap > 0) {
        int64_t g = (cap + 255) / 256;
        if (g > 2048) g = 2048;
        hipLaunchKernelGGL(bgzf_fields, dim3((unsigned)g), dim3(256), 0, st, buf, blk_off, result, bsize, isize,
                           crc, hdr + 2);
        if ((rc = check_launch("bgzf_fields"))) return rc;
        if ((rc = exclusive_scan(LoadI32{isize}, StoreI64{out_off}, result + CBM_SCAN_NBLOCKS, cap, sws,
                                 result + CBM_SCAN_TOTAL_OUT, st)))
            return rc;
    }
    hipLaunchKernelGGL(bgzf_total, dim3(1), dim3(1), 0, st, hdr + 2, result);
    return check_launch("bgzf_total");
}

// ------------------------------------------------------------------ kernel 2
int64_t cbm_inflate_workspace(int64_t n_blocks) {
    if (n_blocks < 0) return -1;
    return INFLATE_WS_HDR + inflate_max_groups() * (int64_t)MLIST_CAP * 8 + tok_ws_bytes(n_blocks);
}

}  // extern "C"

static bool tok_variant(int v) { return v >= CBM_INFLATE_TOK64_S512 && v <= CBM_INFLATE_TOK64_S512_SPLIT; }

// cbm_inflate's work once the workspace is settled: ws holds INFLATE_WS_HDR bytes and max_groups
// match lists (unused by the lane variant), and with `tok` the two-phase inflate's tok_ws_bytes
// after inflate_max_groups() lists; without it a two-phase variant runs the one-phase default
static int inflate_impl(const uint8_t* src, const int64_t* blk_off, const int32_t* bsize, const int32_t* isize,
                        const uint32_t* crc_expect, const int64_t* out_off, int64_t n_blocks, uint8_t* dst,
                        uint32_t* crc, int32_t* status, int64_t* result, const cbm_inflate_tune* tune, void* ws,
                        int64_t max_groups, bool tok, hipStream_t st) {
    int rc = fill(result, 1, 0, st);
    if (rc || n_blocks == 0) return rc;
    ARG(src && blk_off && bsize && isize && crc_expect && out_off && dst && crc && status,
        "cbm_inflate: null pointer");
    int variant = tune ? tune->variant
                       : n_blocks >= CBM_INFLATE_SPLIT_MIN ? CBM_INFLATE_TOK64_S512_SPLIT : CBM_INFLATE_DEFAULT;
    if (tok_variant(variant) && !tok) variant = CBM_INFLATE_WG512_S128;
    const int64_t groups = tune ? tune->groups : 0;
    // two-phase variants: blocks_per_wave = matches a lane of the resolve phase, waves_per_group =
    // its blocks a workgroup
    const int res_mpl = tune && tune->blocks_per_wave > 0 ? tune->blocks_per_wave : 1;
    // (two blocks a workgroup with the split: its parts' resolves share the CUs with decode launches)
    const int res_wpg = tune && tune->waves_per_group > 0 ? tune->waves_per_group
                        : variant == CBM_INFLATE_TOK64_S512_SPLIT ? 2
                                                                  : 4;
    if (variant != CBM_INFLATE_LANES) {
        int64_t* counter = (int64_t*)ws;  // [0] next block, [8..] decoder statistics
        if ((rc = fill(counter, 8 + STAT_N, 0, st))) return rc;
        switch (variant) {
#define WAVE_CASE(V, NT_, SEG_)                                                                               \
    case V:                                                                                                   \
        return launch_wave<NT_, SEG_>(src, blk_off, bsize, isize, crc_expect, out_off, n_blocks, dst, crc, status, \
                                      result, groups, max_groups, counter, st);
            WAVE_CASE(CBM_INFLATE_WG512_S128, 512, 128)
            WAVE_CASE(CBM_INFLATE_WG256, 256, 256)
            WAVE_CASE(CBM_INFLATE_WG128, 128, 256)
            WAVE_CASE(CBM_INFLATE_WG512, 512, 64)
            WAVE_CASE(CBM_INFLATE_WG64, 64, 512)
            WAVE_CASE(CBM_INFLATE_WG256_S128, 256, 128)
            WAVE_CASE(CBM_INFLATE_WG128_S128, 128, 128)
            WAVE_CASE(CBM_INFLATE_WG64_S128, 64, 128)
            WAVE_CASE(CBM_INFLATE_WG64_S256, 64, 256)
            WAVE_CASE(CBM_INFLATE_WG384_S128, 384, 128)
#undef WAVE_CASE
#define TOK_CASE(V, SEG_)                                                                                    \
    case V:                                                                                                  \
        return launch_tok<SEG_>(src, blk_off, bsize, isize, crc_expect, out_off, n_blocks, dst, crc, status, \
                                result, groups, res_mpl, res_wpg, ws, inflate_max_groups(), max_groups, false, st);
            TOK_CASE(CBM_INFLATE_TOK64_S512, 512)
            TOK_CASE(CBM_INFLATE_TOK64_S1024, 1024)
            TOK_CASE(CBM_INFLATE_TOK64_S256, 256)
#undef TOK_CASE
            case CBM_INFLATE_TOK64_S512_FUSED:
                return launch_tok<512, true>(src, blk_off, bsize, isize, crc_expect, out_off, n_blocks, dst, crc,
                                             status, result, groups, res_mpl, res_wpg, ws, inflate_max_groups(),
                                             max_groups, false, st);
            case CBM_INFLATE_TOK64_S512_SPLIT:
                return launch_tok<512>(src, blk_off, bsize, isize, crc_expect, out_off, n_blocks, dst, crc, status,
                                       result, groups, res_mpl, res_wpg, ws, inflate_max_groups(), max_groups, true,
                                       st);
            default: ARG(false, "cbm_inflate: unknown variant");
        }
    }
    int bpw = tune && tune->blocks_per_wave > 0 ? tune->blocks_per_wave : 4;
    int wpg = tune && tune->waves_per_group > 0 ? tune->waves_per_group : 2;
    ARG(bpw >= 1 && bpw <= 64 && wpg >= 1 && wpg <= 16, "cbm_inflate: bad tune");
    size_t lds = (sizeof(FixedTables) + 15) / 16 * 16 + (size_t)wpg * bpw * LANE_BYTES;
    ARG(lds <= 160 * 1024, "cbm_inflate: blocks_per_wave * waves_per_group needs more LDS than a CU has");
    int64_t per_group = (int64_t)bpw * wpg;
    int64_t g = groups > 0 ? groups : (n_blocks + per_group - 1) / per_group;
    const int64_t lane_groups = (int64_t)device_cus() * (int64_t)((160 * 1024) / lds);
    if (g > lane_groups) g = lane_groups;
    if (g < 1) g = 1;
    static bool attr_set = false;
    if (!attr_set) {
        (void)hipFuncSetAttribute((const void*)inflate_lanes, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        attr_set = true;
    }
    hipLaunchKernelGGL(inflate_lanes, dim3((unsigned)g), dim3(64 * wpg), lds, st, src, blk_off, bsize, isize,
                       out_off, n_blocks, dst, status, bpw);
    if ((rc = check_launch("inflate_lanes"))) return rc;
    int64_t cg = (n_blocks + 3) / 4;
    if (cg > (int64_t)device_cus() * 8) cg = (int64_t)device_cus() * 8;
    hipLaunchKernelGGL(crc_blocks, dim3((unsigned)cg), dim3(CRC_NT), 0, st, dst, out_off, isize, crc_expect,
                       n_blocks, crc, status, result);
    return check_launch("crc_blocks");
}

// ------------------------------------------------------------------ inflate without allocation
// The workspace of cbm_inflate_noalloc: the block counter and statistics header, then one match
// list for each of at most STATIC_INFLATE_GROUPS workgroups (89.7 MB). Each GPU holds its own copy
// for the life of the process. Its address in host code comes from hipGetSymbolAddress, once per
// device; the array's name itself is no device address on the host.
constexpr int64_t STATIC_INFLATE_GROUPS = 512;
constexpr int64_t STATIC_INFLATE_BYTES = INFLATE_WS_HDR + STATIC_INFLATE_GROUPS * (int64_t)MLIST_CAP * 8;
static __device__ __align__(256) uint8_t scratch[STATIC_INFLATE_BYTES];

constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[MAX_DEVICES];

// the static workspace of the current device (looked up on the first call on each device)
static int static_scratch(void** out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) {
        set_error("cbm_inflate_noalloc: no current device below %d", MAX_DEVICES);
        return CBM_ERR_LAUNCH;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("cbm_inflate_noalloc: hipGetSymbolAddress of its scratch: %s", hipGetErrorString(e));
            return CBM_ERR_LAUNCH;
        }
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return CBM_OK;
}

extern "C" {

int cbm_inflate(const uint8_t* src, const int64_t* blk_off, const int32_t* bsize, const int32_t* isize,
                const uint32_t* crc_expect, const int64_t* out_off, int64_t n_blocks, uint8_t* dst, uint32_t* crc,
                int32_t* status, int64_t* result, const cbm_inflate_tune* tune, void* ws, int64_t ws_bytes,
                void* stream) {
    ARG(n_blocks >= 0 && result, "cbm_inflate: bad arguments");
    const bool lanes = tune && tune->variant == CBM_INFLATE_LANES;
    ARG(lanes || n_blocks == 0 || (ws && ws_bytes >= cbm_inflate_workspace(n_blocks)),
        "cbm_inflate: workspace too small");
    return inflate_impl(src, blk_off, bsize, isize, crc_expect, out_off, n_blocks, dst, crc, status, result, tune, ws,
                        inflate_max_groups(), true, (hipStream_t)stream);
}

int cbm_inflate_noalloc(const uint8_t* src, const int64_t* blk_off, const int32_t* bsize, const int32_t* isize,
                        const uint32_t* crc_expect, const int64_t* out_off, int64_t n_blocks, uint8_t* dst,
                        uint32_t* crc, int32_t* status, int64_t* result, const cbm_inflate_tune* tune, void* stream) {
    ARG(n_blocks >= 0 && result, "cbm_inflate_noalloc: bad arguments");
    ARG(!tune || tune->groups <= STATIC_INFLATE_GROUPS,
        "cbm_inflate_noalloc: at most 512 workgroups (its static scratch holds 512 match lists)");
    void* ws = nullptr;
    const int rc = static_scratch(&ws);
    if (rc) return rc;
    const int64_t cap = inflate_max_groups() < STATIC_INFLATE_GROUPS ? inflate_max_groups() : STATIC_INFLATE_GROUPS;
    return inflate_impl(src, blk_off, bsize, isize, crc_expect, out_off, n_blocks, dst, crc, status, result, tune, ws,
                        cap, false, (hipStream_t)stream);
}

int cbm_crc32_spans(const uint8_t* data, const int64_t* off, const int64_t* len, int64_t n, uint32_t* out,
                    void* stream) {
    ARG(n >= 0, "cbm_crc32_spans: bad arguments");
    if (n == 0) return CBM_OK;
    ARG(data && off && len && out, "cbm_crc32_spans: null pointer");
    int64_t g = (n + 3) / 4;
    if (g > 65535) g = 65535;
    hipLaunchKernelGGL(crc_spans, dim3((unsigned)g), dim3(CRC_NT), 0, (hipStream_t)stream, data, off, len, n, out);
    return check_launch("crc_spans");
}

// ------------------------------------------------------------------ kernel 3
int64_t cbm_record_chain_workspace(int64_t n, int64_t* seg_bytes) {
    if (seg_bytes) *seg_bytes = REC_SEG;
    return rec_ws_bytes(n);
}

int cbm_record_chain(const uint8_t* data, int64_t start, int64_t n, int32_t n_ref, int64_t* result, void* ws,
                     int64_t ws_bytes, void* stream) {
    hipStream_t st = (hipStream_t)stream;
    ARG(start >= 0 && n >= start && result && ws, "cbm_record_chain: bad arguments");
    ARG(ws_bytes >= rec_ws_bytes(n - start), "cbm_record_chain: workspace too small");
    RecWs w = rec_ws_layout(ws, n - start);
    hipLaunchKernelGGL(rec_result_init, dim3(1), dim3(64), 0, st, result, start, w);
    int rc = check_launch("rec_result_init");
    if (rc || w.nseg == 0) return rc;
    ARG(data, "cbm_record_chain: null data");
    unsigned g = (unsigned)((w.nseg + REC_NT - 1) / REC_NT);
    const unsigned gw = (unsigned)((w.nseg + RW_WAVES - 1) / RW_WAVES);
    hipLaunchKernelGGL(rec_walk_wave<false>, dim3(gw), dim3(RW_WAVES * CBM_WAVE), 0, st, data, start, n, (int)n_ref,
                       SpanTab{}, w);
    if ((rc = check_launch("rec_walk_wave"))) return rc;
    hipLaunchKernelGGL(rec_links, dim3(g), dim3(REC_NT), 0, st, start, w);
    if ((rc = check_launch("rec_links"))) return rc;
    hipLaunchKernelGGL(rec_sum, dim3(g), dim3(REC_NT), 0, st, w, result);
    if ((rc = check_launch("rec_sum"))) return rc;
    hipLaunchKernelGGL(rec_result_links, dim3(1), dim3(1), 0, st, result, w);
    if ((rc = check_launch("rec_result_links"))) return rc;
    return chain_totals(w, result, st);
}

int cbm_record_chain_repair(const uint8_t* data, int64_t start, int64_t n, int32_t n_ref, int64_t* result,
                            void* ws, int64_t ws_bytes, void* stream) {
    hipStream_t st = (hipStream_t)stream;
    ARG(start >= 0 && n >= start && result && ws, "cbm_record_chain_repair: bad arguments");
    ARG(ws_bytes >= rec_ws_bytes(n - start), "cbm_record_chain_repair: workspace too small");
    RecWs w = rec_ws_layout(ws, n - start);
    int64_t ns = w.nseg;
    if (ns == 0) return CBM_OK;
    std::vector<int64_t> f, e, flag;
    std::vector<char> keep;
    int rc = repair_load(w, f, e, flag, keep, st);
    if (rc) return rc;
    if ((rc = repair_span(data, start, n, 0, n_ref, w, f, e, flag, keep, st))) return rc;
    if ((rc = repair_store(w, f, keep, st))) return rc;
    // counts of dropped walks are masked by f < 0 in rec_sum and rec_emit; recount
    hipLaunchKernelGGL(rec_result_init, dim3(1), dim3(64), 0, st, result, start, w);
    hipLaunchKernelGGL(rec_pred_fill, dim3(256), dim3(256), 0, st, w);
    unsigned g = (unsigned)((ns + REC_NT - 1) / REC_NT);
    hipLaunchKernelGGL(rec_sum, dim3(g), dim3(REC_NT), 0, st, w, result);
    if ((rc = check_launch("rec_sum"))) return rc;
    // cnt of dropped segments must not enter the scan
    hipLaunchKernelGGL(rec_zero_dropped, dim3(g), dim3(REC_NT), 0, st, w);
    if ((rc = check_launch("rec_zero_dropped"))) return rc;
    hipLaunchKernelGGL(rec_result_links, dim3(1), dim3(1), 0, st, result, w);
    if ((rc = check_launch("rec_result_links"))) return rc;
    if ((rc = chain_totals(w, result, st))) return rc;
    return hipStreamSynchronize(st) == hipSuccess ? CBM_OK : CBM_ERR_LAUNCH;
}

int cbm_record_candidates(const uint8_t* data, int64_t start, int64_t n, int32_t n_ref, int64_t* count,
                          void* stream) {
    hipStream_t st = (hipStream_t)stream;
    ARG(start >= 0 && n >= start && count, "cbm_record_candidates: bad arguments");
    int rc = fill(count, 1, 0, st);
    if (rc || n == start) return rc;
    ARG(data, "cbm_record_candidates: null data");
    int64_t g = (n - start + 255) / 256;
    const int64_t gmax = (int64_t)device_cus() * 16;
    if (g > gmax) g = gmax;
    hipLaunchKernelGGL(rec_candidates_count, dim3((unsigned)g), dim3(256), 0, st, data, start, n, (int)n_ref,
                       (unsigned long long*)count);
    return check_launch("rec_candidates_count");
}

int cbm_record_offsets(const uint8_t* data, int64_t start, int64_t n, const void* ws, int64_t ws_bytes,
                       int64_t* offsets, int64_t n_rec, void* stream) {
    return cbm_record_offsets_csr(data, start, n, ws, ws_bytes, offsets, nullptr, nullptr, nullptr, n_rec, stream);
}

int cbm_record_offsets_csr(const uint8_t* data, int64_t start, int64_t n, const void* ws, int64_t ws_bytes,
                           int64_t* offsets, int64_t* cigar_off, int64_t* seq_off, int64_t* qual_off, int64_t n_rec,
                           void* stream) {
    hipStream_t st = (hipStream_t)stream;
    ARG(ws && start >= 0 && n >= start && n_rec >= 0, "cbm_record_offsets: bad arguments");
    ARG(ws_bytes >= rec_ws_bytes(n - start), "cbm_record_offsets: workspace too small");
    const CsrOut csr{cigar_off, seq_off, qual_off};
    if (n_rec == 0) {  // the CSR columns hold one zero each
        int rc = CBM_OK;
        for (int64_t* c : {cigar_off, seq_off, qual_off})
            if (c && (rc = fill(c, 1, 0, st))) return rc;
        return rc;
    }
    ARG(data && offsets, "cbm_record_offsets: null pointer");
    return chain_emit(data, start, n, rec_ws_layout((void*)ws, n - start), offsets, n_rec, csr, st);
}

// ------------------------------------------------------------------ kernel 3 over many spans
int64_t cbm_record_spans_workspace(int64_t n_seg) { return n_seg < 0 ? -1 : rec_ws_bytes_seg(n_seg); }

static int spans_finish(SpanTab t, RecWs w, int64_t* result, int64_t* span_rec, hipStream_t st) {
    int rc = chain_totals(w, result, st);
    if (rc) return rc;
    unsigned g = (unsigned)((t.n + 1 + 255) / 256);
    hipLaunchKernelGGL(recs_bases, dim3(g), dim3(256), 0, st, t, w, span_rec);
    return check_launch("recs_bases");
}

int cbm_record_chain_spans(const uint8_t* data, const int64_t* span_start, const int64_t* span_end,
                           const int64_t* seg_base, const int32_t* span_nref, int64_t n_spans, int64_t n_seg,
                           int64_t* result, int64_t* span_rec, void* ws, int64_t ws_bytes, void* stream) {
    hipStream_t st = (hipStream_t)stream;
    ARG(n_spans >= 1 && n_seg >= 0 && result && span_rec && ws, "cbm_record_chain_spans: bad arguments");
    ARG(span_start && span_end && seg_base && span_nref, "cbm_record_chain_spans: null span table");
    ARG(ws_bytes >= rec_ws_bytes_seg(n_seg), "cbm_record_chain_spans: workspace too small");
    RecWs w = rec_ws_layout_seg(ws, n_seg);
    SpanTab t{span_start, span_end, seg_base, span_nref, n_spans};
    hipLaunchKernelGGL(rec_result_init, dim3(1), dim3(64), 0, st, result, (int64_t)0, w);
    int rc = check_launch("rec_result_init");
    if (rc) return rc;
    if (n_seg > 0) {
        ARG(data, "cbm_record_chain_spans: null data");
        unsigned g = (unsigned)((n_seg + REC_NT - 1) / REC_NT);
        const unsigned gw = (unsigned)((n_seg + RW_WAVES - 1) / RW_WAVES);
        hipLaunchKernelGGL(rec_walk_wave<true>, dim3(gw), dim3(RW_WAVES * CBM_WAVE), 0, st, data, (int64_t)0,
                           (int64_t)0, 0, t, w);
        if ((rc = check_launch("rec_walk_wave"))) return rc;
        hipLaunchKernelGGL(recs_links, dim3(g), dim3(REC_NT), 0, st, t, w);
        if ((rc = check_launch("recs_links"))) return rc;
        hipLaunchKernelGGL(recs_sum, dim3(g), dim3(REC_NT), 0, st, t, w, result);
        if ((rc = check_launch("recs_sum"))) return rc;
        hipLaunchKernelGGL(rec_result_links, dim3(1), dim3(1), 0, st, result, w);
        if ((rc = check_launch("rec_result_links"))) return rc;
    }
    return spans_finish(t, w, result, span_rec, st);
}

int cbm_record_chain_spans_repair(const uint8_t* data, const int64_t* span_start, const int64_t* span_end,
                                  const int64_t* seg_base, const int32_t* span_nref, int64_t n_spans, int64_t n_seg,
                                  int64_t* result, int64_t* span_rec, void* ws, int64_t ws_bytes, void* stream) {
    hipStream_t st = (hipStream_t)stream;
    ARG(n_spans >= 1 && n_seg >= 0 && result && span_rec && ws, "cbm_record_chain_spans_repair: bad arguments");
    ARG(span_start && span_end && seg_base && span_nref, "cbm_record_chain_spans_repair: null span table");
    ARG(ws_bytes >= rec_ws_bytes_seg(n_seg), "cbm_record_chain_spans_repair: workspace too small");
    if (n_seg == 0) return CBM_OK;
    RecWs w = rec_ws_layout_seg(ws, n_seg);
    SpanTab t{span_start, span_end, seg_base, span_nref, n_spans};
    std::vector<int64_t> s0(n_spans), s1(n_spans), kb(n_spans + 1);
    std::vector<int32_t> nr(n_spans);
    if (hipMemcpyAsync(s0.data(), span_start, n_spans * 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(s1.data(), span_end, n_spans * 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(kb.data(), seg_base, (n_spans + 1) * 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(nr.data(), span_nref, n_spans * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        set_error("cbm_record_chain_spans_repair: copy failed");
        return CBM_ERR_LAUNCH;
    }
    std::vector<int64_t> f, e, flag;
    std::vector<char> keep;
    int rc = repair_load(w, f, e, flag, keep, st);
    if (rc) return rc;
    for (int64_t s = 0; s < n_spans; s++)
        if ((rc = repair_span(data, s0[s], s1[s], kb[s], nr[s], w, f, e, flag, keep, st))) return rc;
    if ((rc = repair_store(w, f, keep, st))) return rc;
    hipLaunchKernelGGL(rec_result_init, dim3(1), dim3(64), 0, st, result, (int64_t)0, w);
    hipLaunchKernelGGL(rec_pred_fill, dim3(256), dim3(256), 0, st, w);
    unsigned g = (unsigned)((n_seg + REC_NT - 1) / REC_NT);
    hipLaunchKernelGGL(recs_sum, dim3(g), dim3(REC_NT), 0, st, t, w, result);
    if ((rc = check_launch("recs_sum"))) return rc;
    hipLaunchKernelGGL(rec_zero_dropped, dim3(g), dim3(REC_NT), 0, st, w);
    if ((rc = check_launch("rec_zero_dropped"))) return rc;
    hipLaunchKernelGGL(rec_result_links, dim3(1), dim3(1), 0, st, result, w);
    if ((rc = check_launch("rec_result_links"))) return rc;
    if ((rc = spans_finish(t, w, result, span_rec, st))) return rc;
    return hipStreamSynchronize(st) == hipSuccess ? CBM_OK : CBM_ERR_LAUNCH;
}

int cbm_record_offsets_spans(const uint8_t* data, int64_t n_seg, const void* ws, int64_t ws_bytes, int64_t* offsets,
                             int64_t n_rec, void* stream) {
    return cbm_record_offsets_spans_csr(data, n_seg, ws, ws_bytes, offsets, nullptr, nullptr, nullptr, n_rec, stream);
}

int cbm_record_offsets_spans_csr(const uint8_t* data, int64_t n_seg, const void* ws, int64_t ws_bytes,
                                 int64_t* offsets, int64_t* cigar_off, int64_t* seq_off, int64_t* qual_off,
                                 int64_t n_rec, void* stream) {
    hipStream_t st = (hipStream_t)stream;
    ARG(ws && n_seg >= 0 && n_rec >= 0, "cbm_record_offsets_spans: bad arguments");
    ARG(ws_bytes >= rec_ws_bytes_seg(n_seg), "cbm_record_offsets_spans: workspace too small");
    if (n_rec == 0 || n_seg == 0) {
        int rc = CBM_OK;
        for (int64_t* c : {cigar_off, seq_off, qual_off})
            if (c && (rc = fill(c, 1, 0, st))) return rc;
        return rc;
    }
    ARG(data && offsets, "cbm_record_offsets_spans: null pointer");
    return chain_emit(data, 0, 0, rec_ws_layout_seg((void*)ws, n_seg), offsets, n_rec,
                      CsrOut{cigar_off, seq_off, qual_off}, st);
}

// ------------------------------------------------------------------ on-demand record kernels
static unsigned od_grid(int64_t n) { return (unsigned)((n + OD_NT - 1) / OD_NT); }

int cbm_name_hash(const uint8_t* data, const int64_t* off, int64_t n, uint64_t* out, void* stream) {
    ARG(n >= 0, "cbm_name_hash: bad arguments");
    if (n == 0) return CBM_OK;
    ARG(data && off && out, "cbm_name_hash: null pointer");
    hipLaunchKernelGGL(od_name_hash, dim3(od_grid(n)), dim3(OD_NT), 0, (hipStream_t)stream, data, off, n, out);
    return check_launch("od_name_hash");
}

int cbm_hash_spans(const uint8_t* data, const int64_t* voff, const int32_t* vlen, int64_t n, uint64_t* out,
                   void* stream) {
    ARG(n >= 0, "cbm_hash_spans: bad arguments");
    if (n == 0) return CBM_OK;
    ARG(data && voff && vlen && out, "cbm_hash_spans: null pointer");
    hipLaunchKernelGGL(od_hash_spans, dim3(od_grid(n)), dim3(OD_NT), 0, (hipStream_t)stream, data, voff, vlen, n, out);
    return check_launch("od_hash_spans");
}

int cbm_pack_acgt(const uint8_t* data, const int64_t* voff, const int32_t* vlen, int64_t n, uint64_t* out,
                  uint8_t* ok, void* stream) {
    ARG(n >= 0, "cbm_pack_acgt: bad arguments");
    if (n == 0) return CBM_OK;
    ARG(data && voff && vlen && out && ok, "cbm_pack_acgt: null pointer");
    hipLaunchKernelGGL(od_pack_acgt, dim3(od_grid(n)), dim3(OD_NT), 0, (hipStream_t)stream, data, voff, vlen, n, out,
                       ok);
    return check_launch("od_pack_acgt");
}

int cbm_aligned_length(const uint8_t* data, const int64_t* off, int64_t n, int32_t* out, void* stream) {
    ARG(n >= 0, "cbm_aligned_length: bad arguments");
    if (n == 0) return CBM_OK;
    ARG(data && off && out, "cbm_aligned_length: null pointer");
    hipLaunchKernelGGL(od_aligned_len, dim3(od_grid(n)), dim3(OD_NT), 0, (hipStream_t)stream, data, off, n, out);
    return check_launch("od_aligned_len");
}

int cbm_blocks(const uint8_t* data, const int64_t* off, int64_t n, int split_deletions, int fill, const int64_t* first,
               int32_t* cnt, int64_t* rec, int32_t* start, int32_t* end, void* stream) {
    ARG(n >= 0, "cbm_blocks: bad arguments");
    if (n == 0) return CBM_OK;
    ARG(data && off && cnt && (!fill || (first && rec && start && end)), "cbm_blocks: null pointer");
    hipLaunchKernelGGL(od_blocks, dim3(od_grid(n)), dim3(OD_NT), 0, (hipStream_t)stream, data, off, n, split_deletions,
                       fill, first, cnt, rec, start, end);
    return check_launch("od_blocks");
}

int cbm_base_at(const uint8_t* data, const int64_t* off, const int64_t* rec, const int32_t* refpos, int64_t m,
                int8_t* out, void* stream) {
    ARG(m >= 0, "cbm_base_at: bad arguments");
    if (m == 0) return CBM_OK;
    ARG(data && off && rec && refpos && out, "cbm_base_at: null pointer");
    hipLaunchKernelGGL(od_base_at, dim3(od_grid(m)), dim3(OD_NT), 0, (hipStream_t)stream, data, off, rec, refpos, m,
                       out);
    return check_launch("od_base_at");
}

// ------------------------------------------------------------------ kernel 5: BGZF compression
static int64_t deflate_groups(int64_t n_blocks) {
    int64_t g = (n_blocks + DF_WAVES - 1) / DF_WAVES;
    int64_t cap = (int64_t)device_cus() * 2;  // the LDS of a CU holds two workgroups
    return g < cap ? (g < 1 ? 1 : g) : cap;
}

int64_t cbm_deflate_workspace(int64_t n_blocks) {
    if (n_blocks < 0) return -1;
    return deflate_groups(n_blocks) * DF_WAVES * (int64_t)65536 * 4;
}

int cbm_deflate(const uint8_t* src, const int64_t* in_off, const int32_t* in_len, int64_t n_blocks, uint8_t* slots,
                int32_t* bsize, uint32_t* crc, void* ws, int64_t ws_bytes, void* stream) {
    ARG(n_blocks >= 0, "cbm_deflate: bad arguments");
    if (n_blocks == 0) return CBM_OK;
    ARG(src && in_off && in_len && slots && bsize && ws, "cbm_deflate: null pointer");
    ARG(ws_bytes >= cbm_deflate_workspace(n_blocks), "cbm_deflate: workspace too small");
    const int64_t g = deflate_groups(n_blocks);
    hipLaunchKernelGGL(deflate_blocks, dim3((unsigned)g), dim3(DF_NT), 0, (hipStream_t)stream, src, in_off, in_len,
                       n_blocks, slots, bsize, crc, (uint32_t*)ws, (int64_t)65536);
    return check_launch("deflate_blocks");
}

int cbm_copy_spans(const uint8_t* src, const int64_t* src_off, const int64_t* len, const int64_t* dst_off, int64_t n,
                   uint8_t* dst, void* stream) {
    ARG(n >= 0, "cbm_copy_spans: bad arguments");
    if (n == 0) return CBM_OK;
    ARG(src && src_off && len && dst_off && dst, "cbm_copy_spans: null pointer");
    int64_t g = (n + 3) / 4;
    if (g > (int64_t)device_cus() * 16) g = (int64_t)device_cus() * 16;
    hipLaunchKernelGGL(copy_spans, dim3((unsigned)g), dim3(256), 0, (hipStream_t)stream, src, src_off, len, dst_off,
                       n, dst);
    return check_launch("copy_spans");
}

// ------------------------------------------------------------------ kernel 4
// workspace: n_cigar_op and l_seq of every record (int32 each), then the scan's tile sums
int64_t cbm_scatter_workspace(int64_t n_rec) {
    return n_rec < 0 ? -1 : ((8 * n_rec + 255) / 256) * 256 + scan_workspace(n_rec + 1) + 64;
}

int cbm_scatter(const uint8_t* data, const int64_t* offsets, int64_t n_rec, const cbm_columns* cols, void* ws,
                int64_t ws_bytes, void* stream) {
    hipStream_t st = (hipStream_t)stream;
    ARG(n_rec >= 0 && cols, "cbm_scatter: bad arguments");
    ARG(cols->n_tags >= 0 && cols->n_tags <= 4, "cbm_scatter: at most 4 tags");
    ARG(ws && ws_bytes >= cbm_scatter_workspace(n_rec), "cbm_scatter: workspace too small");
    int rc;
    cbm_columns c = *cols;
    int32_t* ncig = (int32_t*)ws;
    int32_t* lseq = ncig + n_rec;
    void* scan_ws = (uint8_t*)ws + ((8 * n_rec + 255) / 256) * 256;
    if (n_rec > 0) ARG(data && offsets, "cbm_scatter: null pointer");
    if (n_rec > 0 && (c.cigar_off || c.seq_off || c.qual_off)) {
        unsigned g = (unsigned)((n_rec + 255) / 256);
        hipLaunchKernelGGL(rec_sizes, dim3(g), dim3(256), 0, st, data, offsets, n_rec, ncig, lseq);
        if ((rc = check_launch("rec_sizes"))) return rc;
    }
    if (c.cigar_off &&
        (rc = exclusive_scan(RecNcig{ncig}, StoreI64{c.cigar_off}, (const int64_t*)nullptr, n_rec, scan_ws,
                             (int64_t*)nullptr, st)))
        return rc;
    if (c.seq_off &&
        (rc = exclusive_scan(RecSeqB{lseq}, StoreI64{c.seq_off}, (const int64_t*)nullptr, n_rec, scan_ws,
                             (int64_t*)nullptr, st)))
        return rc;
    if (c.qual_off &&
        (rc = exclusive_scan(RecQual{lseq}, StoreI64{c.qual_off}, (const int64_t*)nullptr, n_rec, scan_ws,
                             (int64_t*)nullptr, st)))
        return rc;
    if (n_rec > 0) {
        unsigned g = (unsigned)((n_rec + FIELDS_NT - 1) / FIELDS_NT);
        hipLaunchKernelGGL(rec_columns, dim3(g), dim3(FIELDS_THREADS), 0, st, data, offsets, n_rec, c);
        if ((rc = check_launch("rec_columns"))) return rc;
    }
    return CBM_OK;
}

int cbm_scatter_csr(const uint8_t* data, const int64_t* offsets, int64_t n_rec, const cbm_columns* cols,
                    void* stream) {
    ARG(n_rec >= 0 && cols, "cbm_scatter_csr: bad arguments");
    ARG(cols->n_tags >= 0 && cols->n_tags <= 4, "cbm_scatter_csr: at most 4 tags");
    ARG((!cols->cigar || cols->cigar_off) && (!cols->seq || cols->seq_off) && (!cols->qual || cols->qual_off),
        "cbm_scatter_csr: a CSR column needs its offsets");
    if (n_rec == 0) return CBM_OK;
    ARG(data && offsets, "cbm_scatter_csr: null pointer");
    unsigned g = (unsigned)((n_rec + FIELDS_NT - 1) / FIELDS_NT);
    hipLaunchKernelGGL(rec_columns, dim3(g), dim3(FIELDS_THREADS), 0, (hipStream_t)stream, data, offsets, n_rec,
                       *cols);
    return check_launch("rec_columns");
}

}  // extern "C"
