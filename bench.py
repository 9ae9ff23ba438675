"""Benchmark: compressed BAM bytes -> per-record columns on one MI355X.

Workload (default): the ATAC-like paired-end 2 x 50 bp stream of a 200 M-record
coordinate-sorted file (insert ~ lognormal(median 180, sigma 0.6), 2 % unmapped
mates, CB:Z from 10,000 barcodes with Zipf(1.1) weights, BGZF zlib level 6,
65,280-byte payloads), cut into 1 GiB compressed chunks. A step decodes one
whole chunk already resident in HBM: block scan, inflate + CRC32 check, record
chain and the column scatter (fixed fields, CIGAR/seq/qual CSR, CB/UB tags).

Prints one JSON line. `value` is inflated GB/s (U / step time); the roofline
entry is for the inflate kernel, with algorithmic bytes C + U per launch; the
CPU baseline is the oracle (zlib inflate + CRC check, record walk, columns) on
all host cores over a bounded sample of the same chunk. The `serving` entry
compares the served request loop (continuous batching, HBM budget, resident
results) with plain one-request-per-call decode on the same request traces:
closed-loop throughput and p50/p99 latency at swept Poisson arrival rates.
"""

from __future__ import annotations

import argparse
import faulthandler
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_chunk(chunk_bytes: int, seed: int = 1, n_total: int = 200_000_000):
    """The first `chunk_bytes` (or a little more) of the config's BGZF file, cut at a block end."""
    import numpy as np

    from cubam_amd import synth

    # records for about chunk_bytes of compressed data (ratio ~2.2, ~166 bytes a record)
    n_rec = int(chunk_bytes * 2.25 / 166) + 1000
    t = time.time()
    b = synth.make_bam(1, seed, n_rec, n_total=n_total, eof=False)
    sizes = b.block_sizes.astype(np.int64)
    ends = np.cumsum(sizes)
    k = int(np.searchsorted(ends, chunk_bytes, side="right"))
    k = max(1, min(k, len(sizes)))
    used = int(ends[k - 1])
    log(f"generated {b.n_records} records, {b.bgzf.size / 1e9:.2f} GB BGZF in {time.time() - t:.1f} s;"
        f" chunk = {k} blocks, {used / 2**30:.3f} GiB")
    return b, b.bgzf[:used], b.header


def gpu_step(buf, n, start, n_ref, ws, ev=None):
    """One pass of the hot path over the chunk in buf[0:n); returns (U, n_rec)."""
    import torch

    from cubam_amd import _decode

    blocks = _decode.scan_blocks(buf, n, False, ws)
    if ev is not None:
        ev[0].record()
    inf = _decode.inflate_blocks(blocks, None, ws)
    if ev is not None:
        ev[1].record()
    chain = _decode.record_chain(inf.data, start, inf.n, n_ref, ws, inf=inf)
    off, csr = _decode.record_offsets(inf.data, start, inf.n, chain, csr=True)
    cols = _decode.scatter(inf.data, off, chain, tags=("CB", "UB"), ws=ws, csr=csr)
    return inf, chain, off, cols


def h2d_time(chunk, buf, reps: int = 3) -> float:
    """Milliseconds to copy the chunk from pinned host memory into HBM (median of reps)."""
    import numpy as np
    import torch

    host = torch.empty(chunk.size, dtype=torch.uint8, pin_memory=True)
    host.numpy()[:] = chunk
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    times = []
    for _ in range(reps):
        e0.record()
        buf[: chunk.size].copy_(host, non_blocking=True)
        e1.record()
        e1.synchronize()
        times.append(e0.elapsed_time(e1))
    del host
    return float(np.median(times))


def copy_calibration(nbytes: int = 1 << 30, reps: int = 3) -> None:
    """A plain device-to-device copy of known size, for scaling the HBM byte counters."""
    import torch

    src = torch.ones(nbytes // 4, dtype=torch.int32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        torch.mul(src, 1, out=dst)  # an elementwise kernel (a plain copy_ may go to a DMA engine)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "reps": reps}), flush=True)


def column_bytes(chain) -> int:
    """K: bytes of the column schema written (offset 8 + fixed 35 + CSR + 2 tags x 21 per record)."""
    n = chain.n_rec
    return n * (8 + 35 + 3 * 8 + 2 * 21) + 3 * 8 + 4 * chain.n_cigar + chain.n_seq + chain.n_qual


def cpu_baseline(chunk, header_size: int, n_ref: int, cores: int | None = None, sample_mib: int = 256,
                 budget_s: float = 12.0):
    """The oracle over a bounded prefix of the chunk (whole blocks): (GB/s of U, threads, sample).

    zlib inflate with the CRC32 check and the column fill run on ``cores`` threads; the record
    walk is serial, as a BAM reader's is."""
    import numpy as np

    from oracle import oracle

    oracle.build()
    cores = cores or os.cpu_count() or 1
    sample = chunk[: min(chunk.size, sample_mib << 20)]
    reps, times, U, used = 0, [], 0, 0
    t_all = time.time()
    while reps < 3 and (reps == 0 or time.time() - t_all < budget_s):
        t = time.perf_counter()
        off, bs, isz, crc, used, status = oracle.bgzf_scan(sample, False)
        data, _, got, st = oracle.inflate(sample, off, bs, isz, n_threads=cores)
        assert status == 0 and (st == 0).all()
        roff, tail, rst = oracle.record_chain(data[header_size:], n_ref)
        oracle.columns(data[header_size:], roff, ("CB", "UB"), n_threads=cores)
        times.append(time.perf_counter() - t)
        U = int(data.size)
        reps += 1
    med = float(np.median(times))
    return U / med / 1e9, cores, f"first {used / 2**20:.0f} MiB compressed of the chunk ({U / 2**20:.0f} MiB" \
        f" inflated), median of {reps}"


def batched_bench(n_blocks: int, steps: int, seed: int = 2):
    """Config of heterogeneous blocks: every block in one persistent launch, captured in a hipGraph.

    Blocks: payloads uniform 1..65,280 bytes cut from an ATAC record stream, zlib
    level in {0, 1, 6, 9}, 10 % fixed Huffman. The autotuner picks the launch shape
    on a 50,000-block sample; the graph replays the whole batch.
    """
    import numpy as np
    import torch

    from cubam_amd import _decode, synth, tune

    t = time.time()
    hb = synth.hetero_blocks(n_blocks, seed)
    C, U = int(hb.bgzf.size), int(hb.raw.size)
    log(f"batched: {n_blocks} blocks, {C / 1e9:.2f} GB compressed, {U / 1e9:.2f} GB inflated,"
        f" generated in {time.time() - t:.1f} s")
    ws = _decode.Workspace()
    buf = _decode.to_device(hb.bgzf)
    blocks = _decode.scan_blocks(buf, C, True, ws)
    assert blocks.n_blocks == n_blocks
    best, table = tune.autotune(blocks, sample=50_000, log=log, ws=ws)
    default_ms = tune.time_inflate(blocks, None, ws=ws)
    g = tune.GraphInflate(blocks, best)
    g.run()
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    times = []
    for _ in range(steps):
        ev[0].record()
        g.run()
        ev[1].record()
        ev[1].synchronize()
        times.append(ev[0].elapsed_time(ev[1]))
    g.check()  # every block's size and CRC32 against its trailer (zlib's CRC of the payload)
    raw = _decode.to_device(hb.raw)
    exact = bool(torch.equal(g.data[:U], raw[:U]))
    if not exact:
        raise RuntimeError("batched inflate output differs from the payloads")
    ms = float(np.median(times))
    ach = (C + U) / (ms / 1e3) / 1e9
    del raw, g, buf
    return {
        "workload": f"{n_blocks} BGZF blocks, payload uniform 1-65,280 B from an ATAC record stream, zlib level"
                    " in {0,1,6,9}, 10% fixed Huffman; one persistent launch replayed from a hipGraph",
        "blocks": n_blocks,
        "compressed_bytes": C,
        "inflated_bytes": U,
        "ms_per_launch": round(ms, 3),
        "GBps": round(U / (ms / 1e3) / 1e9, 3),
        "bit_exact_bytes_and_crc32": exact,
        "autotune": {"best": list(best), "sample_blocks": 50_000, "candidates_ms": table},
        "default_shape_ms": round(default_ms, 3),
        "roofline": {"bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach / HBM_PEAK_GBS, 5), "traffic": None},
    }


def ingest_bench(bgzf_bytes, inflated: int, chunk_mib: int = 256):
    """read_bam end to end from a file in the page cache (pread into pinned buffers, host-to-device
    copy, decode, columns): with each chunk's copy overlapped with the decode of the chunk before
    it, and serial (copy, then decode, as the reference does). GB/s of inflated bytes."""
    import tempfile

    import torch

    from cubam_amd import bgzf, read_bam
    from cubam_amd.write import EOF_BLOCK

    out = {"file_bytes": int(bgzf_bytes.size) + len(EOF_BLOCK), "inflated_bytes": inflated, "chunk_MiB": chunk_mib}
    with tempfile.NamedTemporaryFile(dir="/tmp", suffix=".bam") as f:
        f.write(bgzf_bytes.tobytes())
        f.write(EOF_BLOCK)
        f.flush()
        for overlap in (True, False):
            bgzf.OVERLAP_UPLOAD = overlap
            best, recs = None, 0
            for _ in range(2):
                torch.cuda.synchronize()
                t = time.perf_counter()
                recs = sum(len(rb) for rb in read_bam(f.name, chunk_bytes=chunk_mib << 20))
                torch.cuda.synchronize()
                dt = time.perf_counter() - t
                best = dt if best is None else min(best, dt)
            out["overlapped" if overlap else "serial"] = {"s": round(best, 4), "GBps": round(inflated / best / 1e9, 3),
                                                          "records": recs}
        bgzf.OVERLAP_UPLOAD = True
    return out


def zipf_cells(codes, ok, n_cells: int = 1000, seed: int = 4):
    """n_cells barcodes (packed codes) drawn without replacement with probability proportional
    to their read counts, which follow the generator's Zipf(1.1) cell weights."""
    import numpy as np
    import torch

    u, cnt = torch.unique(codes.view(torch.int64)[ok], return_counts=True)
    u, cnt = u.cpu().numpy(), cnt.cpu().numpy().astype(np.float64)
    rng = np.random.default_rng(seed)
    pick = rng.choice(u.size, size=min(n_cells, u.size), replace=False, p=cnt / cnt.sum())
    return torch.from_numpy(u[pick]).cuda()


def write_bench(buf, n, start, n_ref, ws, header, reps: int = 3):
    """Write-back: the chunk's records of 1,000 Zipf-selected cells, gathered and BGZF-compressed on
    the GPU (level-1 matching, dynamic Huffman), checked block by block against zlib."""
    import zlib

    import numpy as np
    import torch

    from cubam_amd import _decode
    from cubam_amd.bam import RecordBatch
    from cubam_amd.serve import write_back

    inf, chain, off, cols = gpu_step(buf, n, start, n_ref, ws)
    rb = RecordBatch(inf.data, off, header, cols["fixed"], cols["tags"])
    codes, ok = rb.tag("CB").packed()
    cells = zipf_cells(codes, ok)
    wws = _decode.Workspace()
    res = write_back(rb, cells, ws=wws)  # warm
    torch.cuda.synchronize()
    times = []
    for _ in range(reps):
        t = time.perf_counter()
        res = write_back(rb, cells, ws=wws)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t)
    ms = float(np.median(times)) * 1e3
    # exactness: every block inflates (GPU) to the gathered records, with zlib's CRC32 of its slice
    from cubam_amd.bgzf import _inflate_blocks
    from cubam_amd.write import gather_records

    kept = torch.nonzero(ok & torch.isin(codes.view(torch.int64), cells)).flatten()
    payload = gather_records(rb.data, rb.offset[kept])
    blob = res.blocks
    bb = torch.empty(int(blob.numel()) + 64, dtype=torch.uint8, device=blob.device)
    bb[: blob.numel()].copy_(blob)
    data2, n2, used, inf2 = _inflate_blocks(bb, int(blob.numel()), True, None, _decode.Workspace())
    _decode.check_inflated(inf2)
    exact = n2 == payload.numel() and bool(torch.equal(data2[:n2], payload))
    host = payload.cpu().numpy()
    crc_gpu = inf2.crc.cpu().numpy()
    crc_ok = all(int(crc_gpu[i]) == zlib.crc32(host[a : a + 65280]) for i, a in enumerate(range(0, host.size, 65280)))
    # and zlib itself on a sample of the blocks
    raw = blob.cpu().numpy().tobytes()
    o, k, zl = 0, 0, True
    while o < len(raw) and k < 64:
        bs = (raw[o + 16] | (raw[o + 17] << 8)) + 1
        zl &= zlib.decompress(raw[o + 18 : o + bs - 8], -15) == host[65280 * k : 65280 * (k + 1)].tobytes()
        o += bs
        k += 1
    if not (exact and crc_ok and zl):
        raise RuntimeError("write-back blocks do not round-trip")
    z1 = sum(len(zlib.compress(host[a : a + 65280].tobytes(), 1)) for a in range(0, min(host.size, 64 << 20), 65280))
    sample = min(host.size, 64 << 20)
    ours = 0
    o = 0
    for a in range(0, sample, 65280):
        bs = (raw[o + 16] | (raw[o + 17] << 8)) + 1
        ours += bs - 26
        o += bs
    return {
        "workload": "records of 1,000 Zipf-selected cells (CB) of the 1 GiB ATAC chunk, gathered and BGZF-compressed"
                    " on the GPU (65,280-byte payloads, level-1 greedy matching, dynamic Huffman)",
        "cells": int(cells.numel()),
        "records": int(res.records),
        "payload_bytes": int(res.payload_bytes),
        "bgzf_bytes": int(blob.numel()),
        "ms": round(ms, 3),
        "payload_GBps": round(res.payload_bytes / ms / 1e6, 3),
        "deflate_over_zlib_level1": round(ours / max(z1, 1), 4),
        "bit_exact_payload_and_crc32": bool(exact and crc_ok and zl),
    }


SERVE_FILES = ((0, 30), (0, 31), (0, 32), (0, 33), (1, 34), (1, 35), (1, 36), (1, 37))


def serve_bench(gex_records: int, atac_records: int, n_requests: int, rates, budget_gb: float = 64.0,
                codec: str = "zlib"):
    """Served vs plain region decode over 8 files held in pinned host memory.

    Returns the JSON object of the serving comparison: closed-loop (every request
    queued at once) throughput of both modes, and p50/p99 latency at Poisson
    arrival rates given as multiples of the plain mode's closed-loop request rate.
    """
    import torch

    from cubam_amd import synth
    from cubam_amd.serve import BamSource, Server, make_trace, run_trace, summarize

    t = time.time()
    srcs = []
    for kind, seed in SERVE_FILES:
        n = gex_records if kind == 0 else atac_records
        b = synth.make_bam(kind, seed, n, n_total=n, codec=codec)
        srcs.append(BamSource(b.bgzf, name=f"{'GEX' if kind == 0 else 'ATAC'}-{seed}"))
        del b
        log(f"serving file {len(srcs)} of {len(SERVE_FILES)} ready ({time.time() - t:.0f} s)")
    gen_s = time.time() - t
    c_total = sum(s.size for s in srcs)
    u_total = sum(s.index.stream_bytes for s in srcs)
    log(f"serving: 8 files, {c_total / 1e9:.2f} GB compressed, {u_total / 1e9:.2f} GB inflated,"
        f" generated ({codec} level 6) and indexed in {gen_s:.1f} s")
    budget = int(budget_gb * 1e9)

    def run(mode, trace):
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
        with Server(srcs, budget_bytes=budget, mode=mode) as srv:
            reqs = run_trace(srv, trace)
            out = summarize(reqs)
            out["peak_device_GB"] = round(srv.peak_bytes / 1e9, 3)
        torch.cuda.synchronize()
        if out["errors"]:
            raise RuntimeError(f"{mode}: {out['errors']} requests failed")
        return out

    warm = make_trace(srcs, 24, 0.0, seed=99)
    run("plain", warm)
    run("served", warm)
    closed = make_trace(srcs, n_requests, 0.0, seed=1)
    plain = run("plain", closed)
    served = run("served", closed)
    log(f"closed loop: plain {plain}; served {served}")
    plain_rps = plain["requests"] / plain["wall_s"]
    sweep = []
    for i, m in enumerate(rates):
        tr = make_trace(srcs, n_requests, m * plain_rps, seed=10 + i)
        row = {"rate_rps": round(m * plain_rps, 2), "rate_x_plain_capacity": m,
               "plain": run("plain", tr), "served": run("served", tr)}
        log(f"rate {row['rate_rps']}/s: plain p50 {row['plain']['p50_ms']} p99 {row['plain']['p99_ms']} ms;"
            f" served p50 {row['served']['p50_ms']} p99 {row['served']['p99_ms']} ms")
        sweep.append(row)
    # mixed read/write at the plain mode's capacity: a fifth of the requests write the records of
    # 1,000 Zipf-selected cells back as BGZF blocks
    from cubam_amd.serve import decode_spans

    atac = next(i for i, (kind, _) in enumerate(SERVE_FILES) if kind == 1)
    rb = decode_spans([(srcs[atac], srcs[atac].index.span(0, srcs[atac].index.record_blocks().size))])[0]
    codes, ok = rb.tag("CB").packed()
    cells = zipf_cells(codes, ok)
    del rb, codes, ok
    tr = make_trace(srcs, n_requests, plain_rps, seed=20, write_fraction=0.2, cells=cells)
    mixed = {"rate_rps": round(plain_rps, 2), "write_fraction": 0.2, "plain": run("plain", tr),
             "served": run("served", tr)}
    log(f"mixed read/write: plain p50 {mixed['plain']['p50_ms']} p99 {mixed['plain']['p99_ms']} ms;"
        f" served p50 {mixed['served']['p50_ms']} p99 {mixed['served']['p99_ms']} ms")
    return {
        "mixed_read_write": mixed,
        "workload": (f"8 synthetic BAMs in pinned host memory (4 GEX-like x {gex_records} records, 4 ATAC-like x"
                     f" {atac_records}); region requests, compressed span log-uniform 64 KiB-256 MiB"),
        "compressed_bytes": c_total,
        "inflated_bytes": u_total,
        "compressor": f"{codec} level 6",
        "generate_and_index_s": round(gen_s, 1),
        "budget_GB": budget_gb,
        "requests_per_run": n_requests,
        "closed_loop": {"plain": plain, "served": served,
                        "served_over_plain_GBps": round(served["GBps"] / plain["GBps"], 3)},
        "sweep": sweep,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--chunk-mib", type=int, default=1024)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-write", action="store_true")
    ap.add_argument("--no-ingest", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 1 GiB device-to-device copy (scale for the FETCH_SIZE/WRITE_SIZE counters)")
    ap.add_argument("--batched-blocks", type=int, default=500_000)
    ap.add_argument("--serve-gex-records", type=int, default=2_000_000)
    ap.add_argument("--serve-atac-records", type=int, default=4_000_000)
    ap.add_argument("--serve-codec", choices=("zlib", "libdeflate"), default="zlib",
                    help="compressor of the serving files (libdeflate: same level, about 3x faster to generate)")
    ap.add_argument("--serve-requests", type=int, default=160)
    args = ap.parse_args()
    faulthandler.enable()  # a crash prints every thread's Python stack
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (use --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from cubam_amd import _decode, synth
    from cubam_amd._lib import load

    if args.calibrate_copy:
        copy_calibration()
        return

    load()
    synth.build()
    torch.cuda.set_device(0)
    chunk_bytes = args.chunk_mib << 20
    b, chunk, header = make_chunk(chunk_bytes)
    hdr = _decode_header(header)
    n_ref = len(hdr.names)
    ws = _decode.Workspace()
    buf = _decode.to_device(chunk)
    n = int(chunk.size)
    h2d_ms = h2d_time(chunk, buf)
    ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
    for _ in range(args.warmup):
        out = gpu_step(buf, n, hdr.size, n_ref, ws)
        del out
    torch.cuda.synchronize()
    step_ms, inflate_ms = [], []
    U = K = n_rec = 0
    t0 = time.perf_counter()
    for i in range(args.steps):
        s0 = time.perf_counter()
        inf, chain, off, cols = gpu_step(buf, n, hdr.size, n_ref, ws, ev)
        torch.cuda.synchronize()
        step_ms.append((time.perf_counter() - s0) * 1e3)
        inflate_ms.append(ev[0].elapsed_time(ev[1]))
        U, n_rec, K = inf.n, chain.n_rec, column_bytes(chain)
        n_blocks = int(inf.status.numel())
        del inf, chain, off, cols
        log(f"step {i}: {step_ms[-1]:.2f} ms (inflate {inflate_ms[-1]:.2f} ms)")
    total_s = time.perf_counter() - t0
    ms = total_s * 1e3 / args.steps
    gbs = U / (ms / 1e3) / 1e9
    inf_ms = float(np.mean(inflate_ms))
    achieved = (n + U) / (inf_ms / 1e3) / 1e9
    res = {
        "metric": "uncompressed GB/s BAM->columns on 1 GPU (plain per-chunk decode)",
        "value": round(gbs, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {
            "workload": "ATAC-like 2x50bp paired-end stream of a 200M-record BAM, first 1 GiB compressed chunk",
            "chunk_compressed_bytes": n,
            "inflated_bytes": U,
            "records": n_rec,
            "column_bytes": K,
            "algorithmic_bytes_C_U_K": n + U + K,
            "p50_step_ms": round(float(np.percentile(step_ms, 50)), 3),
            "p99_step_ms": round(float(np.percentile(step_ms, 99)), 3),
            "path_GBps_C_U_K": round((n + U + K) / (ms / 1e3) / 1e9, 2),
            "inflate_tune": list(inflate_tune_used(n_blocks, U)),
            "h2d_ms": round(h2d_ms, 3),
            "pcie_inclusive_GBps": round(U / ((ms + h2d_ms) / 1e3) / 1e9, 3),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": _kernel_name(n_blocks, U),
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": profiled_traffic(_kernel_name(n_blocks, U)),
            "traffic_source": "profiles/hbm_traffic_atac_1GiB_chunk.md (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE)",
            "algorithmic_bytes": n + U,
            "launch_ms": round(inf_ms, 3),
        },
    }
    del buf
    if not args.no_ingest:
        res["ingest"] = ingest_bench(b.bgzf, len(b.header) + b.raw_bytes)
        log(f"ingest: {res['ingest']}")
    if not args.no_write:
        wbuf = _decode.to_device(chunk)
        res["write_back"] = write_bench(wbuf, n, hdr.size, n_ref, ws, hdr)
        log(f"write-back: {res['write_back']}")
        del wbuf
    if not args.no_batched:
        res["batched"] = batched_bench(args.batched_blocks, max(3, min(args.steps, 10)))
        torch.cuda.synchronize()
        torch.cuda.empty_cache()  # the serving runs start from the same allocator state every time
    if not args.no_serve:
        res["serving"] = serve_bench(args.serve_gex_records, args.serve_atac_records, args.serve_requests,
                                     rates=(0.5, 1.0, 2.0), codec=args.serve_codec)
    if not args.no_cpu:
        v, cores, sample = cpu_baseline(chunk, hdr.size, n_ref)
        v1, _, sample1 = cpu_baseline(chunk, hdr.size, n_ref, cores=1, sample_mib=32)
        res["cpu_baseline"] = {"value": round(v, 3), "unit": "GB/s", "cores": cores, "kind": "port",
                               "sample": sample,
                               "single_thread": {"value": round(v1, 3), "unit": "GB/s", "cores": 1,
                                                 "sample": sample1}}
    print(json.dumps(res), flush=True)


def profiled_traffic(shape_label: str):
    """HBM bytes per launch of the inflate kernel from the committed PMC summary (FETCH_SIZE and
    WRITE_SIZE, scaled by a copy kernel), if it is for the same kernel shape; else None."""
    p = ROOT / "profiles" / "hbm_traffic_atac_1GiB_chunk.md"
    try:
        for line in p.read_text().splitlines():
            if line.startswith("traffic_per_launch_bytes "):
                name, value = line[len("traffic_per_launch_bytes "):].rsplit(" ", 1)
                if name in shape_label:
                    return int(value)
    except OSError:
        pass
    return None


def inflate_tune_used(n_blocks: int, inflated: int) -> tuple[int, int, int, int]:
    """(variant, blocks_per_wave, waves_per_group, groups) of the inflate of a batch this size: the
    process's or the autotuned shape, else the library default for the block count."""
    from cubam_amd import _decode
    from cubam_amd._lib import default_variant

    t = _decode.inflate_tune_for(n_blocks, inflated)
    if t is None:
        return (default_variant(n_blocks), 0, 0, 0)
    return (t.variant, t.blocks_per_wave, t.waves_per_group, t.groups)


def _kernel_name(n_blocks: int, inflated: int) -> str:
    from cubam_amd._lib import INFLATE_SHAPES, INFLATE_TOK64_S512_SPLIT, INFLATE_TWO_PHASE

    t = inflate_tune_used(n_blocks, inflated)
    v = t[0]
    shape = INFLATE_SHAPES[v]
    if shape is None:
        return "cbm_inflate: inflate_lanes + crc_blocks"
    if v in INFLATE_TWO_PHASE:
        mpl, wpg = t[1] or 1, t[2] or (2 if v == INFLATE_TOK64_S512_SPLIT else 4)
        split = " over two streams (blocks cut 3:2)" if v == INFLATE_TOK64_S512_SPLIT else ""
        return (f"cbm_inflate (two-phase{split}): inflate_wave<64, {shape[1]}, true> + "
                f"inflate_resolve<{mpl}, {2048 * mpl}, {wpg}> (CRC32 fused in the resolve)")
    return f"cbm_inflate: inflate_wave<{shape[0]}, {shape[1]}> (CRC32 fused)"


def _decode_header(raw: bytes):
    from cubam_amd.bam import BamHeader

    return BamHeader.from_bytes(raw)


if __name__ == "__main__":
    main()
