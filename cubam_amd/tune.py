"""Autotuner of the inflate launch shape, and a hipGraph of a batched inflate.

The inflate kernel comes in launch shapes (``cbm_inflate_tune``):

* workgroup variants, where a workgroup of NT threads decodes one block at a time
  with SEG bits per thread and round. NT and SEG set the LDS layout: the
  compressed-bit staging area is NT * SEG / 8 bytes beside the 64 KiB output
  window and the decode tables.
* two-phase variants (a wavefront per block decodes SEG-bit segments into literals
  and a match list; a resolve pass applies the matches): ``blocks_per_wave`` picks
  the resolve pass's LDS layout (a window of 64 or 128 matches, or a history ring
  of 4 to 32 KiB a wavefront) and ``waves_per_group`` its blocks per workgroup.
* the lane variant, where every lane of a wavefront inflates a block of its own,
  with ``blocks_per_wave`` x ``waves_per_group`` blocks per workgroup.

:func:`autotune` times every candidate on a sample of a batch's blocks and keeps
the fastest per (GPU, batch-size bucket) in a JSON cache
(``$CUBAM_AMD_TUNE_CACHE``, default ``~/.cache/cubam_amd/inflate_tune.json``).
:func:`lookup` returns the cached choice for a batch, and every inflate
(:func:`cubam_amd._decode.inflate_blocks`, so ``read_bam`` and the serving loop)
uses it when one is cached for the batch's bucket on this GPU; ``enable(False)``
or ``CUBAM_AMD_AUTOTUNE=0`` turns that off (the library default is used).

:class:`GraphInflate` captures the inflate of one fixed set of blocks (a
batched heterogeneous decode) in a hipGraph, so each later decode of new bytes
in the same buffers is one graph launch.
"""

from __future__ import annotations

import json
import math
import os
import time
from pathlib import Path

import torch

from cubam_amd import _decode, _lib

#: workgroup variants; the two-phase variants by segment bits, resolve layout (window of 64 or 128
#: matches, or a history ring of 4 to 32 KiB) and resolve blocks per workgroup; then the lane
#: variant at 1, 2, 4 and 8 blocks per workgroup
CANDIDATES = tuple(
    [(v, 0, 0, 0) for v in (0, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15)]
    + [(11, 1, wpg, 0) for wpg in (2, 8)]
    + [(11, 2, wpg, 0) for wpg in (2, 4)]
    + [(11, ring, wpg, 0) for ring, wpg in ((4, 1), (4, 4), (8, 1), (8, 2), (16, 1), (32, 1))]
    + [(v, 1, 2, 0) for v in (12, 13)]
    + [(15, 1, wpg, 0) for wpg in (4, 8)]
    + [(_lib.INFLATE_LANES, bpw, 1, 0) for bpw in (1, 2, 4, 8)]
    + [(_lib.INFLATE_LANES, 4, 2, 0), (_lib.INFLATE_LANES, 8, 2, 0)]
)

_enabled = os.environ.get("CUBAM_AMD_AUTOTUNE", "1") != "0"
_cache: dict | None = None


def cache_path() -> Path:
    p = os.environ.get("CUBAM_AMD_TUNE_CACHE")
    return Path(p) if p else Path.home() / ".cache" / "cubam_amd" / "inflate_tune.json"


def enable(on: bool = True) -> None:
    """Let inflate launches use the cached autotuned shape for their bucket."""
    global _enabled
    _enabled = on


def enabled() -> bool:
    return _enabled


def _load() -> dict:
    global _cache
    if _cache is None:
        try:
            _cache = json.loads(cache_path().read_text())
        except (OSError, ValueError):
            _cache = {}
    return _cache


def _save() -> None:
    p = cache_path()
    try:
        p.parent.mkdir(parents=True, exist_ok=True)
        tmp = p.with_suffix(f".tmp{os.getpid()}")
        tmp.write_text(json.dumps(_load(), indent=1, sort_keys=True))
        os.replace(tmp, p)
    except OSError:
        pass  # a read-only home: the choice stays in this process


def device_key() -> str:
    prop = torch.cuda.get_device_properties(torch.cuda.current_device())
    arch = getattr(prop, "gcnArchName", prop.name).split(":")[0]
    return f"{arch}/{prop.multi_processor_count}cu"


def bucket(n_blocks: int, inflated: int) -> str:
    """Batch-size bucket: powers of two of the block count and of the mean block size."""
    nb = max(1, n_blocks)
    mean = max(1, inflated // nb)
    return f"b{int(math.log2(nb))}-m{int(math.log2(mean))}"


def lookup(n_blocks: int, inflated: int) -> _lib.InflateTune | None:
    """The autotuned shape for a batch, if one is cached for its bucket on this GPU."""
    e = _load().get(device_key(), {}).get(bucket(n_blocks, inflated))
    return _lib.InflateTune(*e["tune"]) if e else None


def record(n_blocks: int, inflated: int, tune, ms: float, table: dict | None = None) -> None:
    d = _load().setdefault(device_key(), {})
    d[bucket(n_blocks, inflated)] = {"tune": list(tune), "ms": ms, "candidates": table or {}}
    _save()


def time_inflate(blocks: _decode.Blocks, tune, reps: int = 2, ws: _decode.Workspace | None = None) -> float:
    """Milliseconds of one inflate launch over ``blocks`` with shape ``tune`` (best of reps); None
    times the library's default shape."""
    t = _lib.InflateTune(*tune) if tune is not None else None
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    best = float("inf")
    for _ in range(reps):
        ev0.record()
        inf = _decode.inflate_blocks(blocks, None, ws, tune=t)
        ev1.record()
        ev1.synchronize()
        _decode.check_inflated(inf)
        best = min(best, ev0.elapsed_time(ev1))
        del inf
    return best


def sample_blocks(blocks: _decode.Blocks, n: int) -> _decode.Blocks:
    """The first n blocks of a scan, as a batch of their own."""
    n = min(n, blocks.n_blocks)
    out_off = blocks.out_off[: n + 1]
    return _decode.Blocks(blocks.buf, n, int(blocks.blk_off[n - 1] + blocks.bsize[n - 1]) if n else 0,
                          int(out_off[-1]), blocks.blk_off[:n], blocks.bsize[:n], blocks.isize[:n],
                          blocks.crc[:n], out_off)


def autotune(blocks: _decode.Blocks, sample: int = 50_000, candidates=CANDIDATES, log=None,
             ws: _decode.Workspace | None = None) -> tuple[tuple, dict]:
    """Time every candidate shape on the first ``sample`` blocks; cache and return the fastest.

    Returns ``(best_tune, {candidate: ms})``. The cache entry is keyed by the
    bucket of the whole batch, so later batches of that size use it.
    """
    s = sample_blocks(blocks, sample)
    table = {}
    t0 = time.time()
    for c in candidates:
        ms = time_inflate(s, c, ws=ws)
        table[",".join(map(str, c))] = round(ms, 3)
        if log:
            log(f"autotune {c}: {ms:.2f} ms for {s.n_blocks} blocks")
    best = min(candidates, key=lambda c: table[",".join(map(str, c))])
    record(blocks.n_blocks, blocks.total_out, best, table[",".join(map(str, best))], table)
    if log:
        log(f"autotune: best {best} in {time.time() - t0:.1f} s")
    return best, table


class GraphInflate:
    """The inflate of a fixed block table, captured once in a hipGraph and replayed.

    The compressed buffer, the block table and the outputs stay at fixed device
    addresses: write new compressed bytes of the same block layout into
    ``blocks.buf`` and call :meth:`run`. The graph holds only kernels (the
    workspace counters are cleared by a fill kernel, not a memset), and with the
    two-stream inflate the fork to its side stream and the join back.
    """

    def __init__(self, blocks: _decode.Blocks, tune=None, ws: _decode.Workspace | None = None):
        L = _lib.load()
        self.blocks = blocks
        self.tune = _decode.inflate_tune_for(blocks.n_blocks, blocks.total_out,
                                             _lib.InflateTune(*tune) if tune else None)
        dev = blocks.buf.device
        nb = blocks.n_blocks
        self.data = torch.empty(blocks.total_out + _lib.PAD, dtype=torch.uint8, device=dev)
        self.crc = torch.empty(max(nb, 1), dtype=torch.uint32, device=dev)
        self.status = torch.empty(max(nb, 1), dtype=torch.int32, device=dev)
        self.bad = torch.empty(1, dtype=torch.int64, device=dev)
        self.ws = torch.empty(L.cbm_inflate_workspace(nb), dtype=torch.uint8, device=dev)
        # eager once on the stream the capture uses: kernel attributes are set, and the side stream a
        # two-stream inflate forks to is made, outside the capture
        self.stream = torch.cuda.Stream()
        self.stream.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(self.stream):
            self._launch()
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph, stream=self.stream):
            self._launch()
        torch.cuda.synchronize()

    def _launch(self) -> None:
        import ctypes

        L = _lib.load()
        b = self.blocks
        _lib.check(L.cbm_inflate(b.buf.data_ptr(), b.blk_off.data_ptr(), b.bsize.data_ptr(), b.isize.data_ptr(),
                                 b.crc.data_ptr(), b.out_off.data_ptr(), b.n_blocks, self.data.data_ptr(),
                                 self.crc.data_ptr(), self.status.data_ptr(), self.bad.data_ptr(),
                                 _decode.tune_arg(self.tune), self.ws.data_ptr(), self.ws.numel(), _lib.stream_ptr()),
                   "cbm_inflate")

    def run(self) -> None:
        """Replay the captured launch on the current stream (asynchronous)."""
        self.graph.replay()

    def check(self) -> None:
        """Synchronise and raise if any block did not inflate cleanly."""
        _decode.raise_block_errors(self, int(self.bad.item()))
