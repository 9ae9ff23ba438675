"""Served region decode: a request loop that batches region requests into shared launches.

BAM files are held in pinned host memory with their :class:`~cubam_amd.index.BlockIndex`.
A request asks for the records of a region of one file; the index turns it into
a run of BGZF blocks that starts where a record starts (a :class:`~cubam_amd.index.Span`),
so no block scan and no host round trip is needed to find the blocks.

:class:`Server` runs the loop (continuous batching):

* **Admission.** Requests wait in arrival order. A request is admitted when its
  estimated device footprint fits the HBM budget next to the batches in flight
  and the results kept resident; results are evicted oldest first to make room.
* **Batching.** A worker that is free takes every admitted request (up to a cap
  on compressed bytes) as one batch: one copy of the block tables, one inflate
  launch over the blocks of all of them (the persistent kernel steals blocks
  across requests), one record-chain launch sequence over all their spans
  (:func:`decode_spans`), one scatter. The batch synchronises once, after the
  record chain, to size the columns.
* **Overlap.** Each worker has its own HIP stream and workspace, so requests
  that arrive while a batch is on the GPU form the next batch and run beside it.
  While no worker is idle, a worker takes its next batch as soon as the current
  one is launched and copies it to the device on a second stream
  (:func:`stage_spans`) while the current one decodes (:func:`decode_staged`).
* **Residency.** Decoded results stay on the device until evicted or released.

``mode="plain"`` serves the same requests one per call on one stream with the
per-chunk path of :func:`~cubam_amd.bam.read_bam` (block scan, inflate, record
chain, scatter: the reference's src/cubam/bam.py:421-446 per request), the
baseline the served mode is compared with.

A result is the same alone and in a batch: every record of a span is decoded by
the same kernels from the same bytes; only where it lands in device memory
differs.
"""

from __future__ import annotations

import ctypes
import threading
import time
from collections import OrderedDict, deque
from dataclasses import dataclass, field
from types import SimpleNamespace

import numpy as np
import torch

from cubam_amd import _decode, _lib
from cubam_amd._lib import PAD, check, ptr, stream_ptr
from cubam_amd.bam import RecordBatch
from cubam_amd.index import BlockIndex, Span, build_index, host_bytes

_U8, _I32, _I64 = torch.uint8, torch.int32, torch.int64


def segment_bytes() -> int:
    """The record chain's segment size, as the library reports it."""
    seg = ctypes.c_int64(0)
    _lib.load().cbm_record_chain_workspace(0, ctypes.byref(seg))
    return int(seg.value)


class BamSource:
    """A BAM file held in pinned host memory, with its block index."""

    def __init__(self, src, index: BlockIndex | None = None, name: str = ""):
        raw = host_bytes(src)
        self.name = name or (str(src) if not isinstance(src, np.ndarray) else "BAM")
        self.host = torch.empty(int(raw.size), dtype=_U8, pin_memory=True)
        self.host.numpy()[:] = raw
        self.index = index if index is not None else build_index(raw)
        self.header = self.index.header
        self.n_ref = len(self.header.names)

    @property
    def size(self) -> int:
        return int(self.host.numel())


# ------------------------------------------------------------------ batched decode
class _Pinned:
    """A grow-only pinned host buffer. ``event``, when set, marks the end of the last copy out of
    it; :meth:`get` waits for it before the buffer is written again."""

    def __init__(self):
        self.t = torch.empty(0, dtype=_U8)
        self.event: torch.cuda.Event | None = None

    def get(self, n: int) -> torch.Tensor:
        if self.event is not None:
            self.event.synchronize()
            self.event = None
        if self.t.numel() < n:
            self.t = torch.empty(max(n, 2 * self.t.numel(), 1 << 16), dtype=_U8, pin_memory=True)
        return self.t[:n]


def estimate_bytes(sp: Span, n_tags: int = 2) -> int:
    """Device bytes a decoded span holds at most while it is decoded: compressed + inflated
    + offsets, fixed and tag columns + the CIGAR/seq/qual columns (at most the inflated size)."""
    return sp.compressed_bytes + 2 * (sp.end - sp.u0) + sp.records * (8 + 35 + 24 + 21 * n_tags) + 4 * PAD


def result_bytes(rb: RecordBatch) -> int:
    seen, total = set(), 0
    ts = [rb.data, rb.offset, *rb.columns.values()]
    for c in rb._tags.values():
        ts.extend(c.values())
    for t in ts:
        key = t.untyped_storage().data_ptr()
        if key not in seen:
            seen.add(key)
            total += t.untyped_storage().nbytes()
    return total


@dataclass
class Staged:
    """A batch whose tables and compressed bytes are on their way to the device (see
    :func:`stage_spans`); :func:`decode_staged` decodes it."""

    items: list
    plan: SimpleNamespace  # host-side sizes of the batch
    tab: torch.Tensor | None  # block and span tables (device)
    staging: torch.Tensor | None  # compressed bytes of every span, side by side (device)
    event: torch.cuda.Event | None  # recorded after the copies


def stage_spans(items: list[tuple[BamSource, Span]], pinned: _Pinned | None = None,
                stream: torch.cuda.Stream | None = None) -> Staged:
    """The host plan of a batch and its host-to-device copies (one packed table, then the
    compressed bytes of each span), issued on ``stream`` (the current stream by default), so
    that a batch can be copied while the one before it decodes."""
    pinned = pinned or _Pinned()
    dev = _decode.device()
    n = len(items)
    if n == 0:
        return Staged(items, SimpleNamespace(n=0), None, None, None)
    # ---- host plan: block table, span table, staging layout
    nbs = np.array([sp.k1 - sp.k0 for _, sp in items], np.int64)
    NB = int(nbs.sum())
    clen = np.array([sp.compressed_bytes for _, sp in items], np.int64)
    cpos = np.zeros(n + 1, np.int64)
    np.cumsum((clen + 15) // 16 * 16, out=cpos[1:])
    blk_off = np.empty(NB, np.int64)
    bsize = np.empty(NB, np.int32)
    isize = np.empty(NB, np.int32)
    crc = np.empty(NB, np.uint32)
    j = 0
    for r, (src, sp) in enumerate(items):
        ix = src.index
        m = sp.k1 - sp.k0
        blk_off[j : j + m] = ix.coff[sp.k0 : sp.k1] - sp.c0 + cpos[r]
        bsize[j : j + m] = ix.bsize[sp.k0 : sp.k1]
        isize[j : j + m] = ix.isize[sp.k0 : sp.k1]
        crc[j : j + m] = ix.crc[sp.k0 : sp.k1]
        j += m
    out_off = np.zeros(NB + 1, np.int64)
    np.cumsum(isize, out=out_off[1:])
    first_blk = np.zeros(n, np.int64)
    np.cumsum(nbs[:-1], out=first_blk[1:])
    base = out_off[first_blk]
    s_start = np.array([sp.start - sp.u0 for _, sp in items], np.int64) + base
    s_end = np.array([sp.end - sp.u0 for _, sp in items], np.int64) + base
    seg_base = np.zeros(n + 1, np.int64)
    seg_bytes = segment_bytes()
    np.cumsum((s_end - s_start + seg_bytes - 1) // seg_bytes, out=seg_base[1:])
    n_seg = int(seg_base[-1])
    nref = np.array([src.n_ref for src, _ in items], np.int32)
    total_out = int(out_off[-1])
    # one pinned table, one copy: int64 parts first, then int32 parts
    parts64 = [blk_off, out_off, s_start, s_end, seg_base]
    parts32 = [bsize, isize, crc.view(np.int32), nref]
    n64 = sum(p.size for p in parts64)
    n32 = sum(p.size for p in parts32)
    nbytes = 8 * n64 + 4 * n32
    host = pinned.get(nbytes)
    h64 = host[: 8 * n64].view(_I64).numpy()
    h32 = host[8 * n64 : nbytes].view(_I32).numpy()
    o = 0
    for p in parts64:
        h64[o : o + p.size] = p
        o += p.size
    o = 0
    for p in parts32:
        h32[o : o + p.size] = p
        o += p.size
    stream = stream or torch.cuda.current_stream()
    with torch.cuda.stream(stream):
        tab = torch.empty(nbytes, dtype=_U8, device=dev)
        tab.copy_(host, non_blocking=True)
        # ---- compressed bytes of every span, side by side
        staging = torch.empty(int(cpos[-1]) + PAD, dtype=_U8, device=dev)
        for r, (src, sp) in enumerate(items):
            staging[cpos[r] : cpos[r] + clen[r]].copy_(src.host[sp.c0 : sp.c1], non_blocking=True)
        ev = torch.cuda.Event()
        ev.record(stream)
    pinned.event = ev
    plan = SimpleNamespace(n=n, NB=NB, n64=n64, n32=n32, n_seg=n_seg, total_out=total_out,
                           cuts64=np.cumsum([0] + [p.size for p in parts64]),
                           cuts32=np.cumsum([0] + [p.size for p in parts32]))
    return Staged(items, plan, tab, staging, ev)


def one_stream(t: _lib.InflateTune | None) -> _lib.InflateTune:
    """The inflate shape of a served batch: ``t`` (or the library default) on the calling stream
    alone. The served loop overlaps batches on its workers' streams already; splitting each batch
    over a side stream as well only adds streams to share the device's hardware queues (measured:
    79 against 60-76 GB/s in the closed loop)."""
    if t is None:
        return _lib.InflateTune(_lib.INFLATE_TOK64_S512, 0, 0, 0)
    if t.variant == _lib.INFLATE_TOK64_S512_SPLIT:
        return _lib.InflateTune(_lib.INFLATE_TOK64_S512, t.blocks_per_wave, t.waves_per_group, t.groups)
    return t


def decode_spans(items: list[tuple[BamSource, Span]], tags=("CB", "UB"), ws: _decode.Workspace | None = None,
                 pinned: _Pinned | None = None) -> list[RecordBatch]:
    """Decode the records of many spans (of one or more files) in one batch, on the current stream.

    Returns one :class:`RecordBatch` per span, in order. They share the batch's
    device buffers: ``offset`` points into the shared ``data`` and the CSR offsets
    (``cigar_off``, ``seq_off``, ``qual_off``) into the shared CSR arrays. Raises
    on corrupt data as the per-chunk path does.
    """
    return decode_staged(stage_spans(items, pinned), tags, ws)


def decode_staged(staged: Staged, tags=("CB", "UB"), ws: _decode.Workspace | None = None) -> list[RecordBatch]:
    """Decode a staged batch on the current stream (after its copies): the kernels of
    :func:`decode_spans`."""
    L = _lib.load()
    ws = ws or _decode.default_workspace()
    dev = _decode.device()
    st = stream_ptr()
    items, pl = staged.items, staged.plan
    n = pl.n
    if n == 0:
        return []
    cur = torch.cuda.current_stream()
    cur.wait_event(staged.event)
    tab, staging = staged.tab, staged.staging
    tab.record_stream(cur)
    staging.record_stream(cur)
    NB, n_seg, total_out, n64, cuts64, cuts32 = pl.NB, pl.n_seg, pl.total_out, pl.n64, pl.cuts64, pl.cuts32
    d64 = tab[: 8 * n64].view(_I64)
    d32 = tab[8 * n64 :].view(_I32)
    blk_off_d, out_off_d, ss_d, se_d, sb_d = (d64[cuts64[i] : cuts64[i + 1]] for i in range(5))
    bsize_d, isize_d, crc_d, nref_d = (d32[cuts32[i] : cuts32[i + 1]] for i in range(4))
    # ---- kernel 2: every block of the batch in one launch
    data = torch.empty(total_out + PAD, dtype=_U8, device=dev)
    crc_out = torch.empty(max(NB, 1), dtype=torch.uint32, device=dev)
    status = torch.empty(max(NB, 1), dtype=_I32, device=dev)
    inf_bad = torch.empty(1, dtype=_I64, device=dev)
    tune = one_stream(_decode.inflate_tune_for(NB, total_out))
    iws = ws.dev("inflate_ws", L.cbm_inflate_workspace(NB))
    check(L.cbm_inflate(ptr(staging), ptr(blk_off_d), ptr(bsize_d), ptr(isize_d), crc_d.data_ptr(), ptr(out_off_d),
                        NB, ptr(data), ptr(crc_out), ptr(status), ptr(inf_bad), _decode.tune_arg(tune), ptr(iws),
                        iws.numel(), st), "cbm_inflate")
    # ---- kernel 3 over all spans
    rws = ws.dev("spans_ws", L.cbm_record_spans_workspace(n_seg))
    result = ws.dev("spans_result", 64)
    span_rec = torch.empty(n + 1, dtype=_I64, device=dev)
    args = (ptr(data), ptr(ss_d), ptr(se_d), ptr(sb_d), ptr(nref_d), n, n_seg, ptr(result), ptr(span_rec), ptr(rws),
            rws.numel(), st)
    check(L.cbm_record_chain_spans(*args), "cbm_record_chain_spans")
    h = ws.host("spans_host", 10 + n)
    h[:8].copy_(result[:64].view(_I64), non_blocking=True)
    h[8:9].copy_(inf_bad, non_blocking=True)
    h[9 : 10 + n].copy_(span_rec, non_blocking=True)
    torch.cuda.current_stream().synchronize()
    del staging
    staged.staging = None  # the compressed bytes are no longer needed
    if int(h[8]):
        _decode.raise_block_errors(SimpleNamespace(status=status[:NB]), int(h[8]))
    if int(h[_lib.REC_BAD_LINKS]):
        check(L.cbm_record_chain_spans_repair(*args), "cbm_record_chain_spans_repair")
        h[:8].copy_(result[:64].view(_I64))
        h[9 : 10 + n].copy_(span_rec)
    if int(h[_lib.REC_INVALID]):
        raise ValueError(_decode.MSG_NO_BLOCK)
    n_rec = int(h[_lib.REC_N])
    bounds = h[9 : 10 + n].numpy().copy()
    # ---- offsets and kernel 4 over all records
    off = torch.empty(max(n_rec, 1), dtype=_I64, device=dev)
    csr = {k: torch.empty(n_rec + 1, dtype=_I64, device=dev) for k in _decode.CSR_OFFSETS}
    check(L.cbm_record_offsets_spans_csr(ptr(data), n_seg, ptr(rws), rws.numel(), ptr(off),
                                         *(ptr(csr[k]) for k in _decode.CSR_OFFSETS), n_rec, st),
          "cbm_record_offsets_spans_csr")
    off = off[:n_rec]
    sizes = SimpleNamespace(n_cigar=int(h[_lib.REC_CIGAR]), n_seq=int(h[_lib.REC_SEQ]), n_qual=int(h[_lib.REC_QUAL]))
    out = _decode.scatter(data, off, sizes, tags=tags, ws=ws, csr=csr)
    fixed, tag_cols = out["fixed"], out["tags"]
    res = []
    for r, (src, sp) in enumerate(items):
        a, b = int(bounds[r]), int(bounds[r + 1])
        if b - a != sp.records:
            raise ValueError(_decode.MSG_NO_BLOCK)
        cols = {}
        for k, v in fixed.items():
            if k in RecordBatch.FIELDS:
                cols[k] = v[a:b]
            elif k.endswith("_off"):
                cols[k] = v[a : b + 1]
            else:
                cols[k] = v  # CSR values: shared, addressed by the *_off columns
        tg = {t: {k: v[a:b] for k, v in c.items()} for t, c in tag_cols.items()}
        res.append(RecordBatch(data, off[a:b], src.header, cols, tg))
    return res


def decode_plain(src: BamSource, sp: Span, tags=("CB", "UB"), ws: _decode.Workspace | None = None) -> RecordBatch:
    """One span decoded by the per-chunk path (block scan, inflate, record chain, scatter) on the
    current stream, as the reference decodes a chunk (src/cubam/bam.py:421-446): the baseline."""
    ws = ws or _decode.default_workspace()
    dev = _decode.device()
    c = sp.compressed_bytes
    buf = torch.empty(c + PAD, dtype=_U8, device=dev)
    buf[:c].copy_(src.host[sp.c0 : sp.c1], non_blocking=True)
    blocks = _decode.scan_blocks(buf, c, True, ws)
    inf = _decode.inflate_blocks(blocks, None, ws)
    start, end = sp.start - sp.u0, sp.end - sp.u0
    chain = _decode.record_chain(inf.data, start, end, src.n_ref, ws, inf=inf)
    if chain.tail != end or (sp.rec1 >= 0 and chain.n_rec != sp.records):  # rec1 < 0: count unknown
        raise ValueError(_decode.MSG_NO_BLOCK)
    off, csr = _decode.record_offsets(inf.data, start, end, chain, csr=True)
    return RecordBatch.decode(inf.data, off, src.header, chain, tags, csr=csr)


# ------------------------------------------------------------------ the request loop
@dataclass
class WriteResult:
    """The records of a write request that passed its filter, as BGZF blocks on the device (no
    header, no EOF block: :meth:`to_bam` makes a whole file of them)."""

    blocks: torch.Tensor  # uint8: BGZF blocks back to back
    payload_bytes: int  # record bytes they hold
    records: int

    def to_bam(self, header) -> bytes:
        from cubam_amd.write import EOF_BLOCK, bgzf_compress

        hb = header.to_bytes() if hasattr(header, "to_bytes") else bytes(header)
        return bgzf_compress(hb).cpu().numpy().tobytes() + self.blocks.cpu().numpy().tobytes() + EOF_BLOCK


def write_back(rb: RecordBatch, cells: torch.Tensor, tag: str = "CB", ws: _decode.Workspace | None = None) -> WriteResult:
    """The records of ``rb`` whose ``tag`` barcode is one of ``cells`` (packed codes, int64; see
    :meth:`~cubam_amd.bam.Tag.packed`), gathered and BGZF-compressed on the current stream."""
    from cubam_amd.write import bgzf_compress, gather_records

    codes, ok = rb.tag(tag).packed()
    keep = ok & torch.isin(codes.view(torch.int64), cells)
    idx = torch.nonzero(keep).flatten()
    rec = gather_records(rb.data, rb.offset[idx])
    n = int(rec.numel())
    blocks = bgzf_compress(rec, ws=ws) if n else torch.empty(0, dtype=_U8, device=rec.device)
    return WriteResult(blocks, n, int(idx.numel()))


@dataclass(eq=False)
class Request:
    file: int
    span: Span
    cells: torch.Tensor | None = None  # a write request: keep these barcodes, return BGZF blocks
    arrival: float = 0.0  # time.perf_counter() when it arrived
    id: int = 0
    result: RecordBatch | None = None
    error: BaseException | None = None
    admitted: float = 0.0
    done_at: float = 0.0
    batch_size: int = 0
    est: int = 0
    done: threading.Event = field(default_factory=threading.Event)

    @property
    def latency(self) -> float:
        return self.done_at - self.arrival


class Server:
    """Continuous-batching region decode over a set of :class:`BamSource` files.

    ``budget_bytes``: device bytes that batches in flight and resident results may
    hold together. ``workers``: batches on the GPU at once (one stream each).
    ``batch_bytes``: compressed bytes per batch at most (a larger request forms a
    batch of its own). ``mode="plain"`` serves one request per call on one stream
    with the per-chunk path, as the baseline.
    """

    def __init__(self, sources: list[BamSource], budget_bytes: int = 64 << 30, workers: int = 3,
                 batch_bytes: int = 512 << 20, tags=("CB", "UB"), mode: str = "served", keep_results: bool = True):
        if mode not in ("served", "plain"):
            raise ValueError(f"mode is 'served' or 'plain', not {mode!r}")
        self.sources = sources
        self.budget = int(budget_bytes)
        self.batch_bytes = int(batch_bytes)
        self.tags = tuple(tags)
        self.mode = mode
        self.keep = keep_results
        self._cv = threading.Condition()
        self._pending: deque[Request] = deque()
        self._resident: OrderedDict[int, tuple[Request, int]] = OrderedDict()
        self.resident_bytes = 0
        self.in_flight_bytes = 0
        self.peak_bytes = 0
        self._closing = False
        self._waiting = 0  # workers waiting for requests
        self._ids = 0
        self.batches: list[int] = []
        n_workers = 1 if mode == "plain" else max(1, workers)
        self._threads = [threading.Thread(target=self._work, args=(i,), daemon=True) for i in range(n_workers)]
        for t in self._threads:
            t.start()

    # ---- client side
    def submit(self, file: int, span: Span | None = None, region: tuple[int, int, int] | None = None,
               arrival: float | None = None, cells: torch.Tensor | None = None) -> Request:
        """Queue a request for a span (or a region ``(ref, beg, end)``) of file ``file``.

        With ``cells`` (packed barcode codes, int64 on the device) it is a write request:
        its result is a :class:`WriteResult`, the records of those cells as BGZF blocks.
        """
        if span is None:
            if region is None:
                raise ValueError("a request names a span or a region")
            span = self.sources[file].index.region(*region)
        if cells is not None and "CB" not in self.tags:
            raise ValueError("write requests filter on CB: serve with CB among the tags")
        req = Request(file, span, cells, time.perf_counter() if arrival is None else arrival)
        req.est = estimate_bytes(span, len(self.tags)) * (2 if cells is not None else 1)
        if req.est > self.budget:
            raise ValueError(f"request needs about {req.est} device bytes, more than the budget {self.budget}")
        with self._cv:
            self._ids += 1
            req.id = self._ids
            self._pending.append(req)
            self._cv.notify_all()
        return req

    def release(self, req: Request) -> None:
        """Drop a resident result (its memory returns to the budget)."""
        with self._cv:
            item = self._resident.pop(req.id, None)
            if item is not None:
                self.resident_bytes -= item[1]
                item[0].result = None
                self._cv.notify_all()

    def close(self) -> None:
        with self._cv:
            self._closing = True
            self._cv.notify_all()
        for t in self._threads:
            t.join()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- admission and batching
    def _evict_for(self, need: int) -> bool:
        while self.in_flight_bytes + self.resident_bytes + need > self.budget and self._resident:
            _, (old, nb) = self._resident.popitem(last=False)
            old.result = None
            self.resident_bytes -= nb
        return self.in_flight_bytes + self.resident_bytes + need <= self.budget

    def _take(self, block: bool = True) -> list[Request] | None:
        """The next batch of admitted requests; None when the server closes (or, with block
        False, when nothing can be admitted now)."""
        with self._cv:
            while True:
                if self._pending:
                    batch, cbytes, est = [], 0, 0
                    while self._pending:
                        r = self._pending[0]
                        c = r.span.compressed_bytes
                        if batch and (self.mode == "plain" or cbytes + c > self.batch_bytes):
                            break
                        if not self._evict_for(est + r.est):
                            break
                        batch.append(self._pending.popleft())
                        cbytes += c
                        est += r.est
                    if batch:
                        self.in_flight_bytes += est
                        self.peak_bytes = max(self.peak_bytes, self.in_flight_bytes + self.resident_bytes)
                        now = time.perf_counter()
                        for r in batch:
                            r.admitted = now
                            r.batch_size = len(batch)
                        return batch
                if (self._closing and not self._pending) or not block:
                    return None
                self._waiting += 1
                self._cv.wait()
                self._waiting -= 1

    def _stage(self, batch: list[Request], pinned: _Pinned, stream) -> Staged | BaseException:
        try:
            return stage_spans([(self.sources[r.file], r.span) for r in batch], pinned, stream)
        except Exception as e:  # noqa: BLE001
            return e

    def _work(self, i: int) -> None:
        """A worker: decodes one batch at a time on its stream. In served mode it takes the next
        batch as soon as the current one is launched and copies it on a second stream while the
        current one decodes."""
        stream = torch.cuda.Stream()
        copy = torch.cuda.Stream() if self.mode != "plain" else None
        ws = _decode.Workspace()
        pins = [_Pinned(), _Pinned()]
        slot = 0
        nxt = None
        with torch.cuda.stream(stream):
            while True:
                if nxt is None:
                    batch = self._take()
                    if batch is None:
                        return
                    staged = self._stage(batch, pins[slot], copy) if copy is not None else None
                    slot ^= 1
                else:
                    batch, staged = nxt
                    nxt = None
                if copy is not None and not self._waiting:  # no worker is idle
                    nb = self._take(block=False)
                    if nb is not None:
                        nxt = (nb, self._stage(nb, pins[slot], copy))
                        slot ^= 1
                results = self._run(batch, ws, pins[slot], staged)
                now = time.perf_counter()
                with self._cv:
                    self.in_flight_bytes -= sum(r.est for r in batch)
                    self.batches.append(len(batch))
                    for r, res in zip(batch, results):
                        if isinstance(res, BaseException):
                            r.error = res
                        else:
                            r.result = res
                            if self.keep:
                                nb = (result_bytes(res) if len(batch) == 1 else r.est) if isinstance(
                                    res, RecordBatch) else int(res.blocks.numel())
                                self._resident[r.id] = (r, nb)
                                self.resident_bytes += nb
                        r.done_at = now
                        r.done.set()
                    self._cv.notify_all()

    def _run(self, batch: list[Request], ws, pinned, staged) -> list:
        try:
            if self.mode == "plain":
                r = batch[0]
                out = [decode_plain(self.sources[r.file], r.span, self.tags, ws)]
            else:
                if isinstance(staged, BaseException):
                    raise staged
                out = decode_staged(staged, self.tags, ws)
            for k, r in enumerate(batch):
                if r.cells is not None:
                    out[k] = write_back(out[k], r.cells, ws=ws)
            torch.cuda.current_stream().synchronize()
            return out
        except Exception as e:  # noqa: BLE001
            if len(batch) == 1:
                return [e]
        # a batch that failed: decode each request alone, so an error stays with its request
        outs = []
        for r in batch:
            try:
                res = decode_spans([(self.sources[r.file], r.span)], self.tags, ws, pinned)[0]
                if r.cells is not None:
                    res = write_back(res, r.cells, ws=ws)
                outs.append(res)
                torch.cuda.current_stream().synchronize()
            except Exception as e:  # noqa: BLE001
                outs.append(e)
        return outs


# ------------------------------------------------------------------ traces
@dataclass
class TraceItem:
    t: float  # arrival, seconds after the start
    file: int
    span: Span
    cells: torch.Tensor | None = None  # a write request


def make_trace(sources: list[BamSource], n: int, rate: float, seed: int = 0, min_bytes: int = 64 << 10,
               max_bytes: int = 256 << 20, write_fraction: float = 0.0, cells=None) -> list[TraceItem]:
    """Poisson arrivals at ``rate`` per second; each request a region of a random file whose
    compressed span is log-uniform in [min_bytes, max_bytes] (clipped to the file). A
    ``write_fraction`` of them are write requests for the barcodes ``cells``."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n)) if rate > 0 else np.zeros(n)
    out = []
    for i in range(n):
        f = int(rng.integers(len(sources)))
        ix = sources[f].index
        rb = ix.record_blocks()
        want = float(np.exp(rng.uniform(np.log(min_bytes), np.log(max_bytes))))
        nblk = rb.size
        # record blocks from a random start until the compressed span reaches `want`
        i0 = int(rng.integers(nblk))
        c0 = ix.coff[rb[i0]]
        i1 = int(np.searchsorted(ix.coff[rb], c0 + want, side="left"))
        i1 = min(max(i1, i0 + 1), nblk)
        if i1 == nblk and ix.coff[-1] + ix.bsize[-1] - c0 < want:  # ran off the end: move the start back
            back = int(np.searchsorted(ix.coff[rb], ix.coff[-1] + ix.bsize[-1] - want, side="left"))
            i0 = max(0, min(i0, back))
        w = cells if (write_fraction > 0 and rng.random() < write_fraction) else None
        out.append(TraceItem(float(t[i]), f, ix.span(i0, i1), w))
    return out


def run_trace(server: Server, trace: list[TraceItem]) -> list[Request]:
    """Submit each request at its arrival time, wait for all, return them (with latencies)."""
    t0 = time.perf_counter()
    reqs = []
    for it in trace:
        delay = t0 + it.t - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        reqs.append(server.submit(it.file, it.span, arrival=t0 + it.t, cells=it.cells))
    for r in reqs:
        r.done.wait()
    return reqs


def summarize(reqs: list[Request]) -> dict:
    lat = np.array([r.latency for r in reqs]) * 1e3
    u = sum(r.span.inflated_bytes for r in reqs)
    c = sum(r.span.compressed_bytes for r in reqs)
    t_first = min(r.arrival for r in reqs)
    t_last = max(r.done_at for r in reqs)
    wall = t_last - t_first
    errors = sum(r.error is not None for r in reqs)
    writes = [r for r in reqs if r.cells is not None]
    return {
        "write_requests": len(writes),
        "requests": len(reqs),
        "errors": errors,
        "p50_ms": round(float(np.percentile(lat, 50)), 3),
        "p99_ms": round(float(np.percentile(lat, 99)), 3),
        "mean_ms": round(float(lat.mean()), 3),
        "GBps": round(u / wall / 1e9, 3),
        "compressed_GBps": round(c / wall / 1e9, 3),
        "inflated_bytes": int(u),
        "wall_s": round(wall, 4),
        "mean_batch": round(float(np.mean([r.batch_size for r in reqs])), 2),
    }
