"""The BAI on-disk index (SAMv1 section 5.2) of a coordinate-sorted BAM, built on the GPU.

The reference has no on-disk index (its docs/reading.md:52-53: a region query reads the
whole file). :func:`index_bam` decodes the file chunk by chunk with the GPU read path
(:func:`cubam_amd.index.build_index`), keeps every record's stream offset, reference,
position, end and flag on the device, and derives there what a BAI holds:

* each record's virtual offset (the file offset of its BGZF block << 16 | its offset in the
  block's data) and the virtual offset just past it (the next record's);
* its bin, ``reg2bin(pos, end)`` over [pos, end) with an end of at least pos + 1;
* the chunks of each bin: maximal runs of consecutive records of one reference and bin,
  [virtual offset of the run's first record, virtual offset past its last);
* the linear index: for each 16 kbp window of a reference, the smallest virtual offset of the
  mapped records overlapping it (empty windows take the value of the window before);
* the pseudo-bin 37450 of each reference (its first and past-last virtual offsets, mapped and
  placed-unmapped record counts) and the count of records with no coordinate.

:class:`Bai` writes and reads the binary layout ("BAI\\1", n_ref, per reference the bins with
their chunks and the linear index, then n_no_coor) and answers region queries with the merged
chunks to read (:meth:`Bai.query`); :func:`read_region` decodes them on the GPU.
"""

from __future__ import annotations

import struct
from dataclasses import dataclass
from pathlib import Path

import numpy as np
import torch

MIN_SHIFT = 14  # 16 kbp windows and leaf bins
META_BIN = 37450  # the pseudo-bin of each reference's summary
MAGIC = b"BAI\x01"
CSI_MAGIC = b"CSI\x01"
# first bin of each level of the binning scheme (SAMv1 5.3): ((1 << 3 * level) - 1) / 7
_LEVEL_FIRST = (0, 1, 9, 73, 585, 4681)
_LEVEL_SHIFT = (29, 26, 23, 20, 17, 14)


def reg2bin(beg, end):
    """The bin of [beg, end) (0-based, end exclusive, end > beg): the smallest bin wholly
    holding it (SAMv1 5.3). Works on ints, numpy arrays and torch tensors alike."""
    e = end - 1
    out = beg * 0  # level 0: bin 0
    for first, shift in zip(_LEVEL_FIRST[1:], _LEVEL_SHIFT[1:]):  # coarse to fine: the last match wins
        same = (beg >> shift) == (e >> shift)
        val = first + (beg >> shift)
        if isinstance(out, torch.Tensor):
            out = torch.where(same, val, out)
        elif isinstance(out, np.ndarray):
            out = np.where(same, val, out)
        elif same:
            out = val
    return out


def _bin_first_window(b: int) -> int:
    """The 16 kbp window where bin ``b``'s region begins."""
    level = 0
    while level + 1 < len(_LEVEL_FIRST) and b >= _LEVEL_FIRST[level + 1]:
        level += 1
    return (b - _LEVEL_FIRST[level]) << (_LEVEL_SHIFT[level] - MIN_SHIFT)


def reg2bins(beg: int, end: int) -> list[int]:
    """Every bin that may hold a record overlapping [beg, end) (SAMv1 5.3)."""
    e = end - 1
    bins = [0]
    for first, shift in zip(_LEVEL_FIRST[1:], _LEVEL_SHIFT[1:]):
        bins.extend(range(first + (beg >> shift), first + (e >> shift) + 1))
    return bins


@dataclass
class RefIndex:
    """One reference's part of a BAI."""

    bins: dict  # bin -> uint64 array (n_chunk, 2) of [begin, end) virtual offsets, in file order
    linear: np.ndarray  # uint64, one virtual offset per 16 kbp window
    meta: tuple | None = None  # (first virtual offset, past-last, mapped, placed unmapped)


@dataclass
class Bai:
    refs: list  # RefIndex per reference of the header
    n_no_coor: int | None = None

    # ---- binary layout (SAMv1 5.2)
    def to_bytes(self) -> bytes:
        out = [MAGIC, struct.pack("<i", len(self.refs))]
        for r in self.refs:
            nb = len(r.bins) + (r.meta is not None)
            out.append(struct.pack("<i", nb))
            for b in sorted(r.bins):
                ch = np.ascontiguousarray(r.bins[b], dtype="<u8")
                out.append(struct.pack("<Ii", b, ch.shape[0]))
                out.append(ch.tobytes())
            if r.meta is not None:
                beg, end, mapped, unmapped = r.meta
                out.append(struct.pack("<IiQQQQ", META_BIN, 2, beg, end, mapped, unmapped))
            lin = np.ascontiguousarray(r.linear, dtype="<u8")
            out.append(struct.pack("<i", lin.size))
            out.append(lin.tobytes())
        if self.n_no_coor is not None:
            out.append(struct.pack("<Q", self.n_no_coor))
        return b"".join(out)

    def write(self, path) -> Path:
        p = Path(path)
        p.write_bytes(self.to_bytes())
        return p

    @classmethod
    def from_bytes(cls, raw: bytes) -> Bai:
        if raw[:4] != MAGIC:
            raise ValueError("not a BAI index (bad magic)")
        at = 4

        def take(fmt):
            nonlocal at
            v = struct.unpack_from(fmt, raw, at)
            at += struct.calcsize(fmt)
            return v

        (n_ref,) = take("<i")
        if n_ref < 0:
            raise ValueError("corrupt BAI index (negative reference count)")
        refs = []
        for _ in range(n_ref):
            (n_bin,) = take("<i")
            bins, meta = {}, None
            for _ in range(n_bin):
                b, n_chunk = take("<Ii")
                ch = np.frombuffer(raw, "<u8", 2 * n_chunk, at).reshape(n_chunk, 2).astype(np.uint64)
                at += 16 * n_chunk
                if b == META_BIN and n_chunk == 2:
                    meta = (int(ch[0, 0]), int(ch[0, 1]), int(ch[1, 0]), int(ch[1, 1]))
                else:
                    bins[b] = ch
            (n_intv,) = take("<i")
            lin = np.frombuffer(raw, "<u8", n_intv, at).astype(np.uint64)
            at += 8 * n_intv
            refs.append(RefIndex(bins, lin, meta))
        n_no_coor = take("<Q")[0] if at + 8 <= len(raw) else None
        return cls(refs, n_no_coor)

    @classmethod
    def read(cls, path) -> Bai:
        raw = Path(path).read_bytes()
        return cls.from_csi_bytes(raw) if raw[:4] == CSI_MAGIC else cls.from_bytes(raw)

    # ---- the CSI layout (the coordinate-sorted index of the CSIv1 specification): the same
    # bins at min_shift 14 and depth 5, each with its first offset ("loffset": the smallest
    # virtual offset of a mapped record overlapping the bin's first 16 kbp window, from the
    # linear index) instead of the linear index
    def to_csi_bytes(self) -> bytes:
        out = [CSI_MAGIC, struct.pack("<iii", MIN_SHIFT, 5, 0), struct.pack("<i", len(self.refs))]
        for r in self.refs:
            nb = len(r.bins) + (r.meta is not None)
            out.append(struct.pack("<i", nb))
            for b in sorted(r.bins):
                ch = np.ascontiguousarray(r.bins[b], dtype="<u8")
                w = _bin_first_window(b)
                loff = int(r.linear[w]) if w < r.linear.size else 0
                out.append(struct.pack("<IQi", b, loff, ch.shape[0]))
                out.append(ch.tobytes())
            if r.meta is not None:
                beg, end, mapped, unmapped = r.meta
                out.append(struct.pack("<IQiQQQQ", META_BIN, 0, 2, beg, end, mapped, unmapped))
        if self.n_no_coor is not None:
            out.append(struct.pack("<Q", self.n_no_coor))
        return b"".join(out)

    @classmethod
    def from_csi_bytes(cls, raw: bytes) -> Bai:
        """A CSI at min_shift 14 and depth 5 as a :class:`Bai` whose linear index is the
        bins' first offsets, window by window (a lower bound for each window, as queries need)."""
        if raw[:4] != CSI_MAGIC:
            raise ValueError("not a CSI index (bad magic)")
        min_shift, depth, l_aux = struct.unpack_from("<iii", raw, 4)
        if (min_shift, depth) != (MIN_SHIFT, 5):
            raise ValueError(f"CSI with min_shift {min_shift} and depth {depth}: only 14 and 5 are read")
        at = 16 + l_aux
        (n_ref,) = struct.unpack_from("<i", raw, at)
        at += 4
        refs = []
        for _ in range(n_ref):
            (n_bin,) = struct.unpack_from("<i", raw, at)
            at += 4
            bins, meta, loffs = {}, None, {}
            for _ in range(n_bin):
                b, loff, n_chunk = struct.unpack_from("<IQi", raw, at)
                at += 16
                ch = np.frombuffer(raw, "<u8", 2 * n_chunk, at).reshape(n_chunk, 2).astype(np.uint64)
                at += 16 * n_chunk
                if b == META_BIN and n_chunk == 2:
                    meta = (int(ch[0, 0]), int(ch[0, 1]), int(ch[1, 0]), int(ch[1, 1]))
                else:
                    bins[b] = ch
                    if b >= _LEVEL_FIRST[-1]:  # leaf bins: one window each
                        loffs[b - _LEVEL_FIRST[-1]] = loff
            n_win = max(loffs) + 1 if loffs else 0
            lin = np.zeros(n_win, np.uint64)
            last = 0
            for w in range(n_win):  # windows without a leaf bin keep the bound before them
                last = loffs.get(w, last)
                lin[w] = last
            refs.append(RefIndex(bins, lin, meta))
        n_no_coor = struct.unpack_from("<Q", raw, at)[0] if at + 8 <= len(raw) else None
        return cls(refs, n_no_coor)

    def write_csi(self, path) -> Path:
        p = Path(path)
        p.write_bytes(self.to_csi_bytes())
        return p

    # ---- queries
    def query(self, ref: int, beg: int, end: int) -> np.ndarray:
        """The chunks ([begin, end) virtual offsets, uint64 (n, 2), sorted, overlaps merged)
        that hold every record of reference ``ref`` overlapping [beg, end): the chunks of the
        bins that may hold such a record, less those ending at or before the linear index's
        bound for the window of ``beg``."""
        if not 0 <= ref < len(self.refs) or end <= beg:
            return np.empty((0, 2), np.uint64)
        r = self.refs[ref]
        beg = max(beg, 0)
        w = beg >> MIN_SHIFT
        min_off = int(r.linear[w]) if w < r.linear.size else (int(r.linear[-1]) if r.linear.size else 0)
        parts = [r.bins[b] for b in reg2bins(beg, end) if b in r.bins]
        if not parts:
            return np.empty((0, 2), np.uint64)
        ch = np.concatenate(parts)
        ch = ch[ch[:, 1] > np.uint64(min_off)]
        if ch.shape[0] == 0:
            return ch
        ch = ch[np.argsort(ch[:, 0], kind="stable")]
        out = [list(ch[0])]
        for b0, e0 in ch[1:]:
            if b0 <= out[-1][1]:
                out[-1][1] = max(out[-1][1], e0)
            else:
                out.append([b0, e0])
        return np.asarray(out, np.uint64)


def build(ref, pos, end, flag, vbeg, vend, ref_lengths) -> Bai:
    """A BAI from per-record arrays in file order (torch tensors on any device, or numpy):
    reference id, 0-based position, end (position + reference-consuming CIGAR length),
    flag, the record's virtual offset and the one just past it."""
    t = [torch.as_tensor(np.asarray(x) if not isinstance(x, torch.Tensor) else x) for x in
         (ref, pos, end, flag, vbeg, vend)]
    dev = t[0].device
    ref, pos, end, flag, vbeg, vend = (x.to(dev, torch.int64) for x in t)
    n_ref = len(ref_lengths)
    n = int(ref.numel())
    placed = (ref >= 0) & (pos >= 0)
    mapped = placed & ((flag & 4) == 0)
    endx = torch.maximum(end, pos + 1)
    bins = torch.where(placed, reg2bin(pos, endx), torch.full_like(pos, -1))
    refs = [RefIndex({}, np.empty(0, np.uint64)) for _ in range(n_ref)]
    n_no_coor = int((~placed).sum().item())
    if n == 0:
        return Bai(refs, 0)
    # chunks: runs of consecutive records of one (reference, bin); unplaced records break runs
    key = torch.where(placed, ref * (1 << 20) + bins, torch.full_like(ref, -1))
    start = torch.ones(n, dtype=torch.bool, device=dev)
    start[1:] = key[1:] != key[:-1]
    first = torch.nonzero(start).flatten()
    last = torch.cat([first[1:] - 1, torch.tensor([n - 1], device=dev)])
    rk = key[first]
    keep = rk >= 0
    runs = torch.stack([rk[keep] >> 20, rk[keep] & ((1 << 20) - 1), vbeg[first[keep]], vend[last[keep]]], 1)
    runs = runs.cpu().numpy()
    # linear index: each mapped record's windows, the smallest virtual offset a window
    n_win = torch.tensor([(int(L) >> MIN_SHIFT) + 1 for L in ref_lengths], dtype=torch.int64, device=dev)
    base = torch.cumsum(n_win, 0) - n_win
    lin = torch.full((int(n_win.sum().item()) if n_ref else 0,), -1, dtype=torch.int64, device=dev)
    inref = mapped & (ref < n_ref)
    inref &= (pos >> MIN_SHIFT) < n_win[torch.where(inref, ref, torch.zeros_like(ref))]  # inside the reference
    mi = torch.nonzero(inref).flatten()
    if mi.numel():
        w0 = pos[mi] >> MIN_SHIFT
        w1 = torch.minimum((endx[mi] - 1) >> MIN_SHIFT, n_win[ref[mi]] - 1)
        cnt = (w1 - w0 + 1).clamp(min=1)
        rep = torch.repeat_interleave(torch.arange(mi.numel(), device=dev), cnt)
        within = torch.arange(rep.numel(), device=dev) - (torch.cumsum(cnt, 0) - cnt)[rep]
        slot = base[ref[mi]][rep] + w0[rep] + within
        big = torch.iinfo(torch.int64).max
        lin_min = torch.full_like(lin, big)
        lin_min.scatter_reduce_(0, slot, vbeg[mi][rep], reduce="amin")
        lin = torch.where(lin_min == big, lin, lin_min)
    lin = lin.cpu().numpy()
    # the pseudo-bin: first and past-last virtual offsets, mapped and placed-unmapped counts
    pr = torch.nonzero(placed & (ref < n_ref)).flatten()
    big = torch.iinfo(torch.int64).max
    rbeg = torch.full((n_ref,), big, dtype=torch.int64, device=dev)
    rend = torch.full((n_ref,), -1, dtype=torch.int64, device=dev)
    if pr.numel():
        rbeg.scatter_reduce_(0, ref[pr], vbeg[pr], reduce="amin")
        rend.scatter_reduce_(0, ref[pr], vend[pr], reduce="amax")
    nm = torch.bincount(ref[mapped & (ref < n_ref)], minlength=n_ref)[:n_ref]
    nu = torch.bincount(ref[placed & ~mapped & (ref < n_ref)], minlength=n_ref)[:n_ref]
    rbeg, rend, nm, nu = (x.cpu().numpy() for x in (rbeg, rend, nm, nu))
    base_h = base.cpu().numpy()
    n_win_h = n_win.cpu().numpy()
    order = np.lexsort((runs[:, 2], runs[:, 1], runs[:, 0])) if runs.size else np.empty(0, np.int64)
    runs = runs[order]
    cut = np.flatnonzero(np.diff(runs[:, 0] * (1 << 20) + runs[:, 1])) + 1 if runs.size else np.empty(0, np.int64)
    for grp in np.split(runs, cut) if runs.size else []:
        r, b = int(grp[0, 0]), int(grp[0, 1])
        if r < n_ref:
            refs[r].bins[b] = grp[:, 2:4].astype(np.uint64)
    for r in range(n_ref):
        w = lin[base_h[r] : base_h[r] + n_win_h[r]]
        have = np.flatnonzero(w >= 0)
        if have.size:
            w = w[: have[-1] + 1].copy()
            # windows with no record take the value of the window before (leading ones stay 0)
            idx = np.where(w >= 0, np.arange(w.size), -1)
            np.maximum.accumulate(idx, out=idx)
            w = np.where(idx >= 0, w[np.maximum(idx, 0)], 0)
            refs[r].linear = w.astype(np.uint64)
        if rend[r] >= 0:
            refs[r].meta = (int(rbeg[r]), int(rend[r]), int(nm[r]), int(nu[r]))
    return Bai(refs, n_no_coor)


def index_bam(src, chunk_bytes: int = 256 << 20):
    """Decode a BAM file (a path or a host byte array) on the GPU; returns its
    :class:`~cubam_amd.index.BlockIndex` and its :class:`Bai`."""
    from cubam_amd import _decode
    from cubam_amd.index import build_index

    ws = _decode.Workspace()
    parts = []

    def collect(data, off, stream, chain):
        cols = _decode.scatter(data, off, chain, tags=(), var=False, ws=ws)["fixed"]
        parts.append((stream.clone(), cols["ref"].to(torch.int64), cols["pos"].to(torch.int64),
                      cols["end"].to(torch.int64), cols["flag"].to(torch.int64)))

    bix = build_index(src, chunk_bytes, ws, on_records=collect)
    dev = _decode.device()
    if parts:
        stream, ref, pos, end, flag = (torch.cat([p[k] for p in parts]) for k in range(5))
    else:
        stream = ref = pos = end = flag = torch.empty(0, dtype=torch.int64, device=dev)
    uoff = torch.from_numpy(bix.uoff).to(dev)
    coff = torch.from_numpy(bix.coff).to(dev)
    vbeg = voffsets(stream, uoff, coff)
    # past a record: the next record's start; past the last (which ends the data): the start of
    # the block after the last one with data
    if stream.numel():
        k = int(np.searchsorted(bix.uoff, bix.stream_bytes - 1, side="right")) - 1
        if bix.stream_bytes == int(bix.uoff[k] + bix.isize[k]):
            tail = int(bix.coff[k] + bix.bsize[k]) << 16
        else:
            tail = (int(bix.coff[k]) << 16) | (bix.stream_bytes - int(bix.uoff[k]))
        vend = torch.cat([vbeg[1:], torch.tensor([tail], dtype=torch.int64, device=dev)])
    else:
        vend = vbeg
    return bix, build(ref, pos, end, flag, vbeg, vend, bix.header.lengths)


def voffsets(stream: torch.Tensor, uoff: torch.Tensor, coff: torch.Tensor) -> torch.Tensor:
    """Virtual offsets of inflated-stream offsets, given each block's stream and file offset
    (an offset at a block's end is the next block's start)."""
    k = torch.searchsorted(uoff, stream, right=True) - 1
    return (coff[k] << 16) | (stream - uoff[k])


def read_region(source, bai: Bai, ref: int, beg: int, end: int, tags=("CB", "UB")):
    """The records of reference ``ref`` overlapping [beg, end), decoded on the GPU from the
    chunks ``bai`` names; ``source`` is a :class:`cubam_amd.serve.BamSource` (its host bytes and
    block index). Returns a list of :class:`~cubam_amd.bam.RecordBatch`, one a chunk, each
    filtered to the overlapping records."""
    from cubam_amd.serve import decode_plain

    ix = source.index
    out = []
    for vb, ve in bai.query(ref, beg, end):
        sp = ix.voffset_span(int(vb), int(ve))
        rb = decode_plain(source, sp, tags)
        rlen_end = torch.maximum(rb.end, rb.pos + 1)
        keep = (rb.ref == ref) & (rb.pos < end) & (rlen_end > beg)
        out.append(rb.take(torch.nonzero(keep).flatten()))
    return out
