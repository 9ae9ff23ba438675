"""Block and record index of a BAM file, built with the GPU read path.

For every BGZF block the index keeps where the block is in the file, where its
data is in the inflated stream, its trailer CRC32, and the first record that
starts in it (its offset in the inflated stream, its index among all records,
its reference id and position). That is what a BAM linear index holds (the
virtual offset of the first record of each block). The reference has no index
(its docs/reading.md:52-53: region queries read the whole file); region
requests use this one to decode only the blocks they need, from a record
boundary to a record boundary, with no block scan.
"""

from __future__ import annotations

import gzip
import io
from dataclasses import dataclass
from pathlib import Path

import numpy as np
import torch

from cubam_amd import _decode
from cubam_amd._lib import PAD
from cubam_amd.bam import BamHeader, _Cursor


class _ArrayFile(io.RawIOBase):
    """A read-only file over a host byte array (no copy)."""

    def __init__(self, arr: np.ndarray):
        self._a, self._p = arr, 0

    def readable(self) -> bool:
        return True

    def readinto(self, b) -> int:
        n = min(len(b), self._a.size - self._p)
        b[:n] = self._a[self._p : self._p + n].tobytes()
        self._p += n
        return n


def header_of(src) -> BamHeader:
    """The header of a BAM file given as a path or a host byte array."""
    if isinstance(src, np.ndarray):
        with gzip.GzipFile(fileobj=io.BufferedReader(_ArrayFile(src))) as fh:
            return BamHeader._parse(_Cursor(fh.read, "BAM"))
    return BamHeader.read(src)


def host_bytes(src) -> np.ndarray:
    """A BAM file as a host byte array (a read-only map of a path)."""
    if isinstance(src, np.ndarray):
        return src
    return np.memmap(src, np.uint8, mode="r")


@dataclass
class Span:
    """The records starting in a run of blocks: blocks [k0, k1) hold them, and they are the
    inflated-stream bytes [start, end) and records [rec0, rec1)."""

    k0: int
    k1: int
    start: int
    end: int
    rec0: int
    rec1: int
    c0: int  # file bytes [c0, c1) of blocks k0..k1-1
    c1: int
    u0: int  # inflated-stream offset of block k0's data

    @property
    def compressed_bytes(self) -> int:
        return self.c1 - self.c0

    @property
    def inflated_bytes(self) -> int:
        return self.end - self.start

    @property
    def records(self) -> int:
        return self.rec1 - self.rec0


@dataclass
class BlockIndex:
    header: BamHeader
    coff: np.ndarray  # int64: block start in the file
    bsize: np.ndarray  # int32
    uoff: np.ndarray  # int64: block data start in the inflated stream (the header counts)
    isize: np.ndarray  # int32
    crc: np.ndarray  # uint32
    first_rec: np.ndarray  # int64: first record starting in the block (inflated-stream offset), -1 if none
    first_idx: np.ndarray  # int64: its index among the file's records, -1 if none
    first_ref: np.ndarray  # int32: its reference id
    first_pos: np.ndarray  # int32: its position
    n_records: int
    stream_bytes: int  # inflated bytes of the file

    def __len__(self) -> int:
        return int(self.coff.size)

    def record_blocks(self) -> np.ndarray:
        """Blocks in which a record starts: the places a request can start."""
        return np.flatnonzero(self.first_rec >= 0)

    def span(self, i0: int, i1: int) -> Span:
        """The records starting in record blocks [i0, i1) (indices into :meth:`record_blocks`)."""
        rb = self.record_blocks()
        if not 0 <= i0 < i1 <= rb.size:
            raise IndexError(f"record blocks [{i0}, {i1}) out of range 0..{rb.size}")
        k0 = int(rb[i0])
        s, r0 = int(self.first_rec[k0]), int(self.first_idx[k0])
        if i1 < rb.size:
            e, r1 = int(self.first_rec[rb[i1]]), int(self.first_idx[rb[i1]])
        else:
            e, r1 = self.stream_bytes, self.n_records
        ends = self.uoff + self.isize
        k1 = int(np.searchsorted(ends, e - 1, side="right")) + 1  # the block holding byte e - 1, plus one
        k1 = max(k1, k0 + 1)
        c1 = int(self.coff[k1 - 1] + self.bsize[k1 - 1])
        return Span(k0, k1, s, e, r0, r1, int(self.coff[k0]), c1, int(self.uoff[k0]))

    def voffset_span(self, vbeg: int, vend: int) -> Span:
        """The span of the records from virtual offset ``vbeg`` (a record start) to ``vend`` (a
        record end), as a BAI chunk names them; its record count is unknown (rec0 = rec1 = -1)."""
        def stream_at(v: int) -> tuple[int, int]:
            c, u = v >> 16, v & 0xFFFF
            k = int(np.searchsorted(self.coff, c))
            if k < self.coff.size and int(self.coff[k]) == c:
                return k, int(self.uoff[k]) + u
            last = self.coff.size - 1
            if last >= 0 and u == 0 and c == int(self.coff[last] + self.bsize[last]):
                return last + 1, int(self.uoff[last] + self.isize[last])
            raise ValueError(f"virtual offset {v:#x} is not at a block of this file")

        k0, start = stream_at(vbeg)
        _, end = stream_at(vend)
        if not (k0 < self.coff.size and start < end <= self.stream_bytes):
            raise ValueError(f"virtual offsets [{vbeg:#x}, {vend:#x}) name no records of this file")
        ends = self.uoff + self.isize
        k1 = max(int(np.searchsorted(ends, end - 1, side="right")) + 1, k0 + 1)
        c1 = int(self.coff[k1 - 1] + self.bsize[k1 - 1])
        return Span(k0, k1, start, end, -1, -1, int(self.coff[k0]), c1, int(self.uoff[k0]))

    def _keys(self, rb: np.ndarray) -> np.ndarray:
        ref = self.first_ref[rb].astype(np.int64)
        ref = np.where(ref < 0, len(self.header.names), ref)  # unplaced reads sort last
        return ref * (1 << 32) + self.first_pos[rb].astype(np.int64) + (1 << 31)

    def region(self, ref: int, beg: int, end: int) -> Span:
        """A span holding every record of a coordinate-sorted file that starts in ref:[beg, end):
        from the last record block starting at or before (ref, beg) to the first one starting
        at or after (ref, end)."""
        rb = self.record_blocks()
        key = self._keys(rb)
        lo = int(np.searchsorted(key, ref * (1 << 32) + beg + (1 << 31), side="right")) - 1
        hi = int(np.searchsorted(key, ref * (1 << 32) + end + (1 << 31), side="left"))
        lo = max(lo, 0)
        return self.span(lo, min(max(hi, lo + 1), rb.size))

    def region_of(self, sp: Span) -> tuple[int, int, int]:
        """(ref, beg, end) of the record blocks a span starts and ends at: a region that
        :meth:`region` maps back to a span holding it when the span's records lie on one
        reference (a span that runs into the next reference gets the rest of its first one)."""
        rb = self.record_blocks()
        i0 = int(np.searchsorted(rb, sp.k0))
        i1 = int(np.searchsorted(self.first_rec[rb], sp.end))
        ref = int(self.first_ref[rb[i0]])
        beg = int(self.first_pos[rb[i0]])
        if i1 < rb.size and int(self.first_ref[rb[i1]]) == ref:
            end = int(self.first_pos[rb[i1]])
        else:
            end = (1 << 31) - 1
        return ref, beg, max(end, beg + 1)

    def save(self, path) -> None:
        np.savez(path, coff=self.coff, bsize=self.bsize, uoff=self.uoff, isize=self.isize, crc=self.crc,
                 first_rec=self.first_rec, first_idx=self.first_idx, first_ref=self.first_ref,
                 first_pos=self.first_pos, n_records=self.n_records, stream_bytes=self.stream_bytes,
                 header=np.frombuffer(self.header.to_bytes(), np.uint8))

    @classmethod
    def load(cls, path) -> BlockIndex:
        z = np.load(path, allow_pickle=False)
        h = BamHeader.from_bytes(z["header"].tobytes())
        return cls(h, z["coff"], z["bsize"], z["uoff"], z["isize"], z["crc"], z["first_rec"], z["first_idx"],
                   z["first_ref"], z["first_pos"], int(z["n_records"]), int(z["stream_bytes"]))


def _le32(data: torch.Tensor, o: torch.Tensor) -> torch.Tensor:
    b = data[o.unsqueeze(1) + torch.arange(4, device=data.device)].to(torch.int64)
    v = b[:, 0] | (b[:, 1] << 8) | (b[:, 2] << 16) | (b[:, 3] << 24)
    return torch.where(v >= (1 << 31), v - (1 << 32), v).to(torch.int32)


def build_index(src, chunk_bytes: int = 256 << 20, ws: _decode.Workspace | None = None,
                on_records=None) -> BlockIndex:
    """Decode a BAM file (a path or a host byte array) chunk by chunk on the GPU and index it.

    ``on_records(data, offsets, stream_offsets, chain)``, if given, is called for each chunk's
    records (device tensors: the chunk's inflated data, the records' offsets in it and in the
    file's inflated stream), in file order."""
    raw = host_bytes(src)
    header = header_of(src)
    n_ref = len(header.names)
    dev = _decode.device()
    ws = ws or _decode.Workspace()
    size = int(raw.size)
    cols = {k: [] for k in ("coff", "bsize", "uoff", "isize", "crc")}
    firsts = []  # (block, stream offset, record index, ref, pos) of records that start a block
    starts_dev = torch.empty(0, dtype=torch.int64, device=dev)  # stream offset of every block so far
    carry = torch.empty(0, dtype=torch.uint8, device=dev)  # compressed: the incomplete last block
    rcarry = torch.empty(0, dtype=torch.uint8, device=dev)  # inflated: the incomplete last record
    pos = u_next = n_rec = 0
    skip = header.size
    last_block = -1
    while pos < size or pos == 0:
        nbytes = min(chunk_bytes, size - pos)
        eof = pos + nbytes >= size
        c = int(carry.numel())
        buf = torch.empty(c + nbytes + PAD, dtype=torch.uint8, device=dev)
        if c:
            buf[:c].copy_(carry)
        buf[c : c + nbytes].copy_(torch.from_numpy(np.asarray(raw[pos : pos + nbytes])))
        buf_off = pos - c
        blocks = _decode.scan_blocks(buf, c + nbytes, eof, ws)
        nb, p = blocks.n_blocks, int(rcarry.numel())
        data_base = u_next - p
        if nb:
            oo = blocks.out_off[:-1] + u_next
            cols["coff"].append(blocks.blk_off.cpu().numpy() + buf_off)
            cols["bsize"].append(blocks.bsize.cpu().numpy())
            cols["isize"].append(blocks.isize.cpu().numpy())
            cols["crc"].append(blocks.crc.cpu().numpy())
            cols["uoff"].append(oo.cpu().numpy())
            starts_dev = torch.cat([starts_dev, oo])
            inf = _decode.inflate_blocks(blocks, rcarry if p else None, ws)
            data, n = inf.data, inf.n
            u_next += blocks.total_out
        else:
            inf = None
            data = torch.empty(p + PAD, dtype=torch.uint8, device=dev)
            if p:
                data[:p].copy_(rcarry)
            n = p
        start = 0
        if skip:
            start = min(skip, n)
            skip -= start
        chain = _decode.record_chain(data, start, n, n_ref, ws, inf=inf)
        if chain.n_rec:
            off = _decode.record_offsets(data, start, n, chain)
            stream = off + data_base
            if on_records is not None:
                on_records(data, off, stream, chain)
            blk = torch.searchsorted(starts_dev, stream, right=True) - 1
            new = torch.ones_like(blk, dtype=torch.bool)
            new[1:] = blk[1:] != blk[:-1]
            new[0] = blk[0] != last_block
            sel = torch.nonzero(new).flatten()
            o = off[sel]
            firsts.append(np.stack([blk[sel].cpu().numpy(), stream[sel].cpu().numpy(),
                                    (sel + n_rec).cpu().numpy(), _le32(data, o + 4).cpu().numpy(),
                                    _le32(data, o + 8).cpu().numpy()]))
            last_block = int(blk[-1])
            n_rec += chain.n_rec
        rcarry = data[chain.tail : n].clone()
        carry = buf[blocks.used : c + nbytes].clone()
        del buf, data
        pos += nbytes
        if eof:
            break
    if rcarry.numel():
        raise ValueError("truncated BAM file (incomplete record at the end)")
    cat = {k: (np.concatenate(v) if v else np.empty(0)) for k, v in cols.items()}
    nb = int(cat["coff"].size)
    first_rec = np.full(nb, -1, np.int64)
    first_idx = np.full(nb, -1, np.int64)
    first_ref = np.full(nb, -1, np.int32)
    first_pos = np.full(nb, -1, np.int32)
    if firsts:
        f = np.concatenate(firsts, axis=1)
        b = f[0]
        first_rec[b], first_idx[b], first_ref[b], first_pos[b] = f[1], f[2], f[3], f[4]
    return BlockIndex(header, cat["coff"].astype(np.int64), cat["bsize"].astype(np.int32),
                      cat["uoff"].astype(np.int64), cat["isize"].astype(np.int32), cat["crc"].astype(np.uint32),
                      first_rec, first_idx, first_ref, first_pos, n_rec, u_next)
