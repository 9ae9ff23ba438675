"""Seeded synthetic BAM and BGZF workloads (host C generator in ``gen.c``, zlib).

``make_bam`` writes the GEX-like (kind 0) or ATAC-like (kind 1) file described
in ``gen.c`` as BGZF: the header in blocks of its own, then the records cut
into payloads of at most 65,280 bytes (``BLOCK_DATA``), so records cross block
seams, then the EOF block. ``hetero_blocks`` cuts payloads of random size out
of an ATAC record stream and compresses each at its own zlib level, some with
fixed Huffman codes only.
"""

from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

#: The empty BGZF block that ends a file (SAMv1 section 4.1.2)
EOF_BLOCK = bytes.fromhex("1f8b08040000000000ff0600424302001b0003000000000000000000")
#: Payload bytes per BGZF block, as htslib writes
BLOCK_DATA = 0xFF00

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "libcbm_synth.so"
_lib = None


def build(force: bool = False) -> Path:
    """Compile ``gen.c`` into ``libcbm_synth.so`` next to it."""
    if force or not _LIB_PATH.exists() or _LIB_PATH.stat().st_mtime < (_HERE / "gen.c").stat().st_mtime:
        import subprocess

        tmp = str(_LIB_PATH) + f".tmp{os.getpid()}"
        subprocess.run(
            ["gcc", "-O2", "-shared", "-fPIC", "-o", tmp, str(_HERE / "gen.c"), "-lz", "-lpthread", "-lm"],
            check=True,
        )
        os.replace(tmp, _LIB_PATH)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(_LIB_PATH))
        vp, i64, u64 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_uint64
        L.cbs_open.restype = vp
        L.cbs_open.argtypes = [ctypes.c_int, u64, i64]
        L.cbs_close.argtypes = [vp]
        L.cbs_header.restype = i64
        L.cbs_header.argtypes = [vp, vp, i64]
        L.cbs_records.restype = i64
        L.cbs_records.argtypes = [vp, vp, i64, i64, ctypes.POINTER(i64)]
        L.cbs_contig_length.restype = i64
        L.cbs_contig_length.argtypes = [vp, ctypes.c_int]
        L.cbs_bgzf_blocks_codec.restype = ctypes.c_int
        L.cbs_bgzf_blocks_codec.argtypes = [vp, vp, vp, i64, vp, vp, vp, i64, vp, ctypes.c_int, ctypes.c_int]
        L.cbs_have_libdeflate.restype = ctypes.c_int
        L.cbs_have_libdeflate.argtypes = []
        _lib = L
    return _lib


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data


def default_threads() -> int:
    return max(1, min(16, os.cpu_count() or 1))


#: compressors of ``bgzf_blocks``: zlib, or libdeflate at the same level (about three times faster
#: on these records, compressed sizes within 1 % of zlib's; levels 1-9 without fixed codes only)
CODECS = {"zlib": 0, "libdeflate": 1}


def have_libdeflate() -> bool:
    """Whether ``codec="libdeflate"`` can be used (its shared library is present)."""
    return bool(lib().cbs_have_libdeflate())


class RecordStream:
    """Records of one synthetic file, produced in order (kind 0 GEX-like, kind 1 ATAC-like)."""

    def __init__(self, kind: int, seed: int, n_total: int):
        self.kind, self.seed, self.n_total = kind, seed, n_total
        self._h = lib().cbs_open(kind, seed, n_total)
        if not self._h:
            raise MemoryError("generator state")

    def header(self) -> bytes:
        buf = np.empty(1 << 16, np.uint8)
        n = lib().cbs_header(self._h, _ptr(buf), buf.size)
        if n < 0:
            raise RuntimeError("header buffer too small")
        return buf[:n].tobytes()

    def records(self, max_bytes: int, max_records: int = 1 << 62) -> tuple[np.ndarray, int]:
        """Whole records filling at most ``max_bytes`` bytes; (bytes, count). Empty at the end."""
        buf = np.empty(max_bytes, np.uint8)
        k = ctypes.c_int64()
        n = lib().cbs_records(self._h, _ptr(buf), max_bytes, max_records, ctypes.byref(k))
        return buf[:n], int(k.value)

    def contig_lengths(self) -> list[int]:
        return [int(lib().cbs_contig_length(self._h, c)) for c in range(25)]

    def close(self) -> None:
        if self._h and _lib is not None:
            _lib.cbs_close(self._h)
        self._h = None

    def __del__(self):
        self.close()


def bgzf_blocks(
    data: np.ndarray,
    off: np.ndarray,
    size: np.ndarray,
    level: np.ndarray | int = 6,
    fixed: np.ndarray | None = None,
    n_threads: int | None = None,
    codec: str = "zlib",
) -> tuple[np.ndarray, np.ndarray]:
    """BGZF blocks of the payloads ``data[off[i] : off[i] + size[i]]``; (bytes, block sizes)."""
    off = np.ascontiguousarray(off, np.int64)
    size = np.ascontiguousarray(size, np.int32)
    n = int(off.size)
    if n == 0:
        return np.empty(0, np.uint8), np.empty(0, np.int32)
    if int(size.max()) > BLOCK_DATA:
        raise ValueError("BGZF payloads hold at most 65,280 bytes")
    lv = np.ascontiguousarray(np.broadcast_to(np.asarray(level, np.int8), (n,)))
    fx = None if fixed is None else np.ascontiguousarray(fixed, np.uint8)
    stride = 1 << 16
    batch = min(n, 16384)  # blocks compressed per call: a bounded staging buffer
    tmp = np.empty(batch * stride, np.uint8)
    bsize = np.empty(n, np.int32)
    data = np.ascontiguousarray(data, np.uint8)
    parts = []
    for b0 in range(0, n, batch):
        b1 = min(n, b0 + batch)
        rc = lib().cbs_bgzf_blocks_codec(
            _ptr(data), _ptr(off[b0:b1]), _ptr(size[b0:b1]), b1 - b0, _ptr(lv[b0:b1]),
            None if fx is None else _ptr(fx[b0:b1]), _ptr(tmp), stride, _ptr(bsize[b0:b1]),
            n_threads or default_threads(), CODECS[codec],
        )
        if rc == -3:
            raise RuntimeError("libdeflate is not available")
        if rc != 0:
            raise RuntimeError(f"{codec} failed")
        bs = bsize[b0:b1]
        part = np.empty(int(bs.sum(dtype=np.int64)), np.uint8)
        pos = 0
        for i in range(b1 - b0):
            k = int(bs[i])
            part[pos : pos + k] = tmp[i * stride : i * stride + k]
            pos += k
        parts.append(part)
    out = np.concatenate(parts) if len(parts) > 1 else parts[0]
    return out, bsize


def _cut(n: int, block: int) -> tuple[np.ndarray, np.ndarray]:
    off = np.arange(0, n, block, dtype=np.int64)
    return off, (np.minimum(off + block, n) - off).astype(np.int32)


@dataclass
class SynthBam:
    """A synthetic BAM in memory: its BGZF bytes and what it holds."""

    bgzf: np.ndarray  # uint8, the whole file
    header: bytes  # uncompressed header
    n_records: int
    raw_bytes: int  # uncompressed bytes of the records
    block_sizes: np.ndarray = field(repr=False)  # every block, EOF included
    contig_lengths: list[int] = field(default_factory=list)

    def write(self, path: str | Path) -> Path:
        p = Path(path)
        self.bgzf.tofile(p)
        return p


def make_bam(
    kind: int,
    seed: int,
    n_records: int,
    *,
    n_total: int | None = None,
    level: int = 6,
    block: int = BLOCK_DATA,
    n_threads: int | None = None,
    eof: bool = True,
    piece_bytes: int = 256 << 20,
    codec: str = "zlib",
) -> SynthBam:
    """The first ``n_records`` records of a synthetic file of ``n_total`` records, as BGZF.

    ``kind`` 0 is GEX-like single-end 90 bp, 1 ATAC-like paired-end 2 x 50 bp. ``codec`` names the
    compressor (:data:`CODECS`).
    """
    rs = RecordStream(kind, seed, n_total or n_records)
    try:
        header = rs.header()
        lens = rs.contig_lengths()
        hoff, hsize = _cut(len(header), block)
        parts, sizes = [], []
        hb, hs = bgzf_blocks(np.frombuffer(header, np.uint8), hoff, hsize, level, n_threads=n_threads, codec=codec)
        parts.append(hb)
        sizes.append(hs)
        done, raw = 0, 0
        carry = np.empty(0, np.uint8)
        while done < n_records:
            recs, k = rs.records(piece_bytes, n_records - done)
            if k == 0:
                break
            done += k
            raw += int(recs.size)
            buf = np.concatenate([carry, recs]) if carry.size else recs
            last = done >= n_records
            full = buf.size if last else buf.size // block * block
            off, size = _cut(full, block)
            b, s = bgzf_blocks(buf, off, size, level, n_threads=n_threads, codec=codec)
            parts.append(b)
            sizes.append(s)
            carry = buf[full:].copy()
        if carry.size:
            off, size = _cut(carry.size, block)
            b, s = bgzf_blocks(carry, off, size, level, n_threads=n_threads, codec=codec)
            parts.append(b)
            sizes.append(s)
        if eof:
            parts.append(np.frombuffer(EOF_BLOCK, np.uint8))
            sizes.append(np.asarray([len(EOF_BLOCK)], np.int32))
        return SynthBam(np.concatenate(parts), header, done, raw, np.concatenate(sizes), lens)
    finally:
        rs.close()


@dataclass
class HeteroBlocks:
    """BGZF blocks of mixed size and level, and the payloads they hold."""

    bgzf: np.ndarray  # uint8, the blocks back to back (no EOF block)
    block_off: np.ndarray  # int64, start of each block in bgzf
    block_size: np.ndarray  # int32
    raw: np.ndarray  # uint8, the payloads back to back
    raw_off: np.ndarray  # int64
    raw_size: np.ndarray  # int32
    level: np.ndarray  # int8
    fixed: np.ndarray  # uint8


def hetero_blocks(
    n_blocks: int, seed: int = 2, *, max_size: int = BLOCK_DATA, n_threads: int | None = None
) -> HeteroBlocks:
    """Payloads of uniform 1..``max_size`` bytes cut from an ATAC record stream, each compressed at a
    random zlib level in {0, 1, 6, 9}; 10 % of the compressed ones use fixed Huffman codes only."""
    rng = np.random.default_rng(seed)
    size = rng.integers(1, max_size + 1, n_blocks).astype(np.int32)
    level = rng.choice(np.asarray([0, 1, 6, 9], np.int8), n_blocks)
    fixed = ((rng.random(n_blocks) < 0.10) & (level > 0)).astype(np.uint8)
    total = int(size.sum(dtype=np.int64))
    rs = RecordStream(1, seed, max(1000, total // 150 + 1000))
    try:
        chunks, have = [], 0
        while have < total:
            recs, k = rs.records(min(256 << 20, total - have + (1 << 20)))
            if k == 0:
                break
            chunks.append(recs)
            have += int(recs.size)
    finally:
        rs.close()
    raw = np.concatenate(chunks)[:total]
    raw_off = np.zeros(n_blocks, np.int64)
    np.cumsum(size[:-1], out=raw_off[1:])
    blob, bsize = bgzf_blocks(raw, raw_off, size, level, fixed, n_threads)
    boff = np.zeros(n_blocks, np.int64)
    np.cumsum(bsize[:-1].astype(np.int64), out=boff[1:])
    return HeteroBlocks(blob, boff, bsize, raw, raw_off, size, level, fixed)
