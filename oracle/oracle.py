"""The oracle: a CPU restatement of the reference's read path (test infrastructure only).

Only tests, the smoke check and the CPU-baseline leg of the benchmark use this
module, and only as the checker: the product path (``cubam_amd``) never
imports it. Numerics live in ``cpu_decode.c`` (plain C + zlib 1.2.11); this
file holds the ctypes bindings and the chunk loop with its two carries.

Reference behaviour restated here (paths relative to the reference repo):
  * chunk loop, compressed carry and errors: src/cubam/bgzf.py:542-652
  * header parse and skip, record carry, truncation error: src/cubam/bam.py:41-72, 421-446
  * record chain: src/cubam/bam.py:403-418, src/cubam/_kernels.py:250-289, 676-702
  * fixed columns: src/cubam/_kernels.py:292-319; tags: src/cubam/_kernels.py:334-388

Pinned by tests/test_oracle.py against zlib/gzip, the record lists of an
independent pure-Python BAM writer (tests/bamfixture.py), known-answer
vectors (EOF block, reg2bin, FNV-1a, pack_acgt) and the golden fixtures in
tests/golden/.
"""

from __future__ import annotations

import ctypes
import gzip
import io
import os
import struct
import subprocess
from dataclasses import dataclass
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "liborc.so"
_lib = None

MSG_TRUNCATED = "Truncated BGZF file: the last block is incomplete"
MSG_CORRUPT = "Corrupt BGZF block"
MSG_SIZE = "Corrupt BGZF block: it does not inflate to the size in its trailer"
MSG_CRC = "Corrupt BGZF block: its CRC32 does not match the data"
MSG_NO_BLOCK = "No valid block or record where one was expected: the data is corrupt"
MSG_TOO_MANY = "Too many candidates: the data is corrupt"

FIELDS = ("ref", "pos", "end", "mapq", "flag", "nref", "npos", "tlen", "l_seq", "clip_left", "clip_right")


def build(force: bool = False) -> Path:
    src = _HERE / "cpu_decode.c"
    if force or not _LIB_PATH.exists() or _LIB_PATH.stat().st_mtime < src.stat().st_mtime:
        tmp = str(_LIB_PATH) + f".tmp{os.getpid()}"
        subprocess.run(
            ["gcc", "-O2", "-shared", "-fPIC", "-o", tmp, str(src), "-lz", "-lpthread"], check=True
        )
        os.replace(tmp, _LIB_PATH)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(_LIB_PATH))
        vp, i64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int
        L.orc_bgzf_scan.argtypes = [vp, i64, i32, vp, vp, vp, vp, i64, vp]
        L.orc_inflate.argtypes = [vp, vp, vp, vp, vp, i64, vp, vp, vp, i32]
        L.orc_crc32.restype = ctypes.c_uint32
        L.orc_crc32.argtypes = [vp, i64]
        L.orc_reg2bin.restype = i32
        L.orc_reg2bin.argtypes = [i32, i32]
        L.orc_is_record.restype = i32
        L.orc_is_record.argtypes = [vp, i64, i64, i32]
        L.orc_record_chain.argtypes = [vp, i64, i32, vp, i64, vp]
        L.orc_record_candidates.restype = i64
        L.orc_record_candidates.argtypes = [vp, i64, i32]
        L.orc_var_sizes.argtypes = [vp, vp, i64, vp, vp]
        L.orc_decode_columns.argtypes = [vp, vp, i64] + [vp] * 17
        L.orc_find_tag.argtypes = [vp, vp, i64, ctypes.c_uint8, ctypes.c_uint8, vp, vp, vp, vp]
        L.orc_fnv1a.restype = ctypes.c_uint64
        L.orc_fnv1a.argtypes = [vp, i64]
        L.orc_pack_acgt.restype = ctypes.c_uint64
        L.orc_pack_acgt.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_int)]
        L.orc_columns_mt.argtypes = [vp, vp, i64] + [vp] * 17 + [ctypes.c_char_p, i32, vp, vp, vp, vp, i32]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def _u8(x) -> np.ndarray:
    if isinstance(x, np.ndarray):
        return np.ascontiguousarray(x, np.uint8)
    return np.frombuffer(bytes(x), np.uint8)


# ------------------------------------------------------------------ BGZF


def bgzf_scan(buf, eof: bool):
    """Blocks of ``buf`` (bgzf.py:542-584): (blk_off, bsize, isize, crc, used, status)."""
    b = _u8(buf)
    cap = b.size // 26 + 2
    off = np.empty(cap, np.int64)
    bs = np.empty(cap, np.int32)
    isz = np.empty(cap, np.int32)
    crc = np.empty(cap, np.uint32)
    out = np.zeros(4, np.int64)
    lib().orc_bgzf_scan(_p(b), b.size, int(eof), _p(off), _p(bs), _p(isz), _p(crc), cap, _p(out))
    k = int(out[0])
    return off[:k], bs[:k], isz[:k], crc[:k], int(out[1]), int(out[2])


def inflate(buf, blk_off, bsize, isize, n_threads: int = 1):
    """Inflate blocks with zlib: (data, out_off, crc, status)."""
    b = _u8(buf)
    blk_off = np.ascontiguousarray(blk_off, np.int64)
    bsize = np.ascontiguousarray(bsize, np.int32)
    isize = np.ascontiguousarray(isize, np.int32)
    n = blk_off.size
    out_off = np.zeros(n, np.int64)
    if n > 1:
        np.cumsum(isize[:-1].astype(np.int64), out=out_off[1:])
    total = int(isize.astype(np.int64).sum())
    dst = np.empty(max(total, 1), np.uint8)
    crc = np.empty(n, np.uint32)
    st = np.empty(n, np.int32)
    if n:
        lib().orc_inflate(_p(b), _p(blk_off), _p(bsize), _p(isize), _p(out_off), n, _p(dst), _p(crc), _p(st),
                          n_threads)
    return dst[:total], out_off, crc, st


def crc32(data) -> int:
    b = _u8(data)
    return int(lib().orc_crc32(_p(b), b.size))


def _inflate_blocks(buf, eof: bool, prefix: np.ndarray, n_threads: int = 1):
    """bgzf.py:542-611 on the CPU: (data = prefix + inflated blocks, used)."""
    off, bs, isz, crc, used, status = bgzf_scan(buf, eof)
    if status == 1:
        raise ValueError(MSG_TRUNCATED)
    if status == 2:
        raise ValueError(MSG_NO_BLOCK)
    if status == 3:
        raise ValueError(MSG_CORRUPT)
    if status != 0:
        raise ValueError("Too many candidates: the data is corrupt")
    if off.size == 0:
        return prefix, used
    data, _, got_crc, st = inflate(buf, off, bs, isz, n_threads)
    if np.any((st == 1) | (st == 2)):
        raise ValueError(MSG_SIZE)
    if np.any(st == 3):
        raise ValueError(MSG_CRC)
    return (np.concatenate([prefix, data]) if prefix.size else data), used


def inflate_chunks(file_bytes, chunk_bytes: int, n_threads: int = 1):
    """bgzf.py:615-652: inflated data chunk by chunk (a list of uint8 arrays)."""
    return list(_inflate_iter(_u8(file_bytes), chunk_bytes, None, n_threads))


def _inflate_iter(file_bytes: np.ndarray, chunk_bytes: int, prefix, n_threads: int = 1):
    n = file_bytes.size
    carry = np.empty(0, np.uint8)
    pos = 0
    while True:
        take = min(chunk_bytes, n - pos)
        piece = file_bytes[pos : pos + take]
        pos += take
        eof = pos >= n
        buf = np.concatenate([carry, piece]) if carry.size else piece
        pre = prefix() if prefix is not None else np.empty(0, np.uint8)
        data, used = _inflate_blocks(buf, eof, pre, n_threads)
        carry = buf[used:].copy()
        yield data
        if eof:
            return


# ------------------------------------------------------------------ BAM


@dataclass(frozen=True)
class Header:
    text: str
    names: list
    lengths: list
    size: int


def read_header(file_bytes) -> Header:
    """bam.py:50-72: the SAM header, references and header size."""
    fh = gzip.GzipFile(fileobj=io.BytesIO(bytes(_u8(file_bytes))))

    def take(n):
        b = fh.read(n)
        if len(b) != n:
            raise ValueError("truncated BAM header")
        return b

    if take(4) != b"BAM\1":
        raise ValueError("not a BAM file")
    (l_text,) = struct.unpack("<i", take(4))
    text = take(l_text).split(b"\0", 1)[0].decode()
    (n_ref,) = struct.unpack("<i", take(4))
    names, lengths, size = [], [], 12 + l_text
    for _ in range(n_ref):
        (ln,) = struct.unpack("<i", take(4))
        names.append(take(ln)[:-1].decode())
        (lr,) = struct.unpack("<i", take(4))
        lengths.append(lr)
        size += 8 + ln
    return Header(text, names, lengths, size)


def record_chain(data, n_ref: int):
    """bam.py:403-418: (offsets, tail, status)."""
    b = _u8(data)
    cap = b.size // 36 + 1
    off = np.empty(cap, np.int64)
    out = np.zeros(3, np.int64)
    lib().orc_record_chain(_p(b), b.size, n_ref, _p(off), cap, _p(out))
    return off[: int(out[0])].copy(), int(out[1]), int(out[2])


def record_candidates(data, n_ref: int) -> int:
    """Offsets of ``data`` where a valid record starts (brute force; _kernels.py:280-284)."""
    b = _u8(data)
    return int(lib().orc_record_candidates(_p(b), b.size, n_ref))


def _excl(x: np.ndarray) -> np.ndarray:
    o = np.zeros(x.size + 1, np.int64)
    np.cumsum(x.astype(np.int64), out=o[1:])
    return o


def columns(data, offsets, tags=("CB", "UB"), n_threads: int = 1) -> dict:
    """Every column of the records at ``offsets``: fixed fields, CSR CIGAR/seq/qual, tags."""
    b = _u8(data)
    off = np.ascontiguousarray(offsets, np.int64)
    n = off.size
    ncig = np.empty(n, np.int32)
    lseq = np.empty(n, np.int32)
    lib().orc_var_sizes(_p(b), _p(off), n, _p(ncig), _p(lseq))
    cig_off = _excl(ncig)
    seq_off = _excl((lseq.astype(np.int64) + 1) // 2)
    qual_off = _excl(lseq)
    c = {
        "ref": np.empty(n, np.int32), "pos": np.empty(n, np.int32), "end": np.empty(n, np.int32),
        "mapq": np.empty(n, np.uint8), "flag": np.empty(n, np.uint16), "nref": np.empty(n, np.int32),
        "npos": np.empty(n, np.int32), "tlen": np.empty(n, np.int32), "l_seq": np.empty(n, np.int32),
        "clip_left": np.empty(n, np.uint16), "clip_right": np.empty(n, np.uint16),
    }
    cigar = np.empty(max(int(cig_off[-1]), 1), np.uint32)
    seq = np.empty(max(int(seq_off[-1]), 1), np.uint8)
    qual = np.empty(max(int(qual_off[-1]), 1), np.uint8)
    nt = len(tags)
    t_type = np.zeros(nt * n, np.uint8)
    t_off = np.zeros(nt * n, np.int64)
    t_len = np.zeros(nt * n, np.int32)
    t_val = np.zeros(nt * n, np.int64)
    if n:
        mt = min(nt, 2)  # the threaded filler takes two tags; more go through orc_find_tag
        lib().orc_columns_mt(
            _p(b), _p(off), n, *(_p(c[k]) for k in FIELDS), _p(cig_off), _p(cigar), _p(seq_off), _p(seq),
            _p(qual_off), _p(qual), "".join(tags[:mt]).encode(), mt, _p(t_type), _p(t_off), _p(t_len), _p(t_val),
            n_threads,
        )
        for k in range(mt, nt):
            s = slice(k * n, (k + 1) * n)
            ty, vo, vl, iv = find_tag(b, off, tags[k])
            t_type[s], t_off[s], t_len[s], t_val[s] = ty, vo, vl, iv
    c["offset"] = off
    c["cigar_off"], c["cigar"] = cig_off, cigar[: int(cig_off[-1])]
    c["seq_off"], c["seq"] = seq_off, seq[: int(seq_off[-1])]
    c["qual_off"], c["qual"] = qual_off, qual[: int(qual_off[-1])]
    for k, t in enumerate(tags):
        s = slice(k * n, (k + 1) * n)
        c[f"tag_{t}"] = {"type": t_type[s], "offset": t_off[s], "length": t_len[s], "value": t_val[s]}
    return c


def find_tag(data, offsets, name: str):
    b = _u8(data)
    off = np.ascontiguousarray(offsets, np.int64)
    n = off.size
    ty, vo, vl, iv = np.zeros(n, np.uint8), np.zeros(n, np.int64), np.zeros(n, np.int32), np.zeros(n, np.int64)
    if n:
        lib().orc_find_tag(_p(b), _p(off), n, ord(name[0]), ord(name[1]), _p(ty), _p(vo), _p(vl), _p(iv))
    return ty, vo, vl, iv


def read_bam(file_bytes, chunk_bytes: int, tags=("CB", "UB"), n_threads: int = 1, name: str = "BAM"):
    """bam.py:421-446 on the CPU: a list of (data, columns) per chunk."""
    return list(iter_bam(file_bytes, chunk_bytes, tags, n_threads, name))


def iter_bam(file_bytes, chunk_bytes: int, tags=("CB", "UB"), n_threads: int = 1, name: str = "BAM"):
    """bam.py:421-446 on the CPU, as a generator of (data, columns) per chunk: the chunks before an
    error are yielded before it raises, as the reference's generator does."""
    fb = _u8(file_bytes)
    header = read_header(fb)
    n_ref = len(header.names)
    skip = header.size
    carry = [np.empty(0, np.uint8)]
    for data in _inflate_iter(fb, chunk_bytes, lambda: carry[0], n_threads):
        if skip:
            k = min(skip, data.size)
            data = data[k:]
            skip -= k
        off, tail, status = record_chain(data, n_ref)
        if off.size == 0 and data.size:
            # no record at offset 0: bam.py:410-412 and _kernels.py:668-672, 687-689
            cand = record_candidates(data, n_ref)
            if cand > data.size // 32 + 1024:
                raise ValueError(MSG_TOO_MANY)
            if cand:
                raise ValueError(MSG_NO_BLOCK)
        carry[0] = data[tail:].copy()
        if off.size:
            yield data, columns(data, off, tags, n_threads)
    if carry[0].size:
        raise ValueError(f"{name}: truncated BAM file (incomplete record at the end)")


def fnv1a(data) -> int:
    b = _u8(data)
    return int(lib().orc_fnv1a(_p(b), b.size))


def pack_acgt(value: bytes) -> tuple[int, bool]:
    b = _u8(value)
    ok = ctypes.c_int()
    code = lib().orc_pack_acgt(_p(b) if b.size else None, b.size, ctypes.byref(ok))
    return int(code), bool(ok.value)
