"""Device-side steps of the read path: block scan, inflate + CRC32, record chain, columns.

Each function enqueues HIP work on the current torch stream through the C ABI and
synchronises only where a size must reach the host (block count and output size
after the scan; record count and CSR sizes after the chain). Error messages are
those of the reference (src/cubam/bgzf.py:557-611, src/cubam/_kernels.py:689,
src/cubam/bam.py:446), plus one for a CRC32 mismatch, which it does not check.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import torch

from cubam_amd import _lib
from cubam_amd._lib import PAD, check, ptr, stream_ptr

MSG_TRUNCATED = "Truncated BGZF file: the last block is incomplete"
MSG_CORRUPT = "Corrupt BGZF block"
MSG_SIZE = "Corrupt BGZF block: it does not inflate to the size in its trailer"
MSG_CRC = "Corrupt BGZF block: its CRC32 does not match the data"
MSG_NO_BLOCK = "No valid block or record where one was expected: the data is corrupt"

_U8, _I32, _I64 = torch.uint8, torch.int32, torch.int64


def device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("cubam_amd needs a GPU (HIP device); none is visible")
    return torch.device("cuda", torch.cuda.current_device())


class Workspace:
    """Grow-only device scratch buffers and pinned host result slots, one set per stream user."""

    def __init__(self):
        self._dev: dict[str, torch.Tensor] = {}
        self._host: dict[str, torch.Tensor] = {}

    def dev(self, key: str, nbytes: int) -> torch.Tensor:
        t = self._dev.get(key)
        if t is None or t.numel() < nbytes:
            t = torch.empty(max(nbytes, 64) + (nbytes >> 3), dtype=_U8, device=device())
            self._dev[key] = t
        return t

    def host(self, key: str, n: int) -> torch.Tensor:
        t = self._host.get(key)
        if t is None or t.numel() < n:
            t = torch.empty(n, dtype=_I64, pin_memory=True)
            self._host[key] = t
        return t


_default_ws: Workspace | None = None


def default_workspace() -> Workspace:
    global _default_ws
    if _default_ws is None:
        _default_ws = Workspace()
    return _default_ws


def padded(n: int) -> int:
    return n + PAD


def to_device(host_u8, dev: torch.device | None = None) -> torch.Tensor:
    """A padded device copy of a host byte buffer (numpy/bytes/tensor)."""
    import numpy as np

    if isinstance(host_u8, (bytes, bytearray, memoryview)):
        host_u8 = np.frombuffer(bytes(host_u8), np.uint8)
    if not isinstance(host_u8, torch.Tensor):
        host_u8 = torch.from_numpy(np.require(host_u8, np.uint8, ["C", "W"]))
    n = host_u8.numel()
    out = torch.empty(padded(n), dtype=_U8, device=dev or device())
    out[:n].copy_(host_u8, non_blocking=False)
    return out


@dataclass
class Blocks:
    """The BGZF blocks found in a compressed buffer (device arrays of n entries)."""

    buf: torch.Tensor  # the compressed buffer (padded)
    n_blocks: int
    used: int
    total_out: int
    blk_off: torch.Tensor
    bsize: torch.Tensor
    isize: torch.Tensor
    crc: torch.Tensor
    out_off: torch.Tensor  # n + 1 entries
    n_cand: int = 0  # header candidates the scan found (more than n_blocks: some were not on the chain)


def scan_blocks(buf: torch.Tensor, n: int, eof: bool, ws: Workspace | None = None) -> Blocks:
    """Kernel 1 over buf[0:n): blocks of the chain from offset 0 (synchronises once)."""
    L = _lib.load()
    ws = ws or default_workspace()
    st = stream_ptr()
    cap = n // 2048 + 64
    while True:
        dev = buf.device
        blk_off = torch.empty(cap, dtype=_I64, device=dev)
        bsize = torch.empty(cap, dtype=_I32, device=dev)
        isize = torch.empty(cap, dtype=_I32, device=dev)
        crc = torch.empty(cap, dtype=torch.uint32, device=dev)
        out_off = torch.empty(cap + 1, dtype=_I64, device=dev)
        result = ws.dev("scan_result", 64)
        wsb = L.cbm_bgzf_scan_workspace(n)
        sws = ws.dev("scan_ws", wsb)
        check(
            L.cbm_bgzf_scan(
                ptr(buf), n, int(eof), cap, ptr(blk_off), ptr(bsize), ptr(isize), ptr(crc), ptr(out_off),
                ptr(result), ptr(sws), sws.numel(), st,
            ),
            "cbm_bgzf_scan",
        )
        h = ws.host("scan_host", 8)
        h.copy_(result[:64].view(_I64), non_blocking=True)
        torch.cuda.current_stream().synchronize()
        status = int(h[_lib.SCAN_STATUS])
        if status == _lib.SCAN_CAP:
            cap = int(h[_lib.SCAN_NCAND]) + 64
            continue
        break
    if status == _lib.SCAN_TRUNCATED:
        raise ValueError(MSG_TRUNCATED)
    if status == _lib.SCAN_NO_BLOCK:
        raise ValueError(MSG_NO_BLOCK)
    if status == _lib.SCAN_CORRUPT:
        raise ValueError(MSG_CORRUPT)
    nb = int(h[_lib.SCAN_NBLOCKS])
    return Blocks(buf, nb, int(h[_lib.SCAN_USED]), int(h[_lib.SCAN_TOTAL_OUT]), blk_off[:nb], bsize[:nb],
                  isize[:nb], crc[:nb], out_off[: nb + 1], int(h[_lib.SCAN_NCAND]))


_tune: _lib.InflateTune | None = None  # None: the autotuned shape, else the library default


def set_inflate_tune(variant: int = 0, blocks_per_wave: int = 0, waves_per_group: int = 0, groups: int = 0) -> None:
    """Launch shape of the inflate kernel for this process; all zeros clear it (the autotuned
    shape where one is cached, else the library default)."""
    global _tune
    shape = (variant, blocks_per_wave, waves_per_group, groups)
    _tune = _lib.InflateTune(*shape) if any(shape) else None


def get_inflate_tune() -> tuple[int, int, int, int]:
    t = _tune
    return (0, 0, 0, 0) if t is None else (t.variant, t.blocks_per_wave, t.waves_per_group, t.groups)


def inflate_tune_for(n_blocks: int, inflated: int, tune: _lib.InflateTune | None = None) -> _lib.InflateTune | None:
    """The launch shape of an inflate of this size: ``tune``, else the process's, else the
    autotuned one cached for the size, else None (the library default)."""
    t = tune if tune is not None else _tune
    if t is None:
        from cubam_amd import tune as autotuner

        if autotuner.enabled():
            t = autotuner.lookup(n_blocks, inflated)
    return t


def tune_arg(t: _lib.InflateTune | None):
    """The cbm_inflate tune argument: a pointer to ``t``, or null for the library default."""
    return None if t is None else ctypes.byref(t)


@dataclass
class Inflated:
    data: torch.Tensor  # prefix + inflated bytes, padded; valid length `n`
    n: int
    crc: torch.Tensor
    status: torch.Tensor
    result: torch.Tensor  # device int64[1]: blocks not OK


def inflate_blocks(blocks: Blocks, prefix: torch.Tensor | None = None, ws: Workspace | None = None,
                   tune: _lib.InflateTune | None = None) -> Inflated:
    """Kernel 2: inflate every block after `prefix` (asynchronous; check with `check_inflated`)."""
    L = _lib.load()
    ws = ws or default_workspace()
    p = 0 if prefix is None else int(prefix.numel())
    n = p + blocks.total_out
    dev = blocks.buf.device
    data = torch.empty(padded(n), dtype=_U8, device=dev)
    if p:
        data[:p].copy_(prefix)
    nb = blocks.n_blocks
    crc = torch.empty(max(nb, 1), dtype=torch.uint32, device=dev)
    status = torch.empty(max(nb, 1), dtype=_I32, device=dev)
    result = torch.empty(1, dtype=_I64, device=dev)
    t = inflate_tune_for(nb, blocks.total_out, tune)
    iws = ws.dev("inflate_ws", L.cbm_inflate_workspace(nb))
    check(
        L.cbm_inflate(
            ptr(blocks.buf), ptr(blocks.blk_off), ptr(blocks.bsize), ptr(blocks.isize), ptr(blocks.crc),
            ptr(blocks.out_off), nb, data.data_ptr() + p, ptr(crc), ptr(status), ptr(result),
            tune_arg(t),
            ptr(iws), iws.numel(), stream_ptr(),
        ),
        "cbm_inflate",
    )
    return Inflated(data, n, crc[:nb], status[:nb], result)


def raise_block_errors(inf: Inflated, bad: int) -> None:
    if bad:
        st = inf.status.cpu() & 0xFF  # bits 8-15 name the check that failed
        if bool(((st == _lib.BLOCK_DEFLATE) | (st == _lib.BLOCK_SIZE)).any()):
            raise ValueError(MSG_SIZE)
        raise ValueError(MSG_CRC)


def check_inflated(inf: Inflated) -> None:
    """Synchronise and raise on any block that did not inflate cleanly."""
    raise_block_errors(inf, int(inf.result.item()))


@dataclass
class Chain:
    n_rec: int
    tail: int  # where the chain stopped: the end of the data, an incomplete record or an invalid one
    n_cigar: int
    n_seq: int
    n_qual: int
    ws: torch.Tensor
    invalid: bool = False  # the chain stopped at a record that fails the structural test
    bad_links: int = 0  # segments whose speculative walks did not link (repaired before returning)


def record_chain(data: torch.Tensor, start: int, n: int, n_ref: int, ws: Workspace | None = None,
                 inf: Inflated | None = None, strict: bool = True) -> Chain:
    """Kernel 3: the record chain of data[start:n) (synchronises once).

    A record that fails the structural test ends the chain there. With ``strict`` that raises
    (a region span must hold whole records); otherwise the chain up to it is returned with
    ``invalid`` set, and ``tail`` is where the invalid record starts, as the reference's
    pointer-jumping chain keeps the records before it (src/cubam/_kernels.py:676-702).
    """
    L = _lib.load()
    ws = ws or default_workspace()
    st = stream_ptr()
    wsb = L.cbm_record_chain_workspace(max(n - start, 0), None)
    rws = ws.dev("rec_ws", wsb)
    result = ws.dev("rec_result", 64)
    check(L.cbm_record_chain(ptr(data), start, n, n_ref, ptr(result), ptr(rws), rws.numel(), st),
          "cbm_record_chain")
    h = ws.host("rec_host", 9)
    h[:8].copy_(result[:64].view(_I64), non_blocking=True)
    if inf is not None:
        h[8:9].copy_(inf.result, non_blocking=True)
    torch.cuda.current_stream().synchronize()
    if inf is not None:
        raise_block_errors(inf, int(h[8]))
    bad_links = int(h[_lib.REC_BAD_LINKS])
    if bad_links:
        check(L.cbm_record_chain_repair(ptr(data), start, n, n_ref, ptr(result), ptr(rws), rws.numel(), st),
              "cbm_record_chain_repair")
        h[:8].copy_(result[:64].view(_I64))
    invalid = bool(int(h[_lib.REC_INVALID]))
    if invalid and strict:
        raise ValueError(MSG_NO_BLOCK)
    return Chain(int(h[_lib.REC_N]), int(h[_lib.REC_TAIL]), int(h[_lib.REC_CIGAR]), int(h[_lib.REC_SEQ]),
                 int(h[_lib.REC_QUAL]), rws, invalid, bad_links)


def record_candidates(data: torch.Tensor, start: int, n: int, n_ref: int, ws: Workspace | None = None) -> int:
    """Byte offsets of data[start:n) where a valid record starts (synchronises).

    The reference's record_candidates count (src/cubam/_kernels.py:280-284, 662-673); the
    reader needs it only when the chain from ``start`` holds no record.
    """
    ws = ws or default_workspace()
    cnt = ws.dev("rec_cand", 64)
    check(_lib.load().cbm_record_candidates(ptr(data), start, n, n_ref, ptr(cnt), stream_ptr()),
          "cbm_record_candidates")
    return int(cnt[:8].view(_I64)[0].item())


def record_offsets(data: torch.Tensor, start: int, n: int, chain: Chain, csr: bool = False):
    """The chain's record offsets; with ``csr``, ``(offsets, csr_columns)``: also the CSR offset
    columns ``cigar_off``, ``seq_off`` and ``qual_off`` (``n_rec + 1`` each) from the chain's
    sums, for :func:`scatter`."""
    L = _lib.load()
    dev = data.device
    off = torch.empty(max(chain.n_rec, 1), dtype=_I64, device=dev)
    if not csr:
        check(L.cbm_record_offsets(ptr(data), start, n, ptr(chain.ws), chain.ws.numel(), ptr(off), chain.n_rec,
                                   stream_ptr()), "cbm_record_offsets")
        return off[: chain.n_rec]
    cols = {k: torch.empty(chain.n_rec + 1, dtype=_I64, device=dev) for k in CSR_OFFSETS}
    check(L.cbm_record_offsets_csr(ptr(data), start, n, ptr(chain.ws), chain.ws.numel(), ptr(off),
                                   *(ptr(cols[k]) for k in CSR_OFFSETS), chain.n_rec, stream_ptr()),
          "cbm_record_offsets_csr")
    return off[: chain.n_rec], cols


#: the CSR offset columns, in the order of the C ABI's arguments
CSR_OFFSETS = ("cigar_off", "seq_off", "qual_off")


FIXED_TYPES = {
    "ref": _I32, "pos": _I32, "end": _I32, "mapq": _U8, "flag": torch.uint16, "nref": _I32, "npos": _I32,
    "tlen": _I32, "l_seq": _I32, "clip_left": torch.uint16, "clip_right": torch.uint16,
}


def scatter(data: torch.Tensor, offsets: torch.Tensor, chain: Chain | None, tags=("CB", "UB"), var: bool = True,
            fixed: bool = True, ws: Workspace | None = None, csr: dict | None = None) -> dict:
    """Kernel 4: the columns of the records at `offsets` (asynchronous).

    ``fixed``: the eleven fixed fields; ``var`` (needs the chain's CSR sizes): CIGAR, seq, qual.
    ``csr``: the CSR offset columns of these records, already computed (``record_offsets(...,
    csr=True)``); without it they are computed here from the records.
    """
    L = _lib.load()
    ws = ws or default_workspace()
    n = int(offsets.numel())
    dev = data.device
    cols = {k: torch.empty(n, dtype=t, device=dev) for k, t in FIXED_TYPES.items()} if fixed else {}
    c = _lib.Columns()
    for k in cols:
        setattr(c, k, ptr(cols[k]))
    if var:
        for k in CSR_OFFSETS:
            if csr is not None and csr[k].numel() != n + 1:
                raise ValueError(f"{k} holds {csr[k].numel()} entries for {n} records")
            cols[k] = csr[k] if csr is not None else torch.empty(n + 1, dtype=_I64, device=dev)
        cols["cigar"] = torch.empty(max(chain.n_cigar, 1), dtype=torch.uint32, device=dev)
        cols["seq"] = torch.empty(max(chain.n_seq, 1), dtype=_U8, device=dev)
        cols["qual"] = torch.empty(max(chain.n_qual, 1), dtype=_U8, device=dev)
        for k in ("cigar_off", "cigar", "seq_off", "seq", "qual_off", "qual"):
            setattr(c, k, ptr(cols[k]))
    tags = tuple(tags or ())
    if len(tags) > 4:
        raise ValueError("at most 4 tags are decoded with the columns")
    c.n_tags = len(tags)
    c.tag_names = "".join(tags).encode().ljust(8, b"\0")
    tag_cols = {}
    for i, t in enumerate(tags):
        if len(t) != 2:
            raise ValueError(f"A tag name has two characters, not {t!r}")
        tc = {
            "type": torch.empty(n, dtype=_U8, device=dev),
            "offset": torch.empty(n, dtype=_I64, device=dev),
            "length": torch.empty(n, dtype=_I32, device=dev),
            "value": torch.empty(n, dtype=_I64, device=dev),
        }
        c.tag_type[i], c.tag_offset[i] = ptr(tc["type"]), ptr(tc["offset"])
        c.tag_length[i], c.tag_value[i] = ptr(tc["length"]), ptr(tc["value"])
        tag_cols[t] = tc
    if var and csr is not None:
        check(L.cbm_scatter_csr(ptr(data), ptr(offsets), n, ctypes.byref(c), stream_ptr()), "cbm_scatter_csr")
    else:
        wsb = L.cbm_scatter_workspace(n)
        sws = ws.dev("scatter_ws", wsb)
        check(L.cbm_scatter(ptr(data), ptr(offsets), n, ctypes.byref(c), ptr(sws), sws.numel(), stream_ptr()),
              "cbm_scatter")
    if var:
        cols["cigar"] = cols["cigar"][: chain.n_cigar]
        cols["seq"] = cols["seq"][: chain.n_seq]
        cols["qual"] = cols["qual"][: chain.n_qual]
    return {"fixed": cols, "tags": tag_cols}
