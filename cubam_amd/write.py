"""BGZF compression and BAM writing on the GPU (the reference's src/cubam/bgzf.py:380-540).

:func:`bgzf_compress` cuts device bytes into payloads of 65,280 bytes (as
htslib and the reference do) and compresses every block in one launch of kernel
5 (``cbm_deflate``: greedy level-1 matching, dynamic Huffman codes, CRC32 and
BGZF framing), then packs the blocks back to back. :class:`BgzfWriter` streams
bytes into a file that way, and ends it with the EOF block. :func:`write_bam`
writes a BAM file from record bytes on the GPU: the header in blocks of its own,
then the records, as the reference's split and sort outputs do
(src/cubam/apps/split.py:44-60).

The compressed bytes are this encoder's own; what round-trips exactly is each
block's payload and CRC32, as zlib (or htslib) decodes them.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np
import torch

from cubam_amd import _decode, _lib
from cubam_amd._lib import DEFLATE_MAX_IN, DEFLATE_SLOT, check, ptr, stream_ptr

#: The empty BGZF block that ends a file (src/cubam/bgzf.py:375)
EOF_BLOCK = bytes.fromhex("1f8b08040000000000ff0600424302001b0003000000000000000000")
BLOCK_DATA = DEFLATE_MAX_IN  # 0xFF00

_U8, _I32, _I64 = torch.uint8, torch.int32, torch.int64


def _as_device(data) -> torch.Tensor:
    if isinstance(data, torch.Tensor) and data.is_cuda:
        return data.reshape(-1).view(_U8) if data.dtype != _U8 else data.reshape(-1)
    if isinstance(data, (bytes, bytearray, memoryview)):
        data = np.frombuffer(bytes(data), np.uint8)
    # a writable copy where needed: torch does not take read-only numpy buffers
    return torch.from_numpy(np.require(data, np.uint8, ["C", "W"])).to(_decode.device())


def copy_spans(src: torch.Tensor, src_off: torch.Tensor, length: torch.Tensor, dst_off: torch.Tensor,
               dst: torch.Tensor) -> None:
    """dst[dst_off[i] : + length[i]] = src[src_off[i] : + length[i]] on the current stream (int64 device arrays)."""
    n = int(src_off.numel())
    check(_lib.load().cbm_copy_spans(ptr(src), ptr(src_off), ptr(length), ptr(dst_off), n, ptr(dst), stream_ptr()),
          "cbm_copy_spans")


def deflate_blocks(data: torch.Tensor, in_off: torch.Tensor, in_len: torch.Tensor,
                   ws: _decode.Workspace | None = None) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """BGZF blocks of the payloads data[in_off[i] : + in_len[i]] (each at most 65,280 bytes).

    Returns (blocks packed back to back, block sizes int32, payload CRC32 uint32), on the device.
    """
    L = _lib.load()
    ws = ws or _decode.default_workspace()
    n = int(in_off.numel())
    dev = data.device
    if n == 0:
        return torch.empty(0, dtype=_U8, device=dev), torch.empty(0, dtype=_I32, device=dev), \
            torch.empty(0, dtype=torch.uint32, device=dev)
    slots = torch.empty(n * DEFLATE_SLOT, dtype=_U8, device=dev)
    bsize = torch.empty(n, dtype=_I32, device=dev)
    crc = torch.empty(n, dtype=torch.uint32, device=dev)
    dws = ws.dev("deflate_ws", L.cbm_deflate_workspace(n))
    in_len32 = in_len.to(_I32)  # held until the call is enqueued
    check(L.cbm_deflate(ptr(data), ptr(in_off), ptr(in_len32), n, ptr(slots), ptr(bsize), ptr(crc), ptr(dws),
                        dws.numel(), stream_ptr()), "cbm_deflate")
    sizes = bsize.to(_I64)
    ends = torch.cumsum(sizes, 0)
    total = int(ends[-1])  # synchronises
    if bool((bsize < 0).any()):
        raise RuntimeError("cbm_deflate: a block's encoded size disagrees with its code lengths")
    out = torch.empty(total, dtype=_U8, device=dev)
    slot_off = torch.arange(n, dtype=_I64, device=dev) * DEFLATE_SLOT
    copy_spans(slots, slot_off, sizes, ends - sizes, out)
    return out, bsize, crc


def bgzf_compress(data, block: int = BLOCK_DATA, ws: _decode.Workspace | None = None) -> torch.Tensor:
    """``data`` (device tensor or host bytes) as BGZF blocks of ``block`` payload bytes, without the
    EOF block, as a device uint8 tensor (src/cubam/bgzf.py:380-422)."""
    if not 0 < block <= BLOCK_DATA:
        raise ValueError(f"a BGZF block holds 1 to {BLOCK_DATA} payload bytes")
    d = _as_device(data)
    n = int(d.numel())
    off = torch.arange(0, n, block, dtype=_I64, device=d.device)
    size = torch.clamp(n - off, max=block)
    return deflate_blocks(d, off, size, ws)[0]


def gather_records(data: torch.Tensor, offset: torch.Tensor) -> torch.Tensor:
    """The records at ``offset`` in ``data`` (each block_size + 4 bytes), back to back."""
    dev = data.device
    n = int(offset.numel())
    if n == 0:
        return torch.empty(0, dtype=_U8, device=dev)
    b = data[offset.unsqueeze(1) + torch.arange(4, device=dev)].to(_I64)
    size = (b[:, 0] | (b[:, 1] << 8) | (b[:, 2] << 16) | (b[:, 3] << 24)) + 4
    ends = torch.cumsum(size, 0)
    out = torch.empty(int(ends[-1]), dtype=_U8, device=dev)
    copy_spans(data, offset.to(_I64), size, ends - size, out)
    return out


class BgzfWriter:
    """A BGZF file written from bytes on the GPU (the reference's BgzfWriter contract,
    src/cubam/bgzf.py:425-540: ``write``, ``flush``, ``close``, the EOF block).

    Bytes written (device tensors or host bytes) go into one device staging buffer.
    When it holds ``flush_bytes`` or is full, its whole 65,280-byte blocks are
    compressed in one launch and written out, and the partial block left over moves to
    the front. :meth:`flush` compresses everything staged, so the next byte starts a
    block; :meth:`close` does that and appends the EOF block. Block boundaries thus fall
    every 65,280 bytes from the last :meth:`flush`, whatever the sizes of the writes.
    ``path=None`` keeps the file in memory (:meth:`getvalue`).
    """

    def __init__(self, path: str | Path | None, flush_bytes: int = 1024 * BLOCK_DATA):
        self.path = Path(path) if path is not None else None
        self.flush_bytes = max(int(flush_bytes), 1)
        # whole blocks of at least flush_bytes, plus room for a partial block
        self._cap = (self.flush_bytes + BLOCK_DATA - 1) // BLOCK_DATA * BLOCK_DATA + BLOCK_DATA
        self._stage: torch.Tensor | None = None
        self._fill = 0
        self._fh = open(self.path, "wb") if self.path is not None else None  # noqa: SIM115  (closed in close())
        self._mem: list[bytes] = []
        self.ws = _decode.Workspace()
        self.closed = False

    def _out(self, b: bytes) -> None:
        if self._fh is None:
            self._mem.append(b)
        else:
            self._fh.write(b)

    def write(self, data) -> None:
        src = _as_device(data)
        total, done = int(src.numel()), 0
        if total and self._stage is None:
            self._stage = torch.empty(self._cap, dtype=_U8, device=src.device)
        while done < total:
            take = min(total - done, self._cap - self._fill)
            self._stage[self._fill : self._fill + take].copy_(src[done : done + take])
            self._fill += take
            done += take
            if self._fill >= self.flush_bytes or self._fill == self._cap:
                self._compress(whole_blocks=True)

    def _compress(self, whole_blocks: bool) -> None:
        n = self._fill
        cut = n // BLOCK_DATA * BLOCK_DATA if whole_blocks else n
        if cut:
            self._out(bgzf_compress(self._stage[:cut], ws=self.ws).cpu().numpy().tobytes())
        left = n - cut
        if left:  # the partial block moves to the front (source and target may overlap: copy it first)
            self._stage[:left].copy_(self._stage[cut:n].clone())
        self._fill = left

    def flush(self) -> None:
        """Compress and write everything staged; the next byte written starts a new block."""
        if self._fill:
            self._compress(whole_blocks=False)

    def close(self) -> None:
        if self.closed:
            return
        self.flush()
        self._out(EOF_BLOCK)
        if self._fh is not None:
            self._fh.close()
        self.closed = True

    def getvalue(self) -> bytes:
        return b"".join(self._mem)

    def __enter__(self) -> BgzfWriter:
        return self

    def __exit__(self, *exc) -> None:
        self.close()


def write_bam(path: str | Path | None, header, records) -> bytes | None:
    """A BAM file of ``header`` (a :class:`~cubam_amd.bam.BamHeader` or its bytes) and ``records``
    (a device uint8 tensor of whole records, or an iterable of them). The header goes in blocks of
    its own. With ``path=None`` the file's bytes are returned."""
    hb = header.to_bytes() if hasattr(header, "to_bytes") else bytes(header)
    w = BgzfWriter(path)
    w.write(hb)
    w.flush()
    for r in [records] if isinstance(records, torch.Tensor) else records:
        w.write(r)
    w.close()
    return w.getvalue() if path is None else None
