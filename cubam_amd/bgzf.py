"""BGZF files read and inflated on the GPU (the reference's src/cubam/bgzf.py:1-125, 542-652).

:func:`inflate_chunks` reads a file in chunks and yields the inflated data of
each chunk as a device ``torch.uint8`` tensor:

1. **Read.** A thread pool reads the next chunks into pinned host slots with
   parallel ``preadv`` calls, a slot a chunk, as slots are handed back.
2. **Find blocks.** Kernel 1 finds the chain of BGZF blocks from the start of
   the chunk; the bytes of an incomplete last block stay on the GPU and go in
   front of the next chunk.
3. **Inflate.** Kernel 2 inflates all blocks of the chunk in one launch and
   checks every block's CRC32 and size against its trailer.
"""

from __future__ import annotations

import os
from collections import deque
from collections.abc import Callable, Iterator
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import numpy as np
import torch

from cubam_amd import _decode
from cubam_amd._lib import PAD

#: The empty BGZF block that ends a file (src/cubam/bgzf.py:375)
EOF_BLOCK = bytes.fromhex("1f8b08040000000000ff0600424302001b0003000000000000000000")
#: Data per BGZF block, as samtools and htslib write (src/cubam/bgzf.py:377)
BLOCK_DATA = 0xFF00


class _ChunkReader:
    """Compressed chunks of a file, read ahead into pinned host slots (the reference's
    ``_ChunkReader`` contract, src/cubam/bgzf.py:38-125: slots, ``host``/``release``,
    ``n_buffers``, ``n_threads``, ``piece``).

    Chunk k is the file's bytes [k * chunk_bytes, (k + 1) * chunk_bytes). Its read is
    split into ``piece``-byte ``preadv`` calls that a pool of ``n_threads`` threads runs;
    up to ``n_buffers`` chunks are in flight, one a slot. Iterating yields
    ``(slot, nbytes, eof)`` in file order once a chunk is whole; ``release(slot)``
    hands the slot back and starts the read of the next unread chunk into it. A
    failed read raises when its chunk is reached.
    """

    def __init__(self, path, chunk_bytes: int, *, n_buffers: int = 3, n_threads: int = 8,
                 piece: int = 16 << 20):
        if chunk_bytes <= 0 or piece <= 0 or n_buffers < 1 or n_threads < 1:
            raise ValueError("chunk_bytes, piece, n_buffers and n_threads must be positive")
        self.path, self.chunk, self.piece = Path(path), chunk_bytes, piece
        self._fd = os.open(self.path, os.O_RDONLY)
        self._size = os.fstat(self._fd).st_size
        self._last = max(0, (self._size - 1) // chunk_bytes)  # index of the final chunk (an empty file has one)
        slot_bytes = max(1, min(chunk_bytes, self._size))
        n_slots = min(n_buffers, self._last + 1)
        pin = torch.cuda.is_available()  # page-locked for the device copies (plain memory where no GPU is)
        self._mem = [torch.empty(slot_bytes, dtype=torch.uint8, pin_memory=pin) for _ in range(n_slots)]
        self._pool = ThreadPoolExecutor(n_threads)
        self._queue: deque = deque()  # (chunk index, slot, nbytes, futures) in file order
        self._next = 0
        self._closed = False
        for slot in range(n_slots):
            self._start(slot)

    @property
    def slots(self) -> int:
        return len(self._mem)

    def host(self, slot: int) -> torch.Tensor:
        return self._mem[slot]

    def release(self, slot: int) -> None:
        if not self._closed and self._next <= self._last:
            self._start(slot)

    def close(self) -> None:
        if self._closed:
            return
        self._closed = True
        self._pool.shutdown(wait=True, cancel_futures=True)
        os.close(self._fd)

    def _start(self, slot: int) -> None:
        k = self._next
        self._next += 1
        lo = k * self.chunk
        n = max(0, min(self.chunk, self._size - lo))
        view = memoryview(self._mem[slot].numpy())
        futures = [self._pool.submit(self._fill, view[p : min(p + self.piece, n)], lo + p)
                   for p in range(0, n, self.piece)]
        self._queue.append((k, slot, n, futures))

    def _fill(self, dst: memoryview, at: int) -> None:
        while dst.nbytes:
            got = os.preadv(self._fd, [dst], at)
            if got <= 0:
                raise ValueError("The file is shorter than its size")
            dst, at = dst[got:], at + got

    def __iter__(self) -> Iterator[tuple[int, int, bool]]:
        while self._next <= self._last or self._queue:
            if not self._queue:
                raise RuntimeError("_ChunkReader: every slot is held; release one before reading on")
            k, slot, n, futures = self._queue.popleft()
            for f in futures:
                f.result()
            yield slot, n, k == self._last


def _inflate_blocks(buf: torch.Tensor, n: int, eof: bool, prefix: torch.Tensor | None = None,
                    ws: _decode.Workspace | None = None):
    """Inflate the complete BGZF blocks of buf[0:n) (starting at a block) after ``prefix``.

    Returns (data, n_data, used, inflated) where data[0:n_data) is prefix + the
    inflated bytes (padded device tensor) and ``used`` the bytes of buf the
    blocks covered (src/cubam/bgzf.py:542-611). Block errors are checked by the
    caller (``_decode.check_inflated`` or with the record chain's sync).
    """
    blocks = _decode.scan_blocks(buf, n, eof, ws)
    if blocks.n_blocks == 0:
        if prefix is None:
            return torch.empty(PAD, dtype=torch.uint8, device=buf.device), 0, blocks.used, None
        p = int(prefix.numel())
        data = torch.empty(p + PAD, dtype=torch.uint8, device=buf.device)
        data[:p].copy_(prefix)
        return data, p, blocks.used, None
    inf = _decode.inflate_blocks(blocks, prefix, ws)
    return inf.data, inf.n, blocks.used, inf


def inflate_chunks(path: str | Path, chunk_bytes: int | None = None, **reader_kw) -> Iterator[torch.Tensor]:
    """Inflated data of a BGZF file on the GPU, chunk by chunk (device uint8 tensors).

    Each chunk holds the data of the complete blocks read so far, so chunks
    split the data at block boundaries (src/cubam/bgzf.py:615-622).
    ``chunk_bytes`` (compressed) defaults to :func:`cubam_amd.auto_chunk_bytes`.
    """
    for data, n, inf in _inflate(path, chunk_bytes, None, **reader_kw):
        if inf is not None:
            _decode.check_inflated(inf)
        yield data[:n]


#: Upload the next chunk on a copy stream while the current one decodes (the reference
#: copies each chunk right before decoding it, src/cubam/bgzf.py:640-646)
OVERLAP_UPLOAD = True


def _inflate(path, chunk_bytes: int | None, prefix: Callable[[], torch.Tensor] | None,
             ws: _decode.Workspace | None = None, **reader_kw):
    """Yields (data, n, inflated) per chunk, each after the bytes ``prefix()`` returns.

    As the reference's ``_inflate`` (src/cubam/bgzf.py:625-652): the compressed
    carry (the incomplete last block) stays on the device in front of the next
    chunk; ``prefix`` is called before each chunk is inflated. The reader's pool
    reads chunks ahead into pinned slots; each chunk's host-to-device copy runs
    on a copy stream while the chunk before it is decoded, and the carry and the
    uploaded chunk are joined on the device.
    """
    from cubam_amd._memory import auto_chunk_bytes

    dev = _decode.device()
    chunk_bytes = chunk_bytes or auto_chunk_bytes()
    reader = _ChunkReader(path, chunk_bytes, **reader_kw)
    compute = torch.cuda.current_stream()
    overlap = OVERLAP_UPLOAD and reader.slots > 1  # with one slot the next chunk waits for its release
    copier = torch.cuda.Stream() if overlap else compute
    items = iter(reader)
    carry = torch.empty(0, dtype=torch.uint8, device=dev)

    def upload():
        item = next(items, None)
        if item is None:
            return None
        slot, nbytes, eof = item
        with torch.cuda.stream(copier):
            stage = torch.empty(nbytes, dtype=torch.uint8, device=dev)
            stage.copy_(reader.host(slot)[:nbytes], non_blocking=True)
            done = torch.cuda.Event()
            done.record(copier)
        return slot, nbytes, eof, stage, done

    try:
        pending = upload()
        while pending is not None:
            slot, nbytes, eof, stage, done = pending
            compute.wait_event(done)
            stage.record_stream(compute)
            c = int(carry.numel())
            buf = torch.empty(c + nbytes + PAD, dtype=torch.uint8, device=dev)
            if c:
                buf[:c].copy_(carry)
            buf[c : c + nbytes].copy_(stage)
            del stage
            if overlap:
                pending = None if eof else upload()  # the next chunk's copy overlaps this chunk's decode
            data, n, used, inf = _inflate_blocks(buf, c + nbytes, eof, prefix() if prefix else None, ws)
            reader.release(slot)  # the scan synchronised the compute stream, which waited for the copy
            if not overlap:
                pending = None if eof else upload()
            carry = buf[used : c + nbytes].clone()
            del buf
            yield data, n, inf
            del data
    finally:
        reader.close()


def inflate_bytes(compressed, eof: bool = True, ws: _decode.Workspace | None = None) -> torch.Tensor:
    """Inflate a whole in-memory BGZF buffer (host bytes/array or a device tensor) on the GPU."""
    if isinstance(compressed, torch.Tensor) and compressed.is_cuda:
        n = int(compressed.numel())
        buf = torch.empty(n + PAD, dtype=torch.uint8, device=compressed.device)
        buf[:n].copy_(compressed)
    else:
        arr = np.frombuffer(bytes(compressed), np.uint8) if not isinstance(compressed, np.ndarray) else compressed
        n = int(arr.size)
        buf = _decode.to_device(arr)
    data, m, used, inf = _inflate_blocks(buf, n, eof, None, ws)
    if inf is not None:
        _decode.check_inflated(inf)
    return data[:m]
