"""BGZF files read and inflated on the GPU (the reference's src/cubam/bgzf.py:1-125, 542-652).

:func:`inflate_chunks` reads a file in chunks and yields the inflated data of
each chunk as a device ``torch.uint8`` tensor:

1. **Read.** A background thread reads the next chunks into pinned host buffers
   with parallel ``pread`` calls (as the reference's ``_ChunkReader``).
2. **Find blocks.** Kernel 1 finds the chain of BGZF blocks from the start of
   the chunk; the bytes of an incomplete last block stay on the GPU and go in
   front of the next chunk.
3. **Inflate.** Kernel 2 inflates all blocks of the chunk in one launch and
   checks every block's CRC32 and size against its trailer.
"""

from __future__ import annotations

import os
import queue
import threading
from collections.abc import Callable, Iterator
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
    """Chunks of a file read ahead into pinned buffers by a background thread.

    Each chunk is read by ``n_threads`` parallel ``pread`` calls of ``piece``
    bytes. Iterating yields ``(slot, nbytes, eof)``; the consumer copies
    :meth:`host` ``(slot)[:nbytes]`` and then calls :meth:`release` ``(slot)``.
    Same contract as the reference (src/cubam/bgzf.py:38-125).
    """

    def __init__(self, path, chunk_bytes: int, *, n_buffers: int = 3, n_threads: int = 8,
                 piece: int = 16 << 20):
        self.path, self.chunk, self.n_threads, self.piece = Path(path), chunk_bytes, n_threads, piece
        self._mem = [torch.empty(chunk_bytes, dtype=torch.uint8, pin_memory=True) for _ in range(n_buffers)]
        self._free: queue.Queue = queue.Queue()
        for i in range(n_buffers):
            self._free.put(i)
        self._full: queue.Queue = queue.Queue()
        self._error: BaseException | None = None
        self._stop = False
        self._thread = threading.Thread(target=self._run, daemon=True)
        self._thread.start()

    def host(self, slot: int) -> torch.Tensor:
        return self._mem[slot]

    def release(self, slot: int) -> None:
        self._free.put(slot)

    def close(self) -> None:
        self._stop = True
        self._free.put(-1)

    @staticmethod
    def _pread(fd: int, mv: memoryview, offset: int) -> None:
        while len(mv):
            got = os.preadv(fd, [mv], offset)
            if not got:
                raise ValueError("The file is shorter than its size")
            mv, offset = mv[got:], offset + got

    def _run(self) -> None:
        from concurrent.futures import ThreadPoolExecutor

        try:
            fd = os.open(self.path, os.O_RDONLY)
            try:
                size = os.fstat(fd).st_size
                off = 0
                with ThreadPoolExecutor(self.n_threads) as ex:
                    while True:
                        slot = self._free.get()
                        if self._stop:
                            return
                        n = min(self.chunk, size - off)
                        mv = memoryview(self._mem[slot].numpy())
                        futures = [
                            ex.submit(self._pread, fd, mv[p : min(p + self.piece, n)], off + p)
                            for p in range(0, n, self.piece)
                        ]
                        for f in futures:
                            f.result()
                        off += n
                        eof = off >= size
                        self._full.put((slot, n, eof))
                        if eof:
                            return
            finally:
                os.close(fd)
        except BaseException as e:  # noqa: BLE001  (re-raised in the consumer)
            self._error = e
            self._full.put(None)

    def __iter__(self) -> Iterator[tuple[int, int, bool]]:
        while True:
            item = self._full.get()
            if item is None:
                raise self._error  # type: ignore[misc]
            yield item
            if item[2]:
                return


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
    chunk; ``prefix`` is called before each chunk is inflated. The reader thread
    reads chunks ahead into pinned buffers; each chunk's host-to-device copy runs
    on a copy stream while the chunk before it is decoded, and the carry and the
    uploaded chunk are joined on the device.
    """
    from cubam_amd._memory import auto_chunk_bytes

    dev = _decode.device()
    chunk_bytes = chunk_bytes or auto_chunk_bytes()
    reader = _ChunkReader(path, chunk_bytes, **reader_kw)
    compute = torch.cuda.current_stream()
    copier = torch.cuda.Stream() if OVERLAP_UPLOAD else compute
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
            if OVERLAP_UPLOAD:
                pending = None if eof else upload()  # the next chunk's copy overlaps this chunk's decode
            data, n, used, inf = _inflate_blocks(buf, c + nbytes, eof, prefix() if prefix else None, ws)
            if not OVERLAP_UPLOAD:
                pending = None if eof else upload()
            reader.release(slot)  # the scan synchronised the compute stream, which waited for the copy
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
