"""GPU memory: the pool, and how much to read at a time (the reference's src/cubam/_memory.py)."""

from __future__ import annotations

from contextlib import contextmanager

#: Peak GPU memory of a chunk's buffers is a few times its inflated size; the apps keep more
_CHUNK_SHARE = 64


def auto_chunk_bytes() -> int:
    """Compressed bytes to read per chunk: 256 MiB, halved while 64 chunks do not fit in free GPU memory.

    Same policy as the reference (src/cubam/_memory.py:11-23): never below 32 MiB. On an
    MI355X (288 GB of HBM3E) this is always 256 MiB; larger chunks can be asked for.
    """
    import torch

    free, _ = torch.cuda.mem_get_info()
    size = 256 << 20
    while size > 32 << 20 and size * _CHUNK_SHARE > free:
        size //= 2
    return size


@contextmanager
def memory_pool():
    """The stream-ordered pool of the reference (RMM, src/cubam/_memory.py:26-45) is torch's
    caching allocator here, which every buffer of the read path already goes through: a no-op."""
    yield
