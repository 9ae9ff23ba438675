"""Read BAM files on an AMD Instinct MI355X (gfx950), as columns.

The reader API of cubam (``read_bam``, ``inflate_chunks``, ``RecordBatch``,
``Tag``, ``BamHeader``) over four hand-written HIP kernels reached through a
C ABI (``libcubam_amd.so``): BGZF block scan, DEFLATE inflate with a CRC32
check, BAM record-chain detection and record-to-column scatter.
"""

from cubam_amd._memory import auto_chunk_bytes, memory_pool
from cubam_amd.bam import BamHeader, RecordBatch, Tag, read_bam
from cubam_amd.bgzf import inflate_chunks

__all__ = ["BamHeader", "RecordBatch", "Tag", "auto_chunk_bytes", "inflate_chunks", "memory_pool", "read_bam"]
