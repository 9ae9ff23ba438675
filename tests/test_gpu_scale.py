"""Parity at the stated sizes of the workloads (needs a GPU; a few minutes).

* GEX-like, 1,000,000 records decoded in one chunk: every column against the oracle.
* ATAC-like (paired-end 2 x 50 bp, 10,000-barcode pool), streamed in 1 GiB compressed chunks:
  a file of at least three such chunks, so records cross chunk seams at full size; every chunk
  against the oracle, the first of them being the benchmark's chunk.
The oracle runs on all host cores; the GPU side and the oracle advance chunk by chunk.
"""

import os
import time

import pytest

from gpucheck import same_batch, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cbm():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from cubam_amd import _lib

    _lib.load()
    import cubam_amd

    return cubam_amd


def _threads():
    return max(1, min(16, os.cpu_count() or 1))


def _lockstep(path, raw, chunk, n_expect, tags=("CB", "UB")):
    from cubam_amd import read_bam
    from oracle import oracle

    ora = oracle.iter_bam(raw, chunk, tags=tags, n_threads=_threads(), name=str(path))
    n = chunks = 0
    t0 = time.time()
    for b in read_bam(path, chunk_bytes=chunk, tags=tags):
        data, c = next(ora)
        same_batch(b, data, c, tags)
        n += len(b)
        chunks += 1
        say(f"  chunk {chunks}: {len(b)} records, {data.size / 1e9:.2f} GB inflated, equal ({time.time() - t0:.0f} s)")
        del b, data, c
    assert next(ora, None) is None
    assert n == n_expect
    return chunks


@pytest.mark.timeout(900)
def test_gex_one_million_records_one_chunk(cbm, synth, tmp_path):
    t = time.time()
    b = synth.make_bam(0, 0, 1_000_000)
    p = b.write(tmp_path / "gex.bam")
    say(f"GEX 1M records: {b.bgzf.size / 1e6:.0f} MB compressed, made in {time.time() - t:.0f} s")
    assert _lockstep(p, b.bgzf, 256 << 20, 1_000_000) == 1


@pytest.mark.timeout(1200)
def test_atac_stream_of_1gib_chunks(cbm, synth, tmp_path):
    t = time.time()
    n = 46_000_000  # about 3.4 GB compressed: three full 1 GiB chunks and a partial one
    b = synth.make_bam(1, 1, n, n_total=200_000_000)
    p = b.write(tmp_path / "atac.bam")
    say(f"ATAC {n} records: {b.bgzf.size / 1e9:.2f} GB compressed, made in {time.time() - t:.0f} s")
    assert b.bgzf.size > 3 * (1 << 30)
    assert _lockstep(p, b.bgzf, 1 << 30, n) >= 3
