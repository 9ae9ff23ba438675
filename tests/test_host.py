"""Host logic that needs no GPU: the autotuner's cache and buckets, the write-back framing
constants, and the BGZF EOF block of the writer."""

import gzip
import json
import zlib

import pytest

import bamfixture as bf
from cubam_amd import _lib, tune, write


@pytest.fixture
def tune_cache(tmp_path, monkeypatch):
    path = tmp_path / "tune.json"
    monkeypatch.setenv("CUBAM_AMD_TUNE_CACHE", str(path))
    monkeypatch.setattr(tune, "device_key", lambda: "gfx950/256cu")
    monkeypatch.setattr(tune, "_cache", None)
    return path


def test_buckets_are_powers_of_two():
    assert tune.bucket(1000, 1000 * 65280) == tune.bucket(1023, 1023 * 65280)
    assert tune.bucket(1000, 1000 * 65280) != tune.bucket(1024, 1024 * 65280)
    assert tune.bucket(1000, 1000 * 65280) != tune.bucket(1000, 1000 * 20000)
    assert tune.bucket(0, 0) == "b0-m0"  # empty batches do not raise


def test_tune_cache_round_trip(tune_cache):
    assert tune.lookup(5000, 5000 * 60000) is None
    tune.record(5000, 5000 * 60000, (3, 0, 0, 0), 12.5, {"0,0,0,0": 13.0, "3,0,0,0": 12.5})
    got = tune.lookup(5001, 5001 * 60000)
    assert (got.variant, got.blocks_per_wave, got.waves_per_group, got.groups) == (3, 0, 0, 0)
    assert tune.lookup(50, 50 * 60000) is None  # another bucket
    # persisted: a fresh process state reads it back
    disk = json.loads(tune_cache.read_text())
    assert disk["gfx950/256cu"][tune.bucket(5000, 5000 * 60000)]["tune"] == [3, 0, 0, 0]
    tune._cache = None
    assert tune.lookup(5000, 5000 * 60000).variant == 3


def test_candidates_cover_the_launch_shapes():
    workgroup = {c[0] for c in tune.CANDIDATES if c[0] != _lib.INFLATE_LANES}
    assert workgroup == {v for v in range(_lib.INFLATE_VARIANTS) if v != _lib.INFLATE_LANES}
    lanes = {(c[1], c[2]) for c in tune.CANDIDATES if c[0] == _lib.INFLATE_LANES}
    assert {bpw for bpw, _ in lanes} == {1, 2, 4, 8}  # blocks per workgroup {1, 2, 4, 8}


def test_write_framing_constants():
    assert write.EOF_BLOCK == bf.EOF_BLOCK
    assert gzip.decompress(write.EOF_BLOCK) == b""
    assert write.BLOCK_DATA == 0xFF00  # htslib's BGZF_BLOCK_SIZE payloads
    assert _lib.DEFLATE_SLOT >= 18 + 8 + write.BLOCK_DATA + 5 * ((write.BLOCK_DATA + 65534) // 65535)
    # the EOF block is an empty stored deflate block with the CRC and size of nothing
    assert write.EOF_BLOCK[-8:] == zlib.crc32(b"").to_bytes(4, "little") + (0).to_bytes(4, "little")
