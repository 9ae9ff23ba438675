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
    # two-phase: the resolve pass's LDS layout (window of 64 or 128 matches, history ring of 4 to
    # 32 KiB) and its blocks per workgroup
    two = {(c[1], c[2]) for c in tune.CANDIDATES if c[0] in _lib.INFLATE_TWO_PHASE}
    assert {bpw for bpw, _ in two} == {0, 1, 2, 4, 8, 16, 32}
    assert {wpg for _, wpg in two} >= {0, 1, 2, 4, 8}


def test_write_framing_constants():
    assert write.EOF_BLOCK == bf.EOF_BLOCK
    assert gzip.decompress(write.EOF_BLOCK) == b""
    assert write.BLOCK_DATA == 0xFF00  # htslib's BGZF_BLOCK_SIZE payloads
    assert _lib.DEFLATE_SLOT >= 18 + 8 + write.BLOCK_DATA + 5 * ((write.BLOCK_DATA + 65534) // 65535)
    # the EOF block is an empty stored deflate block with the CRC and size of nothing
    assert write.EOF_BLOCK[-8:] == zlib.crc32(b"").to_bytes(4, "little") + (0).to_bytes(4, "little")


# ------------------------------------------------------------------ file reader and header (no GPU needed)


def _read_all(path, chunk, **kw):
    from cubam_amd.bgzf import _ChunkReader

    r = _ChunkReader(path, chunk, **kw)
    out = []
    try:
        for slot, n, eof in r:
            out.append((bytes(r.host(slot)[:n].numpy()), eof))
            r.release(slot)
    finally:
        r.close()
    return out


@pytest.mark.parametrize("size, chunk, piece, buffers, threads",
                         [(100_000, 7_000, 1_000, 3, 4), (100_000, 7_000, 7_000, 1, 1), (70_000, 7_000, 3_001, 2, 8),
                          (5, 1 << 20, 16 << 20, 3, 8), (0, 4096, 100, 3, 2), (4096, 4096, 4096, 2, 2)])
def test_chunk_reader_reads_every_byte_in_order(tmp_path, size, chunk, piece, buffers, threads):
    import numpy as np

    data = np.random.default_rng(size).integers(0, 256, size, dtype=np.uint8).tobytes()
    p = tmp_path / "f.bin"
    p.write_bytes(data)
    got = _read_all(p, chunk, n_buffers=buffers, n_threads=threads, piece=piece)
    assert b"".join(g for g, _ in got) == data
    assert [e for _, e in got] == [False] * (len(got) - 1) + [True]
    assert all(len(g) == chunk for g, _ in got[:-1])
    assert len(got) == max(1, -(-size // chunk))


def test_chunk_reader_stalls_loudly_without_a_release(tmp_path):
    from cubam_amd.bgzf import _ChunkReader

    p = tmp_path / "f.bin"
    p.write_bytes(b"x" * 10_000)
    r = _ChunkReader(p, 1000, n_buffers=2, n_threads=2, piece=300)
    it = iter(r)
    next(it)
    next(it)
    with pytest.raises(RuntimeError, match="release"):
        next(it)
    r.close()


def test_chunk_reader_reports_a_short_file(tmp_path, monkeypatch):
    import os

    from cubam_amd import bgzf

    p = tmp_path / "f.bin"
    p.write_bytes(b"y" * 5000)
    real = os.preadv
    monkeypatch.setattr(bgzf.os, "preadv", lambda fd, bufs, at: 0 if at >= 3000 else real(fd, bufs, at))
    with pytest.raises(ValueError, match="shorter than its size"):
        _read_all(p, 4000, piece=1000)


def test_bam_header_parse_and_rebuild():
    from cubam_amd import BamHeader

    text = "@HD\tVN:1.6\tSO:coordinate\n@SQ\tSN:chr1\tLN:1000000\n@PG\tID:a\n@PG\tID:b\n"
    raw = bf.header_bytes(text=text)
    h = BamHeader.from_bytes(raw + b"trailing record bytes")
    assert (h.names, h.lengths, h.size) == ([n for n, _ in bf.REFS], [ln for _, ln in bf.REFS], len(raw))
    assert h.sort_order == "coordinate" and h.lines("@PG") == ["@PG\tID:a", "@PG\tID:b"]
    assert h.to_bytes() == raw
    h.require_coordinate_sorted()
    other = BamHeader.from_bytes(bf.header_bytes(text="@HD\tVN:1.6\tSO:queryname\n"))
    assert other.sort_order == "queryname"
    with pytest.raises(ValueError, match="not coordinate-sorted"):
        other.require_coordinate_sorted("x.bam")
    assert BamHeader.from_bytes(bf.header_bytes(text="@SQ\tSN:a\tLN:1\n")).sort_order == "unknown"
    # text is cut at its first NUL; size still counts l_text bytes
    padded = bf.header_bytes(text="@HD\tSO:unsorted\n\0\0\0\0")
    hp = BamHeader.from_bytes(padded)
    assert hp.text == "@HD\tSO:unsorted\n" and hp.size == len(padded)
    with pytest.raises(ValueError, match="truncated BAM header"):
        BamHeader.from_bytes(raw[:-3])
    with pytest.raises(ValueError, match="not a BAM file"):
        BamHeader.from_bytes(b"BAI\1" + raw[4:])


def test_bam_header_read_from_bgzf(tmp_path):
    from cubam_amd import BamHeader

    p = tmp_path / "h.bam"
    p.write_bytes(bf.bgzf(bf.bam_bytes(bf.random_reads(20)), 100))
    h = BamHeader.read(p)
    assert h.size == len(bf.header_bytes()) and h.names == [n for n, _ in bf.REFS]


def test_generator_codecs_give_the_same_records():
    """libdeflate (the faster compressor of the synthetic files) frames the same payloads as zlib:
    every block inflates (zlib) to the same bytes, and its trailer holds their CRC32 and size."""
    import numpy as np

    from cubam_amd import synth

    if not synth.have_libdeflate():
        pytest.skip("libdeflate is absent")
    a = synth.make_bam(1, 5, 20_000, codec="zlib")
    b = synth.make_bam(1, 5, 20_000, codec="libdeflate")
    assert gzip.decompress(a.bgzf.tobytes()) == gzip.decompress(b.bgzf.tobytes())
    assert np.array_equal(a.block_sizes > 0, b.block_sizes > 0) and a.block_sizes.size == b.block_sizes.size
    data, o = b.bgzf.tobytes(), 0
    for bs in b.block_sizes:
        blk = data[o : o + int(bs)]
        raw = zlib.decompress(blk[18:-8], -15)
        assert int.from_bytes(blk[-8:-4], "little") == zlib.crc32(raw)
        assert int.from_bytes(blk[-4:], "little") == len(raw) <= synth.BLOCK_DATA
        o += int(bs)
    assert o == len(data)
