"""Region lookups of the block/record index on the host (no GPU).

The index is built here, without the GPU, from a BAM written by tests/bamfixture.py:
block framing parsed from the BGZF headers and trailers, record starts from the
pure-Python parser. The GPU test (test_gpu_serve.py) checks that build_index gives
the same arrays; these tests check what BlockIndex.span, region and region_of
compute from them, against brute force over the records, and the .npz round trip.
"""

import random
import struct

import numpy as np
import pytest

import bamfixture as bf
from cubam_amd.bam import BamHeader
from cubam_amd.index import BlockIndex


def _index(reads, block: int) -> tuple[BlockIndex, list[dict], int]:
    raw = bf.bam_bytes(reads)
    data = bf.bgzf(raw, block, eof=True)
    coff, bsize, isize, crc = [], [], [], []
    o = 0
    while o < len(data):
        bs = struct.unpack_from("<H", data, o + 16)[0] + 1
        c, n = struct.unpack_from("<II", data, o + bs - 8)
        coff.append(o), bsize.append(bs), isize.append(n), crc.append(c)
        o += bs
    header = BamHeader.from_bytes(bf.header_bytes())
    recs = bf.parse_records(raw, header.size)
    isize = np.array(isize, np.int32)
    uoff = np.zeros(len(isize), np.int64)
    np.cumsum(isize[:-1], out=uoff[1:])
    offs = np.array([r["offset"] for r in recs], np.int64)
    first_rec = np.full(len(isize), -1, np.int64)
    first_idx = np.full(len(isize), -1, np.int64)
    first_ref = np.full(len(isize), -1, np.int32)
    first_pos = np.full(len(isize), -1, np.int32)
    for k in range(len(isize)):
        j = int(np.searchsorted(offs, uoff[k]))
        if j < offs.size and offs[j] < uoff[k] + isize[k]:
            first_rec[k], first_idx[k] = offs[j], j
            first_ref[k], first_pos[k] = recs[j]["ref"], recs[j]["pos"]
    ix = BlockIndex(header, np.array(coff, np.int64), np.array(bsize, np.int32), uoff, isize,
                    np.array(crc, np.uint32), first_rec, first_idx, first_ref, first_pos, len(recs), len(raw))
    return ix, recs, len(data)


@pytest.fixture(scope="module")
def built():
    reads = sorted(bf.random_reads(3000, seed=11), key=bf.sort_key)
    return _index(reads, 700)


def test_spans_start_and_end_at_record_boundaries(built):
    ix, recs, _ = built
    rb = ix.record_blocks()
    assert rb.size > 50
    starts = {r["offset"] for r in recs}
    rng = random.Random(1)
    for _ in range(300):
        i0 = rng.randrange(rb.size)
        i1 = rng.randint(i0 + 1, min(rb.size, i0 + 40))
        sp = ix.span(i0, i1)
        assert sp.start in starts
        assert sp.end in starts or sp.end == ix.stream_bytes
        # its records are exactly those that start in [start, end)
        inside = [j for j, r in enumerate(recs) if sp.start <= r["offset"] < sp.end]
        assert (sp.rec0, sp.rec1) == (inside[0], inside[-1] + 1)
        # the blocks hold every byte of them, and the file range is theirs
        assert ix.uoff[sp.k0] <= sp.start and sp.end <= ix.uoff[sp.k1 - 1] + ix.isize[sp.k1 - 1]
        assert sp.u0 == ix.uoff[sp.k0]
        assert (sp.c0, sp.c1) == (ix.coff[sp.k0], ix.coff[sp.k1 - 1] + ix.bsize[sp.k1 - 1])


def test_region_holds_every_record_that_starts_in_it(built):
    ix, recs, _ = built
    rng = random.Random(2)
    for _ in range(300):
        ref = rng.randrange(len(bf.REFS))
        beg = rng.randrange(bf.REFS[ref][1])
        end = beg + rng.choice([1, 100, 5000, 200_000])
        sp = ix.region(ref, beg, end)
        want = [j for j, r in enumerate(recs) if r["ref"] == ref and beg <= r["pos"] < end]
        got = range(sp.rec0, sp.rec1)
        assert all(j in got for j in want), (ref, beg, end)


def test_region_of_maps_back(built):
    ix, recs, _ = built
    rb = ix.record_blocks()
    rng = random.Random(3)
    n = 0
    for _ in range(200):
        i0 = rng.randrange(rb.size - 1)
        sp = ix.span(i0, rng.randint(i0 + 1, min(rb.size, i0 + 10)))
        ref, beg, end = ix.region_of(sp)
        back = ix.region(ref, beg, end)
        assert back.rec0 <= sp.rec0
        if recs[sp.rec0]["ref"] == recs[sp.rec1 - 1]["ref"]:  # one reference: the whole span
            assert sp.rec1 <= back.rec1
            n += 1
        else:  # into the next reference: the rest of the first one
            last = max(j for j in range(sp.rec0, sp.rec1) if recs[j]["ref"] == ref)
            assert last < back.rec1
    assert n > 100


def test_span_bounds_are_checked(built):
    ix, _, _ = built
    with pytest.raises(IndexError):
        ix.span(0, 0)
    with pytest.raises(IndexError):
        ix.span(0, ix.record_blocks().size + 1)


def test_save_load_round_trip(built, tmp_path):
    ix, _, _ = built
    p = tmp_path / "ix.npz"
    ix.save(p)
    back = BlockIndex.load(p)
    for k in ("coff", "bsize", "uoff", "isize", "crc", "first_rec", "first_idx", "first_ref", "first_pos"):
        assert np.array_equal(getattr(back, k), getattr(ix, k)), k
    assert (back.n_records, back.stream_bytes) == (ix.n_records, ix.stream_bytes)
    assert back.header.names == ix.header.names and back.header.size == ix.header.size
