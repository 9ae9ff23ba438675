"""The BAI index (SAMv1 5.2): bins, chunks, linear index and pseudo-bin against a brute-force
builder written from the specification, the binary layout, region queries against brute force,
and (GPU) the index of a synthetic BAM built on the GPU against one built from the oracle's
columns, with region reads against the oracle's records."""

import random
import struct

import numpy as np
import pytest
import torch

from cubam_amd import bai


def _reg2bin_spec(beg, end):
    # SAMv1 5.3, written out level by level
    end -= 1
    if beg >> 14 == end >> 14:
        return ((1 << 15) - 1) // 7 + (beg >> 14)
    if beg >> 17 == end >> 17:
        return ((1 << 12) - 1) // 7 + (beg >> 17)
    if beg >> 20 == end >> 20:
        return ((1 << 9) - 1) // 7 + (beg >> 20)
    if beg >> 23 == end >> 23:
        return ((1 << 6) - 1) // 7 + (beg >> 23)
    if beg >> 26 == end >> 26:
        return ((1 << 3) - 1) // 7 + (beg >> 26)
    return 0


def test_reg2bin_matches_the_specification():
    rng = random.Random(3)
    spans = [(0, 1), (16383, 16384), (16383, 16385), (0, 1 << 29), (131071, 131073), ((1 << 26) - 1, 1 << 26)]
    spans += [(b, b + rng.choice([1, 2, 100, 20000, 300000, 5000000])) for b in
              (rng.randrange(0, 1 << 28) for _ in range(3000))]
    for b, e in spans:
        assert bai.reg2bin(b, e) == _reg2bin_spec(b, e)
    bb = np.array([b for b, _ in spans], np.int64)
    ee = np.array([e for _, e in spans], np.int64)
    want = [_reg2bin_spec(b, e) for b, e in spans]
    assert bai.reg2bin(bb, ee).tolist() == want
    assert bai.reg2bin(torch.from_numpy(bb), torch.from_numpy(ee)).tolist() == want


def test_reg2bins_holds_every_bin_of_an_overlapping_region():
    rng = random.Random(4)
    for _ in range(300):
        beg = rng.randrange(0, 1 << 27)
        end = beg + rng.randrange(1, 1 << 20)
        bins = set(bai.reg2bins(beg, end))
        for _ in range(20):  # a record overlapping [beg, end) has its bin in the list
            rb = rng.randrange(max(0, beg - 100000), end)
            re = max(rb + 1, rng.randrange(rb + 1, rb + 200000))
            if re > beg:
                assert _reg2bin_spec(rb, re) in bins


def _records(seed, n_ref=3, n=4000, lengths=(3_000_000, 500_000, 2_000_000)):
    """Sorted synthetic records with virtual offsets: (ref, pos, end, flag, vbeg, vend)."""
    rng = np.random.default_rng(seed)
    ref = np.sort(rng.integers(0, n_ref, n))
    pos = np.empty(n, np.int64)
    for r in range(n_ref):
        m = ref == r
        pos[m] = np.sort(rng.integers(0, lengths[r] - 60000, m.sum()))
    rlen = rng.choice([0, 1, 50, 100, 90, 30000], n, p=[0.02, 0.03, 0.4, 0.4, 0.1, 0.05])
    flag = np.where(rlen == 0, 4, 0) | np.where(rng.random(n) < 0.5, 16, 0)
    end = pos + rlen
    # a tail of unplaced records (no coordinate)
    k = 50
    ref = np.concatenate([ref, np.full(k, -1)])
    pos = np.concatenate([pos, np.full(k, -1)])
    end = np.concatenate([end, np.full(k, -1)])
    flag = np.concatenate([flag, np.full(k, 4)])
    size = rng.integers(60, 400, n + k)
    stream = np.concatenate([[0], np.cumsum(size)])  # record starts, then the end
    block = 65280
    # blocks of 65,280 inflated bytes at file offsets of about 30 kB each
    blk = stream // block
    coff = np.cumsum(np.concatenate([[0], rng.integers(20000, 40000, int(blk.max()) + 2)]))
    v = (coff[blk] << 16) | (stream - blk * block)
    return ref, pos, end, flag, v[:-1], v[1:], lengths


def _brute(ref, pos, end, flag, vbeg, vend, lengths):
    """The index written straight from the specification, one record at a time."""
    n_ref = len(lengths)
    bins = [dict() for _ in range(n_ref)]
    lin = [dict() for _ in range(n_ref)]
    meta = [None] * n_ref
    no_coor = 0
    prev = None
    for i in range(len(ref)):
        r, p = int(ref[i]), int(pos[i])
        if r < 0 or p < 0:
            no_coor += 1
            prev = None
            continue
        e = max(int(end[i]), p + 1)
        b = _reg2bin_spec(p, e)
        ch = bins[r].setdefault(b, [])
        if prev == (r, b):
            ch[-1][1] = int(vend[i])
        else:
            ch.append([int(vbeg[i]), int(vend[i])])
        prev = (r, b)
        mapped = not (int(flag[i]) & 4)
        if mapped:
            for w in range(p >> 14, ((e - 1) >> 14) + 1):
                if w <= lengths[r] >> 14:
                    lin[r][w] = min(lin[r].get(w, 1 << 62), int(vbeg[i]))
        m = meta[r] or [int(vbeg[i]), int(vend[i]), 0, 0]
        m[0], m[1] = min(m[0], int(vbeg[i])), max(m[1], int(vend[i]))
        m[2 if mapped else 3] += 1
        meta[r] = m
    out = []
    for r in range(n_ref):
        if lin[r]:
            top = max(lin[r])
            arr, last = [], 0
            for w in range(top + 1):
                last = lin[r].get(w, last)
                arr.append(last)
        else:
            arr = []
        out.append(({b: np.asarray(c, np.uint64) for b, c in bins[r].items()}, np.asarray(arr, np.uint64),
                    tuple(meta[r]) if meta[r] else None))
    return out, no_coor


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_build_matches_the_brute_force_index(seed):
    rec = _records(seed)
    got = bai.build(*(torch.from_numpy(np.asarray(x)) for x in rec[:6]), rec[6])
    want, no_coor = _brute(*rec)
    assert got.n_no_coor == no_coor == 50
    for r, (wb, wl, wm) in enumerate(want):
        g = got.refs[r]
        assert sorted(g.bins) == sorted(wb)
        for b in wb:
            assert np.array_equal(g.bins[b], wb[b]), (r, b)
        assert np.array_equal(g.linear, wl)
        assert g.meta == wm


def test_binary_layout_round_trip_and_fields():
    rec = _records(5)
    idx = bai.build(*(torch.from_numpy(np.asarray(x)) for x in rec[:6]), rec[6])
    raw = idx.to_bytes()
    assert raw[:4] == b"BAI\x01" and struct.unpack_from("<i", raw, 4)[0] == 3
    back = bai.Bai.from_bytes(raw)
    assert back.to_bytes() == raw
    assert back.n_no_coor == 50
    # the first reference's first bin record, parsed by hand: bin, n_chunk, chunks
    n_bin = struct.unpack_from("<i", raw, 8)[0]
    assert n_bin == len(idx.refs[0].bins) + 1  # and the pseudo-bin
    b0, nc0 = struct.unpack_from("<Ii", raw, 12)
    assert b0 == min(idx.refs[0].bins)
    assert np.array_equal(np.frombuffer(raw, "<u8", 2 * nc0, 20).reshape(-1, 2), idx.refs[0].bins[b0])
    with pytest.raises(ValueError, match="magic"):
        bai.Bai.from_bytes(b"BAM\x01" + raw[4:])


def test_csi_layout_round_trip_and_queries():
    ref, pos, end, flag, vbeg, vend, lengths = _records(8)
    idx = bai.build(*(torch.from_numpy(np.asarray(x)) for x in (ref, pos, end, flag, vbeg, vend)), lengths)
    raw = idx.to_csi_bytes()
    assert raw[:4] == b"CSI\x01" and struct.unpack_from("<iiii", raw, 4) == (14, 5, 0, 3)
    # the first bin of reference 0: bin, loffset (the linear index at its first window), chunks
    b0, loff0, nc0 = struct.unpack_from("<IQi", raw, 24)
    assert b0 == min(idx.refs[0].bins)
    first_w = 0 if b0 < 4681 else b0 - 4681
    assert loff0 == (int(idx.refs[0].linear[first_w]) if first_w < idx.refs[0].linear.size else 0)
    back = bai.Bai.from_csi_bytes(raw)
    assert back.to_csi_bytes() == raw
    assert back.n_no_coor == idx.n_no_coor
    for r in range(3):
        assert sorted(back.refs[r].bins) == sorted(idx.refs[r].bins)
        assert back.refs[r].meta == idx.refs[r].meta
    endx = np.maximum(end, pos + 1)
    rng = random.Random(8)
    for _ in range(200):  # the CSI's queries hold every overlapping record too
        r = rng.randrange(3)
        beg = rng.randrange(0, lengths[r])
        e = beg + rng.choice([1, 100, 5000, 100000])
        ch = back.query(r, beg, e)
        for i in np.flatnonzero((ref == r) & (pos >= 0) & (pos < e) & (endx > beg)):
            assert np.any((ch[:, 0] <= vbeg[i]) & (vend[i] <= ch[:, 1]))


@pytest.mark.parametrize("seed", [6, 7])
def test_query_returns_every_overlapping_record(seed):
    ref, pos, end, flag, vbeg, vend, lengths = _records(seed)
    idx = bai.build(*(torch.from_numpy(np.asarray(x)) for x in (ref, pos, end, flag, vbeg, vend)), lengths)
    rng = random.Random(seed)
    endx = np.maximum(end, pos + 1)
    for _ in range(400):
        r = rng.randrange(3)
        beg = rng.randrange(0, lengths[r])
        e = beg + rng.choice([1, 100, 5000, 100000, 1000000])
        ch = idx.query(r, beg, e)
        assert np.all(ch[1:, 0] > ch[:-1, 1])  # sorted and merged
        hit = np.flatnonzero((ref == r) & (pos >= 0) & (pos < e) & (endx > beg))
        for i in hit:
            assert np.any((ch[:, 0] <= vbeg[i]) & (vend[i] <= ch[:, 1])), (r, beg, e, i)
    assert idx.query(9, 0, 10).shape == (0, 2)


# ---------------------------------------------------------------------------- GPU


@pytest.fixture(scope="module")
def cbm():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from cubam_amd import _lib

    _lib.load()
    import cubam_amd

    return cubam_amd


@pytest.mark.gpu
@pytest.mark.parametrize("kind,seed,n", [(0, 11, 60_000), (1, 12, 80_000)])
def test_gpu_index_equals_the_oracle_built_index(cbm, oracle, synth, kind, seed, n):
    """index_bam (GPU decode; virtual offsets, bins, chunks and windows on the device) against
    build() fed with the oracle's records and columns and the virtual offsets of its block
    table; then region reads through the index against the oracle's overlapping records."""
    b = synth.make_bam(kind, seed, n, n_total=n * 3, block=20_000)
    raw = b.bgzf
    bix, got = bai.index_bam(raw, chunk_bytes=300_000)
    # the oracle: blocks, the inflated stream, the records after the header, their columns
    blk_off, bsize, isize, _, _, st = oracle.bgzf_scan(raw, True)
    assert st == 0
    data, uoff, _, ist = oracle.inflate(raw, blk_off, bsize, isize)
    assert (ist == 0).all()
    h = oracle.read_header(raw)
    off, tail, cst = oracle.record_chain(data[h.size:], len(h.names))
    assert cst == 0 and tail == data.size - h.size
    cols = oracle.columns(data[h.size:], off, tags=())
    stream = off.astype(np.int64) + h.size
    uoff = np.asarray(uoff, np.int64)
    coff = np.asarray(blk_off, np.int64)
    keep = np.asarray(isize) > 0  # the EOF block holds no data
    uo, co, bs_, is_ = uoff[keep], coff[keep], np.asarray(bsize)[keep], np.asarray(isize)[keep]
    k = np.searchsorted(uo, stream, side="right") - 1
    vbeg = (co[k] << 16) | (stream - uo[k])
    kl = int(np.searchsorted(uo, data.size - 1, side="right")) - 1
    assert data.size == uo[kl] + is_[kl]
    vend = np.concatenate([vbeg[1:], [int(co[kl] + bs_[kl]) << 16]])
    want = bai.build(*(torch.from_numpy(np.asarray(x).astype(np.int64)) for x in
                       (cols["ref"], cols["pos"], cols["end"], cols["flag"], vbeg, vend)), h.lengths)
    assert got.to_bytes() == want.to_bytes()
    assert len(got.refs) == len(h.names)
    # region reads: every record overlapping the region, and only those
    from cubam_amd.serve import BamSource

    src = BamSource(raw, index=bix)
    rng = random.Random(seed)
    ref = np.asarray(cols["ref"]).astype(np.int64)
    pos = np.asarray(cols["pos"]).astype(np.int64)
    endx = np.maximum(np.asarray(cols["end"]).astype(np.int64), pos + 1)
    for _ in range(12):
        i = rng.randrange(ref.size)
        if ref[i] < 0:
            continue
        r, beg = int(ref[i]), max(0, int(pos[i]) - rng.choice([0, 500, 20000]))
        end = beg + rng.choice([1, 1000, 100000, 3000000])
        batches = bai.read_region(src, got, r, beg, end, tags=())
        got_off = np.sort(np.concatenate([x.offset.cpu().numpy() for x in batches] or [np.empty(0, np.int64)]))
        hit = np.flatnonzero((ref == r) & (pos >= 0) & (pos < end) & (endx > beg))
        # a batch's offsets are in its own decoded span; compare the records' positions and ends
        got_pe = sorted(zip(np.concatenate([x.pos.cpu().numpy() for x in batches] or [[]]).tolist(),
                            np.concatenate([x.end.cpu().numpy() for x in batches] or [[]]).tolist()))
        want_pe = sorted(zip(pos[hit].tolist(), np.asarray(cols["end"])[hit].astype(np.int64).tolist()))
        assert got_pe == want_pe, (r, beg, end)
        assert got_off.size == hit.size
