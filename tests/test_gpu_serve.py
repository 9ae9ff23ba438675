"""Block index, batched span decode and the request loop against the oracle (needs a GPU).

Every comparison is exact. A span's records are compared with the oracle's
decode of the whole file, sliced to the span's records; a request served in a
mixed batch is compared with the same request decoded alone.
"""

import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIXED = ("ref", "pos", "end", "mapq", "flag", "nref", "npos", "tlen", "l_seq", "clip_left", "clip_right")


def _say(*a):
    print(*a, flush=True, file=sys.stderr)


@pytest.fixture(scope="module")
def cbm():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from cubam_amd import _lib

    _lib.load()
    import cubam_amd

    return cubam_amd


@pytest.fixture(scope="module")
def files(cbm, synth, oracle):
    """Two GEX-like and two ATAC-like files, with the oracle's whole-file decode of each."""
    from cubam_amd.serve import BamSource

    out = []
    for kind, seed, n in ((0, 31, 40_000), (1, 32, 60_000), (0, 33, 25_000), (1, 34, 30_000)):
        b = synth.make_bam(kind, seed, n, n_total=20_000_000)
        src = BamSource(b.bgzf, name=f"kind{kind}-seed{seed}")
        (data, cols), = oracle.read_bam(b.bgzf, b.bgzf.size + 1)
        _say("file", kind, seed, n, b.bgzf.size, "bytes", len(src.index), "blocks")
        out.append((b, src, data, cols))
    return out


def _host(rb, tags=("CB", "UB")):
    """A result as host arrays, every offset made relative to its first record."""
    o = rb.offset.cpu().numpy()
    base = int(o[0]) if o.size else 0
    h = {"offset": o - base}
    for k in FIXED:
        h[k] = getattr(rb, k).cpu().numpy().astype(np.int64)
    for k in ("cigar", "seq", "qual"):
        off = getattr(rb, k + "_off").cpu().numpy()
        vals = getattr(rb, k).cpu().numpy()
        h[k + "_off"] = off - off[0]
        h[k] = vals[off[0] : off[-1]].astype(np.int64)
    for t in tags:
        tg = rb.tag(t)
        ty = tg.type.cpu().numpy()
        h[t + "_type"] = ty
        h[t + "_length"] = tg.length.cpu().numpy()
        h[t + "_value"] = tg.value.cpu().numpy()
        h[t + "_offset"] = np.where(ty != 0, tg.offset.cpu().numpy() - o, 0)
    h["bytes"] = rb.data[rb.offset[0] : rb.offset[-1] + 4].cpu().numpy().tobytes() if o.size else b""
    return h


def _oracle_slice(data, cols, r0, r1, tags=("CB", "UB")):
    o = cols["offset"][r0:r1]
    base = int(o[0])
    h = {"offset": o - base}
    for k in FIXED:
        h[k] = cols[k][r0:r1].astype(np.int64)
    for k in ("cigar", "seq", "qual"):
        off = cols[k + "_off"][r0 : r1 + 1]
        h[k + "_off"] = off - off[0]
        h[k] = cols[k][off[0] : off[-1]].astype(np.int64)
    for t in tags:
        c = cols[f"tag_{t}"]
        ty = c["type"][r0:r1]
        h[t + "_type"] = ty
        h[t + "_length"] = c["length"][r0:r1]
        h[t + "_value"] = c["value"][r0:r1]
        h[t + "_offset"] = np.where(ty != 0, c["offset"][r0:r1] - o, 0)
    h["bytes"] = data[base : int(o[-1]) + 4].tobytes()
    return h


def _same(a, b, what=""):
    assert a.keys() == b.keys()
    for k in a:
        if k == "bytes":
            assert a[k] == b[k], (what, k)
        else:
            assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), (what, k)


def test_index_matches_oracle(files, oracle):
    for b, src, data, cols in files:
        ix = src.index
        off, bs, isz, crc, used, status = oracle.bgzf_scan(b.bgzf, True)
        keep = isz > 0
        assert status == 0
        assert np.array_equal(ix.coff[ix.isize > 0], off[keep]) and np.array_equal(ix.bsize[ix.isize > 0], bs[keep])
        assert np.array_equal(ix.crc[ix.isize > 0], crc[keep])
        assert ix.n_records == b.n_records == cols["offset"].size
        h = ix.header.size
        stream_off = cols["offset"] + h
        assert ix.stream_bytes == data.size + h
        # first record of each block, by brute force over all records
        ends = ix.uoff + ix.isize
        blk = np.searchsorted(ends, stream_off, side="right")
        for k in range(len(ix)):
            sel = np.flatnonzero(blk == k)
            if sel.size == 0:
                assert ix.first_rec[k] == -1
                continue
            i = int(sel[0])
            assert ix.first_rec[k] == stream_off[i] and ix.first_idx[k] == i
            assert ix.first_ref[k] == cols["ref"][i] and ix.first_pos[k] == cols["pos"][i]


def test_index_save_load(files, tmp_path):
    from cubam_amd.index import BlockIndex

    ix = files[1][1].index
    ix.save(tmp_path / "i.npz")
    j = BlockIndex.load(tmp_path / "i.npz")
    assert j.header == ix.header and j.n_records == ix.n_records
    for k in ("coff", "bsize", "uoff", "isize", "crc", "first_rec", "first_idx", "first_ref", "first_pos"):
        assert np.array_equal(getattr(j, k), getattr(ix, k)), k


def _random_spans(files, n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        f = int(rng.integers(len(files)))
        ix = files[f][1].index
        rb = ix.record_blocks()
        i0 = int(rng.integers(rb.size))
        i1 = int(min(rb.size, i0 + 1 + rng.integers(0, max(2, rb.size // 3))))
        out.append((f, ix.span(i0, i1)))
    # whole files and single blocks at both ends
    out.append((0, files[0][1].index.span(0, files[0][1].index.record_blocks().size)))
    out.append((1, files[1][1].index.span(0, 1)))
    rb = files[3][1].index.record_blocks()
    out.append((3, files[3][1].index.span(rb.size - 1, rb.size)))
    return out


def test_decode_spans_matches_oracle(files):
    from cubam_amd.serve import decode_spans

    spans = _random_spans(files, 12, 5)
    res = decode_spans([(files[f][1], sp) for f, sp in spans])
    for (f, sp), rb in zip(spans, res):
        b, src, data, cols = files[f]
        assert len(rb) == sp.records
        _same(_host(rb), _oracle_slice(data, cols, sp.rec0, sp.rec1), (f, sp.k0, sp.k1))


def test_alone_equals_mixed_and_plain(files):
    """Each request decodes the same alone, in a mixed batch, and by the per-chunk path."""
    from cubam_amd.serve import decode_plain, decode_spans

    spans = _random_spans(files, 9, 6)
    mixed = decode_spans([(files[f][1], sp) for f, sp in spans])
    for (f, sp), rb in zip(spans, mixed):
        alone = decode_spans([(files[f][1], sp)])[0]
        plain = decode_plain(files[f][1], sp)
        hm = _host(rb)
        _same(hm, _host(alone), ("alone", f))
        _same(hm, _host(plain), ("plain", f))


def test_region_holds_every_record(files):
    b, src, data, cols = files[0]
    ix = src.index
    rng = np.random.default_rng(7)
    for _ in range(20):
        ref = int(rng.integers(len(ix.header.names)))
        beg = int(rng.integers(0, ix.header.lengths[ref]))
        end = beg + int(rng.integers(1, 5_000_000))
        sp = ix.region(ref, beg, end)
        want = np.flatnonzero((cols["ref"] == ref) & (cols["pos"] >= beg) & (cols["pos"] < end))
        if want.size:
            assert sp.rec0 <= want[0] and want[-1] < sp.rec1, (ref, beg, end)


@pytest.mark.parametrize("workers, budget, batch", [(3, 64 << 30, 512 << 20), (2, 40 << 20, 8 << 20)])
def test_server_results_equal_alone(files, workers, budget, batch):
    """Continuous batching under a budget: every served result equals the request decoded alone."""
    from cubam_amd.serve import Server, decode_spans, make_trace, run_trace

    srcs = [f[1] for f in files]
    trace = make_trace(srcs, 40, rate=400.0, seed=8, min_bytes=16 << 10, max_bytes=4 << 20)
    with Server(srcs, budget_bytes=budget, workers=workers, batch_bytes=batch) as srv:
        reqs = run_trace(srv, trace)
        assert all(r.error is None for r in reqs)
        got = []
        for r in reqs:
            if r.result is None:  # evicted under the budget: decode it again to compare
                continue
            got.append((r, _host(r.result)))
        assert max(srv.batches) >= 1 and srv.peak_bytes <= budget
    assert got
    _say("served", len(reqs), "requests in", len(srv.batches), "batches; kept", len(got))
    for r, h in got:
        alone = decode_spans([(srcs[r.file], r.span)])[0]
        _same(h, _host(alone), r.id)


def test_plain_server(files):
    from cubam_amd.serve import Server, make_trace, run_trace, summarize

    srcs = [f[1] for f in files]
    trace = make_trace(srcs, 10, rate=0, seed=9, min_bytes=16 << 10, max_bytes=1 << 20)
    with Server(srcs, mode="plain") as srv:
        reqs = run_trace(srv, trace)
    assert all(r.error is None for r in reqs) and all(r.batch_size == 1 for r in reqs)
    s = summarize(reqs)
    assert s["requests"] == 10 and s["errors"] == 0


def test_bad_request_does_not_poison_batch(files):
    """A span that fails its CRC32 check fails alone; the rest of its batch is served."""
    import dataclasses

    from cubam_amd.serve import BamSource, Server

    b, src, data, cols = files[1]
    sp = src.index.span(3, 6)
    crc = src.index.crc.copy()
    crc[sp.k0 + 1] ^= 1  # the index disagrees with the block's data
    bad = BamSource(src.host.numpy(), index=dataclasses.replace(src.index, crc=crc), name="bad")
    srcs = [src, bad]
    with Server(srcs, workers=1) as srv:
        good = srv.submit(0, src.index.span(0, 4))
        broken = srv.submit(1, sp)
        other = srv.submit(0, src.index.span(5, 9))
        for r in (good, broken, other):
            r.done.wait()
    assert good.error is None and other.error is None
    assert isinstance(broken.error, ValueError) and "CRC32" in str(broken.error)


def test_mixed_read_write_requests(files, oracle):
    """Write requests in the loop: the records of a barcode group as BGZF blocks, exact against
    the oracle's records, the same alone and in a mixed batch."""
    import gzip
    import zlib

    import torch

    from cubam_amd.serve import Server, decode_spans, make_trace, run_trace, write_back

    srcs = [f[1] for f in files]
    b, src, data, cols = files[1]  # ATAC-like
    tb = cols["tag_CB"]
    cb = [data[o : o + n].tobytes().decode() if t == ord("Z") else "" for t, o, n in
          zip(tb["type"], tb["offset"], tb["length"])]
    pool = sorted(set(cb) - {""})
    keep = pool[::5]
    cells = torch.tensor([oracle.pack_acgt(c.encode())[0] for c in keep], dtype=torch.int64).view(torch.int64)
    cells = cells.cuda()
    trace = [t for t in make_trace(srcs, 30, rate=300.0, seed=12, min_bytes=16 << 10, max_bytes=2 << 20,
                                   write_fraction=1.0, cells=cells) if t.file == 1]
    reads = make_trace(srcs, 20, rate=300.0, seed=13, min_bytes=16 << 10, max_bytes=2 << 20)
    mixed = sorted(trace + reads, key=lambda t: t.t)
    with Server(srcs, workers=2) as srv:
        reqs = run_trace(srv, mixed)
    ks = set(keep)
    n_checked = 0
    for r in reqs:
        assert r.error is None
        if r.cells is None:
            continue
        res = r.result
        sp = r.span
        want = b"".join(
            data[cols["offset"][i] : cols["offset"][i] + 4 + int.from_bytes(
                data[cols["offset"][i] : cols["offset"][i] + 4].tobytes(), "little")].tobytes()
            for i in range(sp.rec0, sp.rec1) if cb[i] in ks)
        blob = res.blocks.cpu().numpy().tobytes()
        assert (gzip.decompress(blob) if blob else b"") == want and res.payload_bytes == len(want)
        # blocks cut as htslib cuts them, each with zlib's CRC32
        o = 0
        for a in range(0, len(want), 65280):
            bs = (blob[o + 16] | (blob[o + 17] << 8)) + 1
            assert int.from_bytes(blob[o + bs - 8 : o + bs - 4], "little") == zlib.crc32(want[a : a + 65280])
            o += bs
        alone = write_back(decode_spans([(srcs[1], sp)])[0], cells)
        assert torch.equal(alone.blocks.cpu(), res.blocks.cpu())
        n_checked += 1
    assert n_checked >= 3
