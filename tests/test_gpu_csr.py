"""The CSR offset columns (cigar_off, seq_off, qual_off) that cbm_record_offsets_csr writes from
the record walks' in-segment sums, against the ones cbm_scatter computes record by record and
against the oracle's cumulative sums (src/cubam/bam.py:275-288). GPU."""

import ctypes
import gzip

import numpy as np
import pytest

import bamfixture as bf
from gpucheck import say

pytestmark = pytest.mark.gpu

CSR = ("cigar_off", "seq_off", "qual_off")


@pytest.fixture(scope="module")
def cbm():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from cubam_amd import _lib

    _lib.load()
    import cubam_amd

    return cubam_amd


def _both(data, start, n, n_ref, strict=True):
    """The chain's CSR columns, and those of the record-by-record path over the same offsets."""
    from cubam_amd import _decode

    chain = _decode.record_chain(data, start, n, n_ref, strict=strict)
    off, csr = _decode.record_offsets(data, start, n, chain, csr=True)
    plain = _decode.scatter(data, off, chain, tags=(), fixed=False)["fixed"]
    return chain, off.cpu().numpy(), {k: csr[k].cpu().numpy() for k in CSR}, {k: plain[k].cpu().numpy() for k in CSR}


def _check(raw: np.ndarray, start: int, n_ref: int):
    from cubam_amd import _decode
    from oracle import oracle

    data = _decode.to_device(raw)
    chain, off, csr, plain = _both(data, start, int(raw.size), n_ref)
    ooff, tail, st = oracle.record_chain(raw[start:], n_ref)
    assert off.tolist() == (ooff + start).tolist()
    oc = oracle.columns(raw, off, tags=())
    for k in CSR:
        assert np.array_equal(csr[k], plain[k]), k
        assert np.array_equal(csr[k], oc[k]), k
    assert (csr["cigar_off"][-1], csr["seq_off"][-1], csr["qual_off"][-1]) == (chain.n_cigar, chain.n_seq,
                                                                               chain.n_qual)
    return data, chain, csr


@pytest.mark.parametrize("kind, n", [(0, 30_000), (1, 60_000)])
def test_chain_csr_matches_per_record_scan(cbm, oracle, synth, kind, n):
    b = synth.make_bam(kind, 5 + kind, n, n_total=20_000_000)
    raw = np.frombuffer(gzip.decompress(b.bgzf.tobytes()), np.uint8)
    say("csr synthetic", kind, n, raw.size)
    _check(raw, len(b.header), 25)


def test_chain_csr_of_the_smallest_records(cbm, oracle):
    """Records of 38 to 45 bytes (one-letter names, no CIGAR, 0-5 bases): the most records a walk
    segment holds, odd and even l_seq."""
    reads = [bf.Read(name="abcdefg"[i % 7], seq="ACGTA"[: i % 6]) for i in range(60_000)]
    raw = np.frombuffer(bf.header_bytes() + b"".join(r.encode() for r in reads), np.uint8)
    say("csr smallest records", raw.size)
    _check(raw, len(bf.header_bytes()), len(bf.REFS))


def test_chain_csr_of_long_reads(cbm, oracle):
    """Reads of 3-9 kb with long CIGARs: records longer than a walk segment, segments with no
    record start."""
    rng = np.random.default_rng(4)
    reads = []
    for i in range(300):
        m = int(rng.integers(3000, 9000))
        cig = [("M", 7), ("I", 1)] * (m // 16) + [("M", m - 8 * (m // 16))]
        reads.append(bf.Read(name=f"long{i}", ref=0, pos=1000 + 40 * i, cigar=cig, seq="ACGT" * (m // 4) + "A" * (m % 4)))
    raw = np.frombuffer(bf.header_bytes() + b"".join(r.encode() for r in reads), np.uint8)
    say("csr long reads", raw.size)
    _check(raw, len(bf.header_bytes()), len(bf.REFS))


def test_chain_csr_of_fewer_records_than_the_chain(cbm, synth):
    """n_rec below the chain's count: entries 0..n_rec of each column, as the full columns."""
    import torch

    from cubam_amd import _decode, _lib

    b = synth.make_bam(1, 9, 20_000, n_total=20_000_000)
    raw = np.frombuffer(gzip.decompress(b.bgzf.tobytes()), np.uint8)
    data = _decode.to_device(raw)
    start, n = len(b.header), int(raw.size)
    chain, off, csr, plain = _both(data, start, n, 25)
    L = _lib.load()
    for m in (1, 777, chain.n_rec - 1):
        o = torch.empty(m, dtype=torch.int64, device=data.device)
        cols = [torch.full((m + 1,), -5, dtype=torch.int64, device=data.device) for _ in CSR]
        _lib.check(L.cbm_record_offsets_csr(data.data_ptr(), start, n, chain.ws.data_ptr(), chain.ws.numel(),
                                            o.data_ptr(), *(c.data_ptr() for c in cols), m, _lib.stream_ptr()),
                   "cbm_record_offsets_csr")
        assert o.cpu().numpy().tolist() == off[:m].tolist()
        for k, c in zip(CSR, cols):
            assert c.cpu().numpy().tolist() == csr[k][: m + 1].tolist(), (m, k)


def test_chain_csr_after_repair(cbm, oracle):
    """A walk that does not link is repaired (its records are not recorded): the CSR columns of
    the repaired chain come from the serial walk of that segment. Fake records start segments 2
    to 5, so that some of them start segments whose walk searches for its first record."""
    from cubam_amd import _decode, _lib

    L = _lib.load()
    seg = ctypes.c_int64()
    L.cbm_record_chain_workspace(1 << 20, ctypes.byref(seg))
    seg = seg.value
    reads = bf.random_reads(300, seed=14)
    fake = b"".join(bf.Read(f"fake{i}", 0, 0, 100 + i, 60, [("M", 10)], seq="ACGTACGTAC").encode() for i in range(2))
    recs, o = [], 0
    while o < 2 * seg - 3000:
        r = reads[len(recs)].encode()
        recs.append(r)
        o += len(r)
    host = bf.Read("host", 0, 0, reads[len(recs)].pos, 60, [("M", 10)], seq="ACGTACGTAC")
    val = o + len(host.encode()) + 3 + 1 + 4
    arr = b"\xff" * (2 * seg - val) + (fake + b"\xff" * (seg - len(fake))) * 3 + fake + b"\xff" * 300
    host.tags = [("XB", "B", ("C", list(arr)))]
    recs.append(host.encode())
    recs += [r.encode() for r in reads[len(recs):]]
    raw = np.frombuffer(b"".join(recs), np.uint8)
    assert all(raw[k * seg : k * seg + len(fake)].tobytes() == fake for k in range(2, 6))
    data = _decode.to_device(raw)
    chain, off, csr, plain = _both(data, 0, int(raw.size), len(bf.REFS))
    assert chain.bad_links > 0 and len(off) == len(reads)
    oc = oracle.columns(raw, off, tags=())
    for k in CSR:
        assert np.array_equal(csr[k], plain[k]) and np.array_equal(csr[k], oc[k]), k


def test_chain_csr_of_an_empty_chain(cbm):
    import torch

    from cubam_amd import _decode, _lib

    raw = np.frombuffer(bf.header_bytes(), np.uint8)
    data = _decode.to_device(raw)
    chain = _decode.record_chain(data, int(raw.size), int(raw.size), len(bf.REFS))
    assert chain.n_rec == 0
    off, csr = _decode.record_offsets(data, int(raw.size), int(raw.size), chain, csr=True)
    assert all(csr[k].cpu().numpy().tolist() == [0] for k in CSR)
    torch.cuda.synchronize()
