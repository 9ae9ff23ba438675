"""On-demand RecordBatch / Tag kernels against closed forms over the fixture's reads (needs a GPU).

Expected values come from the reads as written (tests/bamfixture.py), computed here
from their CIGAR, sequence and tags by the rules of the reference
(src/cubam/bam.py:221-248, 322-389; src/cubam/_kernels.py:391-437, 498-571), and
from the oracle's FNV-1a and pack_acgt, which are pinned against published
vectors and the reference's examples in test_oracle.py.
"""

import random

import numpy as np
import pytest

import bamfixture as bf

pytestmark = pytest.mark.gpu

CIGARS = [
    [("M", 50)],
    [("S", 5), ("M", 20), ("I", 3), ("M", 10), ("D", 4), ("M", 15), ("S", 2)],
    [("M", 30), ("N", 500), ("M", 20)],
    [("H", 4), ("=", 10), ("X", 1), ("=", 9), ("D", 2), ("N", 100), ("M", 8)],
    [("M", 5), ("D", 1), ("M", 5), ("N", 10), ("D", 3), ("M", 5)],
    [("S", 10), ("M", 10)],
]
BARCODES = ["ACGTACGTACGTACGT-1", "ACGT", "A" * 24, "A" * 25, "ACGT-999", "ACGT-1234", "ACGT-", "sample_ACGT",
            "acgt-1", "-1", "", "TTTT-0", "GGGG-12"]


def _reads(n=600, seed=9):
    rng = random.Random(seed)
    reads = []
    for i in range(n):
        cig = rng.choice(CIGARS)
        length = sum(k for op, k in cig if op in "MIS=X")
        seq = "".join(rng.choices("ACGTN", weights=[25, 25, 25, 25, 3], k=length))
        tags = [("NM", "i", rng.randint(0, 9))]
        r = rng.random()
        if r < 0.7:
            tags.append(("CB", "Z", rng.choice(BARCODES)))
        elif r < 0.8:
            tags.append(("CB", "i", 7))  # not a string
        ref = rng.randrange(len(bf.REFS))
        reads.append(bf.Read(name=f"q{i}:{rng.randint(0, 99999)}", flag=0, ref=ref,
                             pos=rng.randrange(bf.REFS[ref][1] - 2000), mapq=30, cigar=cig, seq=seq, tags=tags))
    return sorted(reads, key=bf.sort_key)


@pytest.fixture(scope="module")
def batch(tmp_path_factory):
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from cubam_amd import read_bam

    reads = _reads()
    p = tmp_path_factory.mktemp("od") / "od.bam"
    p.write_bytes(bf.bgzf(bf.bam_bytes(reads), 5000))
    (rb,) = list(read_bam(p, chunk_bytes=1 << 22, tags=()))
    assert len(rb) == len(reads)
    return rb, reads


def _tag(r, name):
    for t, ty, v in r.tags:
        if t == name:
            return ty, v
    return None


def test_name_hash(batch, oracle):
    rb, reads = batch
    got = rb.name_hash().view(__import__("torch").int64).cpu().numpy().astype(np.uint64)
    want = np.array([oracle.fnv1a(r.name.encode()) for r in reads], np.uint64)
    assert np.array_equal(got, want)


def test_names(batch):
    rb, reads = batch
    chars, off = rb.names()
    c, o = chars.cpu().numpy().tobytes(), off.cpu().numpy()
    assert [c[o[i] : o[i + 1]].decode() for i in range(len(reads))] == [r.name for r in reads]


def test_tag_hash(batch, oracle):
    import torch

    rb, reads = batch
    got = rb.tag("CB").hash().view(torch.int64).cpu().numpy().astype(np.uint64)
    want = []
    for r in reads:
        t = _tag(r, "CB")
        want.append(oracle.fnv1a(t[1].encode()) if t and t[0] == "Z" else 0)
    assert np.array_equal(got, np.array(want, np.uint64))


def test_tag_packed_and_unpack(batch, oracle):
    import torch

    from cubam_amd.bam import unpack_acgt

    rb, reads = batch
    codes, ok = rb.tag("CB").packed()
    got = codes.view(torch.int64).cpu().numpy().astype(np.uint64)
    okh = ok.cpu().numpy()
    for i, r in enumerate(reads):
        t = _tag(r, "CB")
        if t and t[0] == "Z":
            want, good = oracle.pack_acgt(t[1].encode())
        else:
            want, good = 0, False
        assert (int(got[i]), bool(okh[i])) == (int(want), bool(good)), (t, got[i], want)
    chars, off = unpack_acgt(codes.view(torch.int64)[ok])
    c, o = chars.cpu().numpy().tobytes(), off.cpu().numpy()
    back = [c[o[i] : o[i + 1]].decode() for i in range(len(o) - 1)]
    assert back == [_tag(r, "CB")[1] for r, g in zip(reads, okh) if g]
    assert len(back) > 50


def test_aligned_length(batch):
    rb, reads = batch
    got = rb.aligned_length().cpu().numpy()
    assert got.tolist() == [sum(n for op, n in r.cigar if op in "M=X") for r in reads]


def _blocks_of(r, split_del):
    out, x = [], r.pos
    s = x
    for op, n in r.cigar:
        if op in "M=X" or (op == "D" and not split_del):
            x += n
        elif op == "N" or (op == "D" and split_del):
            if x > s:
                out.append((s, x))
            x += n
            s = x
    if x > s:
        out.append((s, x))
    return out


@pytest.mark.parametrize("split_del", [False, True])
def test_blocks(batch, split_del):
    rb, reads = batch
    rec, start, end = rb.blocks(split_deletions=split_del)
    got = list(zip(rec.cpu().tolist(), start.cpu().tolist(), end.cpu().tolist()))
    want = [(i, s, e) for i, r in enumerate(reads) for s, e in _blocks_of(r, split_del)]
    assert got == want


def _base_at(r, t):
    x, y = r.pos, 0
    for op, n in r.cigar:
        if op in "M=X":
            if t < x + n:
                if t >= x and y + (t - x) < len(r.seq):
                    return "ACGT".index(r.seq[y + t - x]) if r.seq[y + t - x] in "ACGT" else 4
                return -3
            x += n
            y += n
        elif op in "DN":
            if t < x + n:
                return (-1 if op == "D" else -2) if t >= x else -3
            x += n
        elif op in "IS":
            y += n
    return -3


def test_base_at(batch):
    rb, reads = batch
    rng = random.Random(3)
    rec, pos, want = [], [], []
    for _ in range(5000):
        i = rng.randrange(len(reads))
        r = reads[i]
        t = r.pos + rng.randint(-5, r.ref_len + 5)
        rec.append(i)
        pos.append(t)
        want.append(_base_at(r, t))
    got = rb.base_at(rec, pos).cpu().numpy().tolist()
    assert got == want
    assert {-3, -2, -1, 0, 1, 2, 3, 4} <= set(want)
