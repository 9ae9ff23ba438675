"""Committed golden fixtures (tests/golden/, made by make_golden.py from zlib and the pure-Python
parser) against the oracle (CPU) and the HIP path (GPU)."""

import numpy as np
import pytest

from conftest import ROOT

G = ROOT / "tests" / "golden"
FIXED = ("ref", "pos", "end", "mapq", "flag", "nref", "npos", "tlen", "l_seq", "clip_left", "clip_right")


def _load(name):
    return (G / f"{name}.bam").read_bytes(), dict(np.load(G / f"{name}_expect.npz"))


def _strs(chars, off):
    b = chars.tobytes()
    return [b[off[i] : off[i + 1]] for i in range(len(off) - 1)]


@pytest.mark.parametrize("name", ["gex", "atac"])
@pytest.mark.parametrize("chunk", [3000, 1 << 20])
def test_golden_bam_oracle(oracle, name, chunk):
    raw, exp = _load(name)
    batches = oracle.read_bam(raw, chunk, tags=("CB", "UB"))
    for k in FIXED:
        assert np.concatenate([c[k] for _, c in batches]).astype(np.int64).tolist() == exp[k].tolist(), k
    seqs, quals, cigs, cb = [], [], [], []
    for d, c in batches:
        b = d.tobytes()
        seqs += _strs(c["seq"], c["seq_off"])
        quals += _strs(c["qual"], c["qual_off"])
        cigs += [c["cigar"][c["cigar_off"][i] : c["cigar_off"][i + 1]].tolist() for i in range(len(c["offset"]))]
        t = c["tag_CB"]
        cb += [b[o : o + ln] for o, ln in zip(t["offset"], t["length"])]
    assert seqs == _strs(exp["seq"], exp["seq_off"])
    assert quals == _strs(exp["qual"], exp["qual_off"])
    assert cb == _strs(exp["CB_chars"], exp["CB_off"])
    co = exp["cigar_off"]
    assert cigs == [exp["cigar"][co[i] : co[i + 1]].tolist() for i in range(len(co) - 1)]


def test_golden_blocks_oracle(oracle):
    raw = (G / "blocks.bgzf").read_bytes()
    exp = np.load(G / "blocks_expect.npz")
    off, bs, isz, crc, used, status = oracle.bgzf_scan(raw, True)
    assert status == 0 and used == len(raw)
    keep = isz > 0
    data, _, got, st = oracle.inflate(raw, off[keep], bs[keep], isz[keep])
    assert (st == 0).all()
    assert data.tobytes() == exp["payload"].tobytes()
    assert got.astype(np.int64).tolist() == exp["crc"].tolist()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["gex", "atac"])
@pytest.mark.parametrize("chunk", [3000, 1 << 20])
def test_golden_bam_gpu(tmp_path, name, chunk):
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from cubam_amd import read_bam

    raw, exp = _load(name)
    p = tmp_path / f"{name}.bam"
    p.write_bytes(raw)
    batches = list(read_bam(p, chunk_bytes=chunk))
    for k in FIXED:
        got = torch.cat([getattr(b, k).cpu() for b in batches]).numpy().astype(np.int64)
        assert got.tolist() == exp[k].tolist(), k
    seqs, quals, cigs, cb, ub = [], [], [], [], []
    for b in batches:
        seqs += _strs(b.seq.cpu().numpy(), b.seq_off.cpu().numpy())
        quals += _strs(b.qual.cpu().numpy(), b.qual_off.cpu().numpy())
        co, c = b.cigar_off.cpu().numpy(), b.cigar.cpu().numpy()
        cigs += [c[co[i] : co[i + 1]].astype(np.int64).tolist() for i in range(len(b))]
        cb += [s.encode() for s in b.tag("CB").to_list()]
        ub += [s.encode() for s in b.tag("UB").to_list()]
    assert seqs == _strs(exp["seq"], exp["seq_off"])
    assert quals == _strs(exp["qual"], exp["qual_off"])
    assert cb == _strs(exp["CB_chars"], exp["CB_off"])
    assert ub == _strs(exp["UB_chars"], exp["UB_off"])
    co = exp["cigar_off"]
    assert cigs == [exp["cigar"][co[i] : co[i + 1]].astype(np.int64).tolist() for i in range(len(co) - 1)]


@pytest.mark.gpu
def test_golden_blocks_gpu():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from cubam_amd import _decode

    raw = np.frombuffer((G / "blocks.bgzf").read_bytes(), np.uint8)
    exp = np.load(G / "blocks_expect.npz")
    buf = _decode.to_device(raw)
    blocks = _decode.scan_blocks(buf, raw.size, True)
    inf = _decode.inflate_blocks(blocks)
    _decode.check_inflated(inf)
    assert inf.data[: inf.n].cpu().numpy().tobytes() == exp["payload"].tobytes()
    keep = blocks.isize.cpu().numpy() > 0
    assert inf.crc.cpu().numpy().astype(np.int64)[keep].tolist() == exp["crc"].tolist()
