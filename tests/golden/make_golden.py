"""Write the golden fixtures of tests/golden/ (run from the repository root).

Inputs come from the seeded generator (cubam_amd/synth) and from zlib at mixed
levels; expected outputs come from zlib and from the pure-Python SAMv1 parser
in tests/bamfixture.py, never from the oracle or the GPU path.

  gex.bam, atac.bam       small GEX-like / ATAC-like BAMs, 4,000-byte payloads
  *_expect.npz            every fixed column, CSR CIGAR/seq/qual, CB/UB strings
  blocks.bgzf             BGZF blocks at levels {0,1,6,9}, some fixed-Huffman only
  blocks_expect.npz       per-block payload bytes and CRC32
"""

from __future__ import annotations

import gzip
import sys
import zlib
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))
sys.path.insert(0, str(HERE.parent))

import bamfixture as bf  # noqa: E402

FIXED = ("ref", "pos", "end", "mapq", "flag", "nref", "npos", "tlen", "l_seq", "clip_left", "clip_right")


def _strings(vals):
    b = [v if isinstance(v, bytes) else v.encode() for v in vals]
    off = np.zeros(len(b) + 1, np.int64)
    np.cumsum([len(x) for x in b], out=off[1:])
    return np.frombuffer(b"".join(b), np.uint8).copy(), off


def bam_fixture(name: str, kind: int, seed: int, n: int):
    from cubam_amd import synth

    b = synth.make_bam(kind, seed, n, n_total=5_000_000, block=4000)
    (HERE / f"{name}.bam").write_bytes(b.bgzf.tobytes())
    raw = gzip.decompress(b.bgzf.tobytes())
    recs = bf.parse_records(raw, len(b.header))
    assert len(recs) == n
    out = {k: np.asarray([r[k] for r in recs], np.int64) for k in FIXED}
    out["offset"] = np.asarray([r["offset"] for r in recs], np.int64) - len(b.header)
    for k in ("seq", "qual"):
        out[k], out[k + "_off"] = _strings([r[k] for r in recs])
    cig = [np.asarray(r["cigar"], np.uint32) for r in recs]
    out["cigar"] = np.concatenate(cig)
    out["cigar_off"] = np.concatenate([[0], np.cumsum([c.size for c in cig])]).astype(np.int64)
    for t in ("CB", "UB"):
        vals = [bf.aux_tags(r["aux"]).get(t, ("Z", ""))[1] for r in recs]
        out[f"{t}_chars"], out[f"{t}_off"] = _strings(vals)
    out["header_size"] = np.asarray([len(b.header)])
    np.savez_compressed(HERE / f"{name}_expect.npz", **out)


def blocks_fixture():
    rng = np.random.default_rng(9)
    src = bf.bam_bytes(bf.random_reads(900, seed=9))
    payloads, blocks, levels = [], [], []
    pos = 0
    for i in range(24):
        k = int(rng.integers(1, 6000))
        p = src[pos : pos + k] if pos + k <= len(src) else bytes(rng.integers(0, 256, k, dtype=np.uint8))
        pos += k
        lvl = int([0, 1, 6, 9][i % 4])
        strat = zlib.Z_FIXED if (lvl and i % 5 == 0) else zlib.Z_DEFAULT_STRATEGY
        payloads.append(p)
        blocks.append(bf.bgzf_block(p, lvl, strat))
        levels.append(lvl)
    (HERE / "blocks.bgzf").write_bytes(b"".join(blocks) + bf.EOF_BLOCK)
    pl, off = _strings(payloads)
    np.savez_compressed(HERE / "blocks_expect.npz", payload=pl, payload_off=off,
                        crc=np.asarray([zlib.crc32(p) for p in payloads], np.int64),
                        level=np.asarray(levels, np.int64))


if __name__ == "__main__":
    bam_fixture("gex", 0, 100, 300)
    bam_fixture("atac", 1, 101, 400)
    blocks_fixture()
    for p in sorted(HERE.iterdir()):
        print(p.name, p.stat().st_size)
