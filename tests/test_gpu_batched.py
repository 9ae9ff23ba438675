"""Batched heterogeneous-block inflate: the autotuner and the hipGraph replay (needs a GPU).

Blocks come from the generator at zlib levels {0, 1, 6, 9} with 10 % fixed
Huffman; expected bytes are the generator's payloads, expected CRC32 is zlib's.
"""

import sys
import zlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True, file=sys.stderr)


@pytest.fixture(scope="module")
def hetero():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from cubam_amd import _decode, _lib, synth

    _lib.load()
    hb = synth.hetero_blocks(4000, 12)
    buf = _decode.to_device(hb.bgzf)
    blocks = _decode.scan_blocks(buf, int(hb.bgzf.size), True)
    assert blocks.n_blocks == len(hb.block_off)
    crc = np.array([zlib.crc32(hb.raw[o : o + n].tobytes()) for o, n in zip(hb.raw_off, hb.raw_size)], np.uint32)
    return hb, blocks, crc


def _check(data, crc_dev, status, hb, crc):
    assert int((status != 0).sum()) == 0
    assert data[: hb.raw.size].cpu().numpy().tobytes() == hb.raw.tobytes()
    assert np.array_equal(crc_dev[: crc.size].cpu().numpy(), crc)


def test_autotune_picks_and_caches(hetero, tmp_path, monkeypatch):
    from cubam_amd import _decode, _lib, tune

    hb, blocks, crc = hetero
    monkeypatch.setenv("CUBAM_AMD_TUNE_CACHE", str(tmp_path / "t.json"))
    monkeypatch.setattr(tune, "_cache", None)
    cands = [(0, 0, 0, 0), (2, 0, 0, 0), (_lib.INFLATE_LANES, 4, 1, 0)]
    best, table = tune.autotune(blocks, sample=blocks.n_blocks, candidates=cands, log=_say)
    assert best in cands and len(table) == 3
    assert (tmp_path / "t.json").exists()
    got = tune.lookup(blocks.n_blocks, blocks.total_out)
    assert tuple((got.variant, got.blocks_per_wave, got.waves_per_group, got.groups)) == best
    was = tune.enabled()
    tune.enable(True)
    try:
        assert _decode.inflate_tune_for(blocks.n_blocks, blocks.total_out) is not None  # the cached shape
        inf = _decode.inflate_blocks(blocks, None)
        _decode.check_inflated(inf)
        _check(inf.data, inf.crc, inf.status, hb, crc)
    finally:
        tune.enable(was)
        monkeypatch.setattr(tune, "_cache", None)


@pytest.mark.parametrize("shape", [(0, 0, 0, 0), (1, 8, 1, 0), (15, 0, 0, 0)])
def test_graph_replays_match_eager(hetero, shape):
    """Every replay of the captured launch gives the eager result, with other work in between."""
    import torch

    from cubam_amd import _decode
    from cubam_amd.tune import GraphInflate

    hb, blocks, crc = hetero
    g = GraphInflate(blocks, shape)
    for rep in range(4):
        g.data.fill_(0xAB)
        g.crc.zero_()
        g.status.fill_(-1)
        g.run()
        # eager work on the same stream between replays
        other = _decode.inflate_blocks(blocks, None)
        _decode.check_inflated(other)
        torch.cuda.synchronize()
        g.check()
        _check(g.data, g.crc, g.status, hb, crc)
        _say("replay", rep, "ok")
