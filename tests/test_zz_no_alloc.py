"""cbm_inflate_noalloc, the inflate entry point that allocates nothing (needs a GPU).

Called through the C ABI on two batches in a row, the second smaller than the first, and
compared byte for byte (output, CRC32, status, bad-block count) with cbm_inflate on the same
blocks and launch shape; then with a corrupt block, and with more workgroups than its static
scratch holds (rejected before any launch).
"""

import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True, file=sys.stderr)


@pytest.fixture(scope="module")
def lib():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from cubam_amd import _lib

    return _lib.load()


def _blocks(n, seed, corrupt=False):
    from cubam_amd import _decode, synth

    hb = synth.hetero_blocks(n, seed)
    blob = hb.bgzf.copy()
    if corrupt:  # flip a payload byte of the middle block: its CRC32 (or its stream) no longer holds
        k = n // 2
        o = int(hb.block_off[k]) + 18 + (int(hb.block_size[k]) - 26) // 2
        blob[o] ^= 0x5A
    buf = _decode.to_device(blob)
    blocks = _decode.scan_blocks(buf, int(blob.size), True)
    assert blocks.n_blocks == n
    return hb, blocks


def _run(lib, blocks, noalloc, tune):
    import torch

    from cubam_amd import _decode, _lib

    dev = blocks.buf.device
    nb = blocks.n_blocks
    dst = torch.full((blocks.total_out + 64,), 0xAB, dtype=torch.uint8, device=dev)
    crc = torch.zeros(nb, dtype=torch.uint32, device=dev)
    status = torch.full((nb,), -7, dtype=torch.int32, device=dev)
    result = torch.full((1,), -9, dtype=torch.int64, device=dev)
    st = _lib.stream_ptr()
    args = (_lib.ptr(blocks.buf), _lib.ptr(blocks.blk_off), _lib.ptr(blocks.bsize), _lib.ptr(blocks.isize),
            _lib.ptr(blocks.crc), _lib.ptr(blocks.out_off), nb, _lib.ptr(dst), _lib.ptr(crc), _lib.ptr(status),
            _lib.ptr(result), ctypes.byref(tune))
    if noalloc:
        rc = lib.cbm_inflate_noalloc(*args, st)
    else:
        ws = _decode.default_workspace().dev("inflate_ws", lib.cbm_inflate_workspace(nb))
        rc = lib.cbm_inflate(*args, _lib.ptr(ws), ws.numel(), st)
    assert rc == 0, _lib.last_error()
    torch.cuda.synchronize()
    return (dst.cpu().numpy(), crc.view(torch.int32).cpu().numpy(), status.cpu().numpy(), int(result.item()))


@pytest.mark.parametrize("variant", [0, 5])
def test_noalloc_equals_inflate_on_two_calls(lib, variant):
    from cubam_amd import _lib

    tune = _lib.InflateTune(variant, 0, 0, 0)
    for n, seed in ((1500, 41), (300, 42)):  # the second call is smaller: stale scratch must not matter
        hb, blocks = _blocks(n, seed)
        _say("noalloc", variant, n)
        got = _run(lib, blocks, True, tune)
        ref = _run(lib, blocks, False, tune)
        assert got[3] == ref[3] == 0
        assert np.array_equal(got[2], ref[2]) and np.array_equal(got[1], ref[1])
        assert np.array_equal(got[0], ref[0])
        assert got[0][: hb.raw.size].tobytes() == hb.raw.tobytes()


def test_noalloc_reports_a_bad_block_as_inflate_does(lib):
    from cubam_amd import _lib

    tune = _lib.InflateTune(0, 0, 0, 0)
    hb, blocks = _blocks(200, 43, corrupt=True)
    got = _run(lib, blocks, True, tune)
    ref = _run(lib, blocks, False, tune)
    assert got[3] == ref[3] >= 1
    assert np.array_equal(got[2], ref[2]) and np.array_equal(got[1], ref[1])


def test_noalloc_rejects_more_groups_than_its_scratch(lib):
    from cubam_amd import _lib

    tune = _lib.InflateTune(0, 0, 0, 513)
    res = ctypes.c_int64(0)
    rc = lib.cbm_inflate_noalloc(None, None, None, None, None, None, 1, None, None, None, ctypes.addressof(res),
                                 ctypes.byref(tune), None)
    assert rc == -1 and "512" in _lib.last_error()
