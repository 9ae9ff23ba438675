"""Shared checks of the GPU tests: a RecordBatch against the oracle's columns of the same chunk."""

import sys

import numpy as np

FIXED = ("ref", "pos", "end", "mapq", "flag", "nref", "npos", "tlen", "l_seq", "clip_left", "clip_right")


def say(*a):
    print(*a, flush=True, file=sys.stderr)


def same_batch(b, data, c, tags=("CB", "UB"), var=True):
    """Every column of RecordBatch ``b`` equals the oracle's chunk (``data``, columns ``c``), bit for bit."""
    assert len(b) == len(c["offset"])
    assert np.array_equal(b.offset.cpu().numpy(), c["offset"]), "offset"
    for k in FIXED:
        assert np.array_equal(_np(getattr(b, k)).astype(np.int64), c[k].astype(np.int64)), k
    if var:
        for k in ("cigar_off", "seq_off", "qual_off", "seq", "qual"):
            assert np.array_equal(_np(getattr(b, k)), c[k]), k
        assert np.array_equal(_np(b.cigar).astype(np.uint32), c["cigar"]), "cigar"
    assert np.array_equal(b.data[: data.size].cpu().numpy(), data), "data"
    for t in tags:
        tg = b.tag(t)
        oc = c[f"tag_{t}"]
        for f in ("type", "offset", "length", "value"):
            assert np.array_equal(_np(getattr(tg, f)), oc[f]), (t, f)


def _np(t):
    """A device tensor on the host as numpy (unsigned 16/32/64-bit through a signed view)."""
    import torch

    signed = {torch.uint16: torch.int16, torch.uint32: torch.int32, torch.uint64: torch.int64}.get(t.dtype)
    if signed is None:
        return t.cpu().numpy()
    return t.view(signed).cpu().numpy().view({torch.uint16: np.uint16, torch.uint32: np.uint32,
                                                torch.uint64: np.uint64}[t.dtype])


def read_both(path, raw, chunk, tags=("CB", "UB"), **kw):
    """read_bam on the GPU and the oracle's loop over the same file, in step: a list of
    (gpu batch, oracle data, oracle columns) and each side's error text (None without one)."""
    from cubam_amd import read_bam
    from oracle import oracle

    gpu, gerr = [], None
    try:
        for b in read_bam(path, chunk_bytes=chunk, tags=tags, **kw):
            gpu.append(b)
    except ValueError as e:
        gerr = str(e)
    ora, oerr = [], None
    try:
        for data, c in oracle.iter_bam(raw, chunk, tags=tags, name=str(path)):
            ora.append((data, c))
    except ValueError as e:
        oerr = str(e)
    return gpu, ora, gerr, oerr
