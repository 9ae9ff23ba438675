"""GPU BGZF compression and BAM write-back against zlib (needs a GPU).

Every block written must decompress with zlib to exactly its payload, carry
zlib's CRC32 of it and its size, and the blocks must cut the data where htslib
cuts it (65,280-byte payloads, the header in blocks of its own), so that the
payload and every CRC32 equal those of a zlib re-encode of the same bytes.
"""

import gzip
import sys
import zlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


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


def _blocks(blob: bytes):
    """(payload offset, bsize, isize, crc) of each BGZF block, parsed on the host."""
    out, o = [], 0
    while o < len(blob):
        assert blob[o : o + 4] == b"\x1f\x8b\x08\x04" and blob[o + 12 : o + 14] == b"BC"
        bs = blob[o + 16] | (blob[o + 17] << 8)
        bs += 1
        crc = int.from_bytes(blob[o + bs - 8 : o + bs - 4], "little")
        isz = int.from_bytes(blob[o + bs - 4 : o + bs], "little")
        out.append((o, bs, isz, crc))
        o += bs
    assert o == len(blob)
    return out


def _check_blocks(blob: bytes, raw: bytes, cuts):
    """Every block inflates (zlib) to raw's slice at `cuts`, with zlib's CRC32 and the size."""
    blocks = _blocks(blob)
    assert len(blocks) == len(cuts)
    for (o, bs, isz, crc), (a, b) in zip(blocks, cuts):
        payload = zlib.decompress(blob[o + 18 : o + bs - 8], -15)
        assert payload == raw[a:b] and isz == b - a and crc == zlib.crc32(raw[a:b])
    assert gzip.decompress(blob) == raw


def _cuts(n, block=65280):
    return [(a, min(a + block, n)) for a in range(0, n, block)]


def _payloads():
    rng = np.random.default_rng(4)
    text = b"".join(b"read%06d\tchr%d\t%d\tACGTTGCA\n" % (i, i % 25, 1000 * i) for i in range(9000))
    return {
        "one": b"x",
        "four": b"abcd",
        "zeros": bytes(65280),
        "zeros_plus": bytes(65281),
        "random": rng.integers(0, 256, 200_000, dtype=np.uint8).tobytes(),
        "text": text,
        "runs": (b"A" * 300 + b"B" * 7 + b"CD" * 500) * 40,
        "small_alphabet": rng.integers(35, 75, 150_000, dtype=np.uint8).tobytes(),
    }


@pytest.mark.parametrize("name", list(_payloads().keys()))
def test_bgzf_compress_round_trips_with_zlib(cbm, name):
    from cubam_amd.write import bgzf_compress

    raw = _payloads()[name]
    blob = bgzf_compress(raw).cpu().numpy().tobytes()
    _say(name, len(raw), "->", len(blob))
    _check_blocks(blob, raw, _cuts(len(raw)))


def test_synthetic_bam_bytes_compress(cbm, synth):
    """The records of an ATAC-like file: compression near zlib level 1, exact round trip."""
    from cubam_amd.write import bgzf_compress

    b = synth.make_bam(1, 41, 60_000, n_total=60_000)
    raw = gzip.decompress(b.bgzf.tobytes())
    blob = bgzf_compress(raw).cpu().numpy().tobytes()
    z1 = sum(len(zlib.compress(raw[a:c], 1)) for a, c in _cuts(len(raw)))
    _say("ATAC", len(raw), "bytes: GPU", len(blob), "zlib level 1", z1)
    _check_blocks(blob, raw, _cuts(len(raw)))
    assert len(blob) < 0.75 * len(raw)


def test_gpu_inflate_reads_gpu_deflate(cbm, synth):
    from cubam_amd import bgzf
    from cubam_amd.write import EOF_BLOCK, bgzf_compress

    b = synth.make_bam(0, 42, 20_000, n_total=20_000)
    raw = gzip.decompress(b.bgzf.tobytes())
    blob = bgzf_compress(raw).cpu().numpy().tobytes() + EOF_BLOCK
    back = bgzf.inflate_bytes(np.frombuffer(blob, np.uint8))
    assert back.cpu().numpy().tobytes() == raw


def test_write_filtered_bam(cbm, synth, oracle, tmp_path):
    """Records of a barcode group, gathered on the GPU and written; read back exactly."""
    import torch

    from cubam_amd import read_bam
    from cubam_amd.write import gather_records, write_bam

    b = synth.make_bam(1, 43, 40_000, n_total=40_000)
    src = b.write(tmp_path / "in.bam")
    batches = list(read_bam(src, chunk_bytes=1 << 20))
    cbs = sorted({s for rb in batches for s in rb.tag("CB").to_list()})
    keep = set(cbs[::7])
    parts, want = [], []
    for rb in batches:
        vals = rb.tag("CB").to_list()
        mask = torch.tensor([v in keep for v in vals], device=rb.offset.device)
        idx = torch.nonzero(mask).flatten()
        rec = gather_records(rb.data, rb.offset[idx])
        parts.append(rec)
        o = rb.offset[idx].cpu().numpy()
        host = rb.data.cpu().numpy()
        for x in o:
            n = int.from_bytes(host[x : x + 4].tobytes(), "little") + 4
            want.append(host[x : x + n].tobytes())
    out = tmp_path / "out.bam"
    write_bam(out, batches[0].header, parts)
    blob = out.read_bytes()
    assert blob.endswith(bytes.fromhex("1f8b08040000000000ff0600424302001b0003000000000000000000"))
    hb = batches[0].header.to_bytes()
    recs = b"".join(want)
    # the header in blocks of its own, then the records: htslib's cuts
    cuts = _cuts(len(hb)) + [(len(hb) + a, len(hb) + c) for a, c in _cuts(len(recs))]
    _check_blocks(blob[:-28], hb + recs, cuts)
    (data, cols), = oracle.read_bam(np.frombuffer(blob, np.uint8), len(blob) + 1)
    assert cols["offset"].size == len(want) > 0
    _say("kept", len(want), "of", sum(len(x) for x in batches), "records")
