"""The HIP path at the edges the oracle's known answers pin (needs a GPU).

Each file here goes through ``cubam_amd.read_bam`` and through the oracle's chunk loop; every
batch the GPU yields must equal the oracle's chunk column for column, and an error must be the
same error with the same text, raised after the same batches:
  * records that fail the structural test, one per reject predicate of is_record
    (src/cubam/_kernels.py:250-278), in the only chunk, mid-file and first;
  * find_tag's malformed, B-array, A, f and H cases and decode_fixed's clip edges
    (src/cubam/_kernels.py:292-388);
  * the slow paths: a stored block that holds a whole BGZF block (a header candidate off the
    chain, src/cubam/_kernels.py:676-702) and a record pair inside a B array at a walk
    segment's start (the walks do not link; the chain is repaired).
"""

import ctypes
import struct

import numpy as np
import pytest

import bamfixture as bf
from gpucheck import read_both, same_batch, say

pytestmark = pytest.mark.gpu

BS, REF, POS, LRN, BIN, NCIG, FLAG, LSEQ, NREF, NPOS, NAME = 0, 4, 8, 12, 14, 16, 18, 20, 24, 28, 36


@pytest.fixture(scope="module")
def cbm():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from cubam_amd import _lib

    _lib.load()
    import cubam_amd

    return cubam_amd


def _patch(b: bytes, off: int, fmt: str, v) -> bytes:
    out = bytearray(b)
    struct.pack_into(fmt, out, off, v)
    return bytes(out)


def _same_run(gpu, ora, gerr, oerr, tags):
    assert gerr == oerr, (gerr, oerr)
    assert len(gpu) == len(ora)
    for b, (data, c) in zip(gpu, ora):
        same_batch(b, data, c, tags)


# a record that fails is_record, by the predicate that fails
BREAK = {
    "block_size<33": lambda r: _patch(r, BS, "<i", 32),
    "refID<-1": lambda r: _patch(r, REF, "<i", -2),
    "refID>=n_ref": lambda r: _patch(r, REF, "<i", len(bf.REFS)),
    "pos<-1": lambda r: _patch(r, POS, "<i", -2),
    "next_refID>=n_ref": lambda r: _patch(r, NREF, "<i", len(bf.REFS)),
    "next_pos<-1": lambda r: _patch(r, NPOS, "<i", -5),
    "l_read_name<2": lambda r: _patch(r, LRN, "<B", 1),
    "flag>=4096": lambda r: _patch(r, FLAG, "<H", 4096),
    "l_seq>4Mi": lambda r: _patch(r, LSEQ, "<I", (1 << 22) + 1),
    "fields>block_size": lambda r: _patch(r, NCIG, "<H", 4000),
    "name has @": lambda r: _patch(r, NAME, "<B", ord("@")),
    "name without NUL": lambda r: _patch(r, NAME + r[LRN] - 1, "<B", ord("x")),
    "CIGAR op 9": lambda r: _patch(r, NAME + r[LRN], "<I", (struct.unpack_from("<I", r, NAME + r[LRN])[0] & ~15) | 9),
    "bin": lambda r: _patch(r, BIN, "<H", (struct.unpack_from("<H", r, BIN)[0] + 1) & 0xFFFF),
}


@pytest.mark.parametrize("where", ["only chunk", "mid-file", "first record"])
@pytest.mark.parametrize("pred", sorted(BREAK))
def test_invalid_record_matches_oracle(cbm, tmp_path, pred, where):
    reads = bf.random_reads(240, seed=31)
    recs = [r.encode() for r in reads]
    at = {"only chunk": 100, "mid-file": 100, "first record": 0}[where]
    recs[at] = BREAK[pred](recs[at])
    raw = bf.bgzf(bf.header_bytes() + b"".join(recs), 2000)
    p = tmp_path / "bad.bam"
    p.write_bytes(raw)
    chunk = 1 << 20 if where == "only chunk" else 5000
    say("invalid record", pred, where, len(raw), chunk)
    gpu, ora, gerr, oerr = read_both(p, raw, chunk, tags=("CB", "NM"))
    _same_run(gpu, ora, gerr, oerr, ("CB", "NM"))
    assert oerr is not None
    assert sum(len(b) for b in gpu) <= at


def test_invalid_record_without_candidates_is_carried(cbm, tmp_path):
    reads = bf.random_reads(1, seed=3)
    raw = bf.bgzf(bf.header_bytes() + _patch(reads[0].encode(), FLAG, "<H", 4096), 4096)
    p = tmp_path / "one.bam"
    p.write_bytes(raw)
    gpu, ora, gerr, oerr = read_both(p, raw, 1 << 20, tags=())
    assert gpu == [] and ora == []
    assert gerr == oerr == f"{p}: truncated BAM file (incomplete record at the end)"


def _tag_reads():
    tags = [("XA", "A", "Q"), ("XF", "f", 1.5), ("XH", "H", "1AE3"), ("XB", "B", ("f", [1.0, 2.0, 3.0])),
            ("XC", "c", -5), ("XD", "C", 200), ("XS", "s", -300), ("XT", "S", 65000), ("XI", "i", -70000),
            ("XU", "I", 4000000000), ("XZ", "Z", "hello"), ("XE", "B", ("c", [])), ("NM", "i", 1)]
    raws = [b"", b"XXq\0\0UBZGGGG\0", b"CBZACGT", b"XBBs" + struct.pack("<I", 1000) + b"\1\0\2\0UBZAC\0",
            b"XBBs\x01\x00", b"XIi\x01\x02", b"ab", b"UBZTTTT\0", b"XBBx" + struct.pack("<I", 2) + b"\1" * 8 + b"NMi\1\0\0\0"]
    out = []
    for i in range(64):
        raw = raws[i % len(raws)]
        t = tags[: i % len(tags)] + ([("CB", "Z", "AAAC-1")] if i % 3 else [])
        out.append(bf.Read(f"t{i}", 0, 0, 1000 + i, 60, [("M", 10)], seq="ACGTACGTAC", tags=t, aux_raw=raw))
    return out


def test_tags_and_clips_match_oracle(cbm, tmp_path):
    reads = _tag_reads()
    reads += [
        bf.Read("c0", 0, 1, 500, 60, [("S", 70000), ("M", 10), ("S", 100000)], seq="A" * 170010),
        bf.Read("c1", 0, 1, 600, 60, [("S", 65535), ("M", 5), ("S", 65536)], seq="C" * 131076),
        bf.Read("c2", 0, 1, 700, 60, [("H", 5), ("S", 3), ("M", 10), ("S", 4), ("H", 2)], seq="G" * 17),
        bf.Read("c3", 0, 1, 800, 60, [], seq="ACG"),
        bf.Read("c4", 4, -1, -1, 0, [], seq="AC"),
    ]
    raw = bf.bgzf(bf.bam_bytes(reads), 7000)
    p = tmp_path / "tags.bam"
    p.write_bytes(raw)
    names = ["XA", "XF", "XH", "XB", "XC", "XD", "XS", "XT", "XI", "XU", "XZ", "XE", "NM", "CB", "UB", "ZZ"]
    for chunk in (1 << 20, 9000):
        gpu, ora, gerr, oerr = read_both(p, raw, chunk, tags=tuple(names[:4]))
        _same_run(gpu, ora, gerr, oerr, tuple(names[:4]))
        from oracle import oracle

        for b, (data, c) in zip(gpu, ora):
            for t in names[4:]:  # on demand: one scatter pass a tag
                ty, vo, vl, iv = oracle.find_tag(data, c["offset"], t)
                tg = b.tag(t)
                assert np.array_equal(tg.type.cpu().numpy(), ty), t
                assert np.array_equal(tg.offset.cpu().numpy(), vo), t
                assert np.array_equal(tg.length.cpu().numpy(), vl), t
                assert np.array_equal(tg.value.cpu().numpy(), iv), t
    types = set(int(x) for b in gpu for x in b.tag("UB").type.cpu().numpy())
    assert 255 in types and 0 in types and ord("Z") in types
    clips = [(int(a), int(b)) for bt in gpu for a, b in zip(bt.clip_left.cpu().numpy(), bt.clip_right.cpu().numpy())]
    assert clips[-5:] == [(65535, 65535), (65535, 65535), (0, 4), (0, 0), (0, 0)]


def test_take_keeps_the_csr_columns(cbm, tmp_path):
    import torch

    reads = bf.random_reads(300, seed=8)
    p = tmp_path / "t.bam"
    raw = bf.bgzf(bf.bam_bytes(reads), 4096)
    p.write_bytes(raw)
    from cubam_amd import read_bam

    (b,) = list(read_bam(p, tags=("CB",)))
    idx = torch.tensor([5, 0, 299, 17, 17, 120], device=b.offset.device)
    sub = b.take(idx)
    for values, offs in b.CSR:
        o = getattr(b, offs).cpu().numpy()
        v = getattr(b, values).cpu().numpy()
        so = getattr(sub, offs).cpu().numpy()
        sv = getattr(sub, values).cpu().numpy()
        for j, i in enumerate(idx.cpu().tolist()):
            assert sv[so[j] : so[j + 1]].tobytes() == v[o[i] : o[i + 1]].tobytes(), (values, i)
    mask = b.has_flags(all_of=0x10)
    sub2 = b.take(mask)
    assert len(sub2) == int(mask.sum()) and sub2.cigar_off.numel() == len(sub2) + 1
    assert sub.tag("CB").to_list() == [b.tag("CB").to_list()[i] for i in idx.cpu().tolist()]


def test_spurious_bgzf_header_takes_the_serial_chain(cbm, oracle, tmp_path):
    """A stored (level 0) block whose payload holds a whole BGZF block: the scan finds a header
    candidate inside it, the candidates do not chain, and the chain from offset 0 is walked."""
    from cubam_amd import _decode, inflate_chunks

    inner = bf.bgzf_block(b"a block inside a block " * 20, 6)
    payloads = [b"x" * 3000, b"y" * 100 + inner + b"z" * 200, b"w" * 5000]
    blob = bf.bgzf_block(payloads[0], 6) + bf.bgzf_block(payloads[1], 0) + bf.bgzf_block(payloads[2], 6)
    blob += bf.EOF_BLOCK
    buf = _decode.to_device(np.frombuffer(blob, np.uint8))
    blocks = _decode.scan_blocks(buf, len(blob), True)
    assert blocks.n_blocks == 4 and blocks.n_cand > blocks.n_blocks
    off, _, _, _, used, status = oracle.bgzf_scan(blob, True)
    assert status == 0 and used == len(blob)
    nonempty = blocks.isize.cpu().numpy() > 0  # the oracle's scan leaves out the empty EOF block
    assert blocks.blk_off.cpu().numpy()[nonempty].tolist() == off.tolist()
    p = tmp_path / "s.gz"
    p.write_bytes(blob)
    got = b"".join(c.cpu().numpy().tobytes() for c in inflate_chunks(p, chunk_bytes=1 << 20))
    assert got == b"".join(payloads)


def test_fake_records_at_a_segment_start_are_repaired(cbm, oracle, tmp_path):
    """A B:C array holding two well-formed records, placed at the starts of walk segments 2 to 5
    (a wavefront walks consecutive segments and searches only in its first, so some of them are
    searched): the segment's walk starts on them, its links do not match, and the chain is
    repaired to the oracle's."""
    import gzip

    from cubam_amd import _decode, _lib

    L = _lib.load()
    seg = ctypes.c_int64()
    L.cbm_record_chain_workspace(1 << 20, ctypes.byref(seg))
    seg = seg.value
    reads = bf.random_reads(200, seed=12)
    fake = b"".join(bf.Read(f"fake{i}", 0, 0, 100 + i, 60, [("M", 10)], seq="ACGTACGTAC").encode() for i in range(2))
    recs, o = [], 0
    while o < 2 * seg - 3000:  # ordinary records up to a little before segment 2
        r = reads[len(recs)].encode()
        recs.append(r)
        o += len(r)
    # a record whose B:C array covers segments 2 to 5, with the fake pair starting exactly on each
    # of their starts
    host = bf.Read("host", 0, 0, reads[len(recs)].pos, 60, [("M", 10)], seq="ACGTACGTAC")
    head = len(host.encode())  # the record without aux
    val = o + head + 3 + 1 + 4  # where the array's values start
    pad = 2 * seg - val
    assert 0 < pad < 4000
    arr = b"\xff" * pad + (fake + b"\xff" * (seg - len(fake))) * 3 + fake + b"\xff" * 300
    host.tags = [("XB", "B", ("C", list(arr)))]
    recs.append(host.encode())
    recs += [r.encode() for r in reads[len(recs) :]]
    body = b"".join(recs)
    assert all(body[k * seg : k * seg + len(fake)] == fake for k in range(2, 6))
    raw = bf.bgzf(bf.header_bytes() + body, 6000)
    hdr = len(bf.header_bytes())
    inflated = np.frombuffer(gzip.decompress(raw), np.uint8)
    data = _decode.to_device(inflated[hdr:])
    chain = _decode.record_chain(data, 0, int(inflated.size - hdr), len(bf.REFS))
    assert chain.bad_links > 0
    got = _decode.record_offsets(data, 0, int(inflated.size - hdr), chain).cpu().numpy()
    ooff, tail, st = oracle.record_chain(inflated[hdr:], len(bf.REFS))
    assert got.tolist() == ooff.tolist() and len(got) == len(reads) and st == 0
    p = tmp_path / "fake.bam"
    p.write_bytes(raw)
    gpu, ora, gerr, oerr = read_both(p, raw, 1 << 20, tags=("XB",))
    _same_run(gpu, ora, gerr, oerr, ("XB",))
