"""The HIP path against the oracle and the test writer, bit for bit (needs a GPU).

Every comparison is exact: integer columns, inflated bytes, per-block CRC32.
"""

import gzip
import random
import sys
import zlib

import numpy as np
import pytest

import bamfixture as bf

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True, file=sys.stderr)


@pytest.fixture(scope="module")
def cbm():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import cubam_amd
    from cubam_amd import _lib

    _lib.load()
    return cubam_amd


def _hetero(n, seed, max_size=65280):
    from cubam_amd import synth

    return synth.hetero_blocks(n, seed, max_size=max_size)


def _gpu_blocks(blob, tune=None):
    import torch

    from cubam_amd import _decode

    buf = _decode.to_device(blob)
    blocks = _decode.scan_blocks(buf, int(blob.size), True)
    inf = _decode.inflate_blocks(blocks, None, tune=tune)
    torch.cuda.synchronize()
    return blocks, inf


def test_hetero_blocks_bytes_and_crc(cbm, oracle):
    hb = _hetero(3000, 2)
    _say("hetero blocks", hb.bgzf.size, "bytes", len(hb.block_off), "blocks")
    blocks, inf = _gpu_blocks(hb.bgzf)
    assert blocks.n_blocks == len(hb.block_off)
    assert blocks.blk_off.cpu().numpy().tolist() == hb.block_off.tolist()
    assert blocks.isize.cpu().numpy().tolist() == hb.raw_size.tolist()
    assert int(inf.result.item()) == 0
    assert (inf.status.cpu().numpy() == 0).all()
    got = inf.data[: inf.n].cpu().numpy()
    assert got.tobytes() == hb.raw.tobytes()
    crc = inf.crc.cpu().numpy().astype(np.int64)
    expect = [zlib.crc32(hb.raw[o : o + s].tobytes()) for o, s in zip(hb.raw_off, hb.raw_size)]
    assert crc.tolist() == expect
    _, _, ocrc, ost = oracle.inflate(hb.bgzf, hb.block_off, hb.block_size, hb.raw_size)
    assert (ost == 0).all() and ocrc.astype(np.int64).tolist() == expect


@pytest.mark.parametrize("variant,bpw,wpg,groups", [(0, 0, 0, 0), (0, 0, 0, 7), (2, 0, 0, 0), (3, 0, 0, 0),
                                                    (4, 0, 0, 0), (5, 0, 0, 0), (6, 0, 0, 0), (7, 0, 0, 0),
                                                    (8, 0, 0, 0), (9, 0, 0, 0), (10, 0, 0, 0), (10, 0, 0, 3),
                                                    (11, 0, 0, 0), (11, 0, 0, 5), (12, 0, 0, 0), (13, 0, 0, 0),
                                                    (11, 2, 2, 0), (11, 4, 1, 0), (11, 4, 4, 0), (11, 8, 1, 0), (11, 8, 2, 0), (11, 8, 4, 0),
                                                    (11, 16, 1, 0), (11, 16, 2, 0), (11, 32, 1, 0), (12, 8, 4, 0),
                                                    (13, 16, 2, 0), (14, 0, 0, 0), (14, 0, 0, 9),
                                                    (15, 0, 0, 0), (15, 0, 0, 5), (15, 8, 4, 0),
                                                    (1, 1, 1, 0), (1, 4, 2, 0), (1, 16, 3, 0), (1, 32, 1, 0)])
def test_inflate_every_launch_shape(cbm, variant, bpw, wpg, groups):
    from cubam_amd import _lib

    hb = _hetero(700, 3)
    _say("shape", variant, bpw, wpg, groups)
    tune = _lib.InflateTune(variant, bpw, wpg, groups)
    blocks, inf = _gpu_blocks(hb.bgzf, tune)
    assert int(inf.result.item()) == 0
    assert inf.data[: inf.n].cpu().numpy().tobytes() == hb.raw.tobytes()


def test_split_inflates_on_two_streams_at_once(cbm):
    """The two-stream inflate called on two streams together (each forks its own side stream):
    each result is the one a single-stream inflate of the same blocks gives."""
    import torch

    from cubam_amd import _decode, _lib

    hbs = [_hetero(500, 5), _hetero(650, 6)]
    ref = []
    for hb in hbs:
        _, inf = _gpu_blocks(hb.bgzf, _lib.InflateTune(11, 0, 0, 0))
        ref.append((inf.data[: inf.n].cpu().numpy().tobytes(), inf.crc.cpu().numpy().tolist()))
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    bufs = [_decode.to_device(hb.bgzf) for hb in hbs]
    torch.cuda.synchronize()
    for rep in range(3):
        outs = []
        for hb, st, buf in zip(hbs, streams, bufs):
            with torch.cuda.stream(st):
                ws = _decode.Workspace()
                blocks = _decode.scan_blocks(buf, int(hb.bgzf.size), True, ws)
                outs.append(_decode.inflate_blocks(blocks, None, ws, tune=_lib.InflateTune(15, 0, 0, 0)))
        torch.cuda.synchronize()
        for (data, crc), inf in zip(ref, outs):
            assert int(inf.result.item()) == 0
            assert inf.data[: inf.n].cpu().numpy().tobytes() == data
            assert inf.crc.cpu().numpy().tolist() == crc
        _say("two streams, rep", rep, "ok")


@pytest.mark.parametrize("variant", [11, 15])
def test_two_phase_second_pass_overflows_to_the_one_phase_kernel(cbm, variant):
    """Every block has more matches than the first pass's list holds, so more blocks are deferred
    than the second pass's full-size lists fit in the pool (about 28% of the blocks): the first
    ones go through the second pass, the rest through the one-phase kernel, and every block comes
    out as zlib makes it, CRC32 included."""
    from cubam_amd import _lib

    rng = random.Random(73)
    vocab = [bytes(rng.choices(range(256), k=3)) for _ in range(6)]
    payloads, blob = [], b""
    for i in range(40):
        n = 65280 - 97 * i
        p = b"".join(rng.choice(vocab) + bytes([rng.randrange(256)]) for _ in range(16400))[:n]
        payloads.append(p)
        blob += bf.bgzf_block(p, 1, zlib.Z_DEFAULT_STRATEGY)
    blob += bf.EOF_BLOCK
    _say("second pass overflow, variant", variant)
    blocks, inf = _gpu_blocks(np.frombuffer(blob, np.uint8), _lib.InflateTune(variant, 0, 0, 0))
    assert int(inf.result.item()) == 0
    assert (inf.status.cpu().numpy() == 0).all()
    assert inf.data[: inf.n].cpu().numpy().tobytes() == b"".join(payloads)
    assert inf.crc.cpu().numpy().astype(np.int64).tolist()[: len(payloads)] == [zlib.crc32(p) for p in payloads]


def test_tiny_and_empty_blocks(cbm):
    # an empty block must be the 28-byte EOF block (src/cubam/bgzf.py:572), so b"" only at level 6
    payloads = [b"x", b"ab" * 3, bytes(range(256)), b"\0" * 35000, b"A" * 65280]
    blob = bf.bgzf_block(b"", 6) + b"".join(bf.bgzf_block(p, lvl) for p in payloads for lvl in (0, 1, 6, 9))
    blob += bf.EOF_BLOCK
    blocks, inf = _gpu_blocks(np.frombuffer(blob, np.uint8))
    assert int(inf.result.item()) == 0
    assert inf.data[: inf.n].cpu().numpy().tobytes() == b"".join(p for p in payloads for _ in range(4))


def _data(n, seed=0):
    rng = random.Random(seed)
    words = [bytes(rng.choices(b"ACGT\t\n0123456789", k=rng.randint(1, 30))) for _ in range(500)]
    out = bytearray()
    while len(out) < n:
        out += rng.choice(words)
    return bytes(out[:n])


def _read_chunks(path, **kw):
    from cubam_amd import inflate_chunks

    return b"".join(c.cpu().numpy().tobytes() for c in inflate_chunks(path, **kw))


@pytest.mark.parametrize(("block", "chunk"), [(1000, 4096), (65280, 1 << 20), (300, 777)])
def test_inflate_chunks_roundtrip(cbm, tmp_path, block, chunk):
    data = _data(300_000)
    p = tmp_path / "x.gz"
    p.write_bytes(bf.bgzf(data, block))
    assert _read_chunks(p, chunk_bytes=chunk, piece=chunk // 3 + 1) == data


def test_chunks_are_on_the_gpu(cbm, tmp_path):
    from cubam_amd import inflate_chunks

    p = tmp_path / "x.gz"
    p.write_bytes(bf.bgzf(_data(10_000), 1000))
    assert all(c.is_cuda for c in inflate_chunks(p, chunk_bytes=4096))


def test_empty_file(cbm, tmp_path):
    p = tmp_path / "e.gz"
    p.write_bytes(bf.EOF_BLOCK)
    assert _read_chunks(p, chunk_bytes=4096) == b""


def test_truncated(cbm, tmp_path):
    p = tmp_path / "t.gz"
    p.write_bytes(bf.bgzf(_data(50_000), 1000, eof=False)[:-100])
    with pytest.raises(ValueError, match="Truncated"):
        _read_chunks(p, chunk_bytes=8192)


@pytest.mark.parametrize("where", ["btype", "isize", "crc", "payload"])
def test_corrupt_block(cbm, tmp_path, where):
    raw = bytearray(bf.bgzf(_data(50_000), 1000))
    bsize = raw[16] + (raw[17] << 8) + 1
    if where == "btype":
        raw[18] |= 0b110  # reserved Deflate block type
    elif where == "isize":
        raw[bsize - 4] ^= 0x01
    elif where == "crc":
        raw[bsize - 8] ^= 0x01
    else:
        raw[40] ^= 0x55
    p = tmp_path / "c.gz"
    p.write_bytes(bytes(raw))
    with pytest.raises(ValueError):
        _read_chunks(p, chunk_bytes=8192)


def test_not_bgzf(cbm, tmp_path):
    p = tmp_path / "n.gz"
    p.write_bytes(b"hello world, not compressed" * 100)
    with pytest.raises(ValueError):
        _read_chunks(p, chunk_bytes=4096)


def _gpu_read(path, **kw):
    from cubam_amd import read_bam

    return list(read_bam(path, **kw))


def _cat(batches, k):
    import torch

    return torch.cat([getattr(b, k).cpu() for b in batches]).numpy() if batches else np.empty(0)


@pytest.mark.parametrize(("block", "chunk"), [(4096, 1 << 20), (700, 5000), (65280, 64 << 10)])
def test_read_bam_fields(cbm, tmp_path, block, chunk):
    reads = bf.random_reads(3000)
    p = tmp_path / "r.bam"
    p.write_bytes(bf.bgzf(bf.bam_bytes(reads), block))
    batches = _gpu_read(p, chunk_bytes=chunk, piece=chunk // 2 + 1, tags=("CB", "NM", "RG"))
    assert sum(len(b) for b in batches) == len(reads)
    for k in ("ref", "pos", "end", "mapq", "flag", "nref", "npos", "tlen", "l_seq", "clip_left", "clip_right"):
        got = _cat(batches, k).astype(np.int64).tolist()
        exp = [getattr(r, k) if hasattr(r, k) else None for r in reads]
        if k == "l_seq":
            exp = [len(r.seq) for r in reads]
        elif k in ("clip_left", "clip_right"):
            exp = [r.clips()[k == "clip_right"] for r in reads]
        assert got == exp, k
    cb, rg, nm, seqs, quals, cigs = [], [], [], [], [], []
    for b in batches:
        cb += b.tag("CB").to_list()
        rg += b.tag("RG").to_list()
        t = b.tag("NM")
        nm += np.where(t.present.cpu().numpy(), t.value.cpu().numpy(), -1).tolist()
        so, s = b.seq_off.cpu().numpy(), b.seq.cpu().numpy()
        qo, q = b.qual_off.cpu().numpy(), b.qual.cpu().numpy()
        co, c = b.cigar_off.cpu().numpy(), b.cigar.cpu().numpy()
        for i in range(len(b)):
            seqs.append(s[so[i] : so[i + 1]].tobytes())
            quals.append(q[qo[i] : qo[i + 1]].tobytes())
            cigs.append(c[co[i] : co[i + 1]].astype(np.int64).tolist())
    assert cb == [(r.tag("CB") or ("Z", ""))[1] for r in reads]
    assert rg == [r.tag("RG")[1] for r in reads]
    assert nm == [r.tag("NM")[1] for r in reads]
    assert cigs == [[(n << 4) | bf.OPS.index(op) for op, n in r.cigar] for r in reads]
    for r, s, q in zip(reads, seqs, quals):
        codes = [bf.SEQ.index(ch) for ch in r.seq] + [0]
        assert s == bytes((codes[i] << 4) | codes[i + 1] for i in range(0, len(r.seq), 2))
        assert q == (r.qual if r.qual is not None else b"\xff" * len(r.seq))


def test_five_prime_take_and_flags(cbm, tmp_path):
    reads = bf.random_reads(500, seed=3)
    p = tmp_path / "t.bam"
    p.write_bytes(bf.bgzf(bf.bam_bytes(reads), 4096))
    (b,) = _gpu_read(p, tags=("NM",))
    five = b.five_prime().cpu().numpy().tolist()
    exp = [r.end + r.clips()[1] if r.flag & 16 else r.pos - r.clips()[0] for r in reads]
    assert five == exp
    sel = b.has_flags(all_of=1, none_of=0x400)
    sub = b.take(sel)
    ex = [r for r in reads if r.flag & 1 and not r.flag & 0x400]
    assert len(sub) == len(ex)
    assert sub.pos.cpu().numpy().tolist() == [r.pos for r in ex]
    assert sub.tag("NM").value.cpu().numpy().tolist() == [r.tag("NM")[1] for r in ex]
    assert sub.data.is_cuda


def test_long_reads_cross_blocks(cbm, tmp_path):
    reads = [bf.Read(f"long{i}", 0, 0, 1000 * i, 60, [("M", 20000)]) for i in range(20)]
    p = tmp_path / "l.bam"
    p.write_bytes(bf.bgzf(bf.bam_bytes(reads), 4096))
    got = [int(x) for b in _gpu_read(p, chunk_bytes=30_000) for x in b.end.cpu()]
    assert got == [r.end for r in reads]


def test_header_only(cbm, tmp_path):
    from cubam_amd import BamHeader

    text = "@HD\tVN:1.6\tSO:coordinate\n@SQ\tSN:chr1\tLN:1000000\n@PG\tID:bwa\tPN:bwa\n"
    p = tmp_path / "h.bam"
    p.write_bytes(bf.bgzf(bf.bam_bytes([], text=text), 4096))
    h = BamHeader.read(p)
    assert h.names == [n for n, _ in bf.REFS]
    assert h.sort_order == "coordinate"
    assert h.lines("@PG") == ["@PG\tID:bwa\tPN:bwa"]
    assert _gpu_read(p) == []


def test_truncated_record(cbm, tmp_path):
    reads = bf.random_reads(50, seed=4)
    p = tmp_path / "tr.bam"
    p.write_bytes(bf.bgzf(bf.bam_bytes(reads)[:-10], 4096))
    with pytest.raises(ValueError, match="truncated BAM"):
        _gpu_read(p)


def _compare_with_oracle(batches, obatches, tags=("CB", "UB")):
    import torch

    keys = ("ref", "pos", "end", "mapq", "flag", "nref", "npos", "tlen", "l_seq", "clip_left", "clip_right")
    for k in keys:
        g = np.concatenate([getattr(b, k).cpu().numpy() for b in batches])
        o = np.concatenate([c[k] for _, c in obatches])
        assert np.array_equal(g.astype(np.int64), o.astype(np.int64)), k
    for (b, (d, c)) in zip(batches, obatches):
        assert torch.equal(b.offset.cpu(), torch.from_numpy(c["offset"]))
        for k in ("cigar_off", "seq_off", "qual_off", "seq", "qual"):
            assert np.array_equal(getattr(b, k).cpu().numpy(), c[k]), k
        assert np.array_equal(b.cigar.cpu().numpy().astype(np.uint32), c["cigar"])
        assert b.data[: d.size].cpu().numpy().tobytes() == d.tobytes()
        for t in tags:
            tg = b.tag(t)
            oc = c[f"tag_{t}"]
            for f in ("type", "offset", "length", "value"):
                assert np.array_equal(getattr(tg, f).cpu().numpy(), oc[f]), (t, f)


@pytest.mark.parametrize("kind, n, chunk", [(0, 20_000, 1 << 20), (1, 30_000, 300_000), (1, 5000, 1 << 26)])
def test_synthetic_bam_matches_oracle(cbm, oracle, synth, tmp_path, kind, n, chunk):
    b = synth.make_bam(kind, 11 + kind, n, n_total=50_000_000)
    p = b.write(tmp_path / "s.bam")
    _say("synthetic", kind, n, chunk, b.bgzf.size)
    batches = _gpu_read(p, chunk_bytes=chunk)
    obatches = oracle.read_bam(b.bgzf, chunk)
    assert len(batches) == len(obatches)
    assert sum(len(x) for x in batches) == n
    _compare_with_oracle(batches, obatches)


def test_record_chain_repair_is_identity(cbm, oracle, synth):
    """The serial repair walk gives the same chain as the speculative walks."""
    import torch

    from cubam_amd import _decode, _lib

    b = synth.make_bam(1, 21, 4000, n_total=10_000_000)
    raw = np.frombuffer(gzip.decompress(b.bgzf.tobytes()), np.uint8)
    data = _decode.to_device(raw)
    start, n = len(b.header), int(raw.size)
    chain = _decode.record_chain(data, start, n, 25)
    off = _decode.record_offsets(data, start, n, chain).cpu().numpy()
    L = _lib.load()
    res = torch.zeros(8, dtype=torch.int64, device=data.device)
    _lib.check(L.cbm_record_chain_repair(data.data_ptr(), start, n, 25, res.data_ptr(), chain.ws.data_ptr(),
                                         chain.ws.numel(), _lib.stream_ptr()), "repair")
    r = res.cpu().numpy()
    assert r[_lib.REC_N] == chain.n_rec == 4000 and r[_lib.REC_TAIL] == n and r[_lib.REC_BAD_LINKS] == 0
    off2 = _decode.record_offsets(data, start, n, chain).cpu().numpy()
    ooff, tail, st = oracle.record_chain(raw[start:], 25)
    assert off.tolist() == off2.tolist() == (ooff + start).tolist()


def test_crc32_spans_matches_zlib(cbm):
    import torch

    from cubam_amd import _decode, _lib

    rng = np.random.default_rng(3)
    raw = rng.integers(0, 256, 500_000, dtype=np.uint8)
    d = _decode.to_device(raw)
    off = np.arange(0, raw.size, 30_001)
    size = np.minimum(off + 30_001, raw.size) - off
    off_d = torch.from_numpy(off.astype(np.int64)).cuda()
    len_d = torch.from_numpy(size.astype(np.int64)).cuda()
    out = torch.empty(off.size, dtype=torch.uint32, device="cuda")
    _lib.check(_lib.load().cbm_crc32_spans(d.data_ptr(), off_d.data_ptr(), len_d.data_ptr(), off.size,
                                           out.data_ptr(), _lib.stream_ptr()), "crc")
    got = out.cpu().numpy().astype(np.int64).tolist()
    assert got == [zlib.crc32(raw[o : o + s].tobytes()) for o, s in zip(off, size)]


class _Bits:
    """A DEFLATE bit stream: integers least significant bit first, Huffman codes most significant
    bit first (RFC 1951 3.1.1)."""

    def __init__(self):
        self.v, self.n = 0, 0

    def put(self, value, nbits):
        self.v |= value << self.n
        self.n += nbits

    def huff(self, code, nbits):
        self.put(int(format(code, f"0{nbits}b")[::-1], 2), nbits)

    def data(self):
        return self.v.to_bytes((self.n + 7) // 8, "little")


def _fixed_litlen(b, sym):
    """A literal/length symbol in the fixed Huffman code (RFC 1951 3.2.6)."""
    if sym < 144:
        b.huff(0x30 + sym, 8)
    elif sym < 256:
        b.huff(0x190 + sym - 144, 9)
    elif sym < 280:
        b.huff(sym - 256, 7)
    else:
        b.huff(0xC0 + sym - 280, 8)


# RFC 1951 3.2.5: length and distance bases and extra bits, symbol by symbol
_LEN_BASE = [3, 4, 5, 6, 7, 8, 9, 10, 11, 13, 15, 17, 19, 23, 27, 31, 35, 43, 51, 59, 67, 83, 99, 115, 131,
             163, 195, 227, 258]
_LEN_EXTRA = [0] * 8 + [1] * 4 + [2] * 4 + [3] * 4 + [4] * 4 + [5] * 4 + [0]
_DIST_BASE = [1, 2, 3, 4, 5, 7, 9, 13, 17, 25, 33, 49, 65, 97, 129, 193, 257, 385, 513, 769, 1025, 1537,
              2049, 3073, 4097, 6145, 8193, 12289, 16385, 24577]
_DIST_EXTRA = [0, 0, 0, 0] + [k // 2 for k in range(2, 28)]


def _fixed_block(seed, bad=None):
    """One final fixed-Huffman block: 32,768 random literals, a match with every distance symbol
    (0-29) and every length symbol (257-285), each with random extra bits, then end of block.
    bad = ("dist", 30 or 31) or ("litlen", 286 or 287) puts that symbol in the stream instead,
    a few literals in."""
    rng = random.Random(seed)
    b = _Bits()
    b.put(1, 1)  # BFINAL
    b.put(1, 2)  # BTYPE 01: fixed Huffman codes
    nlit = 40 if bad else 32768
    for _ in range(nlit):
        _fixed_litlen(b, rng.randrange(256))
    if bad and bad[0] == "litlen":
        _fixed_litlen(b, bad[1])
    elif bad:
        _fixed_litlen(b, 257)
        b.huff(bad[1], 5)
    else:
        pairs = [(rng.randrange(29), d) for d in range(30)] + [(c, rng.randrange(8)) for c in range(29)]
        rng.shuffle(pairs)
        for lc, dc in pairs:
            _fixed_litlen(b, 257 + lc)
            b.put(rng.randrange(1 << _LEN_EXTRA[lc]), _LEN_EXTRA[lc])
            b.huff(dc, 5)
            b.put(rng.randrange(1 << _DIST_EXTRA[dc]), _DIST_EXTRA[dc])
    _fixed_litlen(b, 256)
    return b.data()


def _canonical(lengths):
    """Canonical Huffman codes of the given code lengths (RFC 1951 3.2.2), 0 for unused symbols."""
    count = [0] * 16
    for n in lengths:
        count[n] += 1
    count[0] = 0
    code, nxt = 0, [0] * 16
    for n in range(1, 16):
        code = (code + count[n - 1]) << 1
        nxt[n] = code
    out = []
    for n in lengths:
        out.append(nxt[n] if n else 0)
        if n:
            nxt[n] += 1
    return out


def _deep_dynamic_block(seed):
    """One final dynamic-Huffman block whose distance code is a chain: 16 of the 30 distance symbols
    (shuffled) take code lengths 1, 2, ..., 14, 15, 15, so the codes of 9 to 15 bits share one
    8-bit root prefix and a 128-entry subtable. Literal/length lengths are 8 and 9 bits; the
    code-length code gives each length 0-15 four bits. Data: 32,768 random literals, then matches
    with every used distance symbol and random length symbols and extra bits, then end of block."""
    rng = random.Random(seed)
    lit = [8] * 286
    for sym in range(196, 256):
        lit[sym] = 9  # 226 / 256 + 60 / 512 = 1
    dsyms = rng.sample(range(30), 16)
    dist = [0] * 30
    for k, sym in enumerate(dsyms):
        dist[sym] = min(k + 1, 15)
    lcode, dcode = _canonical(lit), _canonical(dist)
    b = _Bits()
    b.put(1, 1)  # BFINAL
    b.put(2, 2)  # BTYPE 10: dynamic Huffman codes
    b.put(286 - 257, 5)
    b.put(30 - 1, 5)
    b.put(19 - 4, 4)
    for sym in (16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15):
        b.put(0 if sym > 15 else 4, 3)
    for n in lit + dist:
        b.huff(n, 4)  # sixteen 4-bit codes: each length's code is its value
    for _ in range(32768):
        sym = rng.randrange(256)
        b.huff(lcode[sym], lit[sym])
    pairs = [(rng.randrange(29), d) for d in dsyms for _ in range(3)]
    rng.shuffle(pairs)
    for lc, dc in pairs:
        b.huff(lcode[257 + lc], lit[257 + lc])
        b.put(rng.randrange(1 << _LEN_EXTRA[lc]), _LEN_EXTRA[lc])
        b.huff(dcode[dc], dist[dc])
        b.put(rng.randrange(1 << _DIST_EXTRA[dc]), _DIST_EXTRA[dc])
    b.huff(lcode[256], lit[256])
    return b.data()


@pytest.mark.parametrize("variant", [0, 11, 15])
def test_deep_distance_codes_match_zlib(cbm, variant):
    """Dynamic-Huffman blocks whose distance codes run to 15 bits (a subtable under the 8-bit
    distance root), against zlib's bytes."""
    from cubam_amd import _lib

    blob, want = bytearray(), []
    for seed in range(4):
        comp = _deep_dynamic_block(seed)
        p = zlib.decompressobj(-15).decompress(comp)
        head = b"\x1f\x8b\x08\x04\x00\x00\x00\x00\x00\xff\x06\x00BC\x02\x00" + (18 + len(comp) + 7).to_bytes(2, "little")
        blob += head + comp + (zlib.crc32(p) & 0xFFFFFFFF).to_bytes(4, "little") + len(p).to_bytes(4, "little")
        want.append(p)
    blob += bf.EOF_BLOCK
    _say("deep distance codes: variant", variant, len(want), "blocks")
    blocks, inf = _gpu_blocks(np.frombuffer(bytes(blob), np.uint8), _lib.InflateTune(variant, 0, 0, 0))
    assert blocks.n_blocks == len(want) + 1
    st = inf.status.cpu().numpy()[: len(want)] & 255
    data = inf.data[: inf.n].cpu().numpy()
    o = 0
    for k, p in enumerate(want):
        assert st[k] == _lib.BLOCK_OK, (k, st[k])
        assert data[o : o + len(p)].tobytes() == p, k
        o += len(p)


@pytest.mark.parametrize("variant", [0, 11, 15])
def test_fixed_blocks_every_length_and_distance_code_match_zlib(cbm, variant):
    """Fixed-Huffman blocks written bit by bit: every length symbol and every distance symbol
    0-29 with random extra bits (the largest distances reach back 32 KiB), against zlib's bytes;
    and blocks naming distance symbols 30 or 31 or literal/length symbols 286 or 287, which zlib
    rejects ("invalid distance code", "invalid literal/length code") and which must be rejected
    here."""
    from cubam_amd import _lib

    cases = [(s, None) for s in range(3)] + [(7, ("dist", 30)), (8, ("dist", 31)), (9, ("litlen", 286)),
                                             (10, ("litlen", 287))]
    blob, want = bytearray(), []
    for seed, bad in cases:
        comp = _fixed_block(seed, bad)
        try:
            p = zlib.decompressobj(-15).decompress(comp)
        except zlib.error:
            p = None
        assert (p is None) == (bad is not None), (seed, bad)
        payload = p if p is not None else bytes(64)
        head = b"\x1f\x8b\x08\x04\x00\x00\x00\x00\x00\xff\x06\x00BC\x02\x00" + (18 + len(comp) + 7).to_bytes(2, "little")
        blob += head + comp + (zlib.crc32(payload) & 0xFFFFFFFF).to_bytes(4, "little") + len(payload).to_bytes(4, "little")
        want.append(p)
    blob += bf.EOF_BLOCK
    _say("fixed blocks: variant", variant, len(want), "blocks")
    blocks, inf = _gpu_blocks(np.frombuffer(bytes(blob), np.uint8), _lib.InflateTune(variant, 0, 0, 0))
    assert blocks.n_blocks == len(want) + 1
    st = inf.status.cpu().numpy()[: len(want)] & 255
    data = inf.data[: inf.n].cpu().numpy()
    o = 0
    for k, p in enumerate(want):
        if p is None:
            assert st[k] == _lib.BLOCK_DEFLATE, (k, cases[k], st[k])
            o += 64
        else:
            assert st[k] == _lib.BLOCK_OK, (k, st[k])
            assert data[o : o + len(p)].tobytes() == p, k
            o += len(p)


@pytest.mark.parametrize("variant,bpw,wpg", [(0, 0, 0), (11, 1, 4), (11, 4, 2), (11, 8, 4), (11, 16, 2),
                                              (14, 0, 0), (15, 0, 0)])
@pytest.mark.parametrize("where", ["header", "anywhere"])
def test_corrupt_dynamic_headers_match_zlib(cbm, oracle, where, variant, bpw, wpg):
    """Blocks with a few bits flipped in their dynamic Huffman header (the code-length code and
    the code lengths: "header") or anywhere in the compressed data, with the trailer of the
    intact payload, among intact blocks. Against zlib on
    each block: the same blocks inflate; a block that zlib inflates to other bytes (a CRC32
    mismatch) gives the same bytes here and the same error; every block zlib rejects is
    rejected here too (as a Deflate or a size error, which zlib tells apart by where it stopped)."""
    from cubam_amd import _lib

    rng = random.Random(17 if where == "header" else 18)
    blob, off, size, isz, payloads = bytearray(), [], [], [], []
    for i in range(1200):
        n = rng.randint(200, 65280)
        kind = i % 3
        if kind == 0:
            p = bytes(rng.choices(b"ACGTN", weights=[30, 30, 30, 30, 1], k=n))
        elif kind == 1:
            p = bytes(rng.choices(range(33, 75), k=n))
        else:
            p = _data(n, seed=i)
        c = zlib.compressobj(6, zlib.DEFLATED, -15, 8)
        comp = bytearray(c.compress(p) + c.flush())
        if rng.random() < 0.75:
            for _ in range(rng.randint(1, 3)):
                j = rng.randrange(min(60, len(comp)) if where == "header" else len(comp))
                comp[j] ^= 1 << rng.randrange(8)
        head = b"\x1f\x8b\x08\x04\x00\x00\x00\x00\x00\xff\x06\x00BC\x02\x00" + (18 + len(comp) + 7).to_bytes(2, "little")
        block = head + bytes(comp) + (zlib.crc32(p) & 0xFFFFFFFF).to_bytes(4, "little") + len(p).to_bytes(4, "little")
        off.append(len(blob))
        size.append(len(block))
        isz.append(len(p))
        payloads.append(p)
        blob += block
    blob += bf.EOF_BLOCK
    arr = np.frombuffer(bytes(blob), np.uint8)
    _say("corrupt blocks:", where, len(off), "blocks, shape", variant, bpw, wpg)
    blocks, inf = _gpu_blocks(arr, _lib.InflateTune(variant, bpw, wpg, 0))
    assert blocks.n_blocks == len(off) + 1  # and the EOF block
    st = inf.status.cpu().numpy()[: len(off)] & 255
    data = inf.data[: inf.n].cpu().numpy()
    odata, oout, _, ost = oracle.inflate(arr, np.array(off), np.array(size, np.int32), np.array(isz, np.int32))
    # the oracle's status codes are the C ABI's per-block codes
    ok, crc_err = _lib.BLOCK_OK, _lib.BLOCK_CRC
    assert ((st == ok) == (ost == ok)).all(), np.flatnonzero((st == ok) != (ost == ok))[:10]
    assert ((st == crc_err) == (ost == crc_err)).all(), np.flatnonzero((st == crc_err) != (ost == crc_err))[:10]
    assert (ost == ok).sum() > 200 and (ost == crc_err).sum() > 50 and ((ost != ok) & (ost != crc_err)).sum() > 100
    o = 0
    for k in range(len(off)):
        if ost[k] in (ok, crc_err):  # a stream zlib decodes to isize bytes: the same bytes
            assert data[o : o + isz[k]].tobytes() == odata[oout[k] : oout[k] + isz[k]].tobytes(), k
        if ost[k] == ok:
            assert data[o : o + isz[k]].tobytes() == payloads[k]
        o += isz[k]


@pytest.mark.parametrize("variant,bpw,wpg", [(11, 0, 0), (12, 0, 0), (13, 0, 0), (11, 8, 4), (11, 16, 2),
                                              (14, 0, 0), (15, 0, 0)])
def test_two_phase_defers_blocks_it_cannot_finish(cbm, oracle, variant, bpw, wpg):
    """Blocks with more matches than the decode phase's list holds (3-byte matches all through),
    stored blocks, fixed-Huffman blocks, and ordinary ones, mixed: the two-phase inflate takes the
    first kind again in its second pass and every block comes out as zlib makes it, CRC32
    included."""
    from cubam_amd import _lib

    rng = random.Random(71)
    vocab = [bytes(rng.choices(range(256), k=3)) for _ in range(6)]
    payloads, blob = [], b""
    for i in range(120):
        kind = i % 4
        if kind == 0:  # thousands of short matches
            p = b"".join(rng.choice(vocab) + bytes([rng.randrange(256)]) for _ in range(16000))[:65280]
            lvl, strat = 1, zlib.Z_DEFAULT_STRATEGY
        elif kind == 1:
            p, lvl, strat = _data(rng.randint(100, 65280), seed=i), 0, zlib.Z_DEFAULT_STRATEGY
        elif kind == 2:
            p, lvl, strat = _data(rng.randint(100, 65280), seed=i), 6, zlib.Z_FIXED
        else:
            p, lvl, strat = _data(rng.randint(100, 65280), seed=i), 9, zlib.Z_DEFAULT_STRATEGY
        payloads.append(p)
        blob += bf.bgzf_block(p, lvl, strat)
    blob += bf.EOF_BLOCK
    blocks, inf = _gpu_blocks(np.frombuffer(blob, np.uint8), _lib.InflateTune(variant, bpw, wpg, 0))
    assert int(inf.result.item()) == 0
    assert (inf.status.cpu().numpy() == 0).all()
    assert inf.data[: inf.n].cpu().numpy().tobytes() == b"".join(payloads)
    crc = inf.crc.cpu().numpy().astype(np.int64).tolist()[: len(payloads)]
    assert crc == [zlib.crc32(p) for p in payloads]
