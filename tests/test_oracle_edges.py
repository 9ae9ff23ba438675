"""Known answers for the edges of the oracle (CPU only): the structural record test, decode_fixed,
find_tag, and what the chunk loop does at a record that fails the test.

Every expected value is written here from the reference's rules, never read back from the oracle:
  * is_record: src/cubam/_kernels.py:250-278
  * decode_fixed: src/cubam/_kernels.py:292-319 (end = pos + the M/D/N/=/X lengths; a soft clip at
    op 0 is the left clip, any later S the right clip; clips saturate at 65535)
  * find_tag: src/cubam/_kernels.py:334-388 (0 missing, 255 malformed, Z/H length without the NUL,
    B length from the subtype byte, A as its character code, f as 0)
  * the record chain and its errors: src/cubam/bam.py:403-446, src/cubam/_kernels.py:662-702
"""

import struct

import numpy as np
import pytest

import bamfixture as bf

N_REF = len(bf.REFS)


def _rec(**kw) -> bytes:
    base = dict(name="r1", flag=0, ref=0, pos=100, mapq=60, cigar=[("M", 10)], seq="ACGTACGTAC")
    base.update(kw)
    return bf.Read(**base).encode()


def _is_record(oracle, b: bytes, n_ref: int = N_REF) -> bool:
    a = np.frombuffer(b, np.uint8)
    return bool(oracle.lib().orc_is_record(a.ctypes.data, a.size, 0, n_ref))


def _patch(b: bytes, off: int, fmt: str, v) -> bytes:
    out = bytearray(b)
    struct.pack_into(fmt, out, off, v)
    return bytes(out)


# byte offsets in an encoded record (its 4-byte block_size first)
BS, REF, POS, LRN, MAPQ, BIN, NCIG, FLAG, LSEQ, NREF, NPOS, TLEN, NAME = 0, 4, 8, 12, 13, 14, 16, 18, 20, 24, 28, 32, 36


def test_is_record_accepts_the_base_records(oracle):
    assert _is_record(oracle, _rec())
    assert _is_record(oracle, _rec(tags=[("NM", "i", 1)]))
    assert _is_record(oracle, _rec(ref=-1, pos=-1, cigar=[], seq="ACG", flag=4))  # unmapped: no bin check


@pytest.mark.parametrize(
    "what, off, fmt, value",
    [
        ("block_size below 33", BS, "<i", 32),
        ("block_size past the data", BS, "<i", None),  # one byte more than the record has
        ("refID below -1", REF, "<i", -2),
        ("refID not below n_ref", REF, "<i", N_REF),
        ("pos below -1", POS, "<i", -2),
        ("next_refID below -1", NREF, "<i", -2),
        ("next_refID not below n_ref", NREF, "<i", N_REF),
        ("next_pos below -1", NPOS, "<i", -2),
        ("l_read_name below 2", LRN, "<B", 1),
        ("flag not below 4096", FLAG, "<H", 4096),
        ("l_seq above 4 Mi", LSEQ, "<I", (1 << 22) + 1),
        ("fields longer than block_size", NCIG, "<H", 2),
        ("bin not reg2bin", BIN, "<H", None),
    ],
)
def test_is_record_reject_predicates(oracle, what, off, fmt, value):
    good = _rec()  # no aux: the field sizes add up to block_size exactly
    if value is None and off == BS:
        value = len(good) - 4 + 1
    if value is None and off == BIN:
        value = struct.unpack_from("<H", good, BIN)[0] + 1
    assert _is_record(oracle, good)
    assert not _is_record(oracle, _patch(good, off, fmt, value)), what


def test_is_record_boundaries_that_pass(oracle):
    good = _rec()
    for off, fmt, v in ((REF, "<i", N_REF - 1), (NREF, "<i", -1), (NPOS, "<i", -1), (FLAG, "<H", 4095)):
        assert _is_record(oracle, _patch(good, off, fmt, v))
    # a mapped read's bin is only checked for pos >= 0
    unmapped = _rec(ref=-1, pos=-1, cigar=[], seq="ACG", flag=4)
    assert _is_record(oracle, _patch(unmapped, BIN, "<H", 1234))


def test_is_record_block_size_above_4_mib(oracle):
    bs = (1 << 22) + 1
    base = bytearray(_rec())
    body = bytearray(4 + bs)  # room for the whole record, so only the size limit can fail
    body[: len(base)] = base
    struct.pack_into("<i", body, BS, bs)
    assert not _is_record(oracle, bytes(body))
    struct.pack_into("<i", body, BS, 1 << 22)  # the limit itself passes (size identity: fields <= block_size)
    assert _is_record(oracle, bytes(body[: 4 + (1 << 22)]))


@pytest.mark.parametrize("ch, ok", [(ord("@"), False), (32, False), (127, False), (33, True), (126, True)])
def test_is_record_name_characters(oracle, ch, ok):
    good = _rec(name="abcd")
    assert _is_record(oracle, _patch(good, NAME + 1, "<B", ch)) == ok


def test_is_record_name_must_end_in_nul(oracle):
    good = _rec(name="abcd")
    assert not _is_record(oracle, _patch(good, NAME + 4, "<B", ord("x")))


def test_is_record_cigar_ops(oracle):
    good = _rec()  # name "r1" + NUL: the CIGAR starts at byte 39
    cg = NAME + 3
    assert struct.unpack_from("<I", good, cg)[0] == (10 << 4) | 0
    assert _is_record(oracle, _patch(good, cg, "<I", (10 << 4) | 8))  # X: same reference length
    assert not _is_record(oracle, _patch(good, cg, "<I", (10 << 4) | 9))


def test_is_record_bin_uses_a_length_of_one_without_cigar(oracle):
    # pos 16384, no CIGAR: bin = reg2bin(16384, 16385) = 4682; reg2bin(16384, 16384) would give 585
    good = _rec(pos=16384, cigar=[], seq="ACG")
    assert struct.unpack_from("<H", good, BIN)[0] == 4682
    assert _is_record(oracle, good)
    assert not _is_record(oracle, _patch(good, BIN, "<H", 585))


def test_is_record_needs_36_bytes(oracle):
    assert not _is_record(oracle, _rec()[:35])


# ------------------------------------------------------------------ decode_fixed


def _columns(oracle, recs, tags=()):
    raw = b"".join(recs)
    offs, o = [], 0
    for r in recs:
        offs.append(o)
        o += len(r)
    return oracle.columns(np.frombuffer(raw, np.uint8), np.array(offs, np.int64), tags=tags)


def test_decode_fixed_known_answers(oracle):
    recs = [
        _rec(cigar=[("S", 70000), ("M", 10), ("S", 100000)], seq="A" * 170010),  # both clips saturate
        _rec(cigar=[("S", 65535), ("M", 5), ("S", 65536)], seq="C" * 131076),  # at and past the limit
        _rec(cigar=[("H", 5), ("S", 3), ("M", 10), ("S", 4), ("H", 2)], seq="G" * 17),  # S after H: a right clip
        _rec(cigar=[("S", 2), ("M", 3), ("S", 5), ("I", 1), ("S", 6)], seq="T" * 17),  # the last S wins
        _rec(pos=100, cigar=[], seq="ACG"),  # mapped, no CIGAR: end = pos
        _rec(ref=-1, pos=-1, cigar=[], seq="AC", flag=4),  # unmapped
        _rec(cigar=[("M", 5), ("I", 3), ("D", 4), ("N", 100), ("=", 2), ("X", 1), ("P", 7), ("H", 1)],
             seq="A" * 11, ref=2, mapq=7, flag=0x93, nref=1, npos=55, tlen=-321),
    ]
    c = _columns(oracle, recs)
    assert c["clip_left"].tolist() == [65535, 65535, 0, 2, 0, 0, 0]
    assert c["clip_right"].tolist() == [65535, 65535, 4, 6, 0, 0, 0]
    assert c["end"].tolist() == [110, 105, 110, 103, 100, -1, 100 + 5 + 4 + 100 + 2 + 1]
    assert c["l_seq"].tolist() == [170010, 131076, 17, 17, 3, 2, 11]
    assert c["ref"].tolist()[-1] == 2 and c["mapq"].tolist()[-1] == 7 and c["flag"].tolist()[-1] == 0x93
    assert c["nref"].tolist()[-1] == 1 and c["npos"].tolist()[-1] == 55 and c["tlen"].tolist()[-1] == -321
    assert c["cigar_off"].tolist() == [0, 3, 6, 11, 16, 16, 16, 24]


# ------------------------------------------------------------------ find_tag


def _aux_start(rec: bytes) -> int:
    lrn = rec[LRN]
    ncig, lseq = struct.unpack_from("<H", rec, NCIG)[0], struct.unpack_from("<I", rec, LSEQ)[0]
    return NAME + lrn + 4 * ncig + (lseq + 1) // 2 + lseq


def _tag(oracle, rec: bytes, name: str):
    a = np.frombuffer(rec, np.uint8)
    ty, vo, vl, iv = oracle.find_tag(a, np.zeros(1, np.int64), name)
    return int(ty[0]), int(vo[0]), int(vl[0]), int(iv[0])


def test_find_tag_every_type(oracle):
    tags = [("XA", "A", "Q"), ("XF", "f", 1.5), ("XH", "H", "1AE3"), ("XB", "B", ("f", [1.0, 2.0, 3.0])),
            ("XC", "c", -5), ("XD", "C", 200), ("XS", "s", -300), ("XT", "S", 65000), ("XI", "i", -70000),
            ("XU", "I", 4000000000), ("XZ", "Z", "hello"), ("XE", "B", ("c", [])), ("XG", "B", ("S", [7, 8]))]
    rec = _rec(tags=tags)
    # value sizes after the 3 bytes of name and type, by the SAM grammar
    size = {"XA": 1, "XF": 4, "XH": 5, "XB": 5 + 12, "XC": 1, "XD": 1, "XS": 2, "XT": 2, "XI": 4, "XU": 4,
            "XZ": 6, "XE": 5, "XG": 5 + 4}
    expect = {  # (type, length, value)
        "XA": ("A", 0, ord("Q")), "XF": ("f", 0, 0), "XH": ("H", 4, 0), "XB": ("B", 17, 0), "XC": ("c", 0, -5),
        "XD": ("C", 0, 200), "XS": ("s", 0, -300), "XT": ("S", 0, 65000), "XI": ("i", 0, -70000),
        "XU": ("I", 0, 4000000000), "XZ": ("Z", 5, 0), "XE": ("B", 5, 0), "XG": ("B", 9, 0),
    }
    at = _aux_start(rec)
    for name, _, _ in tags:
        ty, vo, vl, iv = _tag(oracle, rec, name)
        t, length, value = expect[name]
        assert (ty, vo, vl, iv) == (ord(t), at + 3, length, value), name
        at += 3 + size[name]
    assert at == len(rec)
    assert _tag(oracle, rec, "NM") == (0, 0, 0, 0)


def test_find_tag_first_of_two_wins(oracle):
    rec = _rec(tags=[("NM", "i", 1), ("NM", "i", 2)])
    assert _tag(oracle, rec, "NM")[3] == 1


def test_find_tag_malformed_before_and_after(oracle):
    # CB:Z, then a field of an unknown type 'q', then UB:Z
    rec = _rec(tags=[("CB", "Z", "AAAC-1")], aux_raw=b"XXq\0\0" + b"UBZGGGG\0")
    at = _aux_start(rec)
    assert _tag(oracle, rec, "CB") == (ord("Z"), at + 3, 6, 0)  # before the bad field
    assert _tag(oracle, rec, "UB") == (255, 0, 0, 0)  # after it
    assert _tag(oracle, rec, "ZZ") == (255, 0, 0, 0)  # missing, past it


@pytest.mark.parametrize(
    "aux_raw, name, expect",
    [
        (b"CBZACGT", "CB", 255),  # a Z value with no NUL before the record's end
        (b"CBZACGT", "NM", 3),  # a tag before it is found
        (b"XBBs" + struct.pack("<I", 1000) + struct.pack("<2h", 1, 2) + b"UBZAC\0", "UB", 255),  # after a bad B
        (b"XBBs" + struct.pack("<I", 1000) + struct.pack("<2h", 1, 2), "XB", 255),  # count past the end
        (b"XBBs\x01\x00", "XB", 255),  # subtype and count cut short
        (b"XIi\x01\x02", "XI", 255),  # the asked tag itself cut short
        (b"ab", "ZZ", 0),  # under 3 bytes left: the walk ends, the tag is missing
    ],
)
def test_find_tag_malformed_cases(oracle, aux_raw, name, expect):
    rec = _rec(tags=[("NM", "i", 3)], aux_raw=aux_raw)
    ty, vo, vl, iv = _tag(oracle, rec, name)
    if expect == 3:
        assert (ty, iv) == (ord("i"), 3)
    else:
        assert (ty, vo, vl, iv) == (expect, 0, 0, 0)


def test_find_tag_b_array_of_unknown_subtype_takes_4_bytes(oracle):
    raw = b"XBBx" + struct.pack("<I", 2) + b"\1" * 8 + b"NMi" + struct.pack("<i", 9)
    rec = _rec(aux_raw=raw)
    at = _aux_start(rec)
    assert _tag(oracle, rec, "XB") == (ord("B"), at + 3, 13, 0)
    assert _tag(oracle, rec, "NM") == (ord("i"), at + 16 + 3, 0, 9)


# ------------------------------------------------------------------ the chain at an invalid record

NAME_MSG = "truncated BAM file"


def _reads(n, seed=0):
    return bf.random_reads(n, seed=seed)


def _bad(read: bf.Read) -> bytes:
    """The read's record with a flag of 4096: complete, but it fails the structural test."""
    return _patch(read.encode(), FLAG, "<H", 4096)


def _file(reads, bad_at: int | None, block: int = 4096) -> tuple[bytes, list]:
    recs = [(_bad(r) if i == bad_at else r.encode()) for i, r in enumerate(reads)]
    return bf.bgzf(bf.header_bytes() + b"".join(recs), block), recs


def _run(oracle, raw, chunk):
    got, err = [], None
    try:
        for data, c in oracle.iter_bam(raw, chunk, tags=()):
            got.append(c)
    except ValueError as e:
        err = str(e)
    pos = [int(x) for c in got for x in c["pos"]]
    return pos, err


def test_invalid_record_in_the_only_chunk(oracle):
    reads = _reads(40)
    raw, _ = _file(reads, 17)
    pos, err = _run(oracle, raw, 1 << 20)
    # the records before it are yielded; it is carried, and the file ends with it carried
    assert pos == [r.pos for r in reads[:17]]
    assert err == f"BAM: {NAME_MSG} (incomplete record at the end)"


def test_invalid_record_mid_file(oracle):
    reads = _reads(400, seed=1)
    raw, _ = _file(reads, 150, block=2000)
    pos, err = _run(oracle, raw, 6000)
    # records after it are valid candidates of the next chunk, whose first offset is not one: raise
    assert pos == [r.pos for r in reads[:150]]
    assert err == oracle.MSG_NO_BLOCK


def test_invalid_first_record(oracle):
    reads = _reads(60, seed=2)
    raw, _ = _file(reads, 0)
    pos, err = _run(oracle, raw, 1 << 20)
    assert pos == [] and err == oracle.MSG_NO_BLOCK


def test_invalid_record_with_no_candidate_is_carried(oracle):
    # the only record fails the test and no offset passes it: the bytes are carried to the end
    reads = _reads(1, seed=3)
    raw, _ = _file(reads, 0)
    pos, err = _run(oracle, raw, 1 << 20)
    assert pos == [] and err == f"BAM: {NAME_MSG} (incomplete record at the end)"
    assert oracle.record_candidates(np.frombuffer(_bad(reads[0]), np.uint8), N_REF) == 0
