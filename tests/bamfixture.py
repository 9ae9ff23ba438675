"""A small pure-Python BAM writer and parser for tests (SAMv1 sections 4.1-4.2, zlib).

Independent of the oracle's C code: tests check the oracle (and through it the
GPU path) against the records written here. The random-read recipe follows the
reference's tests/test_bam.py:13-57 (soft clips, a deletion, NM:i, XS:C, CB:Z
on 90 % of reads, an optional B-array tag placed first, RG:Z).
"""

from __future__ import annotations

import random
import struct
import zlib
from dataclasses import dataclass, field

EOF_BLOCK = bytes.fromhex("1f8b08040000000000ff0600424302001b0003000000000000000000")
OPS = "MIDNSHP=X"
SEQ = "=ACMGRSVTWYHKDBN"
REFS = [("chr1", 1_000_000), ("chr2", 500_000), ("chrM", 16_569)]


def reg2bin(beg: int, end: int) -> int:
    """SAMv1 section 5.3, written from the bin sizes: 2^14, 2^17, 2^20, 2^23, 2^26."""
    end -= 1
    first = 0
    offsets = []
    for level in range(6):
        offsets.append(first)
        first += 8**level
    # levels 5..1 from the smallest bins up
    for level, shift in ((5, 14), (4, 17), (3, 20), (2, 23), (1, 26)):
        if beg >> shift == end >> shift:
            return offsets[level] + (beg >> shift)
    return 0


@dataclass
class Read:
    name: str
    flag: int = 0
    ref: int = -1
    pos: int = -1
    mapq: int = 60
    cigar: list = field(default_factory=list)  # [(op, len)]
    nref: int = -1
    npos: int = -1
    tlen: int = 0
    seq: str = ""
    qual: bytes | None = None
    tags: list = field(default_factory=list)  # (name, type, value)
    aux_raw: bytes = b""  # bytes appended after the tags (malformed aux fields)

    @property
    def ref_len(self) -> int:
        return sum(n for op, n in self.cigar if op in "MDN=X")

    @property
    def end(self) -> int:
        return self.pos + self.ref_len

    def clips(self) -> tuple[int, int]:
        left = right = 0
        for k, (op, n) in enumerate(self.cigar):
            if op == "S":
                if k == 0:
                    left = n
                else:
                    right = n
        return min(left, 65535), min(right, 65535)

    def tag(self, name: str):
        for t, ty, v in self.tags:
            if t == name:
                return ty, v
        return None

    def encode(self) -> bytes:
        if not self.seq:
            self.seq = "A" * sum(n for op, n in self.cigar if op in "MIS=X")
        name = self.name.encode() + b"\0"
        rl = self.ref_len
        bin_ = reg2bin(self.pos, self.pos + (rl if rl else 1)) if self.pos >= 0 else 4680
        l_seq = len(self.seq)
        out = bytearray()
        out += struct.pack("<iiBBHHHIiii", self.ref, self.pos, len(name), self.mapq, bin_, len(self.cigar),
                           self.flag, l_seq, self.nref, self.npos, self.tlen)
        out += name
        for op, n in self.cigar:
            out += struct.pack("<I", (n << 4) | OPS.index(op))
        codes = [SEQ.index(c) for c in self.seq]
        for i in range(0, l_seq, 2):
            hi = codes[i]
            lo = codes[i + 1] if i + 1 < l_seq else 0
            out.append((hi << 4) | lo)
        out += self.qual if self.qual is not None else b"\xff" * l_seq
        for t, ty, v in self.tags:
            out += t.encode()
            if ty in "ZH":
                out += ty.encode() + v.encode() + b"\0"
            elif ty == "B":
                sub, vals = v
                fmt = {"c": "b", "C": "B", "s": "h", "S": "H", "i": "i", "I": "I", "f": "f"}[sub]
                out += b"B" + sub.encode() + struct.pack("<I", len(vals)) + struct.pack(f"<{len(vals)}{fmt}", *vals)
            elif ty == "A":
                out += b"A" + v.encode()
            else:
                fmt = {"c": "b", "C": "B", "s": "h", "S": "H", "i": "i", "I": "I", "f": "f"}[ty]
                out += ty.encode() + struct.pack("<" + fmt, v)
        out += self.aux_raw
        return struct.pack("<i", len(out)) + bytes(out)


def header_bytes(refs=REFS, text: str | None = None) -> bytes:
    if text is None:
        text = "@HD\tVN:1.6\tSO:coordinate\n" + "".join(f"@SQ\tSN:{n}\tLN:{ln}\n" for n, ln in refs)
    t = text.encode()
    out = b"BAM\1" + struct.pack("<i", len(t)) + t + struct.pack("<i", len(refs))
    for n, ln in refs:
        nb = n.encode() + b"\0"
        out += struct.pack("<i", len(nb)) + nb + struct.pack("<i", ln)
    return out


def bam_bytes(reads, refs=REFS, text: str | None = None) -> bytes:
    return header_bytes(refs, text) + b"".join(r.encode() for r in reads)


def bgzf_block(payload: bytes, level: int = 6, strategy: int = zlib.Z_DEFAULT_STRATEGY) -> bytes:
    c = zlib.compressobj(level, zlib.DEFLATED, -15, 8, strategy)
    comp = c.compress(payload) + c.flush()
    bsize = 18 + len(comp) + 8
    head = b"\x1f\x8b\x08\x04\x00\x00\x00\x00\x00\xff\x06\x00BC\x02\x00" + struct.pack("<H", bsize - 1)
    return head + comp + struct.pack("<II", zlib.crc32(payload) & 0xFFFFFFFF, len(payload))


def bgzf(data: bytes, block: int = 0xFF00, level: int = 6, eof: bool = True) -> bytes:
    out = bytearray()
    for i in range(0, len(data), block):
        out += bgzf_block(data[i : i + block], level)
    if eof:
        out += EOF_BLOCK
    return bytes(out)


def sort_key(r: Read):
    return (r.ref if r.ref >= 0 else 1 << 30, r.pos)


def random_reads(n: int, seed: int = 0, refs=REFS) -> list[Read]:
    rng = random.Random(seed)
    reads = []
    for i in range(n):
        ref = rng.randrange(len(refs))
        cl, m, cr = rng.choice([0, 0, 3]), rng.randint(20, 150), rng.choice([0, 0, 7])
        cigar = ([("S", cl)] if cl else []) + [("M", m // 2), ("D", 2), ("M", m - m // 2)] + ([("S", cr)] if cr else [])
        tags = [("NM", "i", rng.randint(0, 5)), ("XS", "C", rng.randint(0, 255))]
        if rng.random() < 0.9:
            tags.append(("CB", "Z", "".join(rng.choices("ACGT", k=16)) + "-1"))
        if rng.random() < 0.3:
            tags.insert(0, ("ZB", "B", ("s", [rng.randint(-100, 100) for _ in range(rng.randint(0, 9))])))
        tags.append(("RG", "Z", f"group{rng.randint(0, 3)}"))
        length = sum(k for op, k in cigar if op in "MIS=X")
        seq = "".join(rng.choices("ACGTN", weights=[30, 30, 30, 30, 1], k=length))
        qual = bytes(rng.randint(2, 41) for _ in range(length)) if rng.random() < 0.8 else None
        reads.append(Read(name=f"read{i}:{rng.randint(0, 10**6)}", flag=rng.choice([99, 147, 83, 163, 1123, 0, 16]),
                          ref=ref, pos=rng.randrange(refs[ref][1] - 500), mapq=rng.randint(0, 60), cigar=cigar,
                          nref=ref, npos=rng.randrange(refs[ref][1]), tlen=rng.randint(-1000, 1000), seq=seq,
                          qual=qual, tags=tags))
    return sorted(reads, key=sort_key)


# ------------------------------------------------------------------ parser (ground truth for generated files)


def parse_records(raw: bytes, start: int) -> list[dict]:
    """Every record of an uncompressed BAM stream from `start`: its fields and raw slices."""
    out = []
    o = start
    n = len(raw)
    while o + 4 <= n:
        (bs,) = struct.unpack_from("<i", raw, o)
        if o + 4 + bs > n:
            break
        ref, pos, lrn, mapq, _bin, ncig, flag, lseq, nref, npos, tlen = struct.unpack_from("<iiBBHHHIiii", raw, o + 4)
        p = o + 36
        name = raw[p : p + lrn - 1]
        p += lrn
        cig = list(struct.unpack_from(f"<{ncig}I", raw, p))
        p += 4 * ncig
        seq = raw[p : p + (lseq + 1) // 2]
        p += (lseq + 1) // 2
        qual = raw[p : p + lseq]
        p += lseq
        aux = raw[p : o + 4 + bs]
        rl = sum(v >> 4 for v in cig if (v & 15) in (0, 2, 3, 7, 8))
        cl = cr = 0
        for k, v in enumerate(cig):
            if v & 15 == 4:
                if k == 0:
                    cl = v >> 4
                else:
                    cr = v >> 4
        out.append(dict(offset=o, ref=ref, pos=pos, end=pos + rl, mapq=mapq, flag=flag, nref=nref, npos=npos,
                        tlen=tlen, l_seq=lseq, clip_left=min(cl, 65535), clip_right=min(cr, 65535), name=name,
                        cigar=cig, seq=seq, qual=qual, aux=aux))
        o += 4 + bs
    return out


def aux_tags(aux: bytes) -> dict:
    """name -> (type, value) of a well-formed aux area; Z values as str, integers as int."""
    tags = {}
    p = 0
    sizes = {"A": 1, "c": 1, "C": 1, "s": 2, "S": 2, "i": 4, "I": 4, "f": 4}
    fmts = {"c": "<b", "C": "<B", "s": "<h", "S": "<H", "i": "<i", "I": "<I", "f": "<f"}
    while p + 3 <= len(aux):
        name = aux[p : p + 2].decode()
        ty = chr(aux[p + 2])
        v = p + 3
        if ty in "ZH":
            e = aux.index(b"\0", v)
            val, p = aux[v:e].decode(), e + 1
        elif ty == "B":
            sub = chr(aux[v])
            (cnt,) = struct.unpack_from("<I", aux, v + 1)
            val, p = (sub, cnt), v + 5 + cnt * sizes[sub]
        elif ty == "A":
            val, p = aux[v], v + 1
        else:
            (val,) = struct.unpack_from(fmts[ty], aux, v)
            p = v + sizes[ty]
        tags.setdefault(name, (ty, val))
    return tags
