"""The decode tables' sizes: the largest table a canonical Huffman code can need.

The wave decoder keeps a distance table of root WDBITS bits and WDTAB entries in LDS
(``cubam_amd/csrc/inflate_wave.h``). A table is the root plus, for each root prefix that longer
codes share, a subtable of 2^(deepest code under it - root) entries; the table build rejects a code
whose tables would overrun WDTAB, so WDTAB must cover every complete code of 30 distance symbols
of at most 15 bits. ``table_bound`` computes that worst case. It is checked against zlib's
published bound for its 6-bit distance root (592, ``ENOUGH_DISTS`` in zlib's inftrees.h) and
against brute force over all codes of small alphabets.
"""

from __future__ import annotations

import itertools
import re
from functools import lru_cache
from pathlib import Path

import pytest

HEADER = Path(__file__).resolve().parents[1] / "cubam_amd" / "csrc" / "inflate_wave.h"


def table_bound(n: int, root: int, maxlen: int) -> int:
    """Largest root + subtable size of a complete canonical code of at most n symbols.

    In canonical order the internal nodes at each depth are the rightmost ones, so with a_d
    internal nodes at depth d >= root, c_j = ceil(a_{root+j-1} / 2^(j-1)) root prefixes have a
    subtable of 2^j entries or more, and the subtables add sum_j (c_j - c_{j+1}) 2^j. Above the
    root, a_root internal nodes leave 2^root - a_root root entries to codes of root bits or fewer,
    at least popcount(2^root - a_root) of them.
    """
    best = 0
    for ar in range(1, 2**root + 1):
        budget = n - bin(2**root - ar).count("1")
        if budget < 2 * ar:
            continue

        @lru_cache(None)
        def rest(d: int, a: int, used: int):
            j = d - root + 1
            cj = -(-a // 2 ** (j - 1))
            if d == maxlen - 1:  # all 2a children are leaves of maxlen bits
                return cj * 2**j if used + 2 * a <= budget else None
            top = None
            for a2 in range(0, 2 * a + 1):
                leaves = 2 * a - a2
                if used + leaves + 2 * a2 > budget:
                    continue
                v = rest(d + 1, a2, used + leaves)
                if v is None:
                    continue
                v += (cj - -(-a2 // 2**j)) * 2**j
                top = v if top is None else max(top, v)
            return top

        v = rest(root, ar, 0)
        if v is not None:
            best = max(best, 2**root + v)
    return best


def built_size(lengths: list[int], root: int) -> int:
    """Entries the wave decoder's table build uses for these code lengths (canonical codes)."""
    maxlen = max(lengths)
    count = [0] * (maxlen + 1)
    for n in lengths:
        count[n] += 1
    code, nxt = 0, [0] * (maxlen + 2)
    for n in range(1, maxlen + 1):
        code = (code + count[n - 1]) << 1 if n > 1 else 0
        nxt[n] = code
    deepest: dict[int, int] = {}
    for n in sorted(lengths):
        c = nxt[n]
        nxt[n] += 1
        if n > root:
            prefix = c >> (n - root)  # the code's first root bits (as a number, MSB first)
            deepest[prefix] = max(deepest.get(prefix, 0), n - root)
    return 2**root + sum(2**m for m in deepest.values())


def brute_bound(n: int, root: int, maxlen: int) -> int:
    """Largest table over every complete code of at most n symbols (small alphabets only)."""
    best = 0
    for counts in itertools.product(range(n + 1), repeat=maxlen):
        if sum(counts) > n or sum(c * 2 ** (maxlen - 1 - i) for i, c in enumerate(counts)) != 2**maxlen:
            continue
        lengths = [i + 1 for i, c in enumerate(counts) for _ in range(c)]
        best = max(best, built_size(lengths, root))
    return best


def test_bound_matches_zlib_for_its_distance_root():
    assert table_bound(30, 6, 15) == 592


@pytest.mark.parametrize("n,root,maxlen", [(6, 2, 5), (7, 3, 6), (8, 2, 6), (9, 3, 7), (10, 4, 7)])
def test_bound_matches_brute_force(n, root, maxlen):
    assert table_bound(n, root, maxlen) == brute_bound(n, root, maxlen)


def test_wave_decoder_distance_table_holds_every_code():
    text = HEADER.read_text()
    m = re.search(r"constexpr int WDBITS = (\d+), WDTAB = (\d+);", text)
    assert m, "WDBITS / WDTAB not found"
    bits, entries = int(m.group(1)), int(m.group(2))
    assert table_bound(30, bits, 15) <= entries
    # a tighter table would not hold the worst code
    assert table_bound(30, bits, 15) == entries
