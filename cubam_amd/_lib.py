"""ctypes binding of libcubam_amd.so (the C ABI of include/cubam_amd.h).

There is no fallback: without the built library every entry point raises.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

import torch

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "libcubam_amd.so"

PAD = 64  # CBM_PAD: bytes readable past the end of every input buffer
DEFLATE_SLOT = 65536 + 256  # CBM_DEFLATE_SLOT
DEFLATE_MAX_IN = 65280  # CBM_DEFLATE_MAX_IN

# result slots and codes (mirrors of the header's enums)
SCAN_NBLOCKS, SCAN_USED, SCAN_STATUS, SCAN_TOTAL_OUT, SCAN_NCAND = 0, 1, 2, 3, 4
SCAN_OK, SCAN_TRUNCATED, SCAN_NO_BLOCK, SCAN_CORRUPT, SCAN_CAP = 0, 1, 2, 3, 4
BLOCK_OK, BLOCK_DEFLATE, BLOCK_SIZE, BLOCK_CRC = 0, 1, 2, 3
REC_N, REC_TAIL, REC_INVALID, REC_CIGAR, REC_SEQ, REC_QUAL, REC_BAD_LINKS, REC_NSEG = range(8)


INFLATE_WG512_S128, INFLATE_LANES, INFLATE_WG128, INFLATE_WG512, INFLATE_WG64 = 0, 1, 2, 3, 4
INFLATE_TOK64_S512, INFLATE_TOK64_S1024, INFLATE_TOK64_S256, INFLATE_TOK64_S512_FUSED = 11, 12, 13, 14
INFLATE_TOK64_S512_SPLIT = 15
INFLATE_VARIANTS = 16
#: (threads per block, segment bits) of each workgroup variant; None = one block per lane
INFLATE_SHAPES = {0: (512, 128), 1: None, 2: (128, 256), 3: (512, 64), 4: (64, 512), 5: (256, 128), 6: (128, 128),
                  7: (64, 128), 8: (64, 256), 9: (384, 128), 10: (256, 256), 11: (64, 512), 12: (64, 1024),
                  13: (64, 256), 14: (64, 512), 15: (64, 512)}
#: the two-phase variants (decode to literals and match lists, then resolve)
INFLATE_TWO_PHASE = (INFLATE_TOK64_S512, INFLATE_TOK64_S1024, INFLATE_TOK64_S256, INFLATE_TOK64_S512_FUSED,
                     INFLATE_TOK64_S512_SPLIT)
#: the variant a null tune selects below INFLATE_SPLIT_MIN blocks (CBM_INFLATE_DEFAULT), and from there on
INFLATE_DEFAULT = INFLATE_TOK64_S512
INFLATE_SPLIT_MIN = 4096


def default_variant(n_blocks: int) -> int:
    """The variant a null tune selects for an inflate of ``n_blocks`` blocks."""
    return INFLATE_TOK64_S512_SPLIT if n_blocks >= INFLATE_SPLIT_MIN else INFLATE_DEFAULT


class InflateTune(ctypes.Structure):
    _fields_ = [
        ("variant", ctypes.c_int32),
        ("blocks_per_wave", ctypes.c_int32),
        ("waves_per_group", ctypes.c_int32),
        ("groups", ctypes.c_int32),
    ]


class Columns(ctypes.Structure):
    _fields_ = [
        ("ref", ctypes.c_void_p),
        ("pos", ctypes.c_void_p),
        ("end", ctypes.c_void_p),
        ("mapq", ctypes.c_void_p),
        ("flag", ctypes.c_void_p),
        ("nref", ctypes.c_void_p),
        ("npos", ctypes.c_void_p),
        ("tlen", ctypes.c_void_p),
        ("l_seq", ctypes.c_void_p),
        ("clip_left", ctypes.c_void_p),
        ("clip_right", ctypes.c_void_p),
        ("cigar_off", ctypes.c_void_p),
        ("cigar", ctypes.c_void_p),
        ("seq_off", ctypes.c_void_p),
        ("seq", ctypes.c_void_p),
        ("qual_off", ctypes.c_void_p),
        ("qual", ctypes.c_void_p),
        ("n_tags", ctypes.c_int32),
        ("tag_names", ctypes.c_char * 8),
        ("tag_type", ctypes.c_void_p * 4),
        ("tag_offset", ctypes.c_void_p * 4),
        ("tag_length", ctypes.c_void_p * 4),
        ("tag_value", ctypes.c_void_p * 4),
    ]


#: every symbol include/cubam_amd.h declares
SYMBOLS = (
    "cbm_bgzf_scan_workspace",
    "cbm_bgzf_scan",
    "cbm_inflate_workspace",
    "cbm_inflate",
    "cbm_inflate_noalloc",
    "cbm_crc32_spans",
    "cbm_record_chain_workspace",
    "cbm_record_chain",
    "cbm_record_chain_repair",
    "cbm_record_candidates",
    "cbm_record_offsets",
    "cbm_record_offsets_csr",
    "cbm_record_spans_workspace",
    "cbm_record_chain_spans",
    "cbm_record_chain_spans_repair",
    "cbm_record_offsets_spans",
    "cbm_record_offsets_spans_csr",
    "cbm_scatter_workspace",
    "cbm_scatter",
    "cbm_scatter_csr",
    "cbm_name_hash",
    "cbm_hash_spans",
    "cbm_pack_acgt",
    "cbm_aligned_length",
    "cbm_blocks",
    "cbm_base_at",
    "cbm_deflate_workspace",
    "cbm_deflate",
    "cbm_copy_spans",
    "cbm_last_error",
    "cbm_version",
)

_lib = None


def load():
    """The loaded library (raises when it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is not built; run `python -m cubam_amd.build` "
            "(cubam_amd has no CPU fallback)"
        )
    L = ctypes.CDLL(str(LIB_PATH))
    vp, i64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32
    L.cbm_bgzf_scan_workspace.restype = i64
    L.cbm_bgzf_scan_workspace.argtypes = [i64]
    L.cbm_bgzf_scan.restype = ctypes.c_int
    L.cbm_bgzf_scan.argtypes = [vp, i64, ctypes.c_int, i64, vp, vp, vp, vp, vp, vp, vp, i64, vp]
    L.cbm_inflate_workspace.restype = i64
    L.cbm_inflate_workspace.argtypes = [i64]
    L.cbm_inflate.restype = ctypes.c_int
    L.cbm_inflate.argtypes = [vp, vp, vp, vp, vp, vp, i64, vp, vp, vp, vp, ctypes.POINTER(InflateTune), vp, i64, vp]
    L.cbm_inflate_noalloc.restype = ctypes.c_int
    L.cbm_inflate_noalloc.argtypes = [vp, vp, vp, vp, vp, vp, i64, vp, vp, vp, vp, ctypes.POINTER(InflateTune), vp]
    L.cbm_crc32_spans.restype = ctypes.c_int
    L.cbm_crc32_spans.argtypes = [vp, vp, vp, i64, vp, vp]
    L.cbm_record_chain_workspace.restype = i64
    L.cbm_record_chain_workspace.argtypes = [i64, ctypes.POINTER(i64)]
    L.cbm_record_chain.restype = ctypes.c_int
    L.cbm_record_chain.argtypes = [vp, i64, i64, i32, vp, vp, i64, vp]
    L.cbm_record_chain_repair.restype = ctypes.c_int
    L.cbm_record_chain_repair.argtypes = [vp, i64, i64, i32, vp, vp, i64, vp]
    L.cbm_record_candidates.restype = ctypes.c_int
    L.cbm_record_candidates.argtypes = [vp, i64, i64, i32, vp, vp]
    L.cbm_record_offsets.restype = ctypes.c_int
    L.cbm_record_offsets.argtypes = [vp, i64, i64, vp, i64, vp, i64, vp]
    L.cbm_record_offsets_csr.restype = ctypes.c_int
    L.cbm_record_offsets_csr.argtypes = [vp, i64, i64, vp, i64, vp, vp, vp, vp, i64, vp]
    L.cbm_record_spans_workspace.restype = i64
    L.cbm_record_spans_workspace.argtypes = [i64]
    for f in (L.cbm_record_chain_spans, L.cbm_record_chain_spans_repair):
        f.restype = ctypes.c_int
        f.argtypes = [vp, vp, vp, vp, vp, i64, i64, vp, vp, vp, i64, vp]
    L.cbm_record_offsets_spans.restype = ctypes.c_int
    L.cbm_record_offsets_spans.argtypes = [vp, i64, vp, i64, vp, i64, vp]
    L.cbm_record_offsets_spans_csr.restype = ctypes.c_int
    L.cbm_record_offsets_spans_csr.argtypes = [vp, i64, vp, i64, vp, vp, vp, vp, i64, vp]
    L.cbm_scatter_workspace.restype = i64
    L.cbm_scatter_workspace.argtypes = [i64]
    L.cbm_scatter.restype = ctypes.c_int
    L.cbm_scatter.argtypes = [vp, vp, i64, ctypes.POINTER(Columns), vp, i64, vp]
    L.cbm_scatter_csr.restype = ctypes.c_int
    L.cbm_scatter_csr.argtypes = [vp, vp, i64, ctypes.POINTER(Columns), vp]
    for f, at in ((L.cbm_name_hash, [vp, vp, i64, vp, vp]), (L.cbm_hash_spans, [vp, vp, vp, i64, vp, vp]),
                  (L.cbm_pack_acgt, [vp, vp, vp, i64, vp, vp, vp]), (L.cbm_aligned_length, [vp, vp, i64, vp, vp]),
                  (L.cbm_blocks, [vp, vp, i64, ctypes.c_int, ctypes.c_int, vp, vp, vp, vp, vp, vp]),
                  (L.cbm_base_at, [vp, vp, vp, vp, i64, vp, vp])):
        f.restype = ctypes.c_int
        f.argtypes = at
    L.cbm_deflate_workspace.restype = i64
    L.cbm_deflate_workspace.argtypes = [i64]
    L.cbm_deflate.restype = ctypes.c_int
    L.cbm_deflate.argtypes = [vp, vp, vp, i64, vp, vp, vp, vp, i64, vp]
    L.cbm_copy_spans.restype = ctypes.c_int
    L.cbm_copy_spans.argtypes = [vp, vp, vp, vp, i64, vp, vp]
    L.cbm_last_error.restype = ctypes.c_int
    L.cbm_last_error.argtypes = [ctypes.c_char_p, i64]
    L.cbm_version.restype = ctypes.c_char_p
    L.cbm_version.argtypes = []
    _lib = L
    return L


def last_error() -> str:
    buf = ctypes.create_string_buffer(512)
    load().cbm_last_error(buf, 512)
    return buf.value.decode(errors="replace")


def check(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc}): {last_error()}")


def ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else t.data_ptr()


def stream_ptr(stream: torch.cuda.Stream | None = None) -> int:
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream
