"""The C-ABI library loads and exports every symbol include/cubam_amd.h declares (no GPU calls)."""

import ctypes
import re

from conftest import ROOT


def _declared():
    text = (ROOT / "include" / "cubam_amd.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cbm_\w+)\s*\(", text)))


def test_header_declares_the_binding_symbols():
    from cubam_amd import _lib

    assert _declared() == sorted(_lib.SYMBOLS)


def test_library_exports_every_declared_symbol():
    from cubam_amd import _lib

    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [s for s in _declared() if not hasattr(lib, s)]
    assert not missing, missing


def test_version_and_workspace_queries():
    from cubam_amd import _lib

    L = _lib.load()
    assert L.cbm_version().startswith(b"cubam_amd")
    assert L.cbm_bgzf_scan_workspace(1 << 30) > 0
    seg = ctypes.c_int64()
    assert L.cbm_record_chain_workspace(1 << 30, ctypes.byref(seg)) > 0 and seg.value > 0
    assert L.cbm_scatter_workspace(1000) > 0


def test_product_path_does_not_import_the_oracle():
    pkg = ROOT / "cubam_amd"
    for p in pkg.rglob("*.py"):
        src = p.read_text()
        assert "import oracle" not in src and "from oracle" not in src, p
