"""The C-ABI library: it loads without a GPU, exports every function that
include/atsc3_fec.h declares, the ctypes layer binds exactly those, and the
pure size queries answer (no compute call is made)."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "atsc3_fec.h")


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\bint32_t\s+(\w+)\s*\(", txt)))


def _lib():
    from atsc3_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    return _lib


def test_header_declarations_parsed():
    names = declared()
    assert "ldpc_minsum_f32" in names and "ldpc_kernel_abi" in names
    assert len(names) >= 27


def test_library_exports_every_declared_symbol():
    _l = _lib()
    L = _l.lib()
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing


def test_ctypes_layer_binds_the_header():
    _l = _lib()
    assert sorted(_l.exported_symbols()) == declared()


def test_size_queries():
    L = _lib().lib()
    assert L.ldpc_kernel_abi() == 1
    assert L.atsc3_minsum_qc_ctl_words(10) == 42
    assert L.atsc3_demap_ws_doubles(0, 0) >= 0
    assert L.atsc3_ldm_ws_doubles(2, 32400) == 3 * 2 * ((32400 + 255) // 256)
    assert L.atsc3_minsum_batch_ctl_words(100, 10) > 13 * 10 + 100
    assert L.atsc3_minsum_batch_ctl_words(-1, 10) < 0
    assert L.atsc3_soft_demap_bwd_grid(3, 256) == 3
    assert L.atsc3_minsum_batch_status_word() >= 0
