"""CPU checks of the C ABI: libgbk.so loads and exports every gbk_* symbol
that include/gbk.h declares (no device calls)."""
import ctypes
import os

from graphbrew_amd import gbk


def test_library_exports_header_symbols():
    assert os.path.exists(gbk.LIB_PATH), "run __graft_entry__.build() first"
    lib = ctypes.CDLL(gbk.LIB_PATH)
    names = gbk.exported_symbols()
    assert "gbk_relabel" in names and "gbk_pr" in names and len(names) >= 17
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_declares_every_symbol():
    L = gbk.lib()
    for n in gbk.exported_symbols():
        assert getattr(L, n).restype is not None or n == "gbk_ctx_destroy"
