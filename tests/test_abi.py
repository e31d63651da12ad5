"""The C-ABI library loads and exports every symbol include/maos_tomo.h declares."""
import ctypes
import os
import re

from maos_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_header_symbols_exported():
    hdr = open(os.path.join(ROOT, "include", "maos_tomo.h")).read()
    decl = set(re.findall(r"^\s*(?:const\s+\w+\s*\*\s*|int\s+|void\s+)(\w+)\s*\(", hdr, re.M))
    assert decl == set(_abi.EXPORTS)
    lib = ctypes.CDLL(_abi.LIB_PATH)
    for name in decl:
        assert hasattr(lib, name), name


def test_setup_errors_are_codes():
    import ctypes as C
    L = _abi.lib()
    h = C.c_void_p()
    assert L.mt_setup_create(None, C.byref(h)) == _abi.MT_EINVAL
    assert L.gpu_tomo_setup(None, 0, 1, 30, 0, C.byref(h)) < 0
