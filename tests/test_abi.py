"""CPU: liboecl.so loads and exports every entry point that include/oecl.h declares."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, 'openecloud_amd', 'liboecl.so')


def _declared():
    txt = open(os.path.join(ROOT, 'include', 'oecl.h')).read()
    txt = re.sub(r'/\*.*?\*/', '', txt, flags=re.S)
    return sorted(set(re.findall(r'\b(oecl_[a-z0-9_]+)\s*\(', txt)))


def test_header_declares_entry_points():
    names = _declared()
    assert 'oecl_push_fused' in names and 'oecl_pcg_solve' in names and len(names) >= 20


@pytest.mark.skipif(not os.path.exists(LIB), reason='liboecl.so not built')
def test_library_exports_all_declared_symbols():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(['nm', '-D', '--defined-only', LIB], capture_output=True, text=True).stdout
    exported = set(re.findall(r' T (oecl_[a-z0-9_]+)', out))
    assert set(_declared()) <= exported


@pytest.mark.skipif(not os.path.exists(LIB), reason='liboecl.so not built')
def test_size_queries_without_gpu():
    lib = ctypes.CDLL(LIB)
    lib.oecl_scan_work_bytes.restype = ctypes.c_uint64
    lib.oecl_pcg_work_bytes.restype = ctypes.c_uint64
    assert lib.oecl_scan_work_bytes(ctypes.c_uint32(10 ** 6)) > 0
    assert lib.oecl_pcg_work_bytes(ctypes.c_uint32(8580)) >= 5 * 8 * 8580
