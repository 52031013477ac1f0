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
    # the direct solver's box width: the least 2^p - 1 holding the nxExt - 4 interior columns
    lib.oecl_fastpoisson_box_width.restype = ctypes.c_uint32
    widths = {nx: lib.oecl_fastpoisson_box_width(ctypes.c_uint32(nx)) for nx in (4, 5, 6, 8, 11, 12, 259, 260, 2048, 2051)}
    assert widths == {4: 0, 5: 1, 6: 3, 8: 7, 11: 7, 12: 15, 259: 255, 260: 511, 2048: 2047, 2051: 2047}


@pytest.mark.skipif(not os.path.exists(LIB), reason='liboecl.so not built')
def test_invalid_arguments_are_refused_before_any_device_call():
    """Entry points check their arguments first and return OECL_ERR_ARG (1) without touching
    the device, so these calls are safe on a machine without a GPU."""
    lib = ctypes.CDLL(LIB)
    ARG = 1
    vp = ctypes.c_void_p
    assert lib.oecl_step_run(None, None, None, None, None, None) == ARG
    assert lib.oecl_step_run_batch(None, None, None, None, None, None) == ARG
    h = vp()
    assert lib.oecl_fastpoisson_create(ctypes.c_uint32(3), ctypes.c_uint32(3), ctypes.c_double(1.), ctypes.c_double(1.),
                                       ctypes.c_uint32(0), None, None, None, None, None, ctypes.byref(h), None) == ARG
    assert lib.oecl_unet_create(ctypes.c_uint32(7), ctypes.c_uint32(8), ctypes.c_uint32(0), ctypes.byref(h)) == ARG
    assert lib.oecl_unet_create(ctypes.c_uint32(8), ctypes.c_uint32(12), ctypes.c_uint32(0), ctypes.byref(h)) == ARG
    assert lib.oecl_unet_set_conv_rows(None, 0) == ARG
    assert lib.oecl_secem_pass1(None, 0, None, None, ctypes.c_uint32(4), None, 0, 0, 0, None, None, None) == ARG
    assert lib.oecl_furman_pass1(None, 0, None, None, ctypes.c_uint32(4), None, 0, 0, 0, None, None, None) == ARG
    assert lib.oecl_photo_emit(None, 0, None, ctypes.c_uint32(4), 0, 0, 0, ctypes.c_double(7.), ctypes.c_double(5.),
                               ctypes.c_double(1.), ctypes.c_double(9.1e-31), None) == ARG
    assert lib.oecl_sort_runs_by_cell(None, 0, None, None, ctypes.c_uint64(256), ctypes.c_uint32(1), None, None,
                                      None, 0, 0, None, None, None, None) == ARG
    assert lib.oecl_fcg_solve(None, ctypes.c_uint32(8), ctypes.c_uint32(8), None, None, None, None, None,
                              ctypes.c_double(1e-8), ctypes.c_uint32(10), ctypes.c_uint32(1), ctypes.c_double(0.),
                              ctypes.c_uint32(4), None, None, None, None) == ARG


@pytest.mark.skipif(not os.path.exists(LIB), reason='liboecl.so not built')
def test_noalloc_size_limit_without_gpu():
    """The no-allocation saveAbsorbed refuses sizes beyond its scratch before any device call."""
    lib = ctypes.CDLL(LIB)
    lib.oecl_save_absorbed_noalloc_max.restype = ctypes.c_uint32
    lib.oecl_last_error.restype = ctypes.c_char_p
    mx = lib.oecl_save_absorbed_noalloc_max()
    assert mx == (128 * 1024 // 8 - 5) * 4096
    one = ctypes.c_void_p(16)   # never dereferenced: the size check comes first
    assert lib.oecl_save_absorbed_noalloc(0, one, ctypes.c_uint64(mx + 1), ctypes.c_uint32(mx + 1), one, None,
                                          ctypes.c_uint32(0), one, None) == 3
    assert b'scratch limit' in lib.oecl_last_error()
    assert lib.oecl_save_absorbed_noalloc(0, one, ctypes.c_uint64(8), ctypes.c_uint32(8), None, None,
                                          ctypes.c_uint32(0), one, None) == 1


def test_push_block_size_matches_the_scan_layout():
    """Batched run segments and the per-block run table use the library's push block size; the
    scan's host layout must use the same."""
    from openecloud_amd import _lib as L
    from openecloud_amd.scan import PB
    lib = L.lib()
    lib.oecl_push_block_size.restype = ctypes.c_uint32
    assert lib.oecl_push_block_size() == PB


def _c_offsets(struct, fields, tmp_path):
    """sizeof and offsetof(field) of a struct of include/oecl.h, from a C program built by gcc."""
    src = tmp_path / 'off.c'
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "oecl.h"', 'int main(void) {',
             '    printf("%%zu\\n", sizeof(%s));' % struct]
    lines += ['    printf("%%zu\\n", offsetof(%s, %s));' % (struct, f) for f in fields]
    lines += ['    return 0;', '}']
    src.write_text('\n'.join(lines) + '\n')
    exe = tmp_path / 'off'
    subprocess.check_call(['gcc', '-I', os.path.join(ROOT, 'include'), '-o', str(exe), str(src)])
    return [int(v) for v in subprocess.check_output([str(exe)]).split()]


def test_step_args_mirror_matches_the_header(tmp_path):
    """The ctypes mirror of oecl_step_args has the C struct's size and field offsets."""
    from openecloud_amd.buildup import StepArgs
    names = [f[0] for f in StepArgs._fields_]
    got = _c_offsets('oecl_step_args', names, tmp_path)
    assert got[0] == ctypes.sizeof(StepArgs)
    assert got[1:] == [getattr(StepArgs, n).offset for n in names]


def test_push_arithmetic_choices():
    """'mixed' is the default for FP32 particles and refused for FP64 ones."""
    import torch
    from openecloud_amd.buildup import push_arithmetic
    with pytest.raises(ValueError):
        push_arithmetic(torch.float64, 'mixed')
    with pytest.raises(ValueError):
        push_arithmetic(torch.float32, 'fp16')
    assert push_arithmetic(torch.float32, None) == 'mixed'
    assert push_arithmetic(torch.float32, 'fp64') == 'fp64'
    assert push_arithmetic(torch.float64, None) == 'fp64'
