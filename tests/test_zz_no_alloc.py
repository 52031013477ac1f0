"""GPU: oecl_save_absorbed_noalloc (scratch inside the library) against oecl_save_absorbed (a
workspace from the caller) and the oracle's saveAbsorbed (particleBoundary.pyx:38-65), through
the C ABI, on two calls in a row (the second smaller)."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _lib():
    from openecloud_amd import _lib as L
    lib = L.lib()
    lib.oecl_save_absorbed_noalloc_max.restype = ctypes.c_uint32
    return L, lib


def _case(n, cap, dtype, seed):
    rng = np.random.default_rng(seed)
    pd = rng.standard_normal((6, cap))
    ins = (rng.random(cap) > 0.07).astype(np.uint8)
    tdt = torch.float64 if dtype == 0 else torch.float32
    return (torch.as_tensor(pd, dtype=tdt).cuda(), torch.as_tensor(ins).cuda(),
            pd.astype(np.float64 if dtype == 0 else np.float32), ins)


def _save(L, lib, fn, dtype, data, cap, n, ins, absCap, work=None):
    tdt = data.dtype
    out = torch.full((absCap, 6), -7., dtype=tdt, device='cuda')
    cnt = torch.full((1,), 12345, dtype=torch.int32, device='cuda')
    args = [L.i32(dtype), L.ptr(data), L.u64(cap), L.u32(n), L.ptr(ins), L.ptr(out), L.u32(absCap), L.ptr(cnt)]
    if work is not None:
        args.append(L.ptr(work))
    args.append(L.stream())
    L.check(fn(*args), fn.__name__)
    torch.cuda.synchronize()
    return out.cpu().numpy(), int(cnt.item())


@pytest.mark.parametrize('dtype', [0, 1])
def test_noalloc_matches_workspace_entry_point_on_two_calls(dtype, orc):
    L, lib = _lib()
    cap = 400000
    # second call smaller: the scratch keeps the first call's tile counts, which must not leak
    for n, seed in ((cap, 1), (5000, 2)):
        print('noalloc save_absorbed: dtype %d n %d' % (dtype, n), flush=True)
        data, ins, pd, insh = _case(n, cap, dtype, seed)
        absCap = 40000
        work = torch.empty(int(lib.oecl_scan_work_bytes(L.u32(n))) // 4 + 16, dtype=torch.int32, device='cuda')
        ref, cref = _save(L, lib, lib.oecl_save_absorbed, dtype, data, cap, n, ins, absCap, work)
        got, cgot = _save(L, lib, lib.oecl_save_absorbed_noalloc, dtype, data, cap, n, ins, absCap)
        assert cgot == cref
        assert np.array_equal(got, ref)
        # the oracle's saveAbsorbed on the same rows (AoS, index order)
        exp = orc.save_absorbed(pd[:, :n].T.astype(np.float64), insh[:n].astype(np.uint16))
        assert cgot == exp.shape[0]
        assert np.array_equal(got[:cgot].astype(np.float64), exp)


def test_noalloc_rejects_sizes_beyond_its_scratch():
    L, lib = _lib()
    mx = lib.oecl_save_absorbed_noalloc_max()
    dummy = torch.zeros(8, dtype=torch.uint8, device='cuda')
    rc = lib.oecl_save_absorbed_noalloc(L.i32(0), L.ptr(dummy), L.u64(mx + 1), L.u32(mx + 1), L.ptr(dummy),
                                        None, L.u32(0), L.ptr(dummy), L.stream())
    assert rc == 3
    assert b'scratch limit' in lib.oecl_last_error()
