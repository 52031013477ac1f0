"""CPU ORACLE -- test infrastructure only.

NumPy/ctypes front end of the C restatement in oecl_oracle.c plus the Python restatements in
grid_oracle.py and furman_oracle.py.  Only tests/, __graft_entry__.smoke() and bench.py's
CPU-baseline leg import this package, and only as the checker or the timed baseline.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_D = ctypes.POINTER(ctypes.c_double)
_U32 = ctypes.POINTER(ctypes.c_uint32)
_U16 = ctypes.POINTER(ctypes.c_uint16)
_I32 = ctypes.POINTER(ctypes.c_int32)


def build():
    """Compile liboecl_oracle.so with the Makefile beside this file."""
    subprocess.check_call(['make', '-s', '-C', _HERE])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, 'liboecl_oracle.so')
        if not os.path.exists(path):
            build()
        _LIB = ctypes.CDLL(path)
    return _LIB


def _p(a, t):
    return a.ctypes.data_as(t)


def _aos(pd):
    pd = np.ascontiguousarray(pd, dtype=np.float64)
    assert pd.ndim == 2 and pd.shape[1] == 6
    return pd


def calc_grid_weights(pd, dx, dy, lxExt, lyExt, nxExt):
    """particles.pyx:52-56 -> (inCell, wLL, wUL, wLU, wUU)."""
    pd = _aos(pd)
    n = pd.shape[0]
    ic = np.empty(n, np.uint32)
    w = np.empty((4, n))
    lib().orc_calc_grid_weights(_p(pd, _D), ctypes.c_uint32(n), ctypes.c_double(1. / dx),
                                ctypes.c_double(1. / dy), ctypes.c_double(lxExt * 0.5),
                                ctypes.c_double(lyExt * 0.5), ctypes.c_uint32(nxExt),
                                _p(ic, _U32), _p(w[0], _D), _p(w[1], _D), _p(w[2], _D),
                                _p(w[3], _D))
    return ic, w


def charge_to_grid(pd, ic, w, q, nxExt, npExt):
    pd = _aos(pd)
    rho = np.empty(npExt)
    lib().orc_charge_to_grid(_p(pd, _D), _p(ic, _U32), _p(w[0], _D), _p(w[1], _D), _p(w[2], _D),
                             _p(w[3], _D), _p(rho, _D), ctypes.c_double(q),
                             ctypes.c_uint32(nxExt), ctypes.c_uint32(npExt),
                             ctypes.c_uint32(pd.shape[0]))
    return rho


def field_to_particles(field, ic, w, nxExt, npExt, dims=2):
    field = np.ascontiguousarray(field, dtype=np.float64)
    n = ic.shape[0]
    out = np.empty((n, dims))
    fn = lib().orc_field_to_particles if dims == 2 else lib().orc_field_to_particles3
    fn(_p(field, _D), _p(out, _D), _p(ic, _U32), _p(w[0], _D), _p(w[1], _D), _p(w[2], _D),
       _p(w[3], _D), ctypes.c_uint32(nxExt), ctypes.c_uint32(npExt), ctypes.c_uint32(n))
    return out


def boris_push(pd, e_at, dt, qm, typeB, b=None):
    """particles.pyx:79-92; returns the pushed copy of pd.  b: (3,) for typeB 1, (N,3) for 2."""
    pd = _aos(pd).copy()
    e_at = np.ascontiguousarray(e_at, dtype=np.float64)
    n = pd.shape[0]
    if typeB == 0:
        lib().orc_boris_push_no_b(_p(pd, _D), _p(e_at, _D), ctypes.c_double(dt),
                                  ctypes.c_double(qm * dt), ctypes.c_uint32(n))
    elif typeB == 1:
        b = np.ascontiguousarray(b, dtype=np.float64)
        lib().orc_boris_push_const_b(_p(pd, _D), _p(e_at, _D), _p(b, _D), ctypes.c_double(dt),
                                     ctypes.c_double(0.5 * qm * dt), ctypes.c_uint32(n))
    else:
        b = np.ascontiguousarray(b, dtype=np.float64)
        lib().orc_boris_push(_p(pd, _D), _p(e_at, _D), _p(b, _D), ctypes.c_double(dt),
                             ctypes.c_double(0.5 * qm * dt), ctypes.c_uint32(n))
    return pd


def index_inside(pd, kind, lx, ly, grid=None):
    """particleBoundary.pyx indexInside; kind 'rect' | 'ellip' | 'cutcell'.  For 'cutcell' the
    (clipped) particle data is returned as well."""
    pd = _aos(pd).copy()
    n = pd.shape[0]
    out = np.empty(n, np.uint16)
    if kind == 'rect':
        lib().orc_inside_rect(_p(pd, _D), ctypes.c_uint32(n), ctypes.c_double(lx),
                              ctypes.c_double(ly), _p(out, _U16))
    elif kind == 'ellip':
        lib().orc_inside_ellip(_p(pd, _D), ctypes.c_uint32(n), ctypes.c_double(lx),
                               ctypes.c_double(ly), _p(out, _U16))
    else:
        g = grid
        f = np.ascontiguousarray(g.insideFaces, np.uint16)
        bp = np.ascontiguousarray(g.boundaryPoints, np.float64)
        cpi = np.ascontiguousarray(g.cutCellPointsInd, np.int32)
        lib().orc_inside_cutcell(_p(pd, _D), ctypes.c_uint32(n), ctypes.c_double(0.5 * g.lxExt),
                                 ctypes.c_double(0.5 * g.lyExt), ctypes.c_double(g.dx),
                                 ctypes.c_double(g.dy), ctypes.c_uint32(g.nxExt), _p(f, _U16),
                                 _p(bp, _D), _p(cpi, _I32), _p(out, _U16))
    return out, pd


def save_absorbed(pd, inside):
    pd = _aos(pd)
    inside = np.ascontiguousarray(inside, np.uint16)
    n = pd.shape[0]
    out = np.empty((max(1, n), 6))
    na = lib().orc_save_absorbed(_p(pd, _D), ctypes.c_uint32(n), _p(inside, _U16), _p(out, _D))
    return out[:na].copy()


lib_restype_set = False


def _setup_restypes():
    global lib_restype_set
    if not lib_restype_set:
        lib().orc_save_absorbed.restype = ctypes.c_uint32
        lib().orc_interaction_cutcell.restype = ctypes.c_uint32
        lib().orc_inside_cutcell.restype = ctypes.c_int
        lib_restype_set = True


def interaction_and_normals(ap, kind, lx, ly, dx, dy, grid=None, dt=None, cornerFix=False):
    """particleBoundary.pyx calculateInteractionPoint + calculateNormalVectors.
    Returns (moved absorbed copy, remainingTimeStep, normals, notFound).  cornerFix: the
    rectangle's corner cases take the normal of the wall crossed first."""
    _setup_restypes()
    ap = _aos(ap).copy()
    na = ap.shape[0]
    rts = np.zeros(max(na, 1))
    nv = np.zeros((max(na, 1), 2))
    nf = 0
    if kind == 'rect':
        lw = np.zeros(max(na, 1), np.uint16)
        lib().orc_interaction_rect(_p(ap, _D), ctypes.c_uint32(na), ctypes.c_double(lx),
                                   ctypes.c_double(ly), ctypes.c_double(dx), ctypes.c_double(dy),
                                   _p(rts, _D), _p(lw, _U16), ctypes.c_int(1 if cornerFix else 0))
        lib().orc_normals_rect(ctypes.c_uint32(na), _p(lw, _U16), _p(nv, _D))
    elif kind == 'ellip':
        lib().orc_interaction_ellip(_p(ap, _D), ctypes.c_uint32(na), ctypes.c_double(lx),
                                    ctypes.c_double(ly), ctypes.c_double(dx), ctypes.c_double(dy),
                                    _p(rts, _D))
        lib().orc_normals_ellip(_p(ap, _D), ctypes.c_uint32(na), ctypes.c_double(lx),
                                ctypes.c_double(ly), _p(nv, _D))
    else:
        g = grid
        f = np.ascontiguousarray(g.insideFaces, np.uint16)
        bp = np.ascontiguousarray(g.boundaryPoints, np.float64)
        cpi = np.ascontiguousarray(g.cutCellPointsInd, np.int32)
        ccn = np.ascontiguousarray(g.cutCellNormalVectors, np.float64)
        nf = lib().orc_interaction_cutcell(
            _p(ap, _D), ctypes.c_uint32(na), ctypes.c_double(0.5 * g.lxExt),
            ctypes.c_double(0.5 * g.lyExt), ctypes.c_double(g.dx), ctypes.c_double(g.dy),
            ctypes.c_uint32(g.nxExt), ctypes.c_double(dt), _p(f, _U16), _p(bp, _D), _p(cpi, _I32),
            _p(rts, _D))
        lib().orc_normals_cutcell(_p(ap, _D), ctypes.c_uint32(na), ctypes.c_double(0.5 * g.lxExt),
                                  ctypes.c_double(0.5 * g.lyExt), ctypes.c_double(g.dx),
                                  ctypes.c_double(g.dy), ctypes.c_uint32(g.nxExt), _p(ccn, _D),
                                  _p(nv, _D))
    return ap, rts[:na], nv[:na], nf


def add_and_remove(pd, add, keep, capacity):
    """particles.pyx:94-121 with the reference's capacity rule.  capacity = rows of the
    reference's particleData array.  Returns (new particle data, new capacity)."""
    pd = _aos(pd)
    add = np.ascontiguousarray(add, np.float64).reshape(-1, 6)
    keep = np.ascontiguousarray(keep, np.uint16)
    old = pd.shape[0]
    na = add.shape[0]
    kc = int(keep.sum())
    new = kc + na
    if na == 0 and kc == old:
        return pd.copy(), capacity
    if na == 0:
        full = pd.copy()
        lib().orc_remove(_p(full, _D), _p(keep, _U16), ctypes.c_uint32(old), ctypes.c_uint32(new))
        return full[:new].copy(), capacity
    if new > capacity:
        cap2 = int(new * 1.1)
        out = np.empty((cap2, 6))
        lib().orc_add_remove_new_array(_p(pd, _D), _p(add, _D), _p(keep, _U16),
                                       ctypes.c_uint32(na), ctypes.c_uint32(kc), _p(out, _D))
        return out[:new].copy(), cap2
    full = np.empty((capacity, 6))
    full[:old] = pd
    lib().orc_add_remove(_p(full, _D), _p(add, _D), _p(keep, _U16), ctypes.c_uint32(kc),
                         ctypes.c_uint32(na), ctypes.c_uint32(old))
    return full[:new].copy(), capacity


def add_and_remove_fast(pd, add, keep):
    """The hole-filling rule of the product's fast ordering: additions fill the holes in index
    order, extra additions are appended, and holes left below the new count are filled with
    the kept particles taken from the end (the rule of particles.pyx:254-269)."""
    pd = _aos(pd)
    add = np.ascontiguousarray(add, np.float64).reshape(-1, 6)
    keep = np.asarray(keep).astype(bool)
    n, a = pd.shape[0], add.shape[0]
    holes = np.nonzero(~keep)[0]
    h = holes.size
    out = pd.copy()
    m = min(a, h)
    out[holes[:m]] = add[:m]
    if a >= h:
        return np.vstack([out, add[h:]])
    new = n - h + a
    rest = holes[a:]
    rest = rest[rest < new]
    tail = np.nonzero(keep[new:])[0][::-1] + new
    out[rest] = pd[tail]
    return out[:new].copy()


def philox_words(n, c0, c1, c2, c3, k0, k1):
    """n Philox4x32-10 blocks for counters (c0+i, c1, c2, c3), key (k0, k1) -> uint32[n, 4]."""
    out = np.empty((n, 4), np.uint32)
    lib().orc_philox_stream(ctypes.c_uint32(n), ctypes.c_uint32(c0), ctypes.c_uint32(c1),
                            ctypes.c_uint32(c2), ctypes.c_uint32(c3), ctypes.c_uint32(k0),
                            ctypes.c_uint32(k1), _p(out, _U32))
    return out


def words_to_uniforms(w):
    """Two uniforms in [0, 1) per block: 53 high bits of (w1:w0) and of (w3:w2)."""
    w = w.astype(np.uint64)
    a = ((w[:, 1] << np.uint64(32)) | w[:, 0]) >> np.uint64(11)
    b = ((w[:, 3] << np.uint64(32)) | w[:, 2]) >> np.uint64(11)
    return np.stack([a, b], axis=1).astype(np.float64) * 2.0 ** -53
