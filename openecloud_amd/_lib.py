"""ctypes binding of liboecl.so and the small helpers every module uses.

The library must be built (openecloud_amd.build) and a ROCm GPU must be present; nothing in
this package falls back to a CPU path.
"""
import ctypes
import os

import numpy as np
import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'liboecl.so')

OECL_F64, OECL_F32 = 0, 1
WALL_RECT, WALL_ELLIP, WALL_CUTCELL = 0, 1, 2
FURMAN_SLOTS = 35


class OeclError(RuntimeError):
    pass


class GridDesc(ctypes.Structure):
    _fields_ = [('nx', ctypes.c_uint32), ('ny', ctypes.c_uint32), ('np', ctypes.c_uint32),
                ('dx', ctypes.c_double), ('dy', ctypes.c_double),
                ('lxHalfExt', ctypes.c_double), ('lyHalfExt', ctypes.c_double)]


class WallDesc(ctypes.Structure):
    _fields_ = [('kind', ctypes.c_int32), ('lx', ctypes.c_double), ('ly', ctypes.c_double),
                ('dx', ctypes.c_double), ('dy', ctypes.c_double),
                ('lxHalfExt', ctypes.c_double), ('lyHalfExt', ctypes.c_double),
                ('nx', ctypes.c_uint32),
                ('insideFaces', ctypes.c_void_p), ('boundaryPoints', ctypes.c_void_p),
                ('cutCellPointsInd', ctypes.c_void_p), ('cutCellNormalVectors', ctypes.c_void_p),
                ('rectCornerFix', ctypes.c_int32)]


class FurmanDesc(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in
                ('p1EInf', 'p1Ehat', 'eEHat', 'w', 'p', 'e1', 'e2', 'sigmaE',
                 'p1RInf', 'eR', 'r', 'q', 'r1', 'r2',
                 'deltaTSHat', 's', 'eHat0', 't1', 't2', 't3', 't4',
                 'theta0max', 'deltaERmax')] + \
               [('epsN', ctypes.c_double * 10), ('pSmallN', ctypes.c_double * 10),
                ('particleMass', ctypes.c_double), ('scaleSigmaE', ctypes.c_int32)]


class SecemDesc(ctypes.Structure):
    _fields_ = [('model', ctypes.c_int32)] + [(n, ctypes.c_double) for n in
                ('p1EInf', 'p1Ehat', 'eEHat', 'w', 'p', 'e1', 'e2',
                 'p1RInf', 'eR', 'r', 'q', 'r1', 'r2',
                 'deltaTSHat', 's', 'eHat0', 't1', 't2', 't3', 't4',
                 'sigmaETS', 'muETS', 'theta0max', 'deltaERmax', 'particleMass')]


SECEM_SLOTS = 44
SECEM_STAINLESS, SECEM_SIMPLE = 0, 1

_lib = None


def lib():
    """Load liboecl.so; raise loudly when it is missing or no GPU is present."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OeclError('liboecl.so is not built (run openecloud_amd.build or '
                            '__graft_entry__.build())')
        l = ctypes.CDLL(LIB_PATH)
        l.oecl_last_error.restype = ctypes.c_char_p
        for name in ('oecl_scan_work_bytes', 'oecl_compact_work_bytes', 'oecl_pcg_work_bytes',
                     'oecl_sort_work_bytes', 'oecl_merge_work_bytes', 'oecl_step_work_bytes',
                     'oecl_step_batch_work_bytes', 'oecl_fcg_work_bytes', 'oecl_charge_ordered_work_bytes',
                     'oecl_fcg_batch_work_bytes'):
            getattr(l, name).restype = ctypes.c_uint64
        for name in ('oecl_scan_work_bytes', 'oecl_pcg_work_bytes', 'oecl_sort_work_bytes', 'oecl_merge_work_bytes'):
            getattr(l, name).argtypes = [ctypes.c_uint32]
        l.oecl_fcg_work_bytes.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
        l.oecl_fcg_batch_work_bytes.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
        l.oecl_fastpoisson_box_width.restype = ctypes.c_uint32
        l.oecl_fastpoisson_box_width.argtypes = [ctypes.c_uint32]
        l.oecl_compact_work_bytes.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
        l.oecl_charge_ordered_work_bytes.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
        l.oecl_step_work_bytes.argtypes = [ctypes.c_uint64, ctypes.c_uint32]
        l.oecl_step_batch_work_bytes.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32]
        _lib = l
    return _lib


def require_gpu():
    if not torch.cuda.is_available():
        raise OeclError('openecloud_amd needs a ROCm GPU; none is visible')


def check(rc, what=''):
    if rc != 0:
        msg = lib().oecl_last_error().decode()
        raise OeclError('%s failed with code %d %s' % (what, rc, msg))


def ptr(t):
    """Device pointer of a torch tensor (None -> NULL)."""
    if t is None:
        return ctypes.c_void_p(0)
    return ctypes.c_void_p(t.data_ptr())


def stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def u32(v):
    return ctypes.c_uint32(int(v))


def u64(v):
    return ctypes.c_uint64(int(v))


def f64(v):
    return ctypes.c_double(float(v))


def i32(v):
    return ctypes.c_int32(int(v))


def dtype_code(dt):
    if dt == torch.float64:
        return OECL_F64
    if dt == torch.float32:
        return OECL_F32
    raise ValueError('particle dtype must be float64 or float32')


def grid_desc(g):
    return GridDesc(g.getNxExt(), g.getNyExt(), g.getNpExt(), g.getDx(), g.getDy(),
                    0.5 * g.getLxExt(), 0.5 * g.getLyExt())


def device():
    return torch.device('cuda', torch.cuda.current_device())


def to_dev(a, dtype=None):
    t = torch.as_tensor(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.to(device())


# ---- host arrays: reference drivers index the getters' results with numpy (numpy.sum,
# numpy.asarray, numpy.save on getParticleData() and getChargeOnGrid()).  With host arrays on, the
# array getters and the array-returning methods return NumPy copies of the device data instead
# of device tensors (a synchronising copy each call); their inputs take NumPy arrays either way.
_HOST_ARRAYS = False


def set_host_arrays(on=True):
    """Array getters return NumPy copies (on) or aliasing device tensors (off, the default)."""
    global _HOST_ARRAYS
    _HOST_ARRAYS = bool(on)


def host_arrays():
    return _HOST_ARRAYS


def host_out(fn):
    """Decorator of a public array-returning method: a device tensor result becomes a NumPy copy
    while host arrays are on."""
    import functools

    @functools.wraps(fn)
    def wrapped(*args, **kwargs):
        r = fn(*args, **kwargs)
        if _HOST_ARRAYS and isinstance(r, torch.Tensor):
            return r.detach().cpu().numpy()
        return r
    return wrapped

