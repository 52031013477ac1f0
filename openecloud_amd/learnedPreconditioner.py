"""Learned preconditioner for the Poisson solve (no counterpart in the reference, whose solve is
a SuperLU back-substitution, poissonSolver.pyx:59,85).

UNetPreconditioner wraps the device U-Net of oecl_unet_* (4 levels, 16 channels, 3x3
convolutions, bf16 MFMA with FP32 accumulation).  It is trained on the device on pairs
(q, phi = A^-1 q) of the grid's FIT system with the loss 1 - (z^T q)^2 / ((z^T A z)(phi^T q)):
the A-norm error of the best multiple of the preconditioner's output z for q, relative to
phi's A-norm (a preconditioner's scale does not matter to the flexible CG that applies it,
so the loss ignores it too).  It is applied inside a flexible CG that
keeps its vectors and reductions in FP64.  The network sees the residual scaled to unit RMS,
so the preconditioner is homogeneous of degree one.  With omega > 0 (the default) the network
sits between two damped Jacobi smoothings, z1 = omega D^-1 r, and its output for r1 = r - A z1
enters through an A-norm line search: the smoothing handles the high frequencies a network of
3x3 convolutions hardly represents, and the line search keeps the network from ever making a
step worse than the smoothing alone.  It is then trained on the pairs it meets there
(smoothed_pairs).  solve() falls back to the direct solve when the flexible CG does not reach
its tolerance.

Training pairs: operator_pairs() makes (A phi, phi) with phi random at length scales of 0 to 64
cells.  Pairs (q, A^-1 q) for cloud-like charges also work on small grids, but on large ones
their error is mostly a global smooth potential that a 4-level network of 3x3 convolutions
cannot reach, and training on them stalls.
"""
import ctypes

import numpy as np
import torch

from . import _lib as L

NCONV, C = 14, 16


def image_size(gridObj):
    """(H, W): the extended grid rounded up to multiples of 8."""
    ny, nx = gridObj.getNyExt(), gridObj.getNxExt()
    return (ny + 7) // 8 * 8, (nx + 7) // 8 * 8


class UNetPreconditioner:

    def __init__(self, gridObj, seed=0, omega=2. / 3., mDirs=8):
        self.g = gridObj
        self.omega = float(omega)
        self.mDirs = int(mDirs)
        self.nx, self.ny = gridObj.getNxExt(), gridObj.getNyExt()
        self.H, self.W = image_size(gridObj)
        h = ctypes.c_void_p()
        L.check(L.lib().oecl_unet_create(L.u32(self.H), L.u32(self.W), L.u32(seed), ctypes.byref(h)), 'unet create')
        self._h = h
        dev = L.device()
        np_ = self.nx * self.ny
        self._work = torch.empty(2 * np_, dtype=torch.float64, device=dev)
        self._fcgwork = None
        self.history = []
        self.lastIterations = 0
        self.lastRelres = None
        self.fallbacks = 0

    def __del__(self):
        try:
            if self._h is not None and self._h.value:
                L.lib().oecl_unet_destroy(self._h)
        except Exception:
            pass

    def set_conv_rows(self, rows):
        """Output rows per convolution block (1, 2, 4, 8; 0 = chosen per image size)."""
        L.check(L.lib().oecl_unet_set_conv_rows(self._h, L.i32(rows)), 'unet conv rows')

    # ---------------------------------------------------------------- parameters
    def get_params(self):
        """(W [14][9][16][16], b [14][16]) as float32 host arrays."""
        W = np.zeros((NCONV, 9, C, C), np.float32)
        b = np.zeros((NCONV, C), np.float32)
        L.check(L.lib().oecl_unet_get_params(self._h, W.ctypes.data_as(ctypes.c_void_p),
                                             b.ctypes.data_as(ctypes.c_void_p), None, None), 'unet get')
        return W, b

    def set_params(self, W, b):
        W = np.ascontiguousarray(W, np.float32)
        b = np.ascontiguousarray(b, np.float32)
        assert W.shape == (NCONV, 9, C, C) and b.shape == (NCONV, C)
        L.check(L.lib().oecl_unet_set_params(self._h, W.ctypes.data_as(ctypes.c_void_p),
                                             b.ctypes.data_as(ctypes.c_void_p)), 'unet set')

    def get_grads(self):
        gW = np.zeros((NCONV, 9, C, C), np.float32)
        gb = np.zeros((NCONV, C), np.float32)
        L.check(L.lib().oecl_unet_get_params(self._h, None, None, gW.ctypes.data_as(ctypes.c_void_p),
                                             gb.ctypes.data_as(ctypes.c_void_p)), 'unet grads')
        return gW, gb

    def save(self, path):
        W, b = self.get_params()
        np.savez(path, W=W, b=b)

    def load(self, path):
        d = np.load(path)
        self.set_params(d['W'], d['b'])

    # ---------------------------------------------------------------- raw network
    def forward_image(self, img):
        """img: (H, W, 16) bfloat16 device tensor -> (H, W) float32 device tensor."""
        assert img.dtype == torch.bfloat16 and tuple(img.shape) == (self.H, self.W, C)
        y = torch.empty((self.H, self.W), dtype=torch.float32, device=L.device())
        L.check(L.lib().oecl_unet_forward(self._h, L.ptr(img.contiguous()), L.ptr(y), L.stream()), 'unet forward')
        return y

    def backward_image(self, gout, zero=True):
        """After forward_image: gout (H, W, 16) bf16 -> gradient of the input image (H, W, 16)
        bf16; the parameter gradients are read with get_grads()."""
        gin = torch.empty((self.H, self.W, C), dtype=torch.bfloat16, device=L.device())
        L.check(L.lib().oecl_unet_backward(self._h, L.ptr(gout.contiguous()), L.ptr(gin), L.i32(1 if zero else 0),
                                           L.stream()), 'unet backward')
        return gin

    # ---------------------------------------------------------------- on the grid
    def precondition(self, ps, r, omega=None):
        """The preconditioner applied to r on the solver's inside nodes (oecl_unet_precondition)."""
        z = torch.empty_like(r)
        om = self.omega if omega is None else float(omega)
        L.check(L.lib().oecl_unet_precondition(self._h, L.u32(self.nx), L.u32(self.ny), L.ptr(ps.cx), L.ptr(ps.cy),
                                               L.ptr(ps.mask), L.ptr(r), L.ptr(z), L.f64(om), L.stream()),
                'unet precondition')
        return z

    def smoothed_pairs(self, ps, pairs, omega=None):
        """The pairs (r1, e1) the network meets inside the smoothed preconditioner, from pairs
        (q, phi = A^-1 q): r1 = q - A omega D^-1 q, e1 = phi - omega D^-1 q (A e1 = r1)."""
        om = self.omega if omega is None else float(omega)
        out = []
        for q, phi in pairs:
            r1, e1 = torch.empty_like(q), torch.empty_like(q)
            L.check(L.lib().oecl_unet_smooth_pair(self._h, L.u32(self.nx), L.u32(self.ny), L.ptr(ps.cx),
                                                  L.ptr(ps.cy), L.ptr(ps.mask), L.ptr(q), L.ptr(phi), L.f64(om),
                                                  L.ptr(r1), L.ptr(e1), L.stream()), 'unet smooth pair')
            out.append((r1, e1))
        return out

    def operator_pairs(self, ps, count, scales=(0, 1, 2, 4, 8, 16, 32, 64), seed=0):
        """count pairs (q = A phi, phi) with phi white noise smoothed by a Gaussian of standard
        deviation s cells (s cycling through `scales`, 0 = unsmoothed), zero outside the inside
        nodes and of unit norm.  Their errors spread over every length scale up to the largest
        s, which a 4-level network of 3x3 convolutions can represent; right-hand sides with
        global solutions (a cloud's potential) leave most of their error beyond its reach."""
        rng = np.random.default_rng(seed)
        nx, ny = self.nx, self.ny
        mask = ps.mask.cpu().numpy().astype(bool)
        kx2 = np.fft.rfftfreq(nx)[None, :] ** 2
        ky2 = np.fft.fftfreq(ny)[:, None] ** 2
        out = []
        for k in range(int(count)):
            s = float(scales[k % len(scales)])
            f = rng.standard_normal((ny, nx))
            if s > 0:
                f = np.fft.irfft2(np.fft.rfft2(f) * np.exp(-2. * (np.pi * s) ** 2 * (kx2 + ky2)), s=(ny, nx))
            f = np.where(mask, f.ravel(), 0.)
            f /= max(np.linalg.norm(f), 1e-300)
            phi = torch.as_tensor(f).to(L.device())
            out.append((ps.applyA(phi), phi))
        return out

    def train(self, ps, pairs, steps, lr=1e-3, batch=1, lrEnd=None, seed=0, log=None, logEvery=0,
              validate=None, validateEvery=0, validateIter=4000):
        """Adam on samples (q, phi) drawn from `pairs` (device FP64 grid vectors, phi = A^-1 q),
        `batch` samples per update; the learning rate decays geometrically to lrEnd.  Returns
        the list of (step, mean loss over the last logEvery steps) when logEvery > 0.

        validate: right-hand sides (device FP64 grid vectors) kept out of training.  Every
        validateEvery steps and at the end the flexible CG solves them to 1e-8 (at most
        validateIter iterations each, which a solve that does not converge counts), and the
        parameters with the fewest total iterations are kept: single-sample losses are too noisy
        to tell a good preconditioner from a poor one, the iterations are what it is for.  The
        (step, iterations) record is in self.validation."""
        rng = np.random.default_rng(seed)
        lib = L.lib()
        lrEnd = lr if lrEnd is None else lrEnd
        out, acc, cnt = [], 0., 0
        self.validation = []
        best = None

        def check(step):
            nonlocal best
            its = 0
            for q in validate:
                _, it, _, ok = self.fcg(ps, q, tol=1e-8, maxIter=validateIter)
                its += it if ok else validateIter
            self.validation.append((step, its))
            if log:
                log('train step %d: validation %d iterations' % (step, its))
            if best is None or its < best[0]:
                best = (its, self.get_params())
        for k in range(int(steps)):
            lk = lr * (lrEnd / lr) ** (k / max(1, steps - 1))
            for j in range(batch):
                q, phi = pairs[int(rng.integers(len(pairs)))]
                want = logEvery and (k % logEvery == logEvery - 1 or k == steps - 1) and j == 0
                loss = ctypes.c_double(0.)
                L.check(lib.oecl_unet_train_sample(
                    self._h, L.u32(self.nx), L.u32(self.ny), L.ptr(ps.cx), L.ptr(ps.cy), L.ptr(ps.mask), L.ptr(q),
                    L.ptr(phi), L.ptr(self._work), L.i32(1 if j == 0 else 0),
                    ctypes.byref(loss) if want else None, L.stream()), 'unet train')
                if want:
                    acc += loss.value
                    cnt += 1
            L.check(lib.oecl_unet_adam(self._h, ctypes.c_float(lk), ctypes.c_float(0.9), ctypes.c_float(0.999),
                                       ctypes.c_float(1e-8), L.stream()), 'unet adam')
            if logEvery and (k % logEvery == logEvery - 1 or k == steps - 1) and cnt:
                out.append((k + 1, acc / cnt))
                if log:
                    log('train step %d: loss %.4g' % (k + 1, acc / cnt))
                acc, cnt = 0., 0
            if validate and validateEvery and (k % validateEvery == validateEvery - 1) and k != steps - 1:
                check(k + 1)
        if validate:
            check(int(steps))
            if best[1] is not None:
                self.set_params(*best[1])
        self.history.extend(out)
        return out

    def loss(self, ps, q, phi):
        """The training loss of one pair (does not change the parameters' values)."""
        v = ctypes.c_double(0.)
        L.check(L.lib().oecl_unet_train_sample(
            self._h, L.u32(self.nx), L.u32(self.ny), L.ptr(ps.cx), L.ptr(ps.cy), L.ptr(ps.mask), L.ptr(q), L.ptr(phi),
            L.ptr(self._work), L.i32(1), ctypes.byref(v), L.stream()), 'unet loss')
        return v.value

    def fcg(self, ps, q, phi=None, tol=1e-8, maxIter=2000, check=4, omega=None, mDirs=None):
        """Flexible CG with this preconditioner, keeping mDirs directions (default self.mDirs)
        -> (phi, iterations, relative residual, converged)."""
        if phi is None:
            phi = torch.zeros_like(q)
        it = ctypes.c_uint32(0)
        rr = ctypes.c_double(0.)
        om = self.omega if omega is None else float(omega)
        m = self.mDirs if mDirs is None else int(mDirs)
        lib = L.lib()
        need = int(lib.oecl_fcg_work_bytes(self.nx * self.ny, max(1, min(m, 8)))) // 8 + 8
        if self._fcgwork is None or self._fcgwork.numel() < need:
            self._fcgwork = torch.empty(need, dtype=torch.float64, device=L.device())
        rc = lib.oecl_fcg_solve(self._h, L.u32(self.nx), L.u32(self.ny), L.ptr(ps.cx), L.ptr(ps.cy),
                                L.ptr(ps.mask), L.ptr(q), L.ptr(phi), L.f64(tol), L.u32(maxIter), L.u32(check),
                                L.f64(om), L.u32(m), L.ptr(self._fcgwork), ctypes.byref(it), ctypes.byref(rr),
                                L.stream())
        if rc not in (0, 4):
            L.check(rc, 'fcg solve')
        self.lastIterations, self.lastRelres = it.value, rr.value
        return phi, it.value, rr.value, rc == 0

    def reserve(self, R):
        """Sizes the network's inference buffers for batches of R problems (fcg_batch)."""
        L.check(L.lib().oecl_unet_reserve(self._h, L.u32(R)), 'unet reserve')

    def fcg_batch(self, ps, q, phi, R, tol=1e-8, maxIter=2000, check=4, omega=None, mDirs=None):
        """fcg for R problems at once: q and phi hold R grid vectors back to back (phi: initial
        guesses in, solutions out).  Each run's result is bit for bit fcg's on its own problem.
        -> (iterations[R], relative residuals[R]) as numpy arrays; a run converged when its
        relative residual is <= tol."""
        npx = self.nx * self.ny
        if q.numel() < R * npx or phi.numel() < R * npx:
            raise L.OeclError('fcg_batch: q and phi need %d x %d values' % (R, npx))
        om = self.omega if omega is None else float(omega)
        m = self.mDirs if mDirs is None else int(mDirs)
        lib = L.lib()
        self.reserve(R)
        need = int(lib.oecl_fcg_batch_work_bytes(npx, max(1, min(m, 8)), R)) // 8 + 8
        if self._fcgwork is None or self._fcgwork.numel() < need:
            self._fcgwork = torch.empty(need, dtype=torch.float64, device=L.device())
        its = (ctypes.c_uint32 * R)()
        rel = (ctypes.c_double * R)()
        rc = lib.oecl_fcg_solve_batch(self._h, L.u32(self.nx), L.u32(self.ny), L.ptr(ps.cx), L.ptr(ps.cy),
                                      L.ptr(ps.mask), L.u32(R), L.ptr(q), L.ptr(phi), L.f64(tol), L.u32(maxIter),
                                      L.u32(check), L.f64(om), L.u32(m), L.ptr(self._fcgwork), its, rel, L.stream())
        if rc not in (0, 4):
            L.check(rc, 'fcg batch solve')
        return np.frombuffer(its, dtype=np.uint32).copy(), np.frombuffer(rel, dtype=np.float64).copy()

    def solve(self, ps, q, tol=1e-8, maxIter=2000):
        """phi = A^-1 q by the learned flexible CG; the direct solve when it does not converge."""
        phi, it, rr, ok = self.fcg(ps, q, tol=tol, maxIter=maxIter)
        if ok:
            return phi
        self.fallbacks += 1
        if ps._fp is None:
            raise L.OeclError('learned PCG did not converge (relres %.3g after %d iterations)' % (rr, it))
        out = torch.empty_like(phi)
        ps._direct(ps._rhs(q), out)
        return out
