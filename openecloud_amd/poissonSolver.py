"""PoissonSolver (poissonSolver.pyx:20-99) with the solve on the GPU.

Setup (host, once) builds from the grid the same operators as the reference:
  - the FIT operator A = S~ M_eps S~^T on inside nodes, kept as two edge-coefficient arrays
    (cx on x edges, cy on y edges) of the 5-point stencil;
  - phiToEAtGridPoints = edgeToNode . diag(dsi) . S~^T, assembled with scipy.sparse exactly
    as the reference does and moved to the device as CSR.
solve(q) solves A phi = q on the device and then applies the CSR phi -> E product there.
Two device methods:
  'direct' (default)  capacitance-matrix direct solve (DST-I along x by FP64 MFMA products
                      and tridiagonal solves along y for the box of inside nodes, plus a
                      dense correction of the rows next to the wall),
                      the counterpart of the reference's one-off LU factorisation and
                      back-substitution; its residual is checked against tol every
                      `checkEvery` solves and, if too large, refined by CG.
  'pcg'               Jacobi-preconditioned CG to tol * ||q_inside||, warm-started from the
                      previous potential.
  'learned'           flexible CG preconditioned by a trained UNetPreconditioner (passed as
                      `preconditioner` or set with setPreconditioner), to tol * ||q_inside||,
                      warm-started from the previous potential.  The direct solve is its
                      fallback (when the CG does not converge within maxIter) and its judge:
                      every `checkEvery` solves it solves the same q and the relative
                      difference is recorded (judgeMaxErr, judgeCount).
"""
import ctypes

import numpy as np
import scipy.sparse as spsp
import torch

from . import _lib as L
from .constants import epsilon_0


def capacitance_rows(nx, ny, mask, cx, cy, bx=None):
    """Rows of (M - R) for the capacitance method (see csrc/fastpoisson.hip).

    The box holds the nodes 2 <= i < 2 + bx, 2 <= j < ny - 2 (bx >= nx - 4, default nx - 4;
    columns from nx - 2 on hold no inside node).  Returns (cxu, cyu, sIdx, wPtr, wCol, wVal)
    with box indices t = (i-2) + (j-2)*bx."""
    bx = nx - 4 if bx is None else int(bx)
    if bx < nx - 4:
        raise ValueError('the box must hold every interior column')
    by = ny - 4
    m2 = mask.reshape(ny, nx).astype(bool)
    cx2 = cx.reshape(ny, nx)
    cy2 = cy.reshape(ny, nx)
    # coefficients of edges joining two inside nodes are the uniform ones
    ex = m2[:, :-1] & m2[:, 1:]
    ey = m2[:-1, :] & m2[1:, :]
    vx = np.unique(cx2[:, :-1][ex])
    vy = np.unique(cy2[:-1, :][ey])
    if vx.size != 1 or vy.size != 1:
        raise ValueError('inside-inside edges do not share one coefficient per direction')
    cxu, cyu = float(vx[0]), float(vy[0])
    # a neighbour outside the box is a Dirichlet zero for R as well, so it only needs the
    # uniform edge coefficient; a neighbour inside the box must be an inside node
    okn = m2.copy()
    okn[:2, :] = True
    okn[ny - 2:, :] = True
    okn[:, :2] = True
    okn[:, 2 + bx:] = True
    reg = np.zeros_like(m2)
    c = (slice(1, ny - 1), slice(1, nx - 1))
    reg[c] = (m2[c] & okn[1:-1, 2:] & okn[1:-1, :-2] & okn[2:, 1:-1] & okn[:-2, 1:-1] &
              (cx2[c] == cxu) & (cx2[1:-1, :-2] == cxu) & (cy2[c] == cyu) & (cy2[:-2, 1:-1] == cyu))
    J, I = np.nonzero(m2 & ~reg)
    s_idx, ptr, cols, vals = [], [0], [], []
    for j, i in zip(J, I):
        nb = ((i + 1, j, cx2[j, i], cxu), (i - 1, j, cx2[j, i - 1], cxu),
              (i, j + 1, cy2[j, i], cyu), (i, j - 1, cy2[j - 1, i], cyu))
        d = cx2[j, i] + cx2[j, i - 1] + cy2[j, i] + cy2[j - 1, i]
        t = (i - 2) + (j - 2) * bx
        s_idx.append(t)
        row = [(t, d - (2 * cxu + 2 * cyu))]
        for ii, jj, c, u in nb:
            if 2 <= ii < 2 + bx and 2 <= jj < ny - 2:
                diff = (-c if m2[jj, ii] else 0.) + u
                if diff != 0.:
                    row.append(((ii - 2) + (jj - 2) * bx, diff))
        row.sort()
        cols += [r[0] for r in row]
        vals += [r[1] for r in row]
        ptr.append(len(cols))
    return (cxu, cyu, np.asarray(s_idx, np.int32), np.asarray(ptr, np.int32),
            np.asarray(cols, np.int32), np.asarray(vals, np.float64))


def e2n_structure(csr, npx, nx):
    """Row classes of phiToEAtGridPoints (CSR, rows [Ex | Ey], columns sorted) for the step's
    E kernel: an interior row is exactly (c phi[node - o], -c phi[node + o]) with o = 1 (Ex) or
    nx (Ey) and one c per component; other non-empty rows are listed and summed from the CSR.

    Returns (kind [npx] uint8: bits 0-1 the Ex row, bits 2-3 the Ey row, 0 empty, 1 interior
    pattern, 2 listed; listed rows [int32]; cx; cy)."""
    ptr, ind, val = csr.indptr, csr.indices, csr.data
    rows = np.arange(2 * npx)
    nnz = np.diff(ptr)
    node = rows % npx
    off = np.where(rows < npx, 1, nx)
    two = nnz == 2
    k0 = ptr[:-1]
    c0 = np.where(two, ind[np.minimum(k0, ind.size - 1)], -1)
    c1 = np.where(two, ind[np.minimum(k0 + 1, ind.size - 1)], -1)
    v0 = np.where(two, val[np.minimum(k0, val.size - 1)], 0.)
    v1 = np.where(two, val[np.minimum(k0 + 1, val.size - 1)], 0.)
    shape = two & (c0 == node - off) & (c1 == node + off) & (v0 == -v1) & (v0 != 0.)
    cs = []
    for comp in (0, 1):
        sel = shape & ((rows < npx) if comp == 0 else (rows >= npx))
        if sel.any():
            vals, cnt = np.unique(v0[sel], return_counts=True)
            cs.append(float(vals[np.argmax(cnt)]))
        else:
            cs.append(0.)
    cuni = np.where(rows < npx, cs[0], cs[1])
    reg = shape & (v0 == cuni)
    kind = np.where(nnz == 0, 0, np.where(reg, 1, 2)).astype(np.uint8)
    packed = (kind[:npx] | (kind[npx:] << 2)).astype(np.uint8)
    listed = np.nonzero(kind == 2)[0].astype(np.int32)
    return packed, listed, cs[0], cs[1]


class PoissonSolver:

    def __init__(self, gridObj, tol=1e-12, maxIter=200000, method='direct', checkEvery=1, preconditioner=None):
        L.require_gpu()
        nx = gridObj.getNxExt()
        npx = gridObj.getNpExt()
        self.nx, self.ny, self.npExt = nx, gridObj.getNyExt(), npx
        self.tol, self.maxIter = float(tol), int(maxIter)
        if method not in ('direct', 'pcg', 'learned'):
            raise ValueError("method must be 'direct', 'pcg' or 'learned'")
        self.method = method
        self.checkEvery = int(checkEvery)
        ip = np.asarray(gridObj.getInsidePoints())
        self.insidePointsInd = np.nonzero(ip)[0]
        dsi = np.asarray(gridObj.getDsi(), np.float64)
        dst = np.asarray(gridObj.getDst(), np.float64)
        # poissonSolver.pyx:51: meps = eps0 * dsi * roll(dst, np)
        meps = epsilon_0 * dsi * np.roll(dst, npx)
        self.cx_host = meps[:npx].copy()
        self.cy_host = meps[npx:].copy()
        dev = L.device()
        self.cx = torch.as_tensor(self.cx_host).to(dev)
        self.cy = torch.as_tensor(self.cy_host).to(dev)
        self.mask = torch.as_tensor(ip.astype(np.uint8)).to(dev)
        # poissonSolver.pyx:44-48, :67
        pxmt = spsp.identity(npx) - spsp.eye(npx, npx, -1)
        pymt = spsp.identity(npx) - spsp.eye(npx, npx, -nx)
        st = spsp.hstack([pxmt, pymt])
        phiToE = gridObj.getEdgeToNode().dot(spsp.diags(dsi, 0, (2 * npx, 2 * npx)).dot(st.transpose()))
        csr = spsp.csr_matrix(phiToE)
        csr.sort_indices()
        self.phiToE_host = csr
        self.e2nKind, self.e2nIrr, self.e2nCx, self.e2nCy = e2n_structure(csr, npx, nx)
        self.e2nKind = torch.as_tensor(self.e2nKind).to(dev)
        self.e2nIrr = torch.as_tensor(self.e2nIrr).to(dev)
        self.rowPtr = torch.as_tensor(csr.indptr.astype(np.int32)).to(dev)
        self.col = torch.as_tensor(csr.indices.astype(np.int32)).to(dev)
        self.val = torch.as_tensor(csr.data.astype(np.float64)).to(dev)
        self.phi = torch.zeros(npx, dtype=torch.float64, device=dev)
        self.eAtGridPoints = torch.zeros(2 * npx, dtype=torch.float64, device=dev)
        self._q = torch.zeros(npx, dtype=torch.float64, device=dev)
        self._work = torch.empty(int(L.lib().oecl_pcg_work_bytes(npx)) // 8 + 8,
                                 dtype=torch.float64, device=dev)
        self.lastIterations = 0
        self.lastRelRes = 0.
        self.solveCount = 0
        self._fp = None
        self.preconditioner = preconditioner
        self.fallbacks = 0
        self.judgeMaxErr = 0.
        self.judgeCount = 0
        self._judge = None
        if method in ('direct', 'learned'):
            self._setup_direct()

    def _setup_direct(self):
        nx, ny = self.nx, self.ny
        bx = int(L.lib().oecl_fastpoisson_box_width(L.u32(nx)))
        if bx == 0:
            raise L.OeclError('direct solver: grid too small')
        cxu, cyu, sIdx, wPtr, wCol, wVal = capacitance_rows(
            nx, ny, self.mask.cpu().numpy(), self.cx_host, self.cy_host, bx)
        self.capacitanceRows = int(sIdx.size)
        self._fp_keep = (sIdx, wPtr, wCol, wVal)
        h = ctypes.c_void_p(0)
        cap = ctypes.c_double(0.)
        D = ctypes.POINTER(ctypes.c_double)
        I32 = ctypes.POINTER(ctypes.c_int32)
        L.check(L.lib().oecl_fastpoisson_create(
            L.u32(nx), L.u32(ny), L.f64(cxu), L.f64(cyu), L.u32(sIdx.size),
            sIdx.ctypes.data_as(I32), wPtr.ctypes.data_as(I32), wCol.ctypes.data_as(I32),
            wVal.ctypes.data_as(D), L.ptr(self.mask), ctypes.byref(h), ctypes.byref(cap)),
            'PoissonSolver direct setup')
        self._fp = h
        self.capacitanceMax = cap.value

    def __del__(self):
        try:
            if self._fp is not None and self._fp.value:
                L.lib().oecl_fastpoisson_destroy(self._fp)
        except Exception:
            pass

    def residual(self, q):
        """||q - A phi|| / ||q|| over inside nodes for the current phi (device)."""
        qd = self._rhs(q)
        it = ctypes.c_uint32(0)
        rr = ctypes.c_double(0.)
        L.lib().oecl_pcg_solve(
            L.u32(self.nx), L.u32(self.npExt), L.ptr(self.cx), L.ptr(self.cy), L.ptr(self.mask),
            L.ptr(qd), L.ptr(self.phi), L.f64(np.inf), L.u32(0), L.ptr(self._work),
            ctypes.byref(it), ctypes.byref(rr), L.stream())
        return rr.value

    def _rhs(self, q):
        if isinstance(q, torch.Tensor) and q.is_cuda and q.dtype == torch.float64:
            return q.contiguous()
        return L.to_dev(np.asarray(q, np.float64))

    def setPreconditioner(self, net):
        """The trained UNetPreconditioner of method 'learned'."""
        self.preconditioner = net

    def _direct(self, qd, out):
        L.check(L.lib().oecl_fastpoisson_solve(self._fp, L.ptr(qd), L.ptr(out), L.stream()),
                'PoissonSolver.solve (direct)')

    def _learned(self, qd):
        if self.preconditioner is None:
            raise L.OeclError("method 'learned' needs a trained preconditioner (setPreconditioner)")
        _, it, rr, ok = self.preconditioner.fcg(self, qd, phi=self.phi, tol=self.tol, maxIter=self.maxIter)
        self.lastIterations, self.lastRelRes = it, rr
        if not ok:
            self.fallbacks += 1
            self._direct(qd, self.phi)
        self.solveCount += 1
        if self.checkEvery > 0 and self.solveCount % self.checkEvery == 0:
            if self._judge is None:
                self._judge = torch.empty_like(self.phi)
            self._direct(qd, self._judge)
            den = float(torch.linalg.norm(self._judge))
            err = float(torch.linalg.norm(self.phi - self._judge)) / den if den > 0. else 0.
            self.judgeMaxErr = max(self.judgeMaxErr, err)
            self.judgeCount += 1
        return self.phi

    def solvePhi(self, q):
        """phi[inside] = A^-1 q[inside] (poissonSolver.pyx:85) on the device."""
        qd = self._rhs(q)
        if self.method == 'learned':
            return self._learned(qd)
        if self.method == 'direct':
            L.check(L.lib().oecl_fastpoisson_solve(self._fp, L.ptr(qd), L.ptr(self.phi),
                                                   L.stream()), 'PoissonSolver.solve (direct)')
            self.solveCount += 1
            self.lastIterations = 0
            if self.checkEvery > 0 and self.solveCount % self.checkEvery == 0:
                # residual check; CG refinement from the direct solution if it falls short
                self._pcg(qd, self.maxIter)
            return self.phi
        self._pcg(qd, self.maxIter)
        self.solveCount += 1
        return self.phi

    def _pcg(self, qd, maxIter):
        it = ctypes.c_uint32(0)
        rr = ctypes.c_double(0.)
        L.check(L.lib().oecl_pcg_solve(
            L.u32(self.nx), L.u32(self.npExt), L.ptr(self.cx), L.ptr(self.cy), L.ptr(self.mask),
            L.ptr(qd), L.ptr(self.phi), L.f64(self.tol), L.u32(maxIter), L.ptr(self._work),
            ctypes.byref(it), ctypes.byref(rr), L.stream()), 'PoissonSolver.solve (PCG)')
        self.lastIterations = it.value
        self.lastRelRes = rr.value

    def applyA(self, x, out=None):
        """A x for the FIT operator restricted to the inside nodes (poissonSolver.pyx:55): x is
        taken as zero on the other nodes, and the result is zero there (oecl_fit_apply)."""
        xd = self._rhs(x)
        y = torch.empty_like(xd) if out is None else out
        L.check(L.lib().oecl_fit_apply(L.u32(self.nx), L.u32(self.npExt), L.ptr(self.cx), L.ptr(self.cy),
                                       L.ptr(self.mask), L.ptr(xd), L.ptr(y), L.stream()), 'applyA')
        return y

    def phiToE(self):
        """eAtGridPoints = phiToEAtGridPoints . phi (poissonSolver.pyx:87)."""
        L.check(L.lib().oecl_csr_spmv(
            L.u32(2 * self.npExt), L.ptr(self.rowPtr), L.ptr(self.col), L.ptr(self.val),
            L.ptr(self.phi), L.ptr(self.eAtGridPoints), L.stream()), 'phiToE')
        return self.eAtGridPoints

    def _solve(self, q):
        """solve() on the device: E at the nodes as the device tensor eAtGridPoints."""
        self.solvePhi(q)
        return self.phiToE()

    @L.host_out
    def solve(self, q):
        """poissonSolver.pyx:83-88: returns E at the nodes, [Ex | Ey]."""
        return self._solve(q)

    @L.host_out
    def getPhi(self):
        return self.phi

    @L.host_out
    def getEAtGridPoints(self):
        return self.eAtGridPoints
