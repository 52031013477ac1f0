"""PoissonSolver (poissonSolver.pyx:20-99) with the solve on the GPU.

Setup (host, once) builds from the grid the same operators as the reference:
  - the FIT operator A = S~ M_eps S~^T on inside nodes, kept as two edge-coefficient arrays
    (cx on x edges, cy on y edges) of the 5-point stencil;
  - phiToEAtGridPoints = edgeToNode . diag(dsi) . S~^T, assembled with scipy.sparse exactly
    as the reference does and moved to the device as CSR.
solve(q) runs FP64 conjugate gradients on the device to tol * ||q_inside|| (warm-started
from the previous potential) and then the CSR phi -> E product, also on the device.
"""
import ctypes

import numpy as np
import scipy.sparse as spsp
import torch

from . import _lib as L
from .constants import epsilon_0


class PoissonSolver:

    def __init__(self, gridObj, tol=1e-12, maxIter=200000):
        L.require_gpu()
        nx = gridObj.getNxExt()
        npx = gridObj.getNpExt()
        self.nx, self.npExt = nx, npx
        self.tol, self.maxIter = float(tol), int(maxIter)
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

    def _rhs(self, q):
        if isinstance(q, torch.Tensor) and q.is_cuda and q.dtype == torch.float64:
            return q.contiguous()
        return L.to_dev(np.asarray(q, np.float64))

    def solvePhi(self, q):
        """phi[inside] = A^-1 q[inside] (poissonSolver.pyx:85) on the device."""
        qd = self._rhs(q)
        it = ctypes.c_uint32(0)
        rr = ctypes.c_double(0.)
        L.check(L.lib().oecl_pcg_solve(
            L.u32(self.nx), L.u32(self.npExt), L.ptr(self.cx), L.ptr(self.cy), L.ptr(self.mask),
            L.ptr(qd), L.ptr(self.phi), L.f64(self.tol), L.u32(self.maxIter), L.ptr(self._work),
            ctypes.byref(it), ctypes.byref(rr), L.stream()), 'PoissonSolver.solve (PCG)')
        self.lastIterations = it.value
        self.lastRelRes = rr.value
        self.solveCount += 1
        return self.phi

    def phiToE(self):
        """eAtGridPoints = phiToEAtGridPoints . phi (poissonSolver.pyx:87)."""
        L.check(L.lib().oecl_csr_spmv(
            L.u32(2 * self.npExt), L.ptr(self.rowPtr), L.ptr(self.col), L.ptr(self.val),
            L.ptr(self.phi), L.ptr(self.eAtGridPoints), L.stream()), 'phiToE')
        return self.eAtGridPoints

    def solve(self, q):
        """poissonSolver.pyx:83-88: returns E at the nodes, [Ex | Ey]."""
        self.solvePhi(q)
        return self.phiToE()

    def getPhi(self):
        return self.phi

    def getEAtGridPoints(self):
        return self.eAtGridPoints
