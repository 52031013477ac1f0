"""Grid geometry (host, setup time): the reference's Grid class, vectorised with NumPy.

Mirrors grid.pyx:49-1434 attribute for attribute and getter for getter.  Every array is
computed with the same floating-point operations as the reference's loops (the bisection of
cut edges runs over all cut edges at once, one bisection step per NumPy operation), so the
results are bit-identical; the edge-to-node matrices are assembled from the same entries and
multiplied in the same order with scipy.sparse.
"""
import math

import numpy as np
import scipy.sparse as spsp


def _bound_rect(lxh, lyh):
    """grid.pyx:1437-1441."""
    def f(x, y):
        return ((x < lxh) & (x > -lxh) & (y < lyh) & (y > -lyh)).astype(np.uint16)
    return f


def _bound_ellip(a_sq_inv, b_sq_inv):
    """grid.pyx:1443-1447."""
    def f(x, y):
        return (1 > x ** 2 * a_sq_inv + y ** 2 * b_sq_inv).astype(np.uint16)
    return f


def _vectorise(bf):
    """A user boundFunc may take scalars only; evaluate it point by point if needed."""
    def f(x, y):
        x = np.asarray(x, dtype=np.float64)
        y = np.asarray(y, dtype=np.float64)
        try:
            r = np.asarray(bf(x, y))
            if r.shape == np.broadcast(x, y).shape:
                return (r != 0).astype(np.uint16)
        except Exception:
            pass
        xb, yb = np.broadcast_arrays(x, y)
        out = np.fromiter((1 if bf(float(a), float(b)) else 0 for a, b in zip(xb.ravel(), yb.ravel())),
                          dtype=np.uint16, count=xb.size)
        return out.reshape(xb.shape)
    return f


def _wn_pnpoly(px, py, vv):
    """Winding number of points (px, py) against the closed polygon vv (grid.pyx:1464-1480)."""
    wn = np.zeros(np.broadcast(px, py).shape, dtype=np.int64)
    for i in range(vv.shape[0] - 1):
        x0, y0 = vv[i]
        x1, y1 = vv[i + 1]
        is_left = (x1 - x0) * (py - y0) - (px - x0) * (y1 - y0)
        up = (y0 <= py) & (y1 > py) & (is_left > 0)
        down = (y0 > py) & (y1 <= py) & (is_left < 0)
        wn += up.astype(np.int64) - down.astype(np.int64)
    return wn


class Grid:
    """Equidistant grid with one guard cell on each side (grid.pyx:49-105).

    boundType 0 = rectangular staircase, 1 = elliptical cut cell, 2 = arbitrary staircase,
    3 = arbitrary cut cell, 4 = polygon cut cell (boundFunc = polygon points, first quadrant).
    """

    def __init__(self, nx, ny, lx, ly, boundType, boundFunc=None, cutCellAcc=1e-12,
                 scanAroundPoint=1.e-3):
        self.nx, self.ny = int(nx), int(ny)
        self.lx, self.ly = float(lx), float(ly)
        self.dx = self.lx / (self.nx - 1.)
        self.dy = self.ly / (self.ny - 1.)
        self.cutCellAcc = float(cutCellAcc)
        self.scanAroundPoint = float(scanAroundPoint)
        self.nxExt = self.nx + 2
        self.nyExt = self.ny + 2
        self.npExt = self.nxExt * self.nyExt
        self.np = self.nx * self.ny
        self.lxExt = self.lx + 2. * self.dx
        self.lyExt = self.ly + 2. * self.dy
        self.xMesh = np.linspace(-self.lxExt / 2., self.lxExt / 2, self.nxExt)
        self.yMesh = np.linspace(-self.lyExt / 2., self.lyExt / 2, self.nyExt)
        self.boundType = int(boundType)
        self.boundaryPoints = np.empty((0, 2))
        self.boundaryPointsInd = np.empty(0, dtype=np.uint32)
        self.cutCellPointsInd = -np.ones((self.npExt, 2), dtype=np.int32)
        self.cutCellCenter = np.zeros((self.npExt, 2))
        self.cutCellNormalVectors = np.zeros((self.npExt, 2))
        if boundType == 0:
            self.boundFunc = _bound_rect(self.lx * 0.5, self.ly * 0.5)
            self._staircase()
        elif boundType == 1:
            self.boundFunc = _bound_ellip(4. / self.lx ** 2, 4. / self.ly ** 2)
            self._cutcell()
        elif boundType == 2:
            self.boundFunc = _vectorise(boundFunc)
            self._staircase()
        elif boundType == 3:
            self.boundFunc = _vectorise(boundFunc)
            self._cutcell()
        elif boundType == 4:
            self._prepare_polygon(np.asarray(boundFunc, dtype=np.float64))
            self.boundFunc = self._polygon_bound
            self._cutcell()
        else:
            raise NotImplementedError("Not yet implemented.")

    # --------------------------------------------------------------------- polygon
    def _prepare_polygon(self, pts):
        """grid.pyx:125-156 (prepareBoundFuncPoints): drop nearly collinear points."""
        n = pts.shape[0]
        keep = np.ones(n, dtype=np.uint16)
        ii = 0
        while ii + 2 < n:
            inc = 0
            while ii + 2 + inc < n:
                min_area = 1.e-16 * (math.sqrt((pts[ii, 0] - pts[ii + 1, 0]) ** 2 +
                                               (pts[ii, 1] - pts[ii + 1, 1]) ** 2) *
                                     math.sqrt((pts[ii + 1 + inc, 0] - pts[ii + 2 + inc, 0]) ** 2 +
                                               (pts[ii + 1 + inc, 1] - pts[ii + 2 + inc, 1]) ** 2))
                area = (pts[ii, 0] * (pts[ii + 1 + inc, 1] - pts[ii + 2 + inc, 1]) +
                        pts[ii + 1 + inc, 0] * (pts[ii + 2 + inc, 1] - pts[ii, 1]) +
                        pts[ii + 2 + inc, 0] * (pts[ii, 1] - pts[ii + 1 + inc, 1]))
                if abs(min_area) > abs(area):
                    inc += 1
                else:
                    break
            keep[ii + 1:ii + inc + 1] = 0
            ii += inc + 1
        self.boundFuncPoints = pts[keep == 1].copy()
        self.nBoundFuncPoints = int(keep.sum())

    def _polygon_bound(self, x, y):
        """grid.pyx:108-121 (pointInPolygon on |x|, |y|)."""
        wn = _wn_pnpoly(np.abs(x), np.abs(y), self.boundFuncPoints)
        return (wn != 0).astype(np.uint16)

    def pointInPolygon(self, xx, yy):
        return int(self._polygon_bound(np.float64(xx), np.float64(yy)))

    # --------------------------------------------------------------------- helpers
    def _inside_points(self):
        nx, ny = self.nxExt, self.nyExt
        ip = np.zeros((ny, nx), dtype=np.uint16)
        dxMin = self.scanAroundPoint * self.dx
        dyMin = self.scanAroundPoint * self.dy
        X = self.xMesh[2:nx - 2][None, :]
        Y = self.yMesh[2:ny - 2][:, None]
        bf = self.boundFunc
        ok = ((bf(X + dxMin, Y + dyMin) != 0) & (bf(X + dxMin, Y - dyMin) != 0) &
              (bf(X - dxMin, Y + dyMin) != 0) & (bf(X - dxMin, Y - dyMin) != 0))
        ip[2:ny - 2, 2:nx - 2] = ok.astype(np.uint16)
        return ip

    def _inside_edges(self, ip):
        nx, ny = self.nxExt, self.nyExt
        iex = np.zeros((ny, nx), dtype=np.uint16)
        iey = np.zeros((ny, nx), dtype=np.uint16)
        iex[2:ny - 2, 1:nx - 1] = ip[2:ny - 2, 1:nx - 1] + ip[2:ny - 2, 2:nx]
        iey[1:ny - 1, 2:nx - 2] = ip[1:ny - 1, 2:nx - 2] + ip[2:ny, 2:nx - 2]
        return iex, iey

    def _faces(self, iex, iey):
        nx, ny = self.nxExt, self.nyExt
        f = np.zeros((ny, nx), dtype=np.uint16)
        f[1:ny - 2, 1:nx - 2] = (iex[1:ny - 2, 1:nx - 2] + iex[2:ny - 1, 1:nx - 2] +
                                 iey[1:ny - 2, 1:nx - 2] + iey[1:ny - 2, 2:nx - 1])
        return f

    def _dual(self):
        """grid.pyx:332-370 (identical for staircase and cut cell)."""
        nx, ny = self.nxExt, self.nyExt
        dx, dy = self.dx, self.dy
        dstx = np.zeros((ny, nx))
        dsty = np.zeros((ny, nx))
        dstix = np.zeros((ny, nx))
        dstiy = np.zeros((ny, nx))
        dstx[1:ny - 1, 2:nx - 2] = dx
        dstix[1:ny - 1, 2:nx - 2] = 1. / dx
        dsty[2:ny - 2, 1:nx - 1] = dy
        dstiy[2:ny - 2, 1:nx - 1] = 1. / dy
        dstx[1:ny - 1, 1] = 0.5 * dx
        dstx[1:ny - 1, nx - 2] = 0.5 * dx
        dstix[1:ny - 1, 1] = 1. / (0.5 * dx)
        dstix[1:ny - 1, nx - 2] = 1. / (0.5 * dx)
        dsty[1, 1:nx - 1] = 0.5 * dy
        dsty[ny - 2, 1:nx - 1] = 0.5 * dy
        dstiy[1, 1:nx - 1] = 1. / (0.5 * dy)
        dstiy[ny - 2, 1:nx - 1] = 1. / (0.5 * dy)
        self.dst = np.concatenate([dstx.ravel(), dsty.ravel()])
        self.dsti = np.concatenate([dstix.ravel(), dstiy.ravel()])
        self.dat = (dstx * dsty).ravel()
        self.dati = (dstix * dstiy).ravel()

    def _store_common(self, ip, iex, iey, dsx, dsy, dsix, dsiy, f, da, dai):
        self.insidePoints = ip.ravel().copy()
        self.insideEdges = np.concatenate([iex.ravel(), iey.ravel()])
        self.insideFaces = f.ravel().copy()
        self.ds = np.concatenate([dsx.ravel(), dsy.ravel()])
        self.dsi = np.concatenate([dsix.ravel(), dsiy.ravel()])
        self.da = da.ravel().copy()
        self.dai = dai.ravel().copy()

    @staticmethod
    def _coo(rows, cols, vals, n2):
        rows = np.concatenate(rows) if rows else np.empty(0, np.int64)
        cols = np.concatenate(cols) if cols else np.empty(0, np.int64)
        vals = np.concatenate(vals) if vals else np.empty(0)
        return spsp.csc_matrix(spsp.coo_matrix((vals, (rows, cols)), shape=(n2, n2)))

    # --------------------------------------------------------------------- staircase
    def _staircase(self):
        """grid.pyx:166-458."""
        nx, ny, npx = self.nxExt, self.nyExt, self.npExt
        dx, dy = self.dx, self.dy
        ip = self._inside_points()
        iex, iey = self._inside_edges(ip)
        dsx = np.zeros((ny, nx)); dsy = np.zeros((ny, nx))
        dsix = np.zeros((ny, nx)); dsiy = np.zeros((ny, nx))
        sl = (slice(2, ny - 2), slice(1, nx - 2))
        m = iex[sl] != 0
        dsx[sl] = np.where(m, dx, 0.)
        dsix[sl] = np.where(m, 1. / dx, 0.)
        sl = (slice(1, ny - 2), slice(2, nx - 2))
        m = iey[sl] != 0
        dsy[sl] = np.where(m, dy, 0.)
        dsiy[sl] = np.where(m, 1. / dy, 0.)
        f = self._faces(iex, iey)
        da = np.zeros((ny, nx)); dai = np.zeros((ny, nx))
        sl = (slice(1, ny - 2), slice(1, nx - 2))
        m = f[sl] != 0
        da[sl] = np.where(m, dx * dy, 0.)
        dai[sl] = np.where(m, 1. / (dx * dy), 0.)
        self._store_common(ip, iex, iey, dsx, dsy, dsix, dsiy, f, da, dai)
        self._dual()
        ds = self.ds
        ipf = self.insidePoints
        ie = self.insideEdges
        c = np.nonzero(ipf == 1)[0]
        # grid.pyx:373-408 (all four loops visit exactly the inside points).
        e1 = self._coo([c, c, c + npx, c + npx], [c, c - 1, c + npx, c + npx - nx],
                       [ds[c] / (ds[c - 1] + ds[c]), ds[c - 1] / (ds[c - 1] + ds[c]),
                        ds[c + npx] / (ds[c + npx - nx] + ds[c + npx]),
                        ds[c + npx - nx] / (ds[c + npx - nx] + ds[c + npx])], 2 * npx)
        # grid.pyx:411-458: boundary rows (the y cases write into the x row, as there).
        jj, ii = np.meshgrid(np.arange(1, ny - 1), np.arange(1, nx - 1), indexing='ij')
        cc = (ii + jj * nx).ravel()
        ins = ipf[cc] == 1
        out = ~ins
        rows, cols, vals = [cc[ins], cc[ins] + npx], [cc[ins], cc[ins] + npx], \
            [np.ones(ins.sum()), np.ones(ins.sum())]
        condx = (ipf[cc + nx] == 1) & (ipf[cc - nx] == 1)
        a = out & (ie[cc] == 1) & condx
        b = out & (ie[cc] != 1) & (ie[cc - 1] == 1) & condx
        condy = (ipf[cc + 1] == 1) & (ipf[cc - 1] == 1)
        cy = out & (ie[cc + npx] == 1) & condy
        dy_ = out & (ie[cc + npx] != 1) & (ie[cc + npx - nx] == 1) & condy
        for msk, off in ((a, 1), (b, -1), (cy, nx), (dy_, -nx)):
            rows.append(cc[msk]); cols.append(cc[msk] + off); vals.append(np.ones(msk.sum()))
        e2 = self._coo(rows, cols, vals, 2 * npx)
        self.edgeToNode = e2.dot(e1)

    # --------------------------------------------------------------------- cut cell
    def _bisect(self, x1, x2, yfix, horizontal):
        """grid.pyx:617-631 / :646-660, all edges at once."""
        bf = self.boundFunc
        iters = int(math.log2(1. / self.cutCellAcc) + 2.)
        if horizontal:
            t1 = bf(x1, yfix)
            for _ in range(iters):
                x3 = 0.5 * (x1 + x2)
                t3 = bf(x3, yfix)
                same = t1 == t3
                x1 = np.where(same, x3, x1)
                x2 = np.where(same, x2, x3)
        else:
            t1 = bf(yfix, x1)
            for _ in range(iters):
                x3 = 0.5 * (x1 + x2)
                t3 = bf(yfix, x3)
                same = t1 == t3
                x1 = np.where(same, x3, x1)
                x2 = np.where(same, x2, x3)
        return x1, x2

    def _cutcell(self):
        """grid.pyx:468-1330, the active code."""
        nx, ny, npx = self.nxExt, self.nyExt, self.npExt
        dx, dy = self.dx, self.dy
        xm, ym = self.xMesh, self.yMesh
        bf = self.boundFunc
        ip = self._inside_points()
        iex, iey = self._inside_edges(ip)
        dsx = np.zeros((ny, nx)); dsy = np.zeros((ny, nx))
        dsix = np.zeros((ny, nx)); dsiy = np.zeros((ny, nx))
        # x edges, grid.pyx:602-632
        J, I = np.meshgrid(np.arange(2, ny - 2), np.arange(1, nx - 2), indexing='ij')
        e = iex[J, I]
        full = e == 2
        dsx[J[full], I[full]] = dx
        dsix[J[full], I[full]] = 1. / dx
        cut = e == 1
        Jc, Ic = J[cut], I[cut]
        if Jc.size:
            x1 = xm[Ic].copy(); x2 = xm[Ic + 1].copy(); yv = ym[Jc]
            both = (bf(x1, yv) == 1) & (bf(x2, yv) == 1)
            b1, b2 = self._bisect(x1, x2, yv, True)
            mid = 0.5 * (b1 + b2)
            val = np.where(ip[Jc, Ic] == 1, mid - xm[Ic], xm[Ic + 1] - mid)
            val = np.where(both, dx, val)
            dsx[Jc, Ic] = val
            dsix[Jc, Ic] = 1. / val
        # y edges, grid.pyx:634-661
        J, I = np.meshgrid(np.arange(1, ny - 2), np.arange(2, nx - 2), indexing='ij')
        e = iey[J, I]
        full = e == 2
        dsy[J[full], I[full]] = dy
        dsiy[J[full], I[full]] = 1. / dy
        cut = e == 1
        Jc, Ic = J[cut], I[cut]
        if Jc.size:
            y1 = ym[Jc].copy(); y2 = ym[Jc + 1].copy(); xv = xm[Ic]
            both = (bf(xv, y1) == 1) & (bf(xv, y2) == 1)
            b1, b2 = self._bisect(y1, y2, xv, False)
            mid = 0.5 * (b1 + b2)
            val = np.where(ip[Jc, Ic] == 1, mid - ym[Jc], ym[Jc + 1] - mid)
            val = np.where(both, dy, val)
            dsy[Jc, Ic] = val
            dsiy[Jc, Ic] = 1. / val
        f = self._faces(iex, iey)
        da = np.zeros((ny, nx)); dai = np.zeros((ny, nx))
        self._store_common(ip, iex, iey, dsx, dsy, dsix, dsiy, f, da, dai)
        ds = self.ds
        ipf = self.insidePoints
        ie = self.insideEdges
        # grid.pyx:664-719: boundary points in (row, column, x-before-y) order.
        jj, ii = np.meshgrid(np.arange(1, ny - 1), np.arange(1, nx - 1), indexing='ij')
        cc = (ii + jj * nx).ravel()
        iiv, jjv = ii.ravel(), jj.ravel()
        isx = ie[cc] == 1
        isy = ie[cc + npx] == 1
        order_c = np.concatenate([cc[isx], cc[isy]])
        order_t = np.concatenate([np.zeros(isx.sum(), np.int64), np.ones(isy.sum(), np.int64)])
        order_i = np.concatenate([iiv[isx], iiv[isy]])
        order_j = np.concatenate([jjv[isx], jjv[isy]])
        srt = np.lexsort((order_t, order_c))
        pc, pt, pi, pj = order_c[srt], order_t[srt], order_i[srt], order_j[srt]
        nb = pc.size
        bp = np.empty((nb, 2))
        bpi = np.empty(nb, dtype=np.uint32)
        inside_here = ipf[pc] == 1
        tx = pt == 0
        bp[:, 0] = np.where(tx, np.where(inside_here, xm[pi] + ds[pc], xm[np.minimum(pi + 1, nx - 1)] - ds[pc]),
                            xm[pi])
        bp[:, 1] = np.where(tx, ym[pj],
                            np.where(inside_here, ym[pj] + ds[pc + npx],
                                     ym[np.minimum(pj + 1, ny - 1)] - ds[pc + npx]))
        bpi[:] = np.where(tx, np.where(inside_here, pc + 1, pc),
                          np.where(inside_here, pc + nx + npx, pc + npx))
        cpi = self.cutCellPointsInd
        nbr = np.where(tx, pc - nx, pc - 1)
        for kk in range(nb):
            for cell in (pc[kk], nbr[kk]):
                if cpi[cell, 0] == -1:
                    cpi[cell, 0] = kk
                else:
                    cpi[cell, 1] = kk
        self.boundaryPoints = bp
        self.boundaryPointsInd = bpi
        ff = self.insideFaces
        # grid.pyx:736-790: cut-cell centres and counter-clockwise order.
        c = cc[(ff[cc] != 0) & (ff[cc] != 8)]
        ic = c % nx
        jc = c // nx
        x0, x1_ = xm[ic], xm[ic + 1]
        y0, y1_ = ym[jc], ym[jc + 1]
        ipc, ip1, ipn, ipn1 = ipf[c], ipf[c + 1], ipf[c + nx], ipf[c + nx + 1]
        fc = ff[c]
        dsc, dsc1x = ds[c], ds[c + nx]
        dsy0, dsy1 = ds[c + npx], ds[c + npx + 1]
        cx_ = np.zeros(c.size)
        cy_ = np.zeros(c.size)
        done = np.zeros(c.size, bool)

        def put(mask, vx, vy):
            m = mask & ~done
            cx_[m] = np.broadcast_to(vx, c.shape)[m]
            cy_[m] = np.broadcast_to(vy, c.shape)[m]
            done[m] = True

        f2, f4, f6 = fc == 2, fc == 4, fc == 6
        put(f2 & (ipc == 1), x0 + dsc / 3., y0 + dsy0 / 3.)
        put(f2 & (ip1 == 1), x1_ - dsc / 3., y0 + dsy1 / 3.)
        put(f2 & (ipn == 1), x0 + dsc1x / 3., y1_ - dsy0 / 3.)
        put(f2 & (ipn1 == 1), x1_ - dsc1x / 3., y1_ - dsy1 / 3.)
        put(f4 & (ipc == 1) & (ip1 == 1), (x0 + x1_) * 0.5, y0 + (dsy0 + dsy1) * 0.25)
        put(f4 & (ipn == 1) & (ipn1 == 1), (x0 + x1_) * 0.5, y1_ - (dsy0 + dsy1) * 0.25)
        put(f4 & (ipc == 1) & (ipn == 1), x0 + (dsc + dsc1x) * 0.25, (y0 + y1_) * 0.5)
        put(f4 & (ip1 == 1) & (ipn1 == 1), x1_ - (dsc + dsc1x) * 0.25, (y0 + y1_) * 0.5)
        put(f6 & (ipc == 0), (2. * x0 + 3. * x1_ - dsc) * 0.2, (2. * y0 + 3. * y1_ - dsy0) * 0.2)
        put(f6 & (ip1 == 0), (3. * x0 + 2. * x1_ + dsc) * 0.2, (2. * y0 + 3. * y1_ - dsy1) * 0.2)
        put(f6 & (ipn == 0), (2. * x0 + 3. * x1_ - dsc1x) * 0.2, (3. * y0 + 2. * y1_ + dsy0) * 0.2)
        put(f6 & (ipn1 == 0), (3. * x0 + 2. * x1_ + dsc1x) * 0.2, (3. * y0 + 2. * y1_ + dsy1) * 0.2)
        self.cutCellCenter[c, 0] = cx_
        self.cutCellCenter[c, 1] = cy_
        a, b = cpi[c, 0].copy(), cpi[c, 1].copy()
        swap = ((bp[a, 0] - cx_) * (bp[b, 1] - cy_) < (bp[a, 1] - cy_) * (bp[b, 0] - cx_))
        cpi[c[swap], 0] = b[swap]
        cpi[c[swap], 1] = a[swap]
        # grid.pyx:792-805: normals.
        a, b = cpi[c, 0], cpi[c, 1]
        n0 = bp[a, 1] - bp[b, 1]
        n1 = bp[b, 0] - bp[a, 0]
        # a cell whose two cut points coincide gets 0 * inf = NaN, as the reference does (its
        # loop divides with cdivision=True, grid.pyx:4)
        with np.errstate(divide='ignore', invalid='ignore'):
            t = 1. / np.sqrt(n0 ** 2 + n1 ** 2)
            self.cutCellNormalVectors[c, 0] = n0 * t
            self.cutCellNormalVectors[c, 1] = n1 * t
        # grid.pyx:808-849: face areas.
        daf = np.zeros(npx); daif = np.zeros(npx)
        J, I = np.meshgrid(np.arange(1, ny - 2), np.arange(1, nx - 2), indexing='ij')
        c = (I + J * nx).ravel()
        fc = ff[c]
        dxdy = dx * dy
        ipc, ip1, ipn = ipf[c], ipf[c + 1], ipf[c + nx]
        d_c, d_n = ds[c], ds[c + nx]
        dy_c, dy_1 = ds[c + npx], ds[c + 1 + npx]
        area = np.zeros(c.size)
        m8 = fc == 8
        area[m8] = dxdy
        m2 = fc == 2
        v2 = np.where(ipc == 1, 0.5 * d_c * dy_c,
                      np.where(ip1 == 1, 0.5 * d_c * dy_1,
                               np.where(ipn == 1, 0.5 * d_n * dy_c, 0.5 * d_n * dy_1)))
        area[m2] = v2[m2]
        m4 = fc == 4
        area[m4] = (0.5 * (d_c + d_n) * (dy_c + dy_1))[m4]
        m6 = fc == 6
        v6 = np.where(ipc == 0, dxdy - 0.5 * (dx - d_c) * (dy - dy_c),
                      np.where(ip1 == 0, dxdy - 0.5 * (dx - d_c) * (dy - dy_1),
                               np.where(ipn == 0, dxdy - 0.5 * (dx - d_n) * (dy - dy_c),
                                        dxdy - 0.5 * (dx - d_n) * (dy - dy_1))))
        area[m6] = v6[m6]
        daf[c] = area
        inv = np.zeros(c.size)
        inv[m8] = 1. / dxdy
        mm = m2 | m4 | m6
        with np.errstate(divide='ignore'):
            inv[mm] = 1. / area[mm]
        daif[c] = inv
        self.da, self.dai = daf, daif
        self._dual()
        # grid.pyx:893-925: edge -> inside node interpolation.
        c = np.nonzero(ipf == 1)[0]
        e1 = self._coo([c, c, c + npx, c + npx], [c, c - 1, c + npx, c + npx - nx],
                       [ds[c] / (ds[c - 1] + ds[c]), ds[c - 1] / (ds[c - 1] + ds[c]),
                        ds[c + npx] / (ds[c + npx - nx] + ds[c + npx]),
                        ds[c + npx - nx] / (ds[c + npx - nx] + ds[c + npx])], 2 * npx)
        # grid.pyx:1100-1163: first-order extrapolation to boundary nodes.
        cei = np.zeros(npx)
        out = ipf[cc] == 0
        co = cc[out]
        cei[co] = (ie[co].astype(np.float64) + ie[co - 1] + ie[co + npx] + ie[co - nx + npx])
        big = cei > 1.
        cei[big] = 1. / cei[big]
        ins = cc[~out]
        rows = [ins, ins + npx]
        cols = [ins, ins + npx]
        vals = [np.ones(ins.size), np.ones(ins.size)]
        ax_ = co[ie[co] == 1]
        bx_ = co[(ie[co] != 1) & (ie[co - 1] == 1)]
        ay_ = co[ie[co + npx] == 1]
        by_ = co[ie[co + npx - nx] == 1]
        for sel, off in ((ax_, 1), (bx_, -1), (ay_, nx), (by_, -nx)):
            rows += [sel, sel + npx]
            cols += [sel + off, sel + npx + off]
            vals += [cei[sel], cei[sel]]
        e2 = self._coo(rows, cols, vals, 2 * npx)
        # grid.pyx:1166-1274: diagonal-corner correction.
        rows = [cc, cc + npx]
        cols = [cc, cc + npx]
        vals = [np.ones(cc.size), np.ones(cc.size)]
        taken = np.zeros(co.size, bool)
        for s, t in ((-1, nx), (1, nx), (1, -nx), (-1, -nx)):
            m = (~taken) & (ipf[co + s] == 0) & (ipf[co + t] == 0) & (ipf[co + s + t] == 1)
            taken |= m
            sel = co[m]
            k = sel.size
            rows += [sel, sel, sel, sel + npx, sel + npx, sel + npx]
            cols += [sel + s + t, sel + s, sel + t, sel + npx + s + t, sel + s + npx, sel + t + npx]
            vals += [-np.ones(k), np.ones(k), np.ones(k), -np.ones(k), np.ones(k), np.ones(k)]
        e3 = self._coo(rows, cols, vals, 2 * npx)
        self.edgeToNode = e3.dot(e2.dot(e1))

    # --------------------------------------------------------------------- getters
    def getXMesh(self): return self.xMesh
    def getYMesh(self): return self.yMesh
    def getLx(self): return self.lx
    def getLy(self): return self.ly
    def getLxExt(self): return self.lxExt
    def getLyExt(self): return self.lyExt
    def getDx(self): return self.dx
    def getDy(self): return self.dy
    def getNxExt(self): return self.nxExt
    def getNyExt(self): return self.nyExt
    def getNpExt(self): return self.npExt
    def getNx(self): return self.nx
    def getNy(self): return self.ny
    def getNp(self): return self.np
    def getBoundaryPoints(self): return self.boundaryPoints
    def getBoundaryPointsInd(self): return self.boundaryPointsInd
    def getCutCellPointsInd(self): return self.cutCellPointsInd
    def getCutCellCenter(self): return self.cutCellCenter
    def getCutCellNormalVectors(self): return self.cutCellNormalVectors
    def getInsidePoints(self): return self.insidePoints
    def getInsideEdges(self): return self.insideEdges
    def getInsideFaces(self): return self.insideFaces
    def getDs(self): return self.ds
    def getDsi(self): return self.dsi
    def getDa(self): return self.da
    def getDai(self): return self.dai
    def getDst(self): return self.dst
    def getDsti(self): return self.dsti
    def getDat(self): return self.dat
    def getDati(self): return self.dati
    def getEdgeToNode(self): return self.edgeToNode
    def getBoundFunc(self): return self.boundFunc
