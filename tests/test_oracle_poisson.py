"""CPU: the Poisson oracle pinned by the linear system, the DST closed form of the 5-point
Laplacian, and exact gradients of linear potentials."""
import numpy as np
import pytest
import scipy.fft as sfft

from oracle.grid_oracle import EPS0, OraclePoisson
from openecloud_amd.grid import Grid


def test_rect_operator_is_scaled_5point_laplacian_and_dst_solution():
    g = Grid(40, 24, 0.1, 0.05, 0)
    ps = OraclePoisson(g)
    nxi, nyi = g.nxExt - 4, g.nyExt - 4
    assert ps.a.shape == (nxi * nyi, nxi * nyi)
    cx = EPS0 * g.dy / g.dx
    cy = EPS0 * g.dx / g.dy
    rng = np.random.default_rng(0)
    q = np.zeros(g.npExt)
    q[ps.insidePointsInd] = rng.standard_normal(nxi * nyi) * 1e-12
    ps.solve(q)
    # closed form: A = cx*T_x (x) I + cy*I (x) T_y with T = tridiag(-1, 2, -1), DST-I diagonal
    Q = q[ps.insidePointsInd].reshape(nyi, nxi)
    lx = 2 - 2 * np.cos(np.pi * np.arange(1, nxi + 1) / (nxi + 1))
    ly = 2 - 2 * np.cos(np.pi * np.arange(1, nyi + 1) / (nyi + 1))
    Qh = sfft.dstn(Q, type=1)
    Ph = Qh / (cx * lx[None, :] + cy * ly[:, None])
    P = sfft.idstn(Ph, type=1)
    assert np.allclose(ps.phi[ps.insidePointsInd].reshape(nyi, nxi), P, rtol=1e-10,
                       atol=1e-12 * np.abs(P).max())


@pytest.mark.parametrize('args', [(64, 48, 0.09, 0.07, 1), (50, 30, 0.1, 0.05, 0)])
def test_residual_and_symmetry(args):
    g = Grid(*args)
    ps = OraclePoisson(g)
    assert abs(ps.a - ps.a.T).max() <= 1e-12 * abs(ps.a).max()
    rng = np.random.default_rng(1)
    q = rng.standard_normal(g.npExt) * 1e-12
    ps.solve(q)
    r = ps.a.dot(ps.phi[ps.insidePointsInd]) - q[ps.insidePointsInd]
    assert np.linalg.norm(r) <= 1e-12 * np.linalg.norm(q[ps.insidePointsInd])


def test_linear_potential_gives_exact_interior_field():
    g = Grid(40, 30, 0.09, 0.07, 1)
    ps = OraclePoisson(g)
    X, Y = np.meshgrid(g.xMesh, g.yMesh)
    a, b = 2.5, -4.0
    phi = np.where(g.insidePoints.reshape(X.shape) == 1, a * X + b * Y, 0.).ravel()
    e = ps.phiToE.dot(phi)
    nx, npx = g.nxExt, g.npExt
    ip = g.insidePoints
    # inside nodes whose 4 neighbours are inside: centred differences of a linear phi
    idx = np.array([c for c in np.nonzero(ip)[0]
                    if ip[c - 1] and ip[c + 1] and ip[c - nx] and ip[c + nx]])
    assert np.allclose(e[idx], -a, rtol=1e-9)
    assert np.allclose(e[idx + npx], -b, rtol=1e-9)


def test_point_charge_symmetry():
    g = Grid(41, 41, 0.04, 0.04, 1)
    ps = OraclePoisson(g)
    q = np.zeros(g.npExt)
    c = (g.nxExt // 2) + (g.nyExt // 2) * g.nxExt
    q[c] = 1e-12
    ps.solve(q)
    P = ps.phi.reshape(g.nyExt, g.nxExt)
    assert np.allclose(P, P[::-1, :], rtol=1e-10, atol=1e-14 * np.abs(P).max())
    assert np.allclose(P, P[:, ::-1], rtol=1e-10, atol=1e-14 * np.abs(P).max())
    assert np.allclose(P, P.T, rtol=1e-10, atol=1e-14 * np.abs(P).max())
    assert P.max() == P.flat[c] > 0


@pytest.mark.parametrize('args', [(40, 24, 0.1, 0.05, 0), (37, 30, 0.09, 0.07, 1), (64, 48, 0.09, 0.07, 1)])
@pytest.mark.parametrize('pad', [0, 1, 9])
def test_capacitance_rows_rebuild_the_operator_on_padded_boxes(args, pad):
    """M = R + U W over the box of bx = nxExt - 4 + pad columns (R: the uniform 5-point
    operator with zero Dirichlet data outside the box) has exactly the oracle operator's rows on
    inside nodes, including the nodes next to the first padding column."""
    import scipy.sparse as spsp
    from openecloud_amd.poissonSolver import capacitance_rows
    g = Grid(*args)
    ps = OraclePoisson(g)
    nx, ny, npx = g.nxExt, g.nyExt, g.npExt
    ip = np.asarray(g.insidePoints)
    meps = EPS0 * np.asarray(g.dsi, np.float64) * np.roll(np.asarray(g.dst, np.float64), npx)
    bx, by = nx - 4 + pad, ny - 4
    cxu, cyu, sIdx, wPtr, wCol, wVal = capacitance_rows(nx, ny, ip.astype(np.uint8), meps[:npx], meps[npx:], bx)
    T = lambda n: spsp.diags([-np.ones(n - 1), 2 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1])
    R = spsp.kron(spsp.identity(by), cxu * T(bx)) + spsp.kron(cyu * T(by), spsp.identity(bx))
    W = spsp.csr_matrix((wVal, wCol, wPtr), shape=(sIdx.size, bx * by))
    U = spsp.csr_matrix((np.ones(sIdx.size), (sIdx, np.arange(sIdx.size))), shape=(bx * by, sIdx.size))
    M = (R + U @ W).tocsr()
    # box index of each inside node, in the oracle's inside ordering
    J, I = np.divmod(ps.insidePointsInd, nx)
    t = (I - 2) + (J - 2) * bx
    assert (I >= 2).all() and (I < nx - 2).all()
    rows = M[t, :].toarray()
    ref = np.zeros_like(rows)
    ref[:, t] = ps.a.toarray() if spsp.issparse(ps.a) else ps.a
    assert np.abs(rows - ref).max() <= 1e-12 * np.abs(ref).max()
    if pad:
        # the last interior column's inside nodes now see a padding column
        last = t[I == nx - 3]
        assert set(last.tolist()) <= set(sIdx.tolist())


@pytest.mark.parametrize('args', [(64, 48, 0.09, 0.07, 1), (40, 24, 0.1, 0.05, 0), (70, 70, 0.05, 0.05, 4)])
def test_e2n_row_classes_reproduce_the_csr_product(args):
    """The step's row-class evaluation of E = phiToE phi (interior pattern with one coefficient
    per component, empty rows, listed rows from the CSR), emulated in numpy with the kernel's
    order of operations, equals scipy's CSR product bit for bit."""
    import scipy.sparse as spsp
    from openecloud_amd.poissonSolver import e2n_structure
    a = list(args)
    if a[4] == 4:
        g = Grid(*a, boundFunc=np.array([[0.024, 0.], [0.012, 0.0208], [0., 0.0208]]))
    else:
        g = Grid(*a)
    npx, nx = g.nxExt * g.nyExt, g.nxExt
    st = spsp.hstack([spsp.identity(npx) - spsp.eye(npx, npx, -1), spsp.identity(npx) - spsp.eye(npx, npx, -nx)])
    csr = spsp.csr_matrix(g.edgeToNode.dot(spsp.diags(np.asarray(g.dsi, np.float64), 0, (2 * npx, 2 * npx)).dot(st.T)))
    csr.sort_indices()
    kind, listed, cx, cy = e2n_structure(csr, npx, nx)
    k = np.concatenate([kind & 3, kind >> 2])
    assert set(np.nonzero(k == 2)[0].tolist()) == set(listed.tolist())
    assert (k == 1).sum() > 0.5 * k.size
    rng = np.random.default_rng(5)
    phi = rng.standard_normal(npx)
    ref = csr.dot(phi)
    E = np.zeros(2 * npx)
    node = np.arange(npx)
    for comp, c, o in ((0, cx, 1), (1, cy, nx)):
        kc = (kind >> (2 * comp)) & 3
        sel = node[kc == 1]
        s = np.zeros(sel.size)
        s = s + c * phi[sel - o]
        s = s + (-c) * phi[sel + o]
        E[comp * npx + sel] = s
    for r in listed:
        acc = 0.
        for kk in range(csr.indptr[r], csr.indptr[r + 1]):
            acc += csr.data[kk] * phi[csr.indices[kk]]
        E[r] = acc
    assert np.array_equal(E, ref)
