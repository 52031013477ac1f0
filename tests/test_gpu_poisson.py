"""GPU Poisson solve (PCG + phi -> E) against the oracle's SuperLU solve."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _rhs(g, seed):
    rng = np.random.default_rng(seed)
    q = rng.standard_normal(g.getNpExt()) * 1e-12
    # a smooth beam-like term on top of noise
    X, Y = np.meshgrid(g.getXMesh(), g.getYMesh())
    q += (1e-10 * np.exp(-(X ** 2 + Y ** 2) / (2 * 0.004 ** 2))).ravel()
    return q


# Box sizes nxExt - 4 and nyExt - 4 of both parities: the direct solve pads the box width to
# 2^p - 1 (the DST is an FFT of length 2^p), so most of these carry padding columns, whose
# neighbours in the last interior column become capacitance rows.
@pytest.mark.parametrize('args', [(128, 64, 0.1, 0.05, 0), (96, 72, 0.09, 0.07, 1),
                                  (200, 150, 0.09, 0.07, 1), (101, 77, 0.09, 0.07, 1),
                                  (65, 50, 0.1, 0.05, 0), (64, 47, 0.1, 0.05, 0)])
@pytest.mark.parametrize('method', ['direct', 'pcg'])
def test_solve_matches_lu(args, method):
    from openecloud_amd.grid import Grid
    from openecloud_amd.poissonSolver import PoissonSolver
    from oracle.grid_oracle import OraclePoisson
    g = Grid(*args)
    ps = PoissonSolver(g, tol=1e-14, method=method)
    q = _rhs(g, 1)
    e = ps.solve(torch.as_tensor(q).cuda()).cpu().numpy()
    phi = ps.getPhi().cpu().numpy()
    ref = OraclePoisson(g)
    e_ref = ref.solve(q)
    assert np.linalg.norm(phi - ref.phi) <= 1e-10 * np.linalg.norm(ref.phi)
    assert np.linalg.norm(e - e_ref) <= 1e-10 * np.linalg.norm(e_ref)
    # phi -> E alone is bit-identical to scipy's product on the same phi
    ps.phi.copy_(torch.as_tensor(ref.phi).cuda())
    assert np.array_equal(ps.phiToE().cpu().numpy(), ref.phiToE.dot(ref.phi))


@pytest.mark.parametrize('args', [(128, 64, 0.1, 0.05, 0), (200, 150, 0.09, 0.07, 1),
                                  (101, 77, 0.09, 0.07, 1), (65, 50, 0.1, 0.05, 0),
                                  (64, 47, 0.1, 0.05, 0), (70, 70, 0.05, 0.05, 4)])
def test_direct_alone_matches_lu(args):
    """The capacitance solve without CG refinement (checkEvery=0)."""
    from openecloud_amd.grid import Grid
    from openecloud_amd.poissonSolver import PoissonSolver
    from oracle.grid_oracle import OraclePoisson
    a = list(args)
    if a[4] == 4:   # polygon chamber: a hexagon's first-quadrant points
        g = Grid(*a, boundFunc=np.array([[0.024, 0.], [0.012, 0.0208], [0., 0.0208]]))
    else:
        g = Grid(*a)
    ps = PoissonSolver(g, method='direct', checkEvery=0)
    q = _rhs(g, 3)
    ps.solvePhi(torch.as_tensor(q).cuda())
    phi = ps.getPhi().cpu().numpy()
    ref = OraclePoisson(g)
    ref.solve(q)
    assert ps.lastIterations == 0
    assert np.linalg.norm(phi - ref.phi) <= 1e-10 * np.linalg.norm(ref.phi)
    assert ps.residual(q) <= 1e-9


def test_direct_2048_wide_box_rows_beyond_one_per_thread():
    """A flat 2048 x 64 ellipse: the box is 2047 columns wide (the persistent 2048-point DST
    passes), and its top and bottom rows hold more corrected nodes than the sparse-update pass
    has threads (256), so that pass also scatters the entries beyond one per thread."""
    from openecloud_amd.grid import Grid
    from openecloud_amd.poissonSolver import PoissonSolver
    from oracle.grid_oracle import OraclePoisson
    import openecloud_amd._lib as L
    g = Grid(2048, 64, 0.09, 0.003, 1)
    ps = PoissonSolver(g, method='direct', checkEvery=0)
    bx = int(L.lib().oecl_fastpoisson_box_width(L.u32(ps.nx)))
    assert bx + 1 == 2048
    sIdx = ps._fp_keep[0]
    assert np.bincount(sIdx // bx).max() > 256
    q = _rhs(g, 5)
    ps.solvePhi(torch.as_tensor(q).cuda())
    phi = ps.getPhi().cpu().numpy()
    ref = OraclePoisson(g)
    ref.solve(q)
    assert np.linalg.norm(phi - ref.phi) <= 1e-10 * np.linalg.norm(ref.phi)
    assert ps.residual(q) <= 1e-9


def test_pcg_warm_start_and_reproducible():
    from openecloud_amd.grid import Grid
    from openecloud_amd.poissonSolver import PoissonSolver
    g = Grid(128, 96, 0.09, 0.07, 1)
    ps = PoissonSolver(g, tol=1e-12, method='pcg')
    q = torch.as_tensor(_rhs(g, 2)).cuda()
    ps.solve(q)
    it_cold = ps.lastIterations
    phi1 = ps.getPhi().clone()
    assert it_cold > 64
    ps.solve(q)   # warm start from the solution: converges at once
    assert ps.lastIterations <= 64
    ps.phi.zero_()
    ps.solve(q)
    assert ps.lastIterations == it_cold
    assert torch.equal(ps.getPhi(), phi1)


@pytest.mark.parametrize('args', [(128, 64, 0.1, 0.05, 0), (96, 72, 0.09, 0.07, 1), (70, 70, 0.05, 0.05, 4)])
def test_apply_matches_oracle_operator(args):
    """A x on the inside nodes against the oracle's assembled FIT matrix (poissonSolver.pyx:55),
    with x non-zero outside too (the restriction treats it as zero there); and the direct
    solve inverts it."""
    from openecloud_amd.grid import Grid
    from openecloud_amd.poissonSolver import PoissonSolver
    from oracle.grid_oracle import OraclePoisson
    if args[4] == 4:   # polygon chamber: a hexagon's first-quadrant points
        g = Grid(*args, boundFunc=np.array([[0.024, 0.], [0.012, 0.0208], [0., 0.0208]]))
    else:
        g = Grid(*args)
    ps = PoissonSolver(g, method='direct', checkEvery=0)
    ref = OraclePoisson(g)
    ind = ref.insidePointsInd
    x = np.random.default_rng(3).standard_normal(g.getNpExt())
    y = ps.applyA(torch.as_tensor(x).cuda()).cpu().numpy()
    y_ref = np.zeros_like(x)
    y_ref[ind] = ref.a.dot(x[ind])
    assert np.linalg.norm(y - y_ref) <= 1e-13 * np.linalg.norm(y_ref)
    assert not np.any(np.delete(y, ind))
    xin = np.zeros_like(x)
    xin[ind] = x[ind]
    phi = ps.solvePhi(torch.as_tensor(y).cuda()).cpu().numpy()
    assert np.linalg.norm(phi - xin) <= 1e-9 * np.linalg.norm(xin)


# Wide boxes (padded from about 800 to 1023 columns), with even and odd widths, and irregular
# rows next to an elliptical wall.
@pytest.mark.parametrize('args', [(800, 64, 0.1, 0.05, 0), (801, 64, 0.1, 0.05, 0), (790, 100, 0.09, 0.07, 1)])
def test_direct_wide_box_matches_lu(args):
    from openecloud_amd.grid import Grid
    from openecloud_amd.poissonSolver import PoissonSolver
    from oracle.grid_oracle import OraclePoisson
    g = Grid(*args)
    assert g.getNxExt() - 4 >= 768
    ps = PoissonSolver(g, method='direct', checkEvery=0)
    q = _rhs(g, 5)
    ps.solvePhi(torch.as_tensor(q).cuda())
    phi = ps.getPhi().cpu().numpy()
    ref = OraclePoisson(g)
    ref.solve(q)
    assert np.linalg.norm(phi - ref.phi) <= 1e-10 * np.linalg.norm(ref.phi)
    assert ps.residual(q) <= 1e-9


def test_direct_wide_box_in_a_graph_matches_eager():
    """The wide-box solve captured into a graph: every replay, with eager work between the
    replays, gives the eager result to rounding."""
    from openecloud_amd.grid import Grid
    from openecloud_amd.poissonSolver import PoissonSolver
    g = Grid(801, 64, 0.1, 0.05, 0)
    ps = PoissonSolver(g, method='direct', checkEvery=0)
    q = torch.as_tensor(_rhs(g, 6)).cuda()
    eager = ps.solvePhi(q).clone()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.stream(side):
        ps.solvePhi(q)                    # warm-up on the capture stream
        torch.cuda.synchronize()
        with torch.cuda.graph(graph, stream=side):
            ps.solvePhi(q)
    torch.cuda.current_stream().wait_stream(side)
    for k in range(3):
        ps.phi.zero_()
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        err = float(torch.linalg.norm(ps.phi - eager) / torch.linalg.norm(eager))
        assert err < 1e-12, (k, err)
        ps.solvePhi(q * 0.5)              # eager work between the replays
