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


@pytest.mark.parametrize('args', [(128, 64, 0.1, 0.05, 0), (96, 72, 0.09, 0.07, 1),
                                  (200, 150, 0.09, 0.07, 1)])
def test_pcg_matches_lu(args):
    from openecloud_amd.grid import Grid
    from openecloud_amd.poissonSolver import PoissonSolver
    from oracle.grid_oracle import OraclePoisson
    g = Grid(*args)
    ps = PoissonSolver(g, tol=1e-14)
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


def test_pcg_warm_start_and_reproducible():
    from openecloud_amd.grid import Grid
    from openecloud_amd.poissonSolver import PoissonSolver
    g = Grid(128, 96, 0.09, 0.07, 1)
    ps = PoissonSolver(g, tol=1e-12)
    q = torch.as_tensor(_rhs(g, 2)).cuda()
    ps.solve(q)
    it_cold = ps.lastIterations
    phi1 = ps.getPhi().clone()
    ps.solve(q)   # warm start from the solution: converges at once
    assert ps.lastIterations <= 64
    ps.phi.zero_()
    ps.solve(q)
    assert ps.lastIterations == it_cold
    assert torch.equal(ps.getPhi(), phi1)
