"""GPU parity of the particle kernels against the oracle (FP64 bit-exact, FP32 to 1e-5)."""
import numpy as np
import pytest
import torch

from conftest import random_particles

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def grids():
    from openecloud_amd.grid import Grid
    return {'rect': Grid(128, 64, 0.1, 0.05, 0), 'ellip': Grid(96, 72, 0.09, 0.07, 1)}


def _field(g, seed, dims=2, scale=1e5):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(dims * g.getNpExt()) * scale


def _parts(g, pd, dtype=torch.float64):
    from openecloud_amd.particles import Particles
    p = Particles('electrons', 'variableWeight', dtype=dtype)
    p.setParticleData(pd)
    return p


@pytest.mark.parametrize('gname', ['rect', 'ellip'])
def test_weights_deposit_gather_bitexact(grids, orc, gname):
    g = grids[gname]
    pd = random_particles(g, 5000, seed=1, wall='ellip' if gname == 'ellip' else 'rect')
    p = _parts(g, pd)
    p.calcGridWeights(g)
    ic, w = orc.calc_grid_weights(pd, g.getDx(), g.getDy(), g.getLxExt(), g.getLyExt(), g.getNxExt())
    assert np.array_equal(p.getInCell().cpu().numpy().astype(np.uint32), ic)
    for k, get in enumerate((p.getWeightLowXLowY, p.getWeightUpXLowY, p.getWeightLowXUpY,
                             p.getWeightUpXUpY)):
        assert np.array_equal(get().cpu().numpy(), w[k])
    E = _field(g, 2)
    p.eFieldToParticles(g, E)
    eo = orc.field_to_particles(E, ic, w, g.getNxExt(), g.getNpExt(), 2)
    assert np.array_equal(p.getEAtParticles().cpu().numpy(), eo)
    B = _field(g, 3, 3, 0.5)
    p.bFieldToParticles(g, B)
    bo = orc.field_to_particles(B, ic, w, g.getNxExt(), g.getNpExt(), 3)
    assert np.array_equal(p.getBAtParticles()[:5000].cpu().numpy(), bo)
    p.chargeToGrid(g)
    rho_o = orc.charge_to_grid(pd, ic, w, -1.602176565e-19, g.getNxExt(), g.getNpExt())
    rho = p.getChargeOnGrid().cpu().numpy()
    # atomics change the summation order: compare to the sum's own rounding scale
    assert np.allclose(rho, rho_o, rtol=1e-13, atol=1e-13 * np.abs(rho_o).max())
    # total charge: q * sum(w) exactly up to FP64 summation rounding
    assert abs(rho.sum() - (-1.602176565e-19) * pd[:, 5].sum()) <= 1e-12 * abs(rho.sum())


@pytest.mark.parametrize('gname,n,dtype', [('rect', 5000, 'f64'), ('ellip', 200000, 'f64'), ('rect', 1, 'f64'),
                                            ('ellip', 30000, 'f32'), ('rect', 0, 'f64')])
def test_ordered_deposit_is_the_sequential_sum(grids, orc, gname, n, dtype):
    """orderedDeposit: rho bit-identical to the reference's sequential loop
    (particles.pyx:336-351), for many particles per cell, a single particle, none, and FP32
    storage (weights and products in FP64 from the stored FP32 values, as the oracle does
    from the same values)."""
    from openecloud_amd.particles import Particles
    g = grids[gname]
    tdt = torch.float64 if dtype == 'f64' else torch.float32
    pd = random_particles(g, max(n, 1), seed=21, wall='ellip' if gname == 'ellip' else 'rect')[:n]
    if dtype == 'f32':
        pd = pd.astype(np.float32).astype(np.float64)
    p = Particles('electrons', 'variableWeight', dtype=tdt, orderedDeposit=True)
    p.setParticleData(pd)
    p.calcGridWeights(g)
    p.chargeToGrid(g)
    rho = p.getChargeOnGrid().cpu().numpy()
    if n == 0:
        assert np.all(rho == 0.)
        return
    w4 = [p._w4[k, :n].to(torch.float64).cpu().numpy() for k in range(4)]
    ic = p.getInCell().cpu().numpy().astype(np.uint32)
    if dtype == 'f64':
        ic0, w0 = orc.calc_grid_weights(pd, g.getDx(), g.getDy(), g.getLxExt(), g.getLyExt(), g.getNxExt())
        assert np.array_equal(ic, ic0) and all(np.array_equal(a, b) for a, b in zip(w4, w0))
    rho_o = orc.charge_to_grid(pd, ic, w4, -1.602176565e-19, g.getNxExt(), g.getNpExt())
    assert np.array_equal(rho, rho_o)
    # the atomic deposit of the same particles differs only by summation order
    p.orderedDeposit = False
    p.chargeToGrid(g)
    rho_a = p.getChargeOnGrid().cpu().numpy()
    assert np.allclose(rho_a, rho_o, rtol=1e-13, atol=1e-13 * np.abs(rho_o).max())


@pytest.mark.parametrize('typeB', [0, 1, 2])
def test_boris_push_bitexact(grids, orc, typeB):
    g = grids['rect']
    pd = random_particles(g, 4000, seed=10 + typeB)
    p = _parts(g, pd)
    rng = np.random.default_rng(5)
    eat = rng.standard_normal((4000, 2)) * 3e4
    bat = rng.standard_normal((4000, 3)) * 0.4
    bconst = np.array([0.01, 0.5, -0.02])
    p.eAtParticles[:4000] = torch.as_tensor(eat).cuda()
    p.bAtParticles[:4000] = torch.as_tensor(bat).cuda()
    p.setBConst(bconst)
    qm = p.getParticleCharge() / p.getParticleMass()
    p.borisPush(5e-12, typeBField=typeB)
    ref = orc.boris_push(pd, eat, 5e-12, qm, typeB, bconst if typeB == 1 else bat)
    assert np.array_equal(p.getParticleData().cpu().numpy(), ref)


def test_boris_push_fp32(grids, orc):
    g = grids['rect']
    pd = random_particles(g, 4000, seed=21)
    p = _parts(g, pd, torch.float32)
    rng = np.random.default_rng(6)
    eat = rng.standard_normal((4000, 2)) * 3e4
    bconst = np.array([0., 0.5, 0.])
    p.eAtParticles[:4000] = torch.as_tensor(eat).float().cuda()
    p.setBConst(bconst)
    qm = p.getParticleCharge() / p.getParticleMass()
    p.borisPush(5e-12, typeBField=1)
    ref = orc.boris_push(pd.astype(np.float32).astype(np.float64), eat.astype(np.float32), 5e-12,
                         qm, 1, bconst)
    got = p.getParticleData().cpu().numpy().astype(np.float64)
    scale = np.abs(ref).max(axis=0)
    assert np.all(np.abs(got - ref) <= 1e-5 * scale)


@pytest.mark.parametrize('gname,typeB', [('rect', 0), ('ellip', 1), ('ellip', 2)])
def test_fused_push_matches_unfused(grids, orc, gname, typeB):
    """The fused gather + push + wall + deposit kernel gives the reference step's state."""
    from openecloud_amd import _lib as L
    from openecloud_amd.particleBoundary import AbsorbElliptical, AbsorbRectangular
    import ctypes
    g = grids[gname]
    wall = 'ellip' if gname == 'ellip' else 'rect'
    pd = random_particles(g, 6000, seed=30, wall=wall, vscale=4e6)
    p = _parts(g, pd)
    pb = (AbsorbElliptical if wall == 'ellip' else AbsorbRectangular)(g, p)
    E = _field(g, 31)
    B = _field(g, 32, 3, 0.5)
    bconst = np.array([0., 0.5, 0.])
    p.setBConst(bconst)
    dt = 1e-11
    qm = p.getParticleCharge() / p.getParticleMass()
    rho_next = torch.zeros(g.getNpExt(), dtype=torch.float64, device='cuda')
    Ed = torch.as_tensor(E).cuda()
    Bd = torch.as_tensor(B).cuda()
    gd = L.grid_desc(g)
    b = (ctypes.c_double * 3)(*bconst.tolist())
    L.check(L.lib().oecl_push_fused(ctypes.byref(gd), ctypes.byref(pb.desc), L.i32(p.code),
                                    L.ptr(p._data), L.u64(p.capacity), L.u32(6000), L.ptr(Ed),
                                    L.ptr(Bd), b, L.i32(typeB), L.f64(dt), L.f64(qm),
                                    L.f64(p.getParticleCharge()), L.ptr(p.isInside),
                                    L.ptr(rho_next), L.stream()))
    # oracle: weights -> gather -> push -> inside
    ic, w = orc.calc_grid_weights(pd, g.getDx(), g.getDy(), g.getLxExt(), g.getLyExt(), g.getNxExt())
    eat = orc.field_to_particles(E, ic, w, g.getNxExt(), g.getNpExt(), 2)
    bsel = bconst if typeB == 1 else orc.field_to_particles(B, ic, w, g.getNxExt(), g.getNpExt(), 3)
    ref = orc.boris_push(pd, eat, dt, qm, typeB, bsel)
    ins, _ = orc.index_inside(ref, wall, g.getLx(), g.getLy())
    assert np.array_equal(p.getParticleData().cpu().numpy(), ref)
    assert np.array_equal(p.getIsInside().cpu().numpy().astype(np.uint16), ins)
    assert 0 < ins.sum() < 6000
    kept = ref[ins == 1]
    ic2, w2 = orc.calc_grid_weights(kept, g.getDx(), g.getDy(), g.getLxExt(), g.getLyExt(),
                                    g.getNxExt())
    rho_o = orc.charge_to_grid(kept, ic2, w2, p.getParticleCharge(), g.getNxExt(), g.getNpExt())
    r = rho_next.cpu().numpy()
    assert np.allclose(r, rho_o, rtol=1e-13, atol=1e-13 * np.abs(rho_o).max())
