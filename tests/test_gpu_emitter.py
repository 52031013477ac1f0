"""GPU Furman emission and loader against the oracle with the same uniforms."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _impacts(n, seed, emax=600.):
    rng = np.random.default_rng(seed)
    from oracle.furman_oracle import ELEMENTARY_CHARGE, M_E
    e = rng.random(n) ** 2 * emax  # eV, weighted to low energies
    v = np.sqrt(2 * e * ELEMENTARY_CHARGE / M_E)
    ang = rng.random(n) * 2 * np.pi
    nx, ny = np.cos(ang), np.sin(ang)
    # velocity into the wall (against the inward normal) with random incidence
    th = (rng.random(n) - 0.5) * np.pi * 0.95
    dirx = -(np.cos(th) * nx - np.sin(th) * ny)
    diry = -(np.sin(th) * nx + np.cos(th) * ny)
    vz = rng.standard_normal(n) * 0.2
    nrm = np.sqrt(dirx ** 2 + diry ** 2 + vz ** 2)
    ap = np.empty((n, 6))
    ap[:, 0] = 0.045 * nx
    ap[:, 1] = 0.035 * ny
    ap[:, 2] = v * dirx / nrm
    ap[:, 3] = v * diry / nrm
    ap[:, 4] = v * vz / nrm
    ap[:, 5] = 1e4 * (1 + rng.random(n))
    nv = np.stack([nx, ny], axis=1)
    return ap, nv


def _emitter(ap, nv, sey=1.6, eHat0=300., dtype=torch.float64):
    from openecloud_amd.grid import Grid
    from openecloud_amd.particleBoundary import AbsorbElliptical
    from openecloud_amd.particleEmitter import FurmanEmitter
    from openecloud_amd.particles import Particles
    g = Grid(32, 32, 0.09, 0.07, 1)
    p = Particles('electrons', 'variableWeight', dtype=dtype)
    pb = AbsorbElliptical(g, p)
    n = ap.shape[0]
    pb.absorbedParticles = torch.as_tensor(ap).to(dtype).cuda()
    pb.normalVectors = torch.as_tensor(nv).to(dtype).cuda()
    pb.absorbedMacroParticleCount = n
    return FurmanEmitter(pb, p, 'copper', seyMax=sey, eHat0=eHat0)


@pytest.mark.parametrize('sey', [1.3, 1.6, 2.0])
def test_furman_injected_uniforms(orc, sey):
    from oracle.furman_oracle import generate_secondaries, material
    n = 3000
    ap, nv = _impacts(n, seed=int(sey * 10))
    U = np.random.default_rng(99).random((n, 35))
    em = _emitter(ap, nv, sey)
    sec = em.generateSecondaries(uniforms=U).cpu().numpy()
    nsec = em.getNSec().cpu().numpy()
    ref, nsec_o, _ = generate_secondaries(ap, nv, U, material('copper', sey, eHat0=300.))
    assert np.array_equal(nsec, nsec_o)
    assert sec.shape == ref.shape and sec.shape[0] > n // 2
    assert np.array_equal(sec[:, [0, 1, 5]], ref[:, [0, 1, 5]])
    scale = np.abs(ref[:, 2:5]).max()
    assert np.allclose(sec[:, 2:5], ref[:, 2:5], rtol=1e-9, atol=1e-12 * scale)


def _secem(ap, nv, model, sey=2.0, dtype=torch.float64):
    from openecloud_amd.grid import Grid
    from openecloud_amd.particleBoundary import AbsorbElliptical
    from openecloud_amd.particleEmitter import SecElecEmitter, SimpleSecEmitter
    from openecloud_amd.particles import Particles
    g = Grid(32, 32, 0.09, 0.07, 1)
    p = Particles('electrons', 'variableWeight', dtype=dtype)
    pb = AbsorbElliptical(g, p)
    pb.absorbedParticles = torch.as_tensor(ap).to(dtype).cuda()
    pb.normalVectors = torch.as_tensor(nv).to(dtype).cuda()
    pb.absorbedMacroParticleCount = ap.shape[0]
    cls = SecElecEmitter if model == 'stainless' else SimpleSecEmitter
    return cls(pb, p, seyMax=sey)


@pytest.mark.parametrize('model', ['stainless', 'simple'])
@pytest.mark.parametrize('dtype', ['f64', 'f32'])
def test_secem_injected_uniforms(orc, model, dtype):
    """SecElecEmitter / SimpleSecEmitter on the device against the oracle with the same table:
    counts exact, positions and weights exact, velocities to 1e-12 (FP64) / 1e-6 (FP32) of
    the largest.  Impacts include exact normal incidence (the reference's NaN angle)."""
    from oracle import secem_oracle as so
    n = 4000
    ap, nv = _impacts(n, seed=41 if model == 'stainless' else 43, emax=1500.)
    ap[:50, 2] = -np.sqrt(2 * 300. * so.ELEMENTARY_CHARGE / so.M_E) * nv[:50, 0]
    ap[:50, 3] = -np.sqrt(2 * 300. * so.ELEMENTARY_CHARGE / so.M_E) * nv[:50, 1]
    ap[:50, 4] = 0.
    ap[50:100, 0:5] = ap[50:100, 0:5] * 1e-3      # slow impacts: the true-energy fallback
    tdt = torch.float64 if dtype == 'f64' else torch.float32
    if dtype == 'f32':
        ap = ap.astype(np.float32).astype(np.float64)
        nv = nv.astype(np.float32).astype(np.float64)
    U = np.random.default_rng(5).random((n, so.N_UNIFORMS))
    em = _secem(ap, nv, model, dtype=tdt)
    sec = em.generateSecondaries(uniforms=U).cpu().numpy().astype(np.float64)
    P = so.stainless(seyMax=2.0)
    if dtype == 'f64':
        ref, ns = so.generate_secondaries(ap, nv, U, model, P)
        assert np.array_equal(em.getNSec().cpu().numpy(), ns)
        assert sec.shape == ref.shape and sec.shape[0] > n // 2
        assert np.array_equal(sec[:, [0, 1, 5]], ref[:, [0, 1, 5]])
        scale = np.abs(ref[:, 2:5]).max()
        assert np.allclose(sec[:, 2:5], ref[:, 2:5], rtol=1e-12, atol=1e-13 * scale)
    else:
        # FP32 storage: the device computes in FP64 from the FP32 values, as the oracle does
        ref, ns = so.generate_secondaries(ap, nv, U, model, P)
        assert np.array_equal(em.getNSec().cpu().numpy(), ns)
        assert sec.shape == ref.shape
        scale = np.abs(ref[:, 2:5]).max()
        assert np.allclose(sec, ref.astype(np.float32).astype(np.float64), rtol=1e-6, atol=1e-6 * scale)


@pytest.mark.parametrize('model', ['stainless', 'simple'])
def test_secem_philox_stream_matches_oracle(orc, model):
    """Device Philox draws (purpose 7) equal the oracle's words for the same (key, counter)."""
    from oracle import secem_oracle as so
    from openecloud_amd import randomGen
    n = 2000
    ap, nv = _impacts(n, seed=9, emax=1000.)
    randomGen.seed(4321)
    k0, k1 = randomGen.key()
    em = _secem(ap, nv, model)
    sec = em.generateSecondaries(step=23).cpu().numpy()
    U = np.empty((n, so.N_UNIFORMS))
    for b in range(so.N_UNIFORMS // 2):
        U[:, 2 * b:2 * b + 2] = orc.words_to_uniforms(orc.philox_words(n, 0, b, 23, 7, k0, k1))
    ref, ns = so.generate_secondaries(ap, nv, U, model, so.stainless(seyMax=2.0))
    assert np.array_equal(em.getNSec().cpu().numpy(), ns)
    scale = np.abs(ref[:, 2:5]).max()
    assert np.allclose(sec[:, 2:5], ref[:, 2:5], rtol=1e-12, atol=1e-13 * scale)


def test_secem_rejects_yields_beyond_eight_true_secondaries():
    from openecloud_amd.particleEmitter import SecElecEmitter
    ap, nv = _impacts(4, seed=1)
    _secem(ap, nv, 'stainless', sey=5.5)
    with pytest.raises(ValueError):
        _secem(ap, nv, 'stainless', sey=6.0)
    assert SecElecEmitter.stainlessMod(0.6, -1.)['deltaTSHat'] == 1e-6


def test_furman_philox_stream_matches_oracle(orc):
    """Device Philox draws equal the oracle's words for the same (key, counter)."""
    from oracle.furman_oracle import generate_secondaries, material
    from openecloud_amd import randomGen
    n = 1500
    ap, nv = _impacts(n, seed=5)
    randomGen.seed(1234)
    k0, k1 = randomGen.key()
    em = _emitter(ap, nv, 1.6)
    sec = em.generateSecondaries(step=17).cpu().numpy()
    U = np.empty((n, 36))
    for b in range(18):  # block b of every impact j: counter (j, b, step, purpose 1)
        U[:, 2 * b:2 * b + 2] = orc.words_to_uniforms(orc.philox_words(n, 0, b, 17, 1, k0, k1))
    U = U[:, :35]
    ref, nsec_o, _ = generate_secondaries(ap, nv, U, material('copper', 1.6, eHat0=300.))
    assert np.array_equal(em.getNSec().cpu().numpy(), nsec_o)
    scale = np.abs(ref[:, 2:5]).max()
    assert np.allclose(sec[:, 2:5], ref[:, 2:5], rtol=1e-9, atol=1e-12 * scale)


def test_loader_matches_restatement(orc):
    from openecloud_amd import randomGen
    from openecloud_amd.grid import Grid
    from openecloud_amd.particleBoundary import AbsorbElliptical
    from openecloud_amd.particleEmitter import homoLoader
    from openecloud_amd.particles import Particles
    g = Grid(64, 48, 0.09, 0.07, 1)
    p = Particles('electrons', 'variableWeight')
    pb = AbsorbElliptical(g, p)
    randomGen.seed(7)
    k0, k1 = randomGen.key()
    n = 2000
    vs = 4.194e5
    homoLoader(g, p, pb, n, 1e4, vs, randomizeWeight=1)
    got = p.getParticleData().cpu().numpy()
    lx, ly = g.getLx(), g.getLy()
    ref = np.empty((n, 6))
    for i in range(n):
        for t in range(64):
            u = orc.words_to_uniforms(orc.philox_words(1, i, t, 0, 2, k0, k1))[0]
            x, y = (u[0] - 0.5) * lx, (u[1] - 0.5) * ly
            if not (x * x * (4. / (lx * lx)) + y * y * (4. / (ly * ly)) >= 1.):
                break
        a = orc.words_to_uniforms(orc.philox_words(1, i, 64, 0, 2, k0, k1))[0]
        b = orc.words_to_uniforms(orc.philox_words(1, i, 65, 0, 2, k0, k1))[0]
        c = orc.words_to_uniforms(orc.philox_words(1, i, 66, 0, 2, k0, k1))[0]
        r0, r1 = np.sqrt(-2 * np.log(1 - a[0])), np.sqrt(-2 * np.log(1 - b[0]))
        ref[i] = [x, y, r0 * np.cos(2 * np.pi * a[1]) * vs, r0 * np.sin(2 * np.pi * a[1]) * vs,
                  r1 * np.cos(2 * np.pi * b[1]) * vs, 1e4 * (0.5 + c[0])]
    assert np.array_equal(got[:, [0, 1, 5]], ref[:, [0, 1, 5]])
    assert np.allclose(got[:, 2:5], ref[:, 2:5], rtol=1e-13, atol=1e-8)


@pytest.mark.parametrize('kind', ['rect', 'ellip'])
@pytest.mark.parametrize('dtype', ['f64', 'f32'])
def test_photo_emit_matches_oracle(kind, dtype):
    """oecl_photo_emit against the photoelectron restatement with the same Philox draws."""
    import ctypes
    from openecloud_amd import _lib as L
    from openecloud_amd import randomGen
    from openecloud_amd.grid import Grid
    from openecloud_amd.particleBoundary import AbsorbElliptical, AbsorbRectangular
    from openecloud_amd.particles import Particles
    from oracle.photo_oracle import photo_emit
    g = Grid(128, 96, 0.09, 0.07, 1 if kind == 'ellip' else 0)
    tdt = torch.float64 if dtype == 'f64' else torch.float32
    p = Particles('electrons', 'variableWeight', dtype=tdt)
    p.setParticleData(np.zeros((1, 6)))
    pb = (AbsorbElliptical if kind == 'ellip' else AbsorbRectangular)(g, p)
    randomGen.seed(5)
    k0, k1 = randomGen.key()
    n, step = 30000, 17
    out = torch.empty((n, 6), dtype=tdt, device='cuda')
    L.check(L.lib().oecl_photo_emit(ctypes.byref(pb.desc), L.i32(p.code), L.ptr(out), L.u32(n),
                                    L.u32(step), L.u32(k0), L.u32(k1), L.f64(7.), L.f64(5.),
                                    L.f64(3e4), L.f64(9.10938291e-31), L.stream()))
    got = out.cpu().numpy().astype(np.float64)
    ref = photo_emit(1 if kind == 'ellip' else 0, g.getLx(), g.getLy(), g.getDx(), g.getDy(), n,
                     step, k0, k1, 7., 5., 3e4, 9.10938291e-31)
    tol = 1e-12 if dtype == 'f64' else 1e-6
    scale = np.abs(ref).max(axis=0)
    assert (np.abs(got - ref) / scale).max() < tol
    # every photoelectron starts inside the chamber (the walls' inside tests)
    x, y = got[:, 0], got[:, 1]
    if kind == 'ellip':
        assert np.all(x ** 2 * (4. / g.getLx() ** 2) + y ** 2 * (4. / g.getLy() ** 2) < 1.)
    else:
        assert np.all((np.abs(x) <= 0.5 * g.getLx()) & (np.abs(y) <= 0.5 * g.getLy()))
