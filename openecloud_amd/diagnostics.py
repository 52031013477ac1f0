"""Particle-set diagnostics on the device (FP64, fixed reduction order).

weight_sum is the physical line density (electrons per metre) the reference's drivers record
with numpy.sum(getParticleData()[:, 5]) (example01_basic.py:56); mean_energy_eV is
util.averageEnergy (util.pyx:8-20), weighted by the macro-particle weights.
"""
import torch

from . import _lib as L
from .constants import elementary_charge


def moments(p):
    """(sum w, sum w v^2, sum w^2 v^2, sum w x, sum w y, count) of Particles p."""
    dev = L.device()
    out = torch.empty(6, dtype=torch.float64, device=dev)
    work = torch.empty(512 * 6, dtype=torch.float64, device=dev)
    L.check(L.lib().oecl_particle_moments(L.i32(p.code), L.ptr(p._data), L.u64(p.capacity),
                                          L.u32(p.macroParticleCount), L.ptr(out), L.ptr(work),
                                          L.stream()), 'moments')
    return out.cpu().tolist()


def weight_sum(p):
    return moments(p)[0]


def mean_energy_eV(p):
    m = moments(p)
    if m[0] == 0.:
        return 0.
    return 0.5 * p.getParticleMass() * m[1] / m[0] / elementary_charge
