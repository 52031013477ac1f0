"""CPU restatement of the primary photoelectron source (test infrastructure only).

The reference has no photoelectron source (its Readme lists photoemission as an outlook), so
this restates the definition the device code implements (openecloud_amd/csrc/photo_dev.h):
points on the wall inset by 1e-3 cell (rectangle uniform in arc length, ellipse uniform in the
parametric angle), energy |E0 + sigmaE z| eV with z from Box-Muller, direction about the inward
normal with the reference emitter's secondary angles and rotation (particleEmitter.pyx:
289-302).  Draws: Philox blocks t = 0, 1, 2 of counter (k, t, step, 6).  The tests pin it by
the geometry (points on the inset wall, unit inward normals), |v|^2 = 2 e E / m, the angle law
and the folded-normal energy moments.
"""
import numpy as np

from . import philox_words, words_to_uniforms

QE = 1.602176565e-19   # constants.pyx:5
PURPOSE_PHOTO = 6


def photo_emit(kind, lx, ly, dx, dy, n, step, k0, k1, e0, sigmaE, weight, mass, with_normals=False):
    """(n, 6) photoelectrons of time step `step`; kind 0 rectangle, 1 ellipse."""
    u = [words_to_uniforms(philox_words(n, 0, t, step, PURPOSE_PHOTO, k0, k1)) for t in range(3)]
    uPos, uRad = u[0][:, 0], u[0][:, 1]
    uAng, uTh = u[1][:, 0], u[1][:, 1]
    uPh = u[2][:, 0]
    a = 0.5 * lx - 1e-3 * dx
    b = 0.5 * ly - 1e-3 * dy
    if kind == 1:
        t = 2. * np.pi * uPos
        x = a * np.cos(t)
        y = b * np.sin(t)
        gx, gy = x / (a * a), y / (b * b)
        inv = 1. / np.sqrt(gx * gx + gy * gy)
        nx, ny = -gx * inv, -gy * inv
    elif kind == 0:
        s = uPos * 4. * (a + b)
        x = np.empty(n)
        y = np.empty(n)
        nx = np.zeros(n)
        ny = np.zeros(n)
        m0 = s < 2. * a
        m1 = ~m0 & (s < 2. * a + 2. * b)
        m2 = ~m0 & ~m1 & (s < 4. * a + 2. * b)
        m3 = ~(m0 | m1 | m2)
        x[m0], y[m0], ny[m0] = -a + s[m0], -b, 1.
        x[m1], y[m1], nx[m1] = a, -b + (s[m1] - 2. * a), -1.
        x[m2], y[m2], ny[m2] = a - (s[m2] - 2. * a - 2. * b), b, -1.
        x[m3], y[m3], nx[m3] = -a, b - (s[m3] - 4. * a - 2. * b), 1.
    else:
        raise ValueError('rectangular or elliptical wall only')
    z = np.sqrt(-2. * np.log(1. - uRad)) * np.cos(2. * np.pi * uAng)
    e = np.abs(e0 + sigmaE * z)
    v = np.sqrt(e / (0.5 * mass / QE))
    th = np.arcsin(uTh)
    ct, st = np.cos(th), np.sin(th)
    ph = uPh * np.pi * 2.
    cp, sp = np.cos(ph), np.sin(ph)
    t1 = ct * nx - st * ny
    t2 = st * nx + ct * ny
    out = np.empty((n, 6))
    out[:, 0] = x
    out[:, 1] = y
    out[:, 2] = v * ((cp + nx * nx * (1 - cp)) * t1 + nx * ny * (1 - cp) * t2)
    out[:, 3] = v * ((cp + ny * ny * (1 - cp)) * t2 + nx * ny * (1 - cp) * t1)
    out[:, 4] = v * (-sp * ny * t1 + sp * nx * t2)
    out[:, 5] = weight
    if with_normals:
        return out, np.stack([nx, ny], axis=1), e
    return out
