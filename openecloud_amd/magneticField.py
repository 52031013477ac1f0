"""Multipole magnetic field at the grid nodes (magneticField.pyx:23-59), host setup.

B_r = sum_k r^k (-a_k cos((k+1) phi) + b_k sin((k+1) phi)), B_phi likewise, rotated to
(Bx, By); Bz = 0; laid out [Bx | By | Bz] over the extended nodes.  Coefficients are in
T/m^k with no reference radius, as in the reference.
"""
import numpy as np


def multipoleExpansion(gridObj, multipolesIn, skewMultipolesIn=()):
    mult = np.asarray(multipolesIn, dtype=np.float64)
    nm = mult.shape[0]
    skew = np.zeros(nm)
    sk = np.asarray(skewMultipolesIn, dtype=np.float64)
    skew[:min(nm, sk.shape[0])] = sk[:nm] if sk.shape[0] >= nm else sk
    xm = np.asarray(gridObj.getXMesh())
    ym = np.asarray(gridObj.getYMesh())
    nx, ny, npx = gridObj.getNxExt(), gridObj.getNyExt(), gridObj.getNpExt()
    X, Y = np.meshgrid(xm, ym)            # [j, i]
    r = np.sqrt(X ** 2 + Y ** 2)
    phi = np.arctan2(Y, X)
    phi = np.where(phi < 0, phi + 2 * np.pi, phi)
    br = np.zeros_like(r)
    bphi = np.zeros_like(r)
    for k in range(nm):
        rk = r ** k
        br += rk * (-skew[k] * np.cos((k + 1) * phi) + mult[k] * np.sin((k + 1) * phi))
        bphi += rk * (skew[k] * np.sin((k + 1) * phi) + mult[k] * np.cos((k + 1) * phi))
    b = np.empty(3 * npx)
    b[:npx] = (br * np.cos(phi) - bphi * np.sin(phi)).ravel()
    b[npx:2 * npx] = (br * np.sin(phi) + bphi * np.cos(phi)).ravel()
    b[2 * npx:] = 0.
    return b
