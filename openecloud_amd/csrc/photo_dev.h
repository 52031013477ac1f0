// Primary photoelectrons at the chamber wall, shared by the standalone emitter entry point and
// the one-call step.  The reference has no photoelectron source (its Readme lists it as an
// outlook); this one is defined as follows, for a rectangular or elliptical wall:
//   position   a point of the wall inset by 1e-3 of a cell along each semi-axis: rectangle
//              uniform in arc length, ellipse uniform in the parametric angle;
//   energy     |E0 + sigmaE z| eV with z standard normal (Box-Muller);
//   direction  about the inward wall normal n with the reference emitter's secondary angles,
//              theta = asin(u), phi = 2 pi u, rotated exactly as particleEmitter.pyx:289-302;
//   speed      sqrt(E / (m/2e)).
// Draws: Philox block t of counter (k, t, step, PURPOSE_PHOTO): block 0 = (position, Box-Muller
// radius), block 1 = (Box-Muller angle, theta), block 2 = (phi, unused).
#pragma once
#include "common.h"
#include "wall.h"

#define PURPOSE_PHOTO 6u

namespace photo {

constexpr double kPi = 3.141592653589793;  // constants.pyx:9
constexpr double kQe = 1.602176565e-19;    // constants.pyx:5

template <typename T>
__device__ inline void emit_one(const WallDev &w, uint32_t k, uint32_t step, uint32_t s0, uint32_t s1, double e0,
                                double sigmaE, double weight, double mass, T *o)
{
    const Philox4 r0 = philox4x32_10(k, 0u, step, PURPOSE_PHOTO, s0, s1);
    const Philox4 r1 = philox4x32_10(k, 1u, step, PURPOSE_PHOTO, s0, s1);
    const Philox4 r2 = philox4x32_10(k, 2u, step, PURPOSE_PHOTO, s0, s1);
    const double uPos = u01_from_words(r0.v[0], r0.v[1]), uRad = u01_from_words(r0.v[2], r0.v[3]);
    const double uAng = u01_from_words(r1.v[0], r1.v[1]), uTh = u01_from_words(r1.v[2], r1.v[3]);
    const double uPh = u01_from_words(r2.v[0], r2.v[1]);
    const double a = 0.5 * w.lx - 1e-3 * w.dx, b = 0.5 * w.ly - 1e-3 * w.dy;
    double x, y, nx, ny;
    if (w.kind == OECL_WALL_ELLIP) {
        const double t = 2. * kPi * uPos;
        x = a * cos(t);
        y = b * sin(t);
        const double gx = x / (a * a), gy = y / (b * b);
        const double inv = 1. / sqrt(gx * gx + gy * gy);
        nx = -gx * inv;
        ny = -gy * inv;
    } else {
        const double s = uPos * 4. * (a + b);
        if (s < 2. * a) {
            x = -a + s; y = -b; nx = 0.; ny = 1.;
        } else if (s < 2. * a + 2. * b) {
            x = a; y = -b + (s - 2. * a); nx = -1.; ny = 0.;
        } else if (s < 4. * a + 2. * b) {
            x = a - (s - 2. * a - 2. * b); y = b; nx = 0.; ny = -1.;
        } else {
            x = -a; y = b - (s - 4. * a - 2. * b); nx = 1.; ny = 0.;
        }
    }
    const double z = sqrt(-2. * log(1. - uRad)) * cos(2. * kPi * uAng);
    const double e = fabs(e0 + sigmaE * z);
    const double v = sqrt(e / (0.5 * mass / kQe));
    const double th = asin(uTh);
    const double ct = cos(th), st = sin(th);
    const double ph = uPh * kPi * 2.;
    const double cp = cos(ph), sp = sin(ph);
    const double t1 = ct * nx - st * ny;
    const double t2 = st * nx + ct * ny;
    o[0] = (T)x;
    o[1] = (T)y;
    o[2] = (T)(v * ((cp + nx * nx * (1 - cp)) * t1 + nx * ny * (1 - cp) * t2));
    o[3] = (T)(v * ((cp + ny * ny * (1 - cp)) * t2 + nx * ny * (1 - cp) * t1));
    o[4] = (T)(v * (-sp * ny * t1 + sp * nx * t2));
    o[5] = (T)weight;
}

}  // namespace photo
