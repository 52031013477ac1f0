/*
 * CPU ORACLE -- test infrastructure only.
 *
 * A single-threaded C restatement of openECLOUD's per-particle loops, in the
 * reference's own AoS layout (double particleData[N][6]: x, y, vx, vy, vz, w)
 * and the reference's own loop and arithmetic order.  It is compiled with
 * -ffp-contract=off so that no multiply-add is fused: the Cython-generated C
 * of the reference is built for plain x86-64, which has no FMA.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load
 * this file's library, and only as the checker or as the timed CPU baseline.
 * The product (openecloud_amd) never calls it.
 *
 * Every function names the reference lines it restates.  How each is pinned
 * (golden vectors, closed forms, invariants) is listed in oracle/README.md.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#define NC 6

/* particles.pyx:400-418 (_calcGridWeights) */
void orc_calc_grid_weights(const double *pd, uint32_t n, double dxi, double dyi,
                           double lxHalf, double lyHalf, uint32_t nx, uint32_t *inCell,
                           double *wLL, double *wUL, double *wLU, double *wUU)
{
    for (uint32_t ii = 0; ii < n; ++ii) {
        double fx = (pd[NC * ii] + lxHalf) * dxi;
        double fy = (pd[NC * ii + 1] + lyHalf) * dyi;
        uint32_t ix = (uint32_t)fx;
        uint32_t iy = (uint32_t)fy;
        inCell[ii] = ix + iy * nx;
        fx = fx - ix;
        fy = fy - iy;
        wLL[ii] = (1 - fx) * (1 - fy);
        wUL[ii] = fx * (1 - fy);
        wLU[ii] = (1 - fx) * fy;
        wUU[ii] = fx * fy;
    }
}

/* particles.pyx:336-351 (_chargeToGrid) */
void orc_charge_to_grid(const double *pd, const uint32_t *inCell, const double *wLL,
                        const double *wUL, const double *wLU, const double *wUU,
                        double *rho, double q, uint32_t nx, uint32_t np, uint32_t n)
{
    for (uint32_t ii = 0; ii < np; ++ii) rho[ii] = 0.;
    for (uint32_t ii = 0; ii < n; ++ii) {
        uint32_t ind = inCell[ii];
        rho[ind] += wLL[ii] * pd[NC * ii + 5] * q;
        ind = ind + 1;
        rho[ind] += wUL[ii] * pd[NC * ii + 5] * q;
        ind = ind - 1 + nx;
        rho[ind] += wLU[ii] * pd[NC * ii + 5] * q;
        ind = ind + 1;
        rho[ind] += wUU[ii] * pd[NC * ii + 5] * q;
    }
}

/* particles.pyx:354-372 (_fieldToParticles) */
void orc_field_to_particles(const double *f, double *out, const uint32_t *inCell,
                            const double *wLL, const double *wUL, const double *wLU,
                            const double *wUU, uint32_t nx, uint32_t np, uint32_t n)
{
    for (uint32_t ii = 0; ii < n; ++ii) {
        uint32_t ind = inCell[ii];
        out[2 * ii] = wLL[ii] * f[ind];
        out[2 * ii + 1] = wLL[ii] * f[ind + np];
        ind = ind + 1;
        out[2 * ii] += wUL[ii] * f[ind];
        out[2 * ii + 1] += wUL[ii] * f[ind + np];
        ind = ind - 1 + nx;
        out[2 * ii] += wLU[ii] * f[ind];
        out[2 * ii + 1] += wLU[ii] * f[ind + np];
        ind = ind + 1;
        out[2 * ii] += wUU[ii] * f[ind];
        out[2 * ii + 1] += wUU[ii] * f[ind + np];
    }
}

/* particles.pyx:375-397 (_fieldToParticles3Dims) */
void orc_field_to_particles3(const double *f, double *out, const uint32_t *inCell,
                             const double *wLL, const double *wUL, const double *wLU,
                             const double *wUU, uint32_t nx, uint32_t np, uint32_t n)
{
    for (uint32_t ii = 0; ii < n; ++ii) {
        uint32_t ind = inCell[ii];
        out[3 * ii] = wLL[ii] * f[ind];
        out[3 * ii + 1] = wLL[ii] * f[ind + np];
        out[3 * ii + 2] = wLL[ii] * f[ind + 2 * np];
        ind = ind + 1;
        out[3 * ii] += wUL[ii] * f[ind];
        out[3 * ii + 1] += wUL[ii] * f[ind + np];
        out[3 * ii + 2] += wUL[ii] * f[ind + 2 * np];
        ind = ind - 1 + nx;
        out[3 * ii] += wLU[ii] * f[ind];
        out[3 * ii + 1] += wLU[ii] * f[ind + np];
        out[3 * ii + 2] += wLU[ii] * f[ind + 2 * np];
        ind = ind + 1;
        out[3 * ii] += wUU[ii] * f[ind];
        out[3 * ii + 1] += wUU[ii] * f[ind + np];
        out[3 * ii + 2] += wUU[ii] * f[ind + 2 * np];
    }
}

/* particles.pyx:272-295 (_borisPush, per-particle B); cdtmh = 0.5*q/m*dt */
void orc_boris_push(double *pd, const double *e, const double *b, double dt, double cdtmh,
                    uint32_t n)
{
    for (uint32_t ii = 0; ii < n; ++ii) {
        double *p = pd + NC * ii;
        p[2] += cdtmh * e[2 * ii];
        p[3] += cdtmh * e[2 * ii + 1];
        double t0 = cdtmh * b[ii * 3];
        double t1 = cdtmh * b[ii * 3 + 1];
        double t2 = cdtmh * b[ii * 3 + 2];
        double vs0 = p[2] + p[3] * t2 - p[4] * t1;
        double vs1 = p[3] + p[4] * t0 - p[2] * t2;
        double vs2 = p[4] + p[2] * t1 - p[3] * t0;
        double f = 2. / (1. + t0 * t0 + t1 * t1 + t2 * t2);
        t0 = t0 * f;
        t1 = t1 * f;
        t2 = t2 * f;
        p[2] += vs1 * t2 - vs2 * t1 + cdtmh * e[2 * ii];
        p[3] += vs2 * t0 - vs0 * t2 + cdtmh * e[2 * ii + 1];
        p[4] += vs0 * t1 - vs1 * t0;
        p[0] += dt * p[2];
        p[1] += dt * p[3];
    }
}

/* particles.pyx:298-321 (_borisPushConstB) */
void orc_boris_push_const_b(double *pd, const double *e, const double *b, double dt,
                            double cdtmh, uint32_t n)
{
    double t0 = cdtmh * b[0];
    double t1 = cdtmh * b[1];
    double t2 = cdtmh * b[2];
    for (uint32_t ii = 0; ii < n; ++ii) {
        double *p = pd + NC * ii;
        p[2] += cdtmh * e[2 * ii];
        p[3] += cdtmh * e[2 * ii + 1];
        double vs0 = p[2] + p[3] * t2 - p[4] * t1;
        double vs1 = p[3] + p[4] * t0 - p[2] * t2;
        double vs2 = p[4] + p[2] * t1 - p[3] * t0;
        double f = 2. / (1. + t0 * t0 + t1 * t1 + t2 * t2);
        double s0 = t0 * f, s1 = t1 * f, s2 = t2 * f;
        p[2] += vs1 * s2 - vs2 * s1 + cdtmh * e[2 * ii];
        p[3] += vs2 * s0 - vs0 * s2 + cdtmh * e[2 * ii + 1];
        p[4] += vs0 * s1 - vs1 * s0;
        p[0] += dt * p[2];
        p[1] += dt * p[3];
    }
}

/* particles.pyx:324-333 (_borisPushNoB); cdtm = q/m*dt (full kick, particles.pyx:84) */
void orc_boris_push_no_b(double *pd, const double *e, double dt, double cdtm, uint32_t n)
{
    for (uint32_t ii = 0; ii < n; ++ii) {
        double *p = pd + NC * ii;
        p[2] += cdtm * e[2 * ii];
        p[3] += cdtm * e[2 * ii + 1];
        p[0] += dt * p[2];
        p[1] += dt * p[3];
    }
}

/* particleBoundary.pyx:84-98 (AbsorbRectangular.indexInside); lx, ly un-extended */
void orc_inside_rect(const double *pd, uint32_t n, double lx, double ly, uint16_t *isInside)
{
    double lxHalf = lx * 0.5, lyHalf = ly * 0.5;
    for (uint32_t ii = 0; ii < n; ++ii) {
        if (pd[NC * ii] < -lxHalf || pd[NC * ii] > lxHalf || pd[NC * ii + 1] < -lyHalf ||
            pd[NC * ii + 1] > lyHalf)
            isInside[ii] = 0;
        else
            isInside[ii] = 1;
    }
}

/* particleBoundary.pyx:227-242 (AbsorbElliptical.indexInside) */
void orc_inside_ellip(const double *pd, uint32_t n, double lx, double ly, uint16_t *isInside)
{
    double ax = 4. / (lx * lx), ay = 4. / (ly * ly);
    for (uint32_t ii = 0; ii < n; ++ii) {
        if ((pd[NC * ii] * pd[NC * ii] * ax + pd[NC * ii + 1] * pd[NC * ii + 1] * ay) >= 1.)
            isInside[ii] = 0;
        else
            isInside[ii] = 1;
    }
}

/* particleBoundary.pyx:485-508 (_clipParticleDataToGrid); returns the clip flag */
int orc_clip_to_grid(double *pd, uint32_t n, double lxHalf, double lyHalf)
{
    int clip = 0;
    for (uint32_t ii = 0; ii < n; ++ii) {
        if (pd[NC * ii] > lxHalf) { pd[NC * ii] = lxHalf; clip = 1; }
        else if (pd[NC * ii] < -lxHalf) { pd[NC * ii] = -lxHalf; clip = 1; }
        if (pd[NC * ii + 1] > lyHalf) { pd[NC * ii + 1] = lyHalf; clip = 1; }
        else if (pd[NC * ii + 1] < -lyHalf) { pd[NC * ii + 1] = -lyHalf; clip = 1; }
    }
    return clip;
}

/* particleBoundary.pyx:333-370 (AbsorbArbitraryCutCell.indexInside), clip first.
 * lxHalfExt = 0.5*lxExt; bp = boundaryPoints[nb][2]; cpi = cutCellPointsInd[np][2] */
int orc_inside_cutcell(double *pd, uint32_t n, double lxHalfExt, double lyHalfExt, double dx,
                       double dy, uint32_t nx, const uint16_t *insideFaces, const double *bp,
                       const int32_t *cpi, uint16_t *isInside)
{
    int clip = orc_clip_to_grid(pd, n, lxHalfExt - 1.e-6 * dx, lyHalfExt - 1.e-6 * dy);
    double dxi = 1. / dx, dyi = 1. / dy;
    for (uint32_t ii = 0; ii < n; ++ii) {
        uint32_t ix = (uint32_t)((pd[NC * ii] + lxHalfExt) * dxi);
        uint32_t iy = (uint32_t)((pd[NC * ii + 1] + lyHalfExt) * dyi);
        uint32_t c = ix + iy * nx;
        if (insideFaces[c] == 8)
            isInside[ii] = 1;
        else if (insideFaces[c] == 0)
            isInside[ii] = 0;
        else {
            int32_t a = cpi[2 * c], b = cpi[2 * c + 1];
            if ((bp[2 * a] - pd[NC * ii]) * (bp[2 * b + 1] - pd[NC * ii + 1]) >
                (bp[2 * b] - pd[NC * ii]) * (bp[2 * a + 1] - pd[NC * ii + 1]))
                isInside[ii] = 1;
            else
                isInside[ii] = 0;
        }
    }
    return clip;
}

/* particleBoundary.pyx:38-65 (saveAbsorbed); returns absorbed count */
uint32_t orc_save_absorbed(const double *pd, uint32_t n, const uint16_t *isInside, double *abs_out)
{
    uint32_t na = n;
    for (uint32_t ii = 0; ii < n; ++ii) na -= isInside[ii];
    uint32_t ii = 0, jj = 0;
    while (ii < na) {
        if (isInside[jj] == 0) {
            memcpy(abs_out + NC * ii, pd + NC * jj, NC * sizeof(double));
            ii += 1;
        }
        jj += 1;
    }
    return na;
}

/* particleBoundary.pyx:108-192 (AbsorbRectangular.calculateInteractionPoint).  cornerFix != 0
 * relabels the corner cases with the wall crossed first (x wall when tx > ty, else y wall);
 * cornerFix == 0 keeps the reference's labels. */
void orc_interaction_rect(double *ap, uint32_t na, double lx, double ly, double dx, double dy,
                          double *rts, uint16_t *leftWhere, int cornerFix)
{
    double eps = 1e-6;
    double lxHalf = lx * 0.5 - dx * eps, lyHalf = ly * 0.5 - dy * eps;
    for (uint32_t ii = 0; ii < na; ++ii) {
        double *p = ap + NC * ii;
        int xi = 1, yi = 1;
        double tx, ty;
        if (p[0] > lxHalf) xi = 2;
        else if (p[0] < -lxHalf) xi = 0;
        if (p[1] > lyHalf) yi = 2;
        else if (p[1] < -lyHalf) yi = 0;
        if (xi == 0 && yi == 1) { rts[ii] = (p[0] + lxHalf) / p[2]; leftWhere[ii] = 0; }
        else if (xi == 2 && yi == 1) { rts[ii] = (p[0] - lxHalf) / p[2]; leftWhere[ii] = 1; }
        else if (xi == 1 && yi == 0) { rts[ii] = (p[1] + lyHalf) / p[3]; leftWhere[ii] = 2; }
        else if (xi == 1 && yi == 2) { rts[ii] = (p[1] - lyHalf) / p[3]; leftWhere[ii] = 3; }
        else if (xi == 0 && yi == 0) {
            tx = (p[0] + lxHalf) / p[2]; ty = (p[1] + lyHalf) / p[3];
            if (tx > ty) { rts[ii] = tx; leftWhere[ii] = 0; } else { rts[ii] = ty; leftWhere[ii] = 1; }
        } else if (xi == 2 && yi == 0) {
            tx = (p[0] - lxHalf) / p[2]; ty = (p[1] + lyHalf) / p[3];
            if (tx > ty) { rts[ii] = tx; leftWhere[ii] = 2; } else { rts[ii] = ty; leftWhere[ii] = 1; }
        } else if (xi == 0 && yi == 2) {
            tx = (p[0] + lxHalf) / p[2]; ty = (p[1] - lyHalf) / p[3];
            if (tx > ty) { rts[ii] = tx; leftWhere[ii] = 0; } else { rts[ii] = ty; leftWhere[ii] = 3; }
        } else {
            tx = (p[0] - lxHalf) / p[2]; ty = (p[1] - lyHalf) / p[3];
            if (tx > ty) { rts[ii] = tx; leftWhere[ii] = 2; } else { rts[ii] = ty; leftWhere[ii] = 3; }
        }
        if (cornerFix && xi != 1 && yi != 1)
            leftWhere[ii] = (tx > ty) ? (xi == 0 ? 0 : 1) : (yi == 0 ? 2 : 3);
        p[0] -= rts[ii] * p[2];
        p[1] -= rts[ii] * p[3];
    }
}

/* particleBoundary.pyx:195-221 (AbsorbRectangular.calculateNormalVectors) */
void orc_normals_rect(uint32_t na, const uint16_t *leftWhere, double *nv)
{
    for (uint32_t ii = 0; ii < na; ++ii) {
        if (leftWhere[ii] == 0) { nv[2 * ii] = 1.; nv[2 * ii + 1] = 0.; }
        else if (leftWhere[ii] == 1) { nv[2 * ii] = -1.; nv[2 * ii + 1] = 0.; }
        else if (leftWhere[ii] == 2) { nv[2 * ii] = 0.; nv[2 * ii + 1] = 1.; }
        else if (leftWhere[ii] == 3) { nv[2 * ii] = 0.; nv[2 * ii + 1] = -1.; }
    }
}

/* particleBoundary.pyx:252-274 (AbsorbElliptical.calculateInteractionPoint) */
void orc_interaction_ellip(double *ap, uint32_t na, double lx, double ly, double dx, double dy,
                           double *rts)
{
    double eps = 1.e-6;
    double ax = 4. / ((lx - eps * dx) * (lx - eps * dx));
    double ay = 4. / ((ly - eps * dy) * (ly - eps * dy));
    for (uint32_t ii = 0; ii < na; ++ii) {
        double *p = ap + NC * ii;
        double c = p[0] * p[0] * ax + p[1] * p[1] * ay - 1.;
        double b = p[0] * p[2] * ax + p[1] * p[3] * ay;
        double a = p[2] * p[2] * ax + p[3] * p[3] * ay;
        rts[ii] = (b - sqrt(b * b - a * c)) / a;
        p[0] -= rts[ii] * p[2];
        p[1] -= rts[ii] * p[3];
    }
}

/* particleBoundary.pyx:277-297 (AbsorbElliptical.calculateNormalVectors) */
void orc_normals_ellip(const double *ap, uint32_t na, double lx, double ly, double *nv)
{
    double ax = 4. / (lx * lx), ay = 4. / (ly * ly);
    double ax2 = ax * ax, ay2 = ay * ay;
    for (uint32_t ii = 0; ii < na; ++ii) {
        const double *p = ap + NC * ii;
        double ninv = 1. / sqrt(p[0] * p[0] * ax2 + p[1] * p[1] * ay2);
        nv[2 * ii] = -p[0] * ax * ninv;
        nv[2 * ii + 1] = -p[1] * ay * ninv;
    }
}

/* particleBoundary.pyx:373-457 (AbsorbArbitraryCutCell.calculateInteractionPoint).
 * Returns the number of particles whose crossing was not found (the reference prints). */
uint32_t orc_interaction_cutcell(double *ap, uint32_t na, double lxHalfExt, double lyHalfExt,
                                 double dx, double dy, uint32_t nx, double dt,
                                 const uint16_t *insideFaces, const double *bp,
                                 const int32_t *cpi, double *rts)
{
    double dxi = 1. / dx, dyi = 1. / dy, se = 1.e-6;
    uint32_t notFound = 0;
    for (uint32_t ii = 0; ii < na; ++ii) {
        double *p = ap + NC * ii;
        uint32_t ix = (uint32_t)((p[0] + lxHalfExt) * dxi);
        uint32_t iy = (uint32_t)((p[1] + lyHalfExt) * dyi);
        uint32_t inCell = ix + iy * nx;
        double rx = -dt * p[2], ry = -dt * p[3];
        int xDir[2] = {0, rx < 0. ? -1 : 1};
        int yDir[2] = {0, ry < 0. ? -1 : 1};
        int stop = 0;
        for (int a = 0; a < 2 && !stop; ++a) {
            for (int b = 0; b < 2; ++b) {
                uint32_t c = inCell + xDir[a] + yDir[b] * nx;
                if (insideFaces[c] != 0 && insideFaces[c] != 8) {
                    int32_t i0 = cpi[2 * c], i1 = cpi[2 * c + 1];
                    double sx = bp[2 * i1] - bp[2 * i0];
                    double sy = bp[2 * i1 + 1] - bp[2 * i0 + 1];
                    double rs = rx * sy - ry * sx;
                    if (rs != 0.) {
                        double qx = bp[2 * i0] - p[0];
                        double qy = bp[2 * i0 + 1] - p[1];
                        double t = (qx * sy - qy * sx) / rs;
                        double u = (qx * ry - qy * rx) / rs;
                        if (-se <= t && t <= 1. + se && -se <= u && u <= 1. + se) {
                            p[0] += t * rx * (1. + se);
                            p[1] += t * ry * (1. + se);
                            rts[ii] = t * (1. + se) * dt;
                            stop = 1;
                            break;
                        }
                    }
                }
            }
        }
        if (!stop) notFound++;
    }
    return notFound;
}

/* particleBoundary.pyx:460-481 (AbsorbArbitraryCutCell.calculateNormalVectors) */
void orc_normals_cutcell(const double *ap, uint32_t na, double lxHalfExt, double lyHalfExt,
                         double dx, double dy, uint32_t nx, const double *ccn, double *nv)
{
    double dxi = 1. / dx, dyi = 1. / dy;
    for (uint32_t ii = 0; ii < na; ++ii) {
        uint32_t ix = (uint32_t)((ap[NC * ii] + lxHalfExt) * dxi);
        uint32_t iy = (uint32_t)((ap[NC * ii + 1] + lyHalfExt) * dyi);
        uint32_t c = ix + iy * nx;
        nv[2 * ii] = ccn[2 * c];
        nv[2 * ii + 1] = ccn[2 * c + 1];
    }
}

/* particles.pyx:201-219 (_addAndRemoveParticlesNewArray) */
void orc_add_remove_new_array(const double *full, const double *add, const uint16_t *keep,
                              uint32_t addCount, uint32_t keepCount, double *out)
{
    uint32_t ii = 0, jj = 0;
    while (jj < keepCount) {
        if (keep[ii] == 1) {
            memcpy(out + NC * jj, full + NC * ii, NC * sizeof(double));
            jj += 1;
        }
        ii += 1;
    }
    for (ii = 0; ii < addCount; ++ii)
        memcpy(out + NC * (keepCount + ii), add + NC * ii, NC * sizeof(double));
}

/* particles.pyx:222-252 (_addAndRemoveParticles), in place */
void orc_add_remove(double *full, const double *add, const uint16_t *keep, uint32_t keepCount,
                    uint32_t addCount, uint32_t oldCount)
{
    uint32_t ii = 0, jj = 0;
    while (ii < addCount && jj < oldCount) {
        if (keep[jj] == 0) {
            memcpy(full + NC * jj, add + NC * ii, NC * sizeof(double));
            ii += 1;
        }
        jj += 1;
    }
    if (ii == addCount) {
        ii = jj - addCount;
        while (ii < keepCount) {
            if (keep[jj] == 1) {
                memmove(full + NC * (ii + addCount), full + NC * jj, NC * sizeof(double));
                ii += 1;
            }
            jj += 1;
        }
    } else if (jj == oldCount) {
        while (ii < addCount) {
            memcpy(full + NC * jj, add + NC * ii, NC * sizeof(double));
            ii += 1;
            jj += 1;
        }
    }
}

/* particles.pyx:254-269 (_removeParticles), in place */
void orc_remove(double *full, const uint16_t *keep, uint32_t oldCount, uint32_t newCount)
{
    uint32_t ll = oldCount, jj = 0;
    while (jj < newCount) {
        if (keep[jj] == 0) {
            for (;;) {
                ll -= 1;
                if (keep[ll] == 1) break;
            }
            memcpy(full + NC * jj, full + NC * ll, NC * sizeof(double));
        }
        jj += 1;
    }
}

/* One whole reference time step of the particle side with frozen nodal fields, in the
 * order of example01_basic.py:66-97 minus the Poisson solve and emission: weights, gather
 * (E, and B when typeB==2), push, rectangular/elliptical inside test.  Used by the
 * CPU baseline timing.  wall: 0 = rectangular, 1 = elliptical. */
void orc_particle_step(double *pd, uint32_t n, const double *E, const double *B, int typeB,
                       const double *bConst, double dt, double qm, double dxi, double dyi,
                       double lxHalfExt, double lyHalfExt, uint32_t nx, uint32_t np, int wall,
                       double lx, double ly, uint32_t *inCell, double *w4, double *eAt,
                       double *bAt, uint16_t *isInside, double *rho, double q)
{
    double *wLL = w4, *wUL = w4 + n, *wLU = w4 + 2 * (size_t)n, *wUU = w4 + 3 * (size_t)n;
    orc_calc_grid_weights(pd, n, dxi, dyi, lxHalfExt, lyHalfExt, nx, inCell, wLL, wUL, wLU, wUU);
    orc_charge_to_grid(pd, inCell, wLL, wUL, wLU, wUU, rho, q, nx, np, n);
    orc_field_to_particles(E, eAt, inCell, wLL, wUL, wLU, wUU, nx, np, n);
    if (typeB == 0) {
        orc_boris_push_no_b(pd, eAt, dt, qm * dt, n);
    } else if (typeB == 1) {
        orc_boris_push_const_b(pd, eAt, bConst, dt, 0.5 * qm * dt, n);
    } else {
        orc_field_to_particles3(B, bAt, inCell, wLL, wUL, wLU, wUU, nx, np, n);
        orc_boris_push(pd, eAt, bAt, dt, 0.5 * qm * dt, n);
    }
    if (wall == 0) orc_inside_rect(pd, n, lx, ly, isInside);
    else orc_inside_ellip(pd, n, lx, ly, isInside);
}

/* ---------------------------------------------------------------------------------------
 * Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw, "Parallel random
 * numbers: as easy as 1, 2, 3", SC'11).  The product's device generator must reproduce
 * these words bit for bit; the oracle uses them to inject the same uniforms.
 * ------------------------------------------------------------------------------------- */
static inline void philox_round(uint32_t c[4], const uint32_t k[2])
{
    uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t o0 = hi1 ^ c[1] ^ k[0];
    uint32_t o1 = lo1;
    uint32_t o2 = hi0 ^ c[3] ^ k[1];
    uint32_t o3 = lo0;
    c[0] = o0; c[1] = o1; c[2] = o2; c[3] = o3;
}

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    uint32_t k[2] = {key[0], key[1]};
    for (int r = 0; r < 10; ++r) {
        philox_round(c, k);
        k[0] += 0x9E3779B9u;
        k[1] += 0xBB67AE85u;
    }
    out[0] = c[0]; out[1] = c[1]; out[2] = c[2]; out[3] = c[3];
}

/* n blocks of 4 words for counters (c0 + i, c1, c2, c3) */
void orc_philox_stream(uint32_t n, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                       uint32_t k0, uint32_t k1, uint32_t *out)
{
    uint32_t key[2] = {k0, k1};
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t ctr[4] = {c0 + i, c1, c2, c3};
        orc_philox4x32_10(ctr, key, out + 4 * (size_t)i);
    }
}
