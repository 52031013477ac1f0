// Device form of the three particle walls of particleBoundary.pyx.
#pragma once
#include "common.h"

struct WallDev {
    int32_t kind;
    double lx, ly, dx, dy, lxHalfExt, lyHalfExt;
    double lxHalf, lyHalf;      // rectangle half sizes
    double ax, ay;              // 4/lx^2, 4/ly^2 (ellipse)
    double clipX, clipY;        // cut-cell clip limits, 0.5*lxExt - 1e-6*dx
    double dxi, dyi;
    uint32_t nx;
    const uint16_t *insideFaces;
    const double *bp;
    const int32_t *cpi;
    const double *ccn;
    int32_t rectCornerFix;      // rectangle: normal of the wall crossed first at a corner
};

static inline int wall_to_device(const oecl_wall_t *w, WallDev *d)
{
    if (!w) return OECL_ERR_ARG;
    d->kind = w->kind;
    d->lx = w->lx;
    d->ly = w->ly;
    d->dx = w->dx;
    d->dy = w->dy;
    d->lxHalfExt = w->lxHalfExt;
    d->lyHalfExt = w->lyHalfExt;
    d->lxHalf = w->lx * 0.5;                 // particleBoundary.pyx:88
    d->lyHalf = w->ly * 0.5;
    d->ax = 4. / (w->lx * w->lx);            // particleBoundary.pyx:233
    d->ay = 4. / (w->ly * w->ly);
    d->clipX = w->lxHalfExt - 1.e-6 * w->dx; // particleBoundary.pyx:349-351
    d->clipY = w->lyHalfExt - 1.e-6 * w->dy;
    d->dxi = 1. / w->dx;
    d->dyi = 1. / w->dy;
    d->nx = w->nx;
    d->insideFaces = w->insideFaces;
    d->bp = w->boundaryPoints;
    d->cpi = w->cutCellPointsInd;
    d->ccn = w->cutCellNormalVectors;
    d->rectCornerFix = w->rectCornerFix;
    if (w->kind < 0 || w->kind > 2) return OECL_ERR_ARG;
    if (w->kind == OECL_WALL_CUTCELL &&
        (!w->insideFaces || !w->boundaryPoints || !w->cutCellPointsInd ||
         !w->cutCellNormalVectors))
        return OECL_ERR_ARG;
    return OECL_OK;
}

// Cell of a point as particleBoundary.pyx:353-355 computes it.
__device__ __forceinline__ uint32_t wall_cell(const WallDev &w, double x, double y)
{
    uint32_t ix = (uint32_t)((x + w.lxHalfExt) * w.dxi);
    uint32_t iy = (uint32_t)((y + w.lyHalfExt) * w.dyi);
    return ix + iy * w.nx;
}

// Point-in-cut-cell test of particleBoundary.pyx:357-370 for an already clipped point.
__device__ __forceinline__ uint8_t cutcell_inside(const WallDev &w, double x, double y)
{
    uint32_t c = wall_cell(w, x, y);
    uint16_t f = w.insideFaces[c];
    if (f == 8) return 1;
    if (f == 0) return 0;
    int32_t a = w.cpi[2 * c], b = w.cpi[2 * c + 1];
    return ((w.bp[2 * a] - x) * (w.bp[2 * b + 1] - y) >
            (w.bp[2 * b] - x) * (w.bp[2 * a + 1] - y))
               ? 1
               : 0;
}

// indexInside for one particle.  The cut-cell wall first clips the position to the grid
// (particleBoundary.pyx:349-351 via _clipParticleDataToGrid :485-508) and writes it back.
// A point on the wall (the interaction point of an absorbed particle, from which secondaries
// start) in the particle dtype.  FP64 keeps it as computed.  FP32 rounds toward zero: the
// rectangular and elliptical chambers are centred on the origin, so the rounded point stays
// inside whenever the FP64 point is.  (The wall search insets the point by 1e-6 of a cell, 4e-11 m
// at 45 mm, far below FP32 resolution there: rounded to nearest, half of the points would lie
// outside and their secondaries be absorbed again at once.)
template <typename T>
__device__ __forceinline__ T wall_store(double v);
template <>
__device__ __forceinline__ double wall_store<double>(double v)
{
    return v;
}
template <>
__device__ __forceinline__ float wall_store<float>(double v)
{
    return __double2float_rz(v);
}

template <typename T>
__device__ __forceinline__ uint8_t wall_inside(const WallDev &w, T &x, T &y)
{
    if (w.kind == OECL_WALL_RECT) {
        double xd = (double)x, yd = (double)y;
        return (xd < -w.lxHalf || xd > w.lxHalf || yd < -w.lyHalf || yd > w.lyHalf) ? 0 : 1;
    } else if (w.kind == OECL_WALL_ELLIP) {
        double xd = (double)x, yd = (double)y;
        return (xd * xd * w.ax + yd * yd * w.ay) >= 1. ? 0 : 1;
    } else {
        double xd = (double)x, yd = (double)y;
        if (xd > w.clipX) xd = w.clipX;
        else if (xd < -w.clipX) xd = -w.clipX;
        if (yd > w.clipY) yd = w.clipY;
        else if (yd < -w.clipY) yd = -w.clipY;
        x = (T)xd;
        y = (T)yd;
        return cutcell_inside(w, xd, yd);
    }
}

// Rectangle wall hit (particleBoundary.pyx:108-192 and :195-221): moves (x, y) back along
// (vx, vy) by the remaining time r to the wall inset by 1e-6 of a cell and sets the inward
// normal (n0, n1).  Where both coordinates are out, r is the larger of the two crossing
// times, i.e. the wall crossed first.  The reference labels four of those corner cases with
// the other wall's normal (:162-187); that is kept unless rectCornerFix is set, in which case
// the normal is the crossed-first wall's.
__device__ __forceinline__ void wall_rect_hit(const WallDev &w, double &x, double &y, double vx, double vy,
                                              double &r, double &n0, double &n1)
{
    const double eps = 1e-6;
    const double lxh = w.lx * 0.5 - w.dx * eps, lyh = w.ly * 0.5 - w.dy * eps;
    int xi = 1, yi = 1, lw = 0;
    if (x > lxh) xi = 2;
    else if (x < -lxh) xi = 0;
    if (y > lyh) yi = 2;
    else if (y < -lyh) yi = 0;
    if (yi == 1) {
        r = (xi == 0 ? x + lxh : x - lxh) / vx;
        lw = xi == 0 ? 0 : 1;
    } else if (xi == 1) {
        r = (yi == 0 ? y + lyh : y - lyh) / vy;
        lw = yi == 0 ? 2 : 3;
    } else {
        const double tx = (xi == 0 ? x + lxh : x - lxh) / vx;
        const double ty = (yi == 0 ? y + lyh : y - lyh) / vy;
        const bool xFirst = tx > ty;
        r = xFirst ? tx : ty;
        if (w.rectCornerFix) lw = xFirst ? (xi == 0 ? 0 : 1) : (yi == 0 ? 2 : 3);
        else if (xFirst) lw = xi == 0 ? 0 : 2;  // reference codes at a corner
        else lw = yi == 0 ? 1 : 3;
    }
    x -= r * vx;
    y -= r * vy;
    n0 = lw == 0 ? 1. : (lw == 1 ? -1. : 0.);
    n1 = lw == 2 ? 1. : (lw == 3 ? -1. : 0.);
}

// The reference's isInside(x, y) used by the loader and the split (no clipping).
__device__ __forceinline__ uint8_t wall_is_inside_point(const WallDev &w, double x, double y)
{
    if (w.kind == OECL_WALL_RECT)
        return (x < -w.lxHalf || x > w.lxHalf || y < -w.lyHalf || y > w.lyHalf) ? 0 : 1;
    if (w.kind == OECL_WALL_ELLIP) return (x * x * w.ax + y * y * w.ay) >= 1. ? 0 : 1;
    // The reference indexes its cell arrays without a range check; outside the grid the
    // point is reported outside here.
    if (x <= -w.lxHalfExt || x >= w.lxHalfExt || y <= -w.lyHalfExt || y >= w.lyHalfExt) return 0;
    return cutcell_inside(w, x, y);
}
