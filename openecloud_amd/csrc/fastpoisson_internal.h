// Library-internal access to the direct Poisson solver (not part of the C ABI): the one-call
// step writes the masked right-hand side straight into the solver's box vector.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

struct oecl_fastpoisson;

struct FpBox {
    double *b;             // bx * by per right-hand side, column-major (x fastest), sides nb apart
    const uint8_t *mask;   // inside flags of the extended grid
    uint32_t nx, ny, bx;   // extended grid, box width (nx - 4)
};

FpBox fastpoisson_box(const oecl_fastpoisson *h);
// Room for R right-hand sides in the solver's box vectors (reallocates them when R grows, so
// call it before fastpoisson_box and before any graph capture).
int fastpoisson_reserve(oecl_fastpoisson *h, uint32_t R);
// Solve R systems with b already holding q on inside box nodes and 0 elsewhere in the box;
// phi (R potentials np apart) is written on the box nodes only (nodes outside the box must be
// zero already).
int fastpoisson_solve_box(oecl_fastpoisson *h, double *phi, uint32_t R, hipStream_t s);
// The same with the right-hand side gathered from the grid inside the first transform:
// q = rho + table[r nScale + *stepDev] * prof on the inside nodes of run r (prof and table may
// be null: q = rho), and every node of rho zeroed except grid rows 0, 1, ny-2 and ny-1.
int fastpoisson_solve_grid(oecl_fastpoisson *h, double *rho, const double *prof, const double *table, uint32_t nScale,
                           const uint32_t *stepDev, double *phi, uint32_t R, hipStream_t s);

