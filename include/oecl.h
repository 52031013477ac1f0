/*
 * liboecl -- C ABI of the MI355X (gfx950) electron-cloud PIC step.
 *
 * Every entry point is extern "C", takes plain device pointers, sizes and a hipStream_t
 * (passed as void*), returns 0 on success or an OECL_ERR_* code, and never allocates or
 * frees caller memory unless its comment says it owns a workspace.  The text of the last
 * failure is available from oecl_last_error().
 *
 * Particle storage is one structure-of-arrays block of 6 rows (x, y, vx, vy, vz, w), each
 * row `stride` elements long, in FP64 (dtype OECL_F64) or FP32 (OECL_F32).  Node fields are
 * FP64 and laid out like the reference: node index i + j*nxExt, E as [Ex | Ey] and B as
 * [Bx | By | Bz], each block npExt long.  Charge on the grid is always FP64.
 *
 * Each declaration names the reference routine it replaces (file:line in openECLOUD).
 */
#ifndef OECL_H
#define OECL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    OECL_OK = 0,
    OECL_ERR_ARG = 1,        /* invalid argument / shape */
    OECL_ERR_HIP = 2,        /* a HIP call or launch failed */
    OECL_ERR_CAPACITY = 3,   /* output would exceed the capacity passed in */
    OECL_ERR_NOCONV = 4,     /* iterative solve did not reach its tolerance */
    OECL_ERR_RANGE = 5       /* particle outside the grid / loader gave up */
};

enum { OECL_F64 = 0, OECL_F32 = 1 };

/* Wall kinds of particleBoundary.pyx: AbsorbRectangular :77, AbsorbElliptical :225,
 * AbsorbArbitraryCutCell :302. */
enum { OECL_WALL_RECT = 0, OECL_WALL_ELLIP = 1, OECL_WALL_CUTCELL = 2 };

/* Grid.__init__ quantities (grid.pyx:57-68): extended node counts and sizes. */
typedef struct {
    uint32_t nx, ny, np;        /* nxExt, nyExt, npExt */
    double dx, dy;
    double lxHalfExt, lyHalfExt; /* 0.5*lxExt, 0.5*lyExt */
} oecl_grid_t;

/* Wall description (particleBoundary.pyx:21-36 + the grid arrays the cut-cell wall reads). */
typedef struct {
    int32_t kind;
    double lx, ly;               /* un-extended chamber size (getLx/getLy) */
    double dx, dy;
    double lxHalfExt, lyHalfExt;
    uint32_t nx;                 /* nxExt */
    const uint16_t *insideFaces;          /* [npExt]          (cut cell only) */
    const double *boundaryPoints;         /* [nb][2]          (cut cell only) */
    const int32_t *cutCellPointsInd;      /* [npExt][2]       (cut cell only) */
    const double *cutCellNormalVectors;   /* [npExt][2]       (cut cell only) */
    int32_t rectCornerFix;   /* rectangle only: 0 = the reference's normals where a particle
                                left through a corner (particleBoundary.pyx:154-187 label four
                                of those cases with the other wall's normal); 1 = the normal of
                                the wall crossed first.  The interaction point is the same. */
} oecl_wall_t;

/* Furman-Pivi model parameters (particleEmitter.pyx:107-152). */
typedef struct {
    double p1EInf, p1Ehat, eEHat, w, p, e1, e2, sigmaE;
    double p1RInf, eR, r, q, r1, r2;
    double deltaTSHat, s, eHat0, t1, t2, t3, t4;
    double theta0max, deltaERmax;
    double epsN[10], pSmallN[10];
    double particleMass;
    int32_t scaleSigmaE;
} oecl_furman_t;

/* The two simpler emitters of particleEmitter.pyx: SecElecEmitter (:316-482, parameters of
 * stainlessMod :334-355) and SimpleSecEmitter (:489-596, whose values are fixed in its code,
 * so it reads `model` only). */
enum { OECL_SECEM_STAINLESS = 0, OECL_SECEM_SIMPLE = 1 };
typedef struct {
    int32_t model;
    double p1EInf, p1Ehat, eEHat, w, p, e1, e2;
    double p1RInf, eR, r, q, r1, r2;
    double deltaTSHat, s, eHat0, t1, t2, t3, t4;
    double sigmaETS, muETS, theta0max, deltaERmax;
    double particleMass;
} oecl_secem_t;

const char *oecl_last_error(void);
int oecl_device_sync(void);
/* Plain device-to-device copy of `bytes` (pointers and size 16-byte aligned).  No reference
 * counterpart: the known-size kernel against which HBM traffic counters are calibrated. */
int oecl_copy_bytes(void *dst, const void *src, uint64_t bytes, void *stream);

/* ---- particles.pyx ------------------------------------------------------------------ */

/* particles.pyx:52-56 -> _calcGridWeights :400-418.  w4 holds 4 rows (LL, UL, LU, UU) of
 * wstride elements of the particle dtype. */
int oecl_calc_grid_weights(const oecl_grid_t *g, int dtype, const void *data, uint64_t stride,
                           uint32_t n, uint32_t *inCell, void *w4, uint64_t wstride,
                           void *stream);

/* particles.pyx:70-77 -> _chargeToGrid :336-351.  Zeroes rho[npExt] and deposits
 * w_k * weight * q at the four corners (FP64 accumulation). */
int oecl_charge_to_grid(const oecl_grid_t *g, int dtype, const void *data, uint64_t stride,
                        uint32_t n, const uint32_t *inCell, const void *w4, uint64_t wstride,
                        double q, double *rho, void *stream);
/* particles.pyx:70-77 -> _chargeToGrid :336-351 with the reference's summation order: every
 * node's charge is the sum of its contributions in increasing particle index, starting from
 * zero, exactly as the sequential loop forms it, so rho is bit-identical to the reference for
 * the same particles (the atomic deposit above agrees to rounding, 1e-13 relative).  A stable
 * sort by cell plus one ordered merge per node; for parity runs.  work:
 * oecl_charge_ordered_work_bytes(n, npExt) bytes. */
uint64_t oecl_charge_ordered_work_bytes(uint32_t n, uint32_t np);
int oecl_charge_to_grid_ordered(const oecl_grid_t *g, int dtype, const void *data, uint64_t stride,
                                uint32_t n, const uint32_t *inCell, const void *w4, uint64_t wstride,
                                double q, double *rho, void *work, void *stream);

/* particles.pyx:58-68 -> _fieldToParticles :354-372 (dims 2) / _fieldToParticles3Dims
 * :375-397 (dims 3).  out is AoS [n][dims] in the particle dtype. */
int oecl_field_to_particles(const oecl_grid_t *g, int dtype, const double *field,
                            const uint32_t *inCell, const void *w4, uint64_t wstride, uint32_t n,
                            int dims, void *out, void *stream);

/* particles.pyx:79-92 -> _borisPushNoB :324-333 (typeB 0, full kick q/m*dt),
 * _borisPushConstB :298-321 (typeB 1, bConst[3]) and _borisPush :272-295 (typeB 2, bAt
 * AoS [n][3]).  eAt AoS [n][2].  All in the particle dtype except bConst (FP64). */
int oecl_boris_push(int dtype, void *data, uint64_t stride, uint32_t n, const void *eAt,
                    const void *bAt, const double *bConst, int typeB, double dt, double qm,
                    void *stream);

/* Fused form of one particle step (particles.pyx:52-92 + particleBoundary indexInside):
 * bilinear weights in registers, E (and B: typeB 2 from bField[3*npExt]) gathered at the
 * particle, Boris push, wall test.  Writes isInside[n] (1 = keep) and, when rhoNext is not
 * NULL, deposits the charge of every particle that stays inside at its NEW position
 * (the first half of the next step's _chargeToGrid; rhoNext is not zeroed here). */
int oecl_push_fused(const oecl_grid_t *g, const oecl_wall_t *wall, int dtype, void *data,
                    uint64_t stride, uint32_t n, const double *E, const double *bField,
                    const double *bConst, int typeB, double dt, double qm, double q,
                    uint8_t *isInside, double *rhoNext, void *stream);

/* Deposit without zeroing (rho += ...) of `n` particles, weights in registers.  Used for
 * the secondaries and after particle management. */
int oecl_deposit_add(const oecl_grid_t *g, int dtype, const void *data, uint64_t stride,
                     uint32_t n, double q, double *rho, void *stream);
/* The same for an AoS block [n][6] (the emitter's secondaries). */
int oecl_deposit_add_aos(const oecl_grid_t *g, int dtype, const void *aos, uint32_t n, double q,
                         double *rho, void *stream);

/* Moments of the particle set in FP64 with a fixed reduction order: out[6] = (sum w,
 * sum w|v|^2, sum w^2|v|^2, sum w x, sum w y, count).  Replaces getMeanWeight's loop
 * (particles.pyx:39-47), the driver's numpy.sum of the weights (example01_basic.py:56) and
 * util.averageEnergy (util.pyx:8-20).  work: 512*6 doubles. */
int oecl_particle_moments(int dtype, const void *data, uint64_t stride, uint32_t n, double *out,
                          double *work, void *stream);

/* ---- particleBoundary.pyx ----------------------------------------------------------- */

/* indexInside: AbsorbRectangular :84-98, AbsorbElliptical :227-242, AbsorbArbitraryCutCell
 * :333-370 (which first clips to the grid, :485-508; *clipped is set to 1 if it did). */
int oecl_index_inside(const oecl_wall_t *wall, int dtype, void *data, uint64_t stride,
                      uint32_t n, uint8_t *isInside, int32_t *clipped, void *stream);

/* ParticleBoundary.saveAbsorbed :38-65.  Counts absorbed particles (isInside == 0), and
 * copies them in index order to absorbed (AoS [cap][6], particle dtype).  work must hold
 * oecl_scan_work_bytes(n) bytes.  *nAbsorbed is written on the device (uint32). */
uint64_t oecl_scan_work_bytes(uint32_t n);
int oecl_save_absorbed(int dtype, const void *data, uint64_t stride, uint32_t n,
                       const uint8_t *isInside, void *absorbed, uint32_t absCap,
                       uint32_t *nAbsorbed, void *work, void *stream);
/* oecl_save_absorbed without a workspace and without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The scan's scratch is a
 * 128 KiB array inside the library, one per GPU, held until the process ends; n is limited to
 * oecl_save_absorbed_noalloc_max() (67 088 384) and larger n returns OECL_ERR_CAPACITY before
 * any device work.  Calls on one GPU share that array: calls that overlap on one GPU (on
 * different streams or threads) race and must be serialised by the caller. */
uint32_t oecl_save_absorbed_noalloc_max(void);
int oecl_save_absorbed_noalloc(int dtype, const void *data, uint64_t stride, uint32_t n,
                               const uint8_t *isInside, void *absorbed, uint32_t absCap,
                               uint32_t *nAbsorbed, void *stream);

/* calculateInteractionPoint + calculateNormalVectors of the three walls
 * (Rect :108-221, Ellip :252-297, CutCell :373-481) over nAbs absorbed particles (AoS).
 * remaining[nAbs] (particle dtype) and normals [nAbs][2] are outputs; for the cut cell wall
 * dt is the last push's time step and *notFound counts particles whose crossing was not
 * found (the reference prints a warning for each). */
int oecl_wall_interaction(const oecl_wall_t *wall, int dtype, void *absorbed, uint32_t nAbs,
                          double dt, void *remaining, void *normals, uint32_t *notFound,
                          void *stream);

/* ---- particleEmitter.pyx ------------------------------------------------------------ */

/* FurmanEmitter.generateSecondaries :154-308, first loop (:178-229): nSec[nAbs] per impact
 * plus what the second loop needs.  uniforms: NULL to draw from Philox4x32-10 under
 * (seed0, seed1) with counter (impact, slot/2, step, 1), else an injected FP64 table
 * [nAbs][35].  impact[] receives per-impact e0, deltaE, deltaR, pN1 (4 doubles each). */
int oecl_furman_pass1(const oecl_furman_t *prm, int dtype, const void *absorbed,
                      const void *normals, uint32_t nAbs, const double *uniforms,
                      uint32_t seed0, uint32_t seed1, uint32_t step, uint32_t *nSec,
                      double *impact, void *stream);

/* Exclusive scan of nSec -> offsets[nAbs + 1] (offsets[nAbs] = total). */
int oecl_scan_u32(const uint32_t *in, uint32_t n, uint32_t *offsetsOut, void *work,
                  void *stream);

/* Second loop (:231-305): writes the secondaries (AoS [total][6], particle dtype). */
int oecl_furman_pass2(const oecl_furman_t *prm, int dtype, const void *absorbed,
                      const void *normals, uint32_t nAbs, const uint32_t *nSec,
                      const uint32_t *offsets, const double *impact, const double *uniforms,
                      uint32_t seed0, uint32_t seed1, uint32_t step, void *secondaries,
                      uint32_t secCap, void *stream);

/* SecElecEmitter / SimpleSecEmitter .generateSecondaries (particleEmitter.pyx:357-482 /
 * :501-596), first loop: nSec[nAbs] (elastic + rediffused + true secondaries of each impact,
 * at most 10) and impact[nAbs][4] = (e0, nE + 2 nR + 4 nTS, -, -).  uniforms: NULL for
 * Philox4x32-10 under (seed0, seed1) with counter (impact, slot/2, step, 7), else an injected
 * FP64 table [nAbs][44].  The true-secondary count must stay <= 8: the host rejects yields
 * that could exceed it. */
int oecl_secem_pass1(const oecl_secem_t *prm, int dtype, const void *absorbed, const void *normals,
                     uint32_t nAbs, const double *uniforms, uint32_t seed0, uint32_t seed1, uint32_t step,
                     uint32_t *nSec, double *impact, void *stream);
/* Second loop: the secondaries at offsets[] (oecl_scan_u32 of nSec) into AoS [secCap][6]. */
int oecl_secem_pass2(const oecl_secem_t *prm, int dtype, const void *absorbed, const void *normals,
                     uint32_t nAbs, const uint32_t *offsets, const double *impact, const double *uniforms,
                     uint32_t seed0, uint32_t seed1, uint32_t step, void *secondaries, uint32_t secCap,
                     void *stream);

/* homoLoader (particleEmitter.pyx:41-72): n particles uniform in the chamber by rejection
 * against the wall, velocities N(0, vsigma) per component, w = weightInit (times
 * 0.5 + u when randomizeWeight).  Counter-based: particle i uses Philox counter
 * (i, k, 0, 2).  *failed counts particles that found no inside point in 64 attempts. */
int oecl_homo_loader(const oecl_wall_t *wall, int dtype, void *data, uint64_t stride,
                     uint32_t n, double weightInit, double vsigma, int randomizeWeight,
                     uint32_t seed0, uint32_t seed1, uint32_t *failed, void *stream);

/* Primary photoelectrons of time step `step` into out[n][6] (AoS).  No counterpart in the
 * reference (its Readme.md lists photoemission as an outlook); the source of the elliptical
 * dipole workload.  Rectangular or elliptical wall: points on the wall inset by 1e-3 cell,
 * energy |e0 + sigmaE z| eV, direction about the inward normal with the reference emitter's
 * secondary angles (particleEmitter.pyx:289-302), weight `weight`.  Philox counter
 * (k, t, step, 6). */
int oecl_photo_emit(const oecl_wall_t *wall, int dtype, void *out, uint32_t n, uint32_t step,
                    uint32_t seed0, uint32_t seed1, double e0, double sigmaE, double weight,
                    double particleMass, void *stream);

/* ---- Particles.addAndRemoveParticles (particles.pyx:94-121) ------------------------- */

/* Applies the reference's three compaction rules on the SoA block:
 *   mode 0: _removeParticles :254-269 (nAdd must be 0)
 *   mode 1: _addAndRemoveParticles :222-252 (in place, holes first filled by additions)
 *   mode 2: _addAndRemoveParticlesNewArray :201-219 (stable keep, then append)
 *   mode 3: fast form -- additions fill the first holes, extra additions are appended,
 *           remaining holes below the new count are filled from the tail (the rule of
 *           _removeParticles); the same particle set as modes 1 and 2.
 * add: AoS [nAdd][6] (particle dtype).  spare: a second SoA block of the same stride used
 * by modes 1 and 2.  work: oecl_compact_work_bytes(n, nAdd) bytes.  The new count is
 * written to *newCount on the device. */
uint64_t oecl_compact_work_bytes(uint32_t n, uint32_t nAdd);
int oecl_add_and_remove(int dtype, void *data, void *spare, uint64_t stride, uint32_t n,
                        const uint8_t *keep, const void *add, uint32_t nAdd, int mode,
                        uint32_t *newCount, void *work, void *stream);

/* ---- sorting (particles.pyx:431-459 _quicksortByColumn) ----------------------------- */

/* Stable LSD radix sort of keys[n] ascending with 32-bit payloads (8 key bits per pass,
 * ceil(bits/8) passes).  work: oecl_sort_work_bytes(n) bytes. */
uint64_t oecl_sort_work_bytes(uint32_t n);
int oecl_radix_sort_u64(uint64_t *keys, uint32_t *vals, uint32_t n, int bits, void *work,
                        void *stream);
/* Particles.sortByColumn (particles.pyx:49-50): the block sorted ascending by one column
 * (stable) is written to spare; keys[n], perm[n] scratch; work: oecl_sort_work_bytes(n). */
int oecl_sort_by_column(int dtype, const void *data, void *spare, uint64_t stride, uint32_t n,
                        int column, uint64_t *keys, uint32_t *perm, void *work, void *stream);
/* Orders the particles by grid cell (stable) into spare (same stride): consecutive particles
 * then share cells, so the field gather stays in cache and the deposit sums in LDS.  No
 * reference counterpart (the reference's loops are serial); keys[n], perm[n] scratch, work:
 * oecl_sort_work_bytes(n). */
int oecl_sort_by_cell(const oecl_grid_t *g, int dtype, const void *data, void *spare,
                      uint64_t stride, uint32_t n, uint64_t *keys, uint32_t *perm, void *work,
                      void *stream);
/* The cell sort of a batched particle block (oecl_runs_t layout): every slot of every run's
 * segment is ordered by (run, bin) into `spare`, live particles first, so each run stays in
 * its segment.  The bin is the cell (ix, iy) >> binShift in row-major (order 0) or Morton
 * (order 1) order.  keys/perm hold stride entries; work is oecl_sort_work_bytes(stride). */
int oecl_sort_runs_by_cell(const oecl_grid_t *g, int dtype, const void *data, void *spare, uint64_t stride,
                           uint32_t R, const uint32_t *segOff, const uint32_t *blockRun, const uint32_t *nDev,
                           int order, int binShift, uint64_t *keys, uint32_t *perm, void *work, void *stream);
/* dst[:, i] = src[:, perm[i]], i < n. */
int oecl_permute_particles(int dtype, const void *src, void *dst, uint64_t stride,
                           const uint32_t *perm, uint32_t n, void *stream);

/* ---- particleManagement.pyx --------------------------------------------------------- */

/* w *= f for all n particles (particleManagement.pyx:45-46, :58-59). */
int oecl_scale_weights(int dtype, void *data, uint64_t stride, uint32_t n, double f,
                       void *stream);
int oecl_fill_u8(uint8_t *a, uint32_t n, uint8_t v, void *stream);
/* globalRandom's kill set (:47-50): keep[i] = 0 for a uniformly random subset of k of the n
 * particles (the k smallest Philox keys).  keys[n], perm[n] scratch; work:
 * oecl_sort_work_bytes(n). */
int oecl_random_kill(uint32_t n, uint32_t k, uint32_t streamIndex, uint32_t seed0,
                     uint32_t seed1, uint8_t *keep, uint64_t *keys, uint32_t *perm, void *work,
                     void *stream);
/* globalRandom's clones (:60-63): add[nAdd][6] = copies of uniformly random particles. */
int oecl_random_clones(int dtype, const void *data, uint64_t stride, uint32_t n, uint32_t nAdd,
                       uint32_t streamIndex, uint32_t seed0, uint32_t seed1, void *add,
                       void *stream);
/* _searchSorted (:244-261) of three values over the weight column sorted ascending:
 * out[k] = first index whose weight exceeds v_k. */
int oecl_search_sorted_weights(int dtype, const void *data, uint64_t stride, uint32_t n,
                               double v0, double v1, double v2, uint32_t *out, void *stream);
/* localRandom's split of particles [first, first+cnt) (:130-163) into add[cnt][6]. */
int oecl_split_heaviest(const oecl_wall_t *wall, int dtype, void *data, uint64_t stride,
                        uint32_t first, uint32_t cnt, double dx, double dy, double lx, double ly,
                        uint32_t streamIndex, uint32_t seed0, uint32_t seed1, void *add,
                        void *stream);
/* localRandom's merge (particleManagement.pyx:166-236, kdTree.pyx:146-255) over particles
 * [0, nc), sorted by weight ascending (rank = index, lightest first).  The reference's rule:
 * a particle merges with its nearest neighbour in the 5-D space (x, y, vx, vy, vz) scaled by
 * pw[5] (device) when the squared distance is below disLimit (:185 compares the squared
 * distance with factorDisLimit * ds); the survivor is one of the two with probability
 * proportional to weight and takes the pair's weight (scheme 1: velocity rescaled to keep the
 * pair's kinetic energy, :209-232); merged particles leave the pool.  Run as rounds of parallel
 * matching: each pooled particle finds its exact nearest neighbour in the pool through a
 * Morton-ordered tree (Z-order sort of the scaled points, leaves of 8, an implicit binary tree
 * of bounding boxes, branch-and-bound queries), a pair holds both of its particles only when
 * it is the lowest-ranked claim on each, and accepted pairs are taken in rank order until
 * toMerge merges are made or no particle finds a neighbour.  G (2..8): the survivor stays in
 * the pool while its group has fewer than G particles (2: the reference's pairs).  keep[nc]
 * written; *merged (device) and *hostMerged = merges made; *hostRounds = rounds; pairs
 * (optional, [toMerge][2]) receives (survivor, removed) per merge.  work:
 * oecl_merge_work_bytes(nc).  Synchronises the stream once per round. */
uint64_t oecl_merge_work_bytes(uint32_t nc);
int oecl_nn_merge(int dtype, void *data, uint64_t stride, uint32_t nc, const double *pw, int G,
                  double disLimit, uint32_t toMerge, int scheme, uint32_t streamIndex, uint32_t seed0,
                  uint32_t seed1, uint8_t *keep, uint32_t *merged, uint32_t *pairs, uint32_t *hostMerged,
                  uint32_t *hostRounds, void *work, void *stream);

/* ---- poissonSolver.pyx -------------------------------------------------------------- */

/* FIT operator of poissonSolver.pyx:44-55 as a 5-point stencil on the full extended grid:
 * cx[p] = eps0*dsi[p]*dst[p+np] (x-edge p -> p+1), cy[p] = eps0*dsi[p+np]*dst[p]
 * (y-edge p -> p+nx); mask[p] = insidePoints[p].  Solves A phi = q on inside nodes by
 * Jacobi-preconditioned CG in FP64 from the initial guess in phi (non-inside nodes are
 * forced to 0), to ||r|| <= tol*||q_inside||.  work: oecl_pcg_work_bytes(np) bytes.
 * *iters and *relres are host outputs (this call synchronises the stream). */
uint64_t oecl_pcg_work_bytes(uint32_t np);
int oecl_pcg_solve(uint32_t nx, uint32_t np, const double *cx, const double *cy,
                   const uint8_t *mask, const double *q, double *phi, double tol,
                   uint32_t maxIter, void *work, uint32_t *iters, double *relres,
                   void *stream);

/* y = A x with A the FIT operator of oecl_pcg_solve restricted to the inside nodes
 * (poissonSolver.pyx:55: x is taken as zero on non-inside nodes, y is 0 there).  The
 * reference forms A only inside its SuperLU factorisation; this is its product, used to make
 * pairs (A phi, phi) for training the learned preconditioner.  Inside nodes lie off the
 * extended grid's outer ring, as the reference's grids guarantee. */
int oecl_fit_apply(uint32_t nx, uint32_t np, const double *cx, const double *cy,
                   const uint8_t *mask, const double *x, double *y, void *stream);

/* phiToEAtGridPoints.dot(phi) (poissonSolver.pyx:67,87) as a CSR SpMV with the rows' own
 * column order: y[r] = sum_k val[k]*x[col[k]], k from rowPtr[r] to rowPtr[r+1]. */
int oecl_csr_spmv(uint32_t nRows, const int32_t *rowPtr, const int32_t *col,
                  const double *val, const double *x, double *y, void *stream);

/* q_total = rho + scale * profile (beam.py:68-79 / :129-136 getCharge added to the
 * electron charge, example01_basic.py:81). */
int oecl_axpy_rhs(uint32_t np, const double *rho, double scale, const double *profile,
                  double *out, void *stream);

/* Direct solve of the same system (poissonSolver.pyx:59,85 -- the SuperLU factorisation
 * and back-substitution) by the capacitance-matrix method: the box of nodes
 * 2 <= i < 2 + bx, 2 <= j < nyExt - 2 (bx = oecl_fastpoisson_box_width(nxExt) >= nxExt - 4,
 * by = nyExt - 4) with the uniform edge coefficients cxu (x) and cyu (y) is inverted exactly
 * by a type-I DST along x (an FFT of length bx + 1 per line) and tridiagonal solves along y,
 * and the m inside rows whose operator row differs from the box's are corrected through the
 * m x m capacitance matrix, built and inverted on the device at creation.  sIdx [m]: box
 * index (i-2) + (j-2)*bx of each corrected row; W (CSR over box indices): those rows of
 * (M - R) (box columns from nxExt - 2 on hold no inside node, so an inside node next to
 * them is a corrected row).  maskDev (inside flags, device) must outlive the handle.
 * *capMax receives max |C_aa| (a scale check).  nxExt - 4 <= 2047, nyExt - 4 <= 2048. */
typedef struct oecl_fastpoisson oecl_fastpoisson;
/* The solver's box width for an extended grid of nxExt columns: the least 2^p - 1 that is at
 * least nxExt - 4 (0 for nxExt < 5).  No counterpart in the reference (its SuperLU solve has
 * no box). */
uint32_t oecl_fastpoisson_box_width(uint32_t nxExt);
int oecl_fastpoisson_create(uint32_t nxExt, uint32_t nyExt, double cxu, double cyu, uint32_t m,
                            const int32_t *sIdx, const int32_t *wPtr, const int32_t *wCol,
                            const double *wVal, const uint8_t *maskDev, oecl_fastpoisson **out,
                            double *capMax);
/* Room in the handle for R right-hand sides solved together (the batched step,
 * oecl_step_run_batch); allocates, so call it before capturing a batched step into a graph. */
int oecl_fastpoisson_reserve(oecl_fastpoisson *h, uint32_t R);
/* phi (npExt, zero on non-inside nodes) = A^-1 q on inside nodes. */
int oecl_fastpoisson_solve(oecl_fastpoisson *h, const double *q, double *phi, void *stream);
int oecl_fastpoisson_destroy(oecl_fastpoisson *h);

/* ---- learned preconditioner (no counterpart in the reference) ------------------------- */

/* A 4-level U-Net (16 channels, 3x3 convolutions, ReLU, 2x2 average pooling, nearest
 * upsampling with additive skips, 14 convolutions) used as the preconditioner of a flexible
 * CG on the FIT Poisson system, and trained on the device.  Images are H x W (multiples of 8)
 * NHWC bfloat16 with 16 channels; parameters are FP32 masters W [14][9][16][16] (tap
 * 3 (dy + 1) + (dx + 1), out, in) and b [14][16] with bfloat16 copies for the products
 * (bf16 MFMA, FP32 accumulation).  seed: the He initialisation of W (b = 0). */
typedef struct oecl_unet oecl_unet;
int oecl_unet_create(uint32_t H, uint32_t W, uint32_t seed, oecl_unet **out);
int oecl_unet_destroy(oecl_unet *u);
/* Output rows per convolution block: 1, 2, 4 or 8, or 0 (the default) to choose per image size
 * (the tallest tile that leaves at least 1024 blocks).  The result does not depend on it. */
int oecl_unet_set_conv_rows(oecl_unet *u, int rows);
/* Host arrays in / out (any may be NULL in get). */
int oecl_unet_set_params(oecl_unet *u, const float *W, const float *b);
int oecl_unet_get_params(const oecl_unet *u, float *W, float *b, float *gW, float *gb);
/* y[H][W] (FP32) = network(img), img H x W x 16 bf16 with the input in channel 0. */
int oecl_unet_forward(oecl_unet *u, const uint16_t *img, float *y, void *stream);
/* After a forward: gradients for the output gradient gout (H x W x 16 bf16, channel 0 =
 * dL/dy), accumulated into the parameter gradients (zeroed first when zero != 0); gin
 * (optional) receives the gradient of the input image. */
int oecl_unet_backward(oecl_unet *u, const uint16_t *gout, uint16_t *gin, int zero, void *stream);
/* One Adam step (Kingma & Ba) on the accumulated gradients. */
int oecl_unet_adam(oecl_unet *u, float lr, float beta1, float beta2, float eps, void *stream);
/* The preconditioner on the inside nodes (mask) of the nx x ny grid (FP64).  omega = 0: the
 * network alone, z = rms(r) * network(r / rms(r)).  omega > 0: the network between damped
 * Jacobi smoothings, z1 = omega D^-1 r, n = network applied to r1 = r - A z1,
 * z = z1 + c n + omega D^-1 (r1 - c A n) with the A-norm line search c = n.r1 / n.A n of either
 * sign (c = 0 when n.A n is not positive); cx, cy: the stencil arrays of oecl_pcg_solve, D its
 * diagonal. */
int oecl_unet_precondition(oecl_unet *u, uint32_t nx, uint32_t ny, const double *cx, const double *cy,
                           const uint8_t *mask, const double *r, double *z, double omega, void *stream);
/* The pair the network meets inside the smoothed preconditioner: for (q, phi = A^-1 q),
 * r1 = q - A (omega D^-1 q) and e1 = phi - omega D^-1 q (so A e1 = r1). */
int oecl_unet_smooth_pair(oecl_unet *u, uint32_t nx, uint32_t ny, const double *cx, const double *cy,
                          const uint8_t *mask, const double *q, const double *phi, double omega, double *r1,
                          double *e1, void *stream);
/* One training sample (q, phi = A^-1 q): forward on q, the loss L = 1 - (z^T q)^2 /
 * ((z^T A z)(phi^T q)) -- the A-norm error of the best multiple of the preconditioner's output
 * z, relative to phi's A-norm (the flexible CG does not depend on z's scale) -- and backward.
 * cx, cy, mask: the PCG stencil arrays of oecl_pcg_solve.  work: 2 nx ny doubles.  *loss
 * (optional, synchronises) receives L. */
int oecl_unet_train_sample(oecl_unet *u, uint32_t nx, uint32_t ny, const double *cx, const double *cy,
                           const uint8_t *mask, const double *q, const double *phi, double *work, int zero,
                           double *loss, void *stream);
/* Flexible CG (Notay's FCG(m): each new direction is A-orthogonalised against the last mDirs,
 * 1 <= mDirs <= 8; mDirs = 1 is the truncated form) preconditioned as oecl_unet_precondition
 * with omega, FP64 vectors and reductions: phi (initial guess in, solution out) until
 * ||q - A phi|| <= tol ||q||, the norm read every `check` iterations.  Returns OECL_ERR_NOCONV
 * when maxIter is reached.  work: oecl_fcg_work_bytes(np, mDirs) bytes. */
uint64_t oecl_fcg_work_bytes(uint32_t np, uint32_t mDirs);
int oecl_fcg_solve(oecl_unet *u, uint32_t nx, uint32_t ny, const double *cx, const double *cy, const uint8_t *mask,
                   const double *q, double *phi, double tol, uint32_t maxIter, uint32_t check, double omega,
                   uint32_t mDirs, void *work, uint32_t *iters, double *relres, void *stream);
/* The same flexible CG for R problems on one grid in one set of launches (each kernel takes the
 * run as its second grid dimension): q and phi of run b at b nx ny; iters[R] and relres[R]
 * (host) as oecl_fcg_solve gives them.  A run stops at the first check point where it has
 * converged, so each run's phi, iters and relres equal oecl_fcg_solve's on its own problem bit
 * for bit.  Needs oecl_unet_reserve(u, R') with R' >= R first (R <= 4096), which sizes the
 * network's activation and reduction buffers for R images.  work:
 * oecl_fcg_batch_work_bytes(np, mDirs, R) bytes. */
int oecl_unet_reserve(oecl_unet *u, uint32_t R);
uint64_t oecl_fcg_batch_work_bytes(uint32_t np, uint32_t mDirs, uint32_t R);
int oecl_fcg_solve_batch(oecl_unet *u, uint32_t nx, uint32_t ny, const double *cx, const double *cy,
                         const uint8_t *mask, uint32_t R, const double *q, double *phi, double tol, uint32_t maxIter,
                         uint32_t check, double omega, uint32_t mDirs, void *work, uint32_t *iters, double *relres,
                         void *stream);

/* ---- the whole plain step (example01_basic.py:52-97, example02/04 variants) ---------- */

/* Arguments of oecl_step_run.  Counts live on the device: nDev (particles), stepDev (step
 * index, also the Philox stream of the emitter and the index into beamScale).  stats[8]:
 * 0 total absorbed, 1 total emitted secondaries, 2 status bits (1 more hits than absCap,
 * 2 more secondaries and photoelectrons than secCap, 4 capacity exceeded, 8 more
 * photoelectrons in a step than photoCap), 3 cut-cell crossings not found, 4 hits and 5 added
 * particles (secondaries + photoelectrons) of the last step, 6 total photoelectrons.
 * countHist[step] receives the particle count at the start of each step (step < histCap).
 * Photoelectrons (oecl_photo_emit's source, no counterpart in the reference) are generated
 * when photoCount is not NULL: photoCount[step] macro-particles of weight photoWeight[step]
 * (step < nPhoto), added after the step's secondaries. */
typedef struct {
    int32_t dtype;
    void *data;                 /* particle SoA block, 6 rows of `stride` (= capacity) */
    uint64_t stride;
    uint32_t *nDev;
    uint8_t *isInside;          /* [stride] or NULL: the push's inside flag per slot (before the
                                   step's compaction); not written when NULL */
    double *rho, *q;            /* [npExt]: charge of the current particles (in / out), scratch */
    const double *profile;      /* [npExt] beam profile, or NULL */
    const double *beamScale;    /* [nScale] beam scale per step, or NULL */
    uint32_t nScale;
    double *phi, *E;            /* [npExt], [2 npExt] */
    const int32_t *e2nRowPtr, *e2nCol;   /* phiToEAtGridPoints as CSR */
    const double *e2nVal;
    const double *bField;       /* [3 npExt] for typeB 2 */
    double b0, b1, b2;          /* constant B for typeB 1 */
    int32_t typeB;
    double dt, qm, qe;          /* time step, q/m, particle charge */
    void *absorbed, *normals, *remaining;   /* [absCap][6], [absCap][2], [absCap] */
    uint32_t absCap;
    uint32_t *nSec, *secOff;    /* [absCap], [absCap + 1] */
    double *impact;             /* [absCap][4] */
    void *secondaries;          /* [secCap][6] */
    uint32_t secCap;
    int32_t emit;               /* 0: absorb without emission */
    uint32_t seed0, seed1;
    uint32_t *stepDev;
    uint32_t *stats;
    uint32_t *countHist;
    uint32_t histCap;
    void *work;                 /* oecl_step_work_bytes(stride, absCap) */
    void *evPushBegin, *evPushEnd;   /* hipEvent_t recorded around the push kernel, or NULL */
    const uint32_t *photoCount; /* [nPhoto] photoelectron macro-particles per step, or NULL */
    const double *photoWeight;  /* [nPhoto] their weight */
    uint32_t nPhoto;
    double photoE0, photoSigmaE;  /* eV */
    void *photoSlots;           /* [photoCap][6] */
    uint32_t photoCap;
    const oecl_secem_t *secem;  /* NULL: the Furman emitter `prm`; else this simpler emitter */
    double *Eil;                /* [2 npExt] (per run) scratch: E interleaved (Ex, Ey) per node for
                                   the push's gather, or NULL to gather from E */
    int32_t externalPhi;        /* 0: the step solves for phi with the direct solver; 1: phi (per
                                   run) already holds the potential of this step's charge
                                   rho + s(step) * profile, solved by the caller (any method) */
    const double *Bil;          /* [2 npExt] or NULL, typeB 2 with a grid field whose Bz is zero
                                   everywhere: (Bx, By) per node, shared by all runs; the push then
                                   gathers B from it (one 16-byte load per node, the same products
                                   in the same order) and E from Eil, which must be given.  No
                                   reference counterpart (a layout of the same data). */
    const uint8_t *e2nKind;     /* [npExt] or NULL: row classes of phiToEAtGridPoints per node,
                                   bits 0-1 the Ex row, bits 2-3 the Ey row: 0 empty, 1 exactly
                                   (c phi[node - o], -c phi[node + o]) with c = e2nCx, o = 1 (Ex)
                                   or c = e2nCy, o = nxExt (Ey), 2 listed in e2nIrr and summed
                                   from the CSR.  The step then evaluates E without reading the
                                   CSR of the interior rows (the same products and sums). */
    const int32_t *e2nIrr;      /* [nIrr] rows (of 2 npExt) of class 2 */
    uint32_t nIrr;
    double e2nCx, e2nCy;
    /* Arithmetic of the push for FP32 particles (dtype OECL_F32; ignored for FP64), no
     * reference counterpart (the reference computes in FP64 only, particles.pyx:272-418):
     * 0  FP64 registers for the weights, gathers, Boris rotation, wall test and deposit;
     * 1  mixed: weights, gathers and the Boris push in FP32 from FP32 copies of the fields,
     *    the charge split into the four corners, the run sums and rho in FP64; a particle
     *    within nearCells cells of the wall before or after the push (emitted electrons,
     *    grazing orbits) is pushed and tested exactly as in mode 0 (FP64 fields, FP64
     *    rotation, FP64 wall test), so every absorption decision is the FP64 one. */
    int32_t pushArith;
    float *Eil32;               /* [2 npExt] (per run) scratch for pushArith 1: (Ex, Ey) per node in
                                   FP32 (E itself stays FP64) */
    const float *Bil32;         /* [2 npExt] or NULL: Bil in FP32, needed with Bil for pushArith 1 */
    double nearCells;           /* pushArith 1: the wall margin in cells (>= 1) */
    /* The push's deposit tiles (no reference counterpart; a tuning choice that changes only the
     * summation order of rho): depSpan, lanes over which runs of equal cells are summed before
     * the tile (0: 4; else a power of two, 2-64); depTile 0 one LDS tile per push block (per-wave
     * tiles when the block's cells do not fit it), 1 per-wave tiles only. */
    uint32_t depSpan;
    int32_t depTile;
} oecl_step_args;

/* Particles per push block (256): batched run segments are multiples of it, and the per-block
 * run table blockRun of oecl_runs_t has one entry per that many slots.  No reference
 * counterpart; callers that lay out runs check it against their own constant. */
uint32_t oecl_push_block_size(void);

uint64_t oecl_step_work_bytes(uint64_t capacity, uint32_t absCap);

/* A batched step: R independent buildups (a parameter scan) on one grid, one wall and one
 * direct Poisson operator, all advanced by one call (or one replay of a captured graph).  No
 * counterpart in the reference, which runs one buildup per process.  Run r's particles live in
 * [segOff[r], segOff[r] + nDev[r]) of the one SoA block of oecl_step_args (stride = segOff[R]);
 * its segment [segOff[r], segOff[r + 1]) is a multiple of oecl_push_block_size() = 256
 * particles.  blockRun[b] is the run of particles [256 b, 256 b + 256).  Run r's hits use rows [absOff[r], absOff[r + 1]) of
 * absorbed, normals, remaining, nSec, impact (absCap = absOff[R]); secOff holds absOff[R] + R
 * entries (run r's from absOff[r] + r).  Per run, one after another: nDev, stats[8],
 * countHist[histCap], beamScale[nScale], rho, q, phi (npExt each), E (2 npExt).  Emitter: the
 * Furman parameters furman[r] (device), or oecl_step_args.secem for all runs.  Philox keys:
 * keys[2r], keys[2r + 1].  secondaries may be NULL (no copy of the emitted rows); the
 * photoelectron source is not available here. */
typedef struct {
    uint32_t R;
    const uint32_t *segOff;       /* device [R + 1] */
    const uint32_t *blockRun;     /* device [segOff[R] / 256] */
    const uint32_t *absOff;       /* device [R + 1] */
    const oecl_furman_t *furman;  /* device [R] */
    const uint32_t *keys;         /* device [2 R] */
} oecl_runs_t;
uint64_t oecl_step_batch_work_bytes(uint64_t capacity, uint32_t absTotal, uint32_t R);
int oecl_step_run_batch(const oecl_grid_t *g, const oecl_wall_t *wall, const oecl_runs_t *runs,
                        oecl_fastpoisson *fp, const oecl_step_args *a, void *stream);
/* hipEvent helpers for callers without a HIP binding. */
int oecl_event_create(void **ev);
int oecl_event_destroy(void *ev);
int oecl_event_elapsed_ms(void *begin, void *end, float *ms);
/* One full time step with the device-resident counts of `a`; no host synchronisation. */
int oecl_step_run(const oecl_grid_t *g, const oecl_wall_t *wall, const oecl_furman_t *prm,
                  oecl_fastpoisson *fp, const oecl_step_args *a, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* OECL_H */
