/*
 * CPU oracle for the LULESH Lagrange leap-frog cycle -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain C++17 restatement of the reference Julia path (src/domain.jl:34-3080,
 * src/utils.jl:36-164, examples/benchmark.jl:80-106).  Only tests/, the smoke
 * check and the CPU-baseline leg of bench.py may load it, and only as the
 * checker: it is never the thing measured or shipped.
 *
 * How each part is pinned (tests/test_oracle_pins.py; none re-types a
 * formula of this file): cbrt -- exact integer arithmetic; element volume
 * -- unit cube, parallelepipeds, Gauss quadrature; volume derivative --
 * finite differences (slots 2/4 pin quirk Q2); node normals -- zero sum /
 * gradient of the volume, face-5 closed form (quirk Q1); shape derivatives,
 * velocity gradient, characteristic length -- affine/linear closed forms;
 * FB hourglass -- affine motion and pure modes on a cube; MonoQ gradients --
 * box and parallelepiped closed forms; MonoQ limiter -- closed forms of
 * every branch; EOS -- rest state and the compression energy balance;
 * timeIncrement! -- branch table; Domain(prob) -- independent
 * constructions; cycle 1 -- closed form; the whole quirk-free cycle -- the
 * LULESH 2.0 published s=30 result (932 iterations, e = 2.025075e+05);
 * quirk Q3 -- changes results and keeps the reference's non-termination at
 * s=30.  In tests/test_pins_monoq_verify.py: verify's symmetry numbers --
 * hand-computed on an uploaded asymmetric plane; the MonoQ neighbour reads
 * and quirk Q3 -- an exact-in-binary compression layer along each axis and
 * side.  In tests/test_regions.py: the region index sets -- the nr = 1
 * closed form, the invariants the algorithm fixes, hand-derived values of
 * the partial sort, and the generator's seed (SHA-256 of the seed word).
 * Parity unpinned at the last ulp: Julia's cbrt (not correctly rounded) and
 * its @simd reassociation of the <= 8-term corner gathers.  Parity unpinned
 * outright: the region draw stream for nr > 1 (no Julia run to compare
 * with; it never reaches the arithmetic).
 *
 * Index conventions: every array here is 0-based in memory.  Index *values*
 * keep the reference's stored bases: nodelist, the six face-neighbour lists,
 * the symm sets and nodeElemStart hold 0-based ids, nodeElemCornerList and matElemlist hold
 * 1-based ids (src/domain.jl:158-165, 623-661, 773-820).
 */
#ifndef LULESH_ORACLE_H
#define LULESH_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* quirk bits; ORACLE_QUIRKS_REFERENCE (all on) reproduces the reference. */
#define ORACLE_Q1_FACE5_NORMAL 1 /* src/domain.jl:1050-1055 */
#define ORACLE_Q2_DVDX_SWAP 2    /* src/domain.jl:1216-1229 */
#define ORACLE_Q3_XIP_SELF 4     /* src/domain.jl:2349      */
#define ORACLE_QUIRKS_REFERENCE 7

typedef struct oracle_domain oracle_domain;

/* Domain(prob) for one structured rank; nr/balance/cost have no effect on
 * the arithmetic (regions are never read by the compute path). */
oracle_domain *oracle_create(int nx, int quirks, int nthreads);
/* the same with the region index sets of createRegionIndexSets!
 * (src/domain.jl:338-465) for nr regions, balance and cost (oracle_create
 * uses the reference's defaults 11, 1, 1); exported by oracle_ifield as
 * regNumList, regElemlist, regCSR, regReps, regSorted.  NULL for nr < 1 or
 * negative balance/cost, or a cost denominator beyond Int64. */
oracle_domain *oracle_create_regions(int nx, int quirks, int nthreads, int nr, int balance,
                                     int cost);
/* the state Random.seed!(0) gives Julia 1.7's xoshiro256++ generator, as
 * the region builder uses it (the SHA-256 digest of four zero bytes) */
void oracle_julia_rng_state(uint64_t out[4]);
void oracle_destroy(oracle_domain *d);

/* return 0 or one of the error codes below (same numbers as the product) */
int oracle_time_increment(oracle_domain *d);
int oracle_lagrange_leapfrog(oracle_domain *d);
int oracle_lagrange_nodal(oracle_domain *d);
int oracle_lagrange_elements(oracle_domain *d);
int oracle_calc_time_constraints(oracle_domain *d);
/* driver loop of examples/benchmark.jl:82-98; returns error code */
int oracle_run(oracle_domain *d, int64_t num_iters, double *elapsed_s);

/* out[0]=e[1], out[1]=MaxAbsDiff, out[2]=TotalAbsDiff, out[3]=MaxRelDiff */
void oracle_verify(const oracle_domain *d, double out[4]);

double *oracle_dfield(oracle_domain *d, const char *name, int64_t *n);
int64_t *oracle_ifield(oracle_domain *d, const char *name, int64_t *n);
uint32_t *oracle_elemBC(oracle_domain *d, int64_t *n);
int oracle_get_scalar(const oracle_domain *d, const char *name, double *v);
int oracle_set_scalar(oracle_domain *d, const char *name, double v);

/* element-level functions, for the pin tests.  Node order as in the
 * reference: x[0..7] is nodes 1..8. */
double oracle_cbrt(double x);
double oracle_calc_elem_volume(const double *x, const double *y, const double *z);
void oracle_calc_elem_node_normals(const double *x, const double *y, const double *z,
                                   int quirks, double *B /* [8][3] */);
double oracle_calc_elem_shape_function_derivatives(const double *x, const double *y,
                                                   const double *z, double *b /* [8][3] */);
void oracle_calc_elem_volume_derivative(const double *x, const double *y, const double *z,
                                        int quirks, double *dvdx, double *dvdy, double *dvdz);
double oracle_calc_elem_characteristic_length(const double *x, const double *y,
                                              const double *z, double volume);
void oracle_calc_elem_velocity_gradient(const double *xv, const double *yv, const double *zv,
                                        const double *b, double detJ, double *d /* [6] */);
/* Flanagan-Belytschko force of one element: hgf[8][3] */
void oracle_calc_elem_fb_hourglass(const double *x, const double *y, const double *z,
                                   const double *xd, const double *yd, const double *zd,
                                   double determ, double ss, double mass, double hgcoef,
                                   int quirks, double *hgf);
/* calcMonotonicQGradientsForElems of one element:
 * out = {delx_xi, delx_eta, delx_zeta, delv_xi, delv_eta, delv_zeta} */
void oracle_monoq_gradients_elem(const double *X, const double *Y, const double *Z,
                                 const double *XV, const double *YV, const double *ZV,
                                 double volo, double vnew, double *out);
/* calcMonotonicQRegionForElems of one element (limiter mult 2, max slope 1):
 * dv/dm/dp/dx[3] per direction xi, eta, zeta; q = {qlin, qquad} */
void oracle_monoq_region_elem(const double *dv, const double *dm, const double *dp,
                              const double *dx, double vdov, double elemMass, double volo,
                              double vnew, double qlc, double qqc, double *q);
/* one element of evalEOSForElems (+ calcSoundSpeed); out = {p, e, q, ss} */
void oracle_eval_eos_elem(double vnew, double v_unused, double delv, double e_old,
                          double p_old, double q_old, double ql, double qq, double *out);
/* timeIncrement! on explicit scalars; st = {time, deltatime, cycle, dtcourant,
 * dthydro, stoptime, dtmax, dtfixed, multlb, multub} is updated in place */
void oracle_time_increment_scalars(double *st);

#ifdef __cplusplus
}
#endif
#endif
