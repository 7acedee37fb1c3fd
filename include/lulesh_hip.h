/*
 * liblulesh_hip -- C ABI of the MI355X (gfx950) LULESH Lagrange leap-frog
 * cycle.  Plain C types only: a Julia ccall, a Python ctypes stub or C/C++
 * can bind it.  Every entry point returns 0 (LULESH_OK) or an error code;
 * lulesh_last_error(ctx) gives the text, which for the physics errors is
 * exactly the reference's error() message.
 *
 * The reference has no FFI layer: its path is Julia generic functions on a
 * Domain{Float64}.  Each entry point below names the reference function it
 * replaces (file:line in JuliaLabs/LULESH.jl).
 *
 * Data layout: the context owns structure-of-arrays FP64 device buffers, one
 * per Domain field (src/types.jl:4-87), 0-based, element id
 * e = plane*s^2 + row*s + col, node id n = plane*(s+1)^2 + row*(s+1) + col
 * (the numbering of buildMesh!, src/domain.jl:117-174).  Connectivity,
 * symmetry node sets and face neighbours of the structured mesh are computed
 * arithmetically on the device; they equal the arrays Domain(prob) builds
 * (src/domain.jl:615-661, 773-916).  Host buffers passed in are borrowed for
 * the duration of the call only.
 *
 * Threading: a context is not re-entrant; separate contexts are independent.
 * All work runs on the context's own HIP stream and each call returns after
 * that stream is synchronised.
 */
#ifndef LULESH_HIP_H
#define LULESH_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------------
 * 1..9 are the reference's physics errors, numbered in program order; when
 * several occur in one cycle the lowest (earliest in the reference's order)
 * is reported, which is the one the reference would have raised. */
#define LULESH_OK 0
#define LULESH_ERR_EARLY_VOLUME 1   /* "Early Volume Error"                   src/domain.jl:1104 */
#define LULESH_ERR_MID_VOLUME 2     /* "Mid Volume Error"                     src/domain.jl:1758 */
#define LULESH_ERR_HG_VOLUME 3      /* "Volume Error: Volume is negative"     src/domain.jl:1727 */
#define LULESH_ERR_KIN_VOLUME 4     /* "Volume Error: calcKinematicsForElems" src/domain.jl:2118 */
#define LULESH_ERR_LAG_VOLUME 5     /* "Volume Error :2157"                   src/domain.jl:2169 */
#define LULESH_ERR_BC 6             /* "Error" (invalid elemBC mask)          src/domain.jl:2343 */
#define LULESH_ERR_QSTOP 7          /* "QStopError"                           src/domain.jl:2580 */
#define LULESH_ERR_EOS_VOLUME 8     /* "Volume Error :2887"                   src/domain.jl:2905 */
#define LULESH_ERR_UPDATE_VOLUME 9  /* "Volume Error :2908"                   src/domain.jl:2926 */
#define LULESH_ERR_INVALID_ARG 100  /* bad argument (size, name, pointer)  */
#define LULESH_ERR_HIP 101          /* a HIP runtime call failed           */
#define LULESH_ERR_UNSUPPORTED 102  /* valid request this build does not do */

/* ---- arithmetic quirks of the reference --------------------------------
 * On by default so results match the reference; a caller may clear them to
 * get the LULESH 2.0 (C++) arithmetic. */
#define LULESH_Q1_FACE5_NORMAL 1 /* src/domain.jl:1050-1055: pfy[1] += areaY twice, no pfz[1] */
#define LULESH_Q2_DVDX_SWAP 2    /* src/domain.jl:1216-1229: volume-derivative slots 2<->4   */
#define LULESH_Q3_XIP_SELF 4     /* src/domain.jl:2349: delv_xi[lxip[i]] without the +1      */
#define LULESH_QUIRKS_REFERENCE 7

typedef struct lulesh_ctx lulesh_ctx;

/* Domain(prob::LuleshProblem) for one structured rank (src/domain.jl:467-771,
 * src/LULESH.jl:31-49): builds the Sedov problem with nx elements per edge
 * on the device.  nr/balance/cost are accepted as in LuleshProblem but, as in
 * the reference, never reach the arithmetic (the compute loops walk the
 * identity matElemlist, src/domain.jl:657-661).  `device` is a HIP ordinal. */
int lulesh_create(int32_t nx, int32_t nr, int32_t balance, int32_t cost, int32_t quirks,
                  int32_t device, lulesh_ctx **out);
int lulesh_destroy(lulesh_ctx *ctx);
/* back to the state Domain(prob) returns (same nx, quirks) */
int lulesh_reset(lulesh_ctx *ctx);

int64_t lulesh_num_elem(const lulesh_ctx *ctx);
int64_t lulesh_num_node(const lulesh_ctx *ctx);
const char *lulesh_last_error(const lulesh_ctx *ctx);
const char *lulesh_error_string(int code);

/* Domain field access by the reference's field name (src/types.jl:4-87):
 * element fields e p q ql qq v volo delv vdov arealg ss elemMass vnew
 * delv_xi delv_eta delv_zeta delx_xi delx_eta delx_zeta dxx dyy dzz;
 * node fields x y z xd yd zd xdd ydd zdd fx fy fz nodalMass.  n must equal
 * the field length.  Fields the fused kernels do not keep by default
 * (ql qq dxx dyy dzz xdd ydd zdd fx fy fz) are written only while
 * lulesh_set_keep_aux(ctx, 1) is in force. */
int lulesh_download(lulesh_ctx *ctx, const char *field, double *dst, int64_t n);
int lulesh_upload(lulesh_ctx *ctx, const char *field, const double *src, int64_t n);
/* elemBC masks (src/bc.jl bit layout), length numElem */
int lulesh_download_elem_bc(lulesh_ctx *ctx, uint32_t *dst, int64_t n);
int lulesh_upload_elem_bc(lulesh_ctx *ctx, const uint32_t *src, int64_t n);
/* Domain scalars by name (src/types.jl:91-126): time deltatime cycle
 * dtcourant dthydro stoptime dtmax dtfixed deltatimemultlb deltatimemultub
 * u_cut hgcoef qstop monoq_max_slope monoq_limiter_mult e_cut p_cut ss4o3
 * q_cut v_cut qlc_monoq qqc_monoq qqc eosvmax eosvmin pmin emin dvovmax
 * refdens */
int lulesh_get_scalar(lulesh_ctx *ctx, const char *name, double *value);
int lulesh_set_scalar(lulesh_ctx *ctx, const char *name, double value);
int lulesh_set_keep_aux(lulesh_ctx *ctx, int32_t on);

/* ---- the hot path -------------------------------------------------------*/
/* timeIncrement!(domain)            src/domain.jl:920-961 */
int lulesh_time_increment(lulesh_ctx *ctx);
/* lagrangeLeapFrog(domain)          src/domain.jl:3067-3080 */
int lulesh_lagrange_leapfrog(lulesh_ctx *ctx);
/* its three phases, for phase-by-phase checks:
 * lagrangeNodal(domain)             src/domain.jl:1840-1872
 * lagrangeElements(domain)          src/domain.jl:2932-2962
 * calcTimeConstraintsForElems(domain) src/domain.jl:3058-3064 */
int lulesh_lagrange_nodal(lulesh_ctx *ctx);
int lulesh_lagrange_elements(lulesh_ctx *ctx);
int lulesh_calc_time_constraints(lulesh_ctx *ctx);
/* The driver loop of examples/benchmark.jl:82-98:
 *   while time < stoptime: timeIncrement!; lagrangeLeapFrog; break if cycle >= num_iters
 * run on the device (cycles replayed from a captured hipGraph, the loop
 * condition evaluated on the device).  *cycles_run = cycles executed. */
int lulesh_run(lulesh_ctx *ctx, int64_t num_iters, int64_t *cycles_run);
/* verifyAndWriteFinalOutput's numbers (src/utils.jl:124-152):
 * out = {e[1], MaxAbsDiff, TotalAbsDiff, MaxRelDiff} over plane 0 */
int lulesh_verify(lulesh_ctx *ctx, double out[4]);

/* ---- measurement --------------------------------------------------------
 * With timing on, every kernel launch of the context is bracketed by HIP
 * events on the context's stream; lulesh_kernel_time returns the summed
 * device time (ms) and launch count of kernel `id` since the last reset. */
#define LULESH_K_NODAL 0 /* lagrangeNodal: corner forces in LDS, fixed-order gather, node update */
#define LULESH_K_KIN 1   /* kinematics + MonoQ gradients */
#define LULESH_K_QEOS 2  /* MonoQ limiter, EOS, volume update, dt partial minima */
#define LULESH_K_DT 3    /* dt minima + next timeIncrement! */
#define LULESH_K_FLIP 4  /* makes the positions k_nodal wrote current */
#define LULESH_K_COUNT 5
int lulesh_set_timing(lulesh_ctx *ctx, int32_t on);
/* Device arithmetic check (no reference counterpart): evaluates, on device
 * `device`, the scalar routines every kernel relies on, so tests can pin
 * their rounding: op 0 = cbrt(a) (correctly rounded), 1 = sqrt(a),
 * 2 = a / b.  a, b, out are host arrays of n doubles (b unused for 0, 1). */
int lulesh_math_check(int32_t op, const double *a, const double *b, double *out, int64_t n,
                      int32_t device);
int lulesh_kernel_time(lulesh_ctx *ctx, int32_t id, double *total_ms, int64_t *launches);
const char *lulesh_kernel_name(int32_t id);

#ifdef __cplusplus
}
#endif
#endif
