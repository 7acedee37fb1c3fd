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
#define LULESH_ERR_HOST_MEMORY 103  /* a host allocation failed (no C++ exception
                                       ever leaves an entry point); the
                                       context's state is then unspecified */
#define LULESH_ERR_SCRATCH 104      /* the domain does not fit the fixed scratch
                                       array of lulesh_lagrange_leapfrog_noalloc
                                       (nx > LULESH_NOALLOC_MAX_NX) */

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
 * on the device.  nr/balance/cost are those of LuleshProblem: they define
 * the region index sets (lulesh_region_download) and, as in the reference,
 * never reach the arithmetic (the compute loops walk the identity
 * matElemlist, src/domain.jl:657-661).  `device` is a HIP ordinal.  nx must
 * lie in 1..800, nr in 1..LULESH_MAX_REGIONS, balance and cost >= 0 with
 * sum_{r<=nr} r^balance within Int64 (LULESH_ERR_INVALID_ARG otherwise). */
#define LULESH_MAX_REGIONS 4096
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

/* ---- regions ------------------------------------------------------------
 * createRegionIndexSets!(domain, nr, balance) of Domain(prob)
 * (src/domain.jl:313-465), per instance, built on the host the first time
 * one of the two calls below needs them.  lulesh_region_download copies one
 * array as the reference stores it: "regNumList" (numElem region numbers,
 * 1-based), "regElemlist" (numElem element ids, 1-based, grouped by region
 * in sortRegions' order), "regCSR" (nr running end offsets, as the fill loop
 * leaves them), "regReps" and "regSorted" (nr each, as sortRegions' partial
 * bubble sort leaves them); n must equal the length.  The reference draws
 * the regions with Julia 1.7's default generator seeded by Random.seed!(0)
 * (xoshiro256++, seeded from the SHA-256 of the seed's words); this library
 * restates that generator, so nr > 1 matches the reference only as far as
 * that restatement does (nr = 1 draws nothing).
 * lulesh_set_cost_mode(ctx, 1) switches on LULESH 2.0's load-imbalance cost
 * model: k_qeos evaluates each element's energy update regReps times for
 * its region (1, 1+cost or 10(1+cost)), as LULESH 2.0's EvalEOSForElems
 * does; every result stays bit-identical, only the time changes.  Off by
 * default; the first switch-on uploads 4 bytes per element. */
int lulesh_region_download(lulesh_ctx *ctx, int32_t inst, const char *name, int64_t *dst,
                           int64_t n);
int lulesh_set_cost_mode(lulesh_ctx *ctx, int32_t on);

/* ---- one cycle without allocation --------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc -- and that hold the Domain's fields themselves in
 * device memory (a Julia GPUDomain of device arrays, for instance).
 * lulesh_lagrange_leapfrog_noalloc runs one lagrangeLeapFrog
 * (src/domain.jl:3067-3080), with the kernels lulesh_lagrange_leapfrog
 * runs, on those fields in place.  It allocates nothing, on the device or
 * the host: the cycle's scratch (the second node set, the kinematics
 * temporaries, the dt partial minima, the block tables and the device-side
 * scalars) lives in a fixed per-GPU array of LULESH_NOALLOC_SCRATCH_BYTES,
 * reserved when the library's code is loaded on a GPU and held until the
 * process ends.  nx above LULESH_NOALLOC_MAX_NX returns LULESH_ERR_SCRATCH.
 * All calls on one GPU share that GPU's array and leave their contents in
 * it (each call sets up what it reads first): calls that overlap on one GPU
 * (from several host threads or streams) race; the caller serialises them.
 *   dom      nx, quirks and device pointers to the Domain fields, 0-based,
 *            numbered as above: element fields nx^3 doubles (elemBC: nx^3
 *            uint32 masks), node fields (nx+1)^3 doubles; e p q v ss delv
 *            vdov arealg and x y z xd yd zd are updated, the others read.
 *   scalars  host array of the 29 Domain scalars in the order of the name
 *            list of lulesh_get_scalar (cycle as a double): all are read,
 *            and dtcourant (index 3) and dthydro (index 4) are written as
 *            calcTimeConstraintsForElems leaves them.
 *   stream   a hipStream_t (NULL: the null stream); the call returns after
 *            synchronising it.
 *   device   the HIP ordinal the pointers belong to (made current).
 * Returns 0, the lowest physics error 1..9 raised (the fields are then
 * undefined, as after the reference's exception), or LULESH_ERR_*. */
#define LULESH_NOALLOC_SCRATCH_BYTES (32u << 20)
#define LULESH_NOALLOC_MAX_NX 68
#define LULESH_NUM_SCALARS 29
typedef struct lulesh_device_domain {
  int32_t nx, quirks;
  double *e, *p, *q, *v, *volo, *delv, *vdov, *arealg, *ss, *elemMass;
  uint32_t *elemBC;
  double *x, *y, *z, *xd, *yd, *zd, *nodalMass;
} lulesh_device_domain;
int lulesh_lagrange_leapfrog_noalloc(const lulesh_device_domain *dom, double *scalars,
                                     void *stream, int32_t device);

/* ---- ragged batches -----------------------------------------------------
 * No reference counterpart (the reference runs one Domain per process).  A
 * batch holds n independent Sedov domains, laid out back to back in the
 * same flat SoA buffers, each with its own size nx[i], deposited energy
 * einit_scale[i] x the reference's einit (dt0 follows from it), initial
 * hgcoef[i] and region count nr[i] (NULL arrays take the reference values
 * 1.0, 3.0 and 1).  Every cycle-level entry point above advances all
 * instances with one launch per kernel and per-instance dt; an instance
 * whose driver loop has ended is skipped, so its state stays exactly what
 * a lone run leaves.  The single-domain entry points address instance 0. */
/* n <= LULESH_MAX_INSTANCES, each nx[i] in 1..800, and the instances' nodes
 * (sum of (nx[i]+1)^3) below 2^31; LULESH_ERR_INVALID_ARG otherwise */
#define LULESH_MAX_INSTANCES (1 << 20)
int lulesh_batch_create(int32_t n, const int32_t *nx, const double *einit_scale,
                        const double *hgcoef, const int32_t *nr, int32_t quirks, int32_t device,
                        lulesh_ctx **out);
int32_t lulesh_batch_size(const lulesh_ctx *ctx);
int64_t lulesh_batch_num_elem(const lulesh_ctx *ctx, int32_t inst);
int64_t lulesh_batch_num_node(const lulesh_ctx *ctx, int32_t inst);
int lulesh_batch_download(lulesh_ctx *ctx, int32_t inst, const char *field, double *dst,
                          int64_t n);
int lulesh_batch_upload(lulesh_ctx *ctx, int32_t inst, const char *field, const double *src,
                        int64_t n);
int lulesh_batch_download_elem_bc(lulesh_ctx *ctx, int32_t inst, uint32_t *dst, int64_t n);
int lulesh_batch_upload_elem_bc(lulesh_ctx *ctx, int32_t inst, const uint32_t *src, int64_t n);
int lulesh_batch_get_scalar(lulesh_ctx *ctx, int32_t inst, const char *name, double *value);
int lulesh_batch_set_scalar(lulesh_ctx *ctx, int32_t inst, const char *name, double value);
/* the driver loop for every instance at once; cycles_run[i] per instance.
 * An instance that raises one of the reference's errors stops there (its
 * state is then undefined, as after the reference's exception) while the
 * others run on; the call returns the lowest error code raised by any
 * instance and lulesh_batch_error gives each instance's own code (0 if it
 * ran cleanly) for the last run. */
int lulesh_batch_run(lulesh_ctx *ctx, int64_t num_iters, int64_t *cycles_run);
int lulesh_batch_error(const lulesh_ctx *ctx, int32_t inst);
int lulesh_batch_verify(lulesh_ctx *ctx, int32_t inst, double out[4]);

/* ---- differentiation ----------------------------------------------------
 * Replaces the reference's Enzyme reverse-mode call on lagrangeLeapFrog
 * (examples/benchmark.jl:48-51,85-86) with hand-written adjoint kernels.
 * lulesh_grad_run runs `ncycles` cycles of the driver loop from the current
 * state (checkpointing it every `checkpoint_every` cycles), then sweeps back
 * through them, recomputing each segment from its checkpoint, and returns
 *   *loss        = sum_e e[e] * elemMass[e] at the end (total internal energy),
 *   grad_e[E]    = d loss / d e at the start,
 *   grad_params  = d loss / d {hgcoef, qlc_monoq, qqc_monoq},
 *   dt_schedule  = the ncycles time steps the forward pass took (may be NULL).
 * The time step is held fixed at the recorded schedule (an input leaf, as
 * timeIncrement! is outside the differentiated call in the reference); the
 * context is left in the forward pass's end state.  One-domain contexts only
 * (LULESH_ERR_UNSUPPORTED otherwise).  The kernels of this section address
 * nodes and elements by 32-bit byte offsets, which every accepted nx (<= 800)
 * satisfies; the setup still checks it (LULESH_ERR_UNSUPPORTED).
 * lulesh_replay advances the state `ncycles` cycles with those time steps
 * instead of timeIncrement! (for finite-difference checks of the gradient). */
int lulesh_grad_run(lulesh_ctx *ctx, int64_t ncycles, int64_t checkpoint_every, double *grad_e,
                    double *grad_params, double *loss, double *dt_schedule);
/* lulesh_grad_run in bounded memory (the same Enzyme call,
 * examples/benchmark.jl:48-51,85-86, over many cycles): binomial
 * checkpointing, Griewank & Walther's "revolve" (ACM TOMS Algorithm 799),
 * keeps at most `snapshots` saved states (6N + 5E doubles each, the start
 * state included) and two cycle slots (6N + 14E each) whatever ncycles is,
 * and re-runs cycles as the optimal binomial schedule requires: with
 * c = min(snapshots, ncycles) and r the least integer with
 * C(c + r, c) >= ncycles, ncycles + r*ncycles - C(c + r, c + 1) cycles
 * in all (the first pass is the schedule's first descent, each cycle's
 * last run keeps its temporaries for the adjoint).  Outputs are those of
 * lulesh_grad_run, bit for bit; *forward_cycles (may be NULL) returns the
 * cycles run in all. */
int lulesh_grad_run_revolve(lulesh_ctx *ctx, int64_t ncycles, int64_t snapshots, double *grad_e,
                            double *grad_params, double *loss, double *dt_schedule,
                            int64_t *forward_cycles);
/* the cycles lulesh_grad_run_revolve runs for (ncycles, snapshots), first
 * pass included (host-only; -1 for arguments < 1) */
int64_t lulesh_revolve_cycles(int64_t ncycles, int64_t snapshots);
int lulesh_replay(lulesh_ctx *ctx, const double *dt, int64_t ncycles);
/* lulesh_vjp_cycle is the reference's per-cycle call
 * Enzyme.autodiff(lagrangeLeapFrog, Duplicated(domain, shadow))
 * (examples/benchmark.jl:86): it advances the state by one lagrangeLeapFrog
 * (at the current deltatime, as lulesh_lagrange_leapfrog does) and maps the
 * context's shadow -- the adjoint of the state fields x y z xd yd zd e p q v
 * ss -- in place from the adjoint of the state after the cycle to the
 * adjoint of the state before it; grad_params[3] (may be NULL) accumulates
 * (+=) d/d{hgcoef, qlc_monoq, qqc_monoq}.  deltatime, volo, elemMass and
 * nodalMass are held fixed (no adjoint is returned for them) and the
 * cycle's temporaries carry none.  The shadow starts at zero and is moved
 * one field at a time by lulesh_shadow_upload/download (names as for
 * lulesh_download, the eleven state fields only); lulesh_grad_run
 * overwrites it.  One-domain contexts only. */
int lulesh_vjp_cycle(lulesh_ctx *ctx, double *grad_params);
int lulesh_shadow_upload(lulesh_ctx *ctx, const char *field, const double *src, int64_t n);
int lulesh_shadow_download(lulesh_ctx *ctx, const char *field, double *dst, int64_t n);

/* ---- measurement --------------------------------------------------------
 * With timing on, every kernel launch of the context is bracketed by HIP
 * events on the context's stream; lulesh_kernel_time returns the summed
 * device time (ms) and launch count of kernel `id` since the last reset. */
#define LULESH_K_NODAL 0 /* lagrangeNodal: corner forces in LDS, fixed-order gather, node update */
#define LULESH_K_KIN 1   /* kinematics + MonoQ gradients */
#define LULESH_K_QEOS 2  /* MonoQ limiter, EOS, volume update, dt partial minima */
#define LULESH_K_DT 3    /* dt minima + next timeIncrement! */
#define LULESH_K_COUNT 4
int lulesh_set_timing(lulesh_ctx *ctx, int32_t on);
/* k_nodal geometry (no reference counterpart): each workgroup marches
 * `depth` node planes of its node column (16x14 nodes for large domains);
 * in a single-domain context the first chunk takes one plane more, so
 * that every chunk but the last marches depth + 1 element planes.
 * 0 restores the default: per instance, node planes split into equal
 * chunks whose depth minimises (launch rounds x planes per workgroup), a
 * round being CUs x the workgroups one CU holds.  Tile shape: per
 * instance the etx x (256/etx) element tile, etx in [lo, hi], that covers
 * the node columns with the fewest tiles (lulesh_set_nodal_tiles; 0, 0
 * restores the default 16..32; 2 <= lo <= hi <= 128); lulesh_nodal_tile
 * reports an instance's {etx, ety, tiles along x, tiles along y}.  Neither
 * changes any result: every node sums its corners in the same order.
 * lulesh_autotune times k_nodal on the live state for five tile-width
 * ranges x the default and depths 2..32 (with its parity flip off the
 * kernel only writes the inactive node set, so the state is unchanged; the
 * fx.. and xdd.. fields of lulesh_set_keep_aux and every instance's error
 * code also stay as they were) and keeps the fastest pair; *chosen (the
 * depth of instance 0) and *best_ms (per launch) may be NULL. */
int lulesh_set_nodal_depth(lulesh_ctx *ctx, int32_t depth);
int32_t lulesh_nodal_depth(const lulesh_ctx *ctx);
int lulesh_set_nodal_tiles(lulesh_ctx *ctx, int32_t lo, int32_t hi);
int lulesh_nodal_tile(const lulesh_ctx *ctx, int32_t inst, int32_t *shape);
int lulesh_autotune(lulesh_ctx *ctx, int32_t *chosen, double *best_ms);
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
