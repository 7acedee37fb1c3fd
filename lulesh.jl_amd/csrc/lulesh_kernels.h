// Device-side data of one LULESH domain and the launcher interface between
// the kernels and the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lulesh {

constexpr int ERR_NONE = 0x7fffffff;
constexpr int ERR_EARLY_VOLUME = 1, ERR_MID_VOLUME = 2, ERR_HG_VOLUME = 3, ERR_KIN_VOLUME = 4,
              ERR_LAG_VOLUME = 5, ERR_BC = 6, ERR_QSTOP = 7, ERR_EOS_VOLUME = 8,
              ERR_UPDATE_VOLUME = 9;

// Domain scalars (src/types.jl:91-126) plus the device-side driver state.
struct DevScalars {
  double time, deltatime, dtcourant, dthydro;
  double stoptime, dtmax, dtfixed, deltatimemultlb, deltatimemultub;
  double u_cut, hgcoef, qstop, monoq_max_slope, monoq_limiter_mult;
  double e_cut, p_cut, ss4o3, q_cut, v_cut, qlc_monoq, qqc_monoq, qqc;
  double eosvmax, eosvmin, pmin, emin, dvovmax, refdens;
  long long cycle;
  long long num_iters;  // driver loop bound of the current lulesh_run
  int err;              // lowest error code raised (ERR_NONE when clean)
  int active;           // 1 while the current cycle executes (driver loop)
  int parity;           // which NodeSet holds the current x..zd
  int pad_;
};

// node positions and velocities; two sets, ping-ponged by the fused nodal
// kernel (reads set `parity`, writes the other, then parity flips)
struct NodeSet {
  double *x, *y, *z, *xd, *yd, *zd;
};

// structure-of-arrays FP64 fields, one device buffer each
struct DevFields {
  // element
  double *e, *p, *q, *ql, *qq, *v, *volo, *delv, *vdov, *arealg, *ss, *elemMass;
  double *vnew, *delv_xi, *delv_eta, *delv_zeta, *delx_xi, *delx_eta, *delx_zeta;
  double *dxx, *dyy, *dzz;
  uint32_t *elemBC;
  // node
  NodeSet ns[2];
  double *xdd, *ydd, *zdd, *fx, *fy, *fz, *nodalMass;
};

enum KernelId {
  KID_NODAL = 0,
  KID_KIN = 1,
  KID_QEOS = 2,
  KID_DT = 3,
  KID_FLIP = 4,
  KID_DTPART = 5,
  KID_TIME_INCREMENT = 6,
  KID_RUN_BEGIN = 7,
  KID_CLEAR_ERR = 8,
};

struct LaunchArgs {
  DevScalars *sc;
  DevFields f;
  double *dtpart;  // 2 x qeos_blocks(E) time-constraint partial minima
  int s;
  int64_t E, N;
  int quirks;
  int gated;       // kernels exit when sc->active == 0
  int keep_aux;    // also write fx.., xdd.., dxx.., ql, qq
  int loop_mode;   // k_dt evaluates the driver-loop test and the next dt
  int64_t num_iters;
};

hipError_t launch_init(hipStream_t st, int s, int64_t E, int64_t N, const DevFields &f,
                       DevScalars *sc);
hipError_t launch_kernel(int id, hipStream_t st, const LaunchArgs &a);
int64_t qeos_blocks(int64_t E);
hipError_t launch_math(hipStream_t st, int op, const double *a, const double *b, double *out,
                       int64_t n);

}  // namespace lulesh
