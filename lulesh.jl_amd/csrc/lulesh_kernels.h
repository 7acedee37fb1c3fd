// Device-side data of a batch of LULESH domains and the launcher interface
// between the kernels and the C ABI.  A single domain is a batch of one.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lulesh_plan.h"

namespace lulesh {

constexpr int ERR_NONE = 0x7fffffff;
constexpr int ERR_EARLY_VOLUME = 1, ERR_MID_VOLUME = 2, ERR_HG_VOLUME = 3, ERR_KIN_VOLUME = 4,
              ERR_LAG_VOLUME = 5, ERR_BC = 6, ERR_QSTOP = 7, ERR_EOS_VOLUME = 8,
              ERR_UPDATE_VOLUME = 9;

// Domain scalars (src/types.jl:91-126) plus the device-side driver state,
// one per instance.
struct DevScalars {
  double time, deltatime, dtcourant, dthydro;
  double stoptime, dtmax, dtfixed, deltatimemultlb, deltatimemultub;
  double u_cut, hgcoef, qstop, monoq_max_slope, monoq_limiter_mult;
  double e_cut, p_cut, ss4o3, q_cut, v_cut, qlc_monoq, qqc_monoq, qqc;
  double eosvmax, eosvmin, pmin, emin, dvovmax, refdens;
  long long cycle;
  long long num_iters;  // driver loop bound of the current run
  int err;              // lowest error code raised (ERR_NONE when clean)
  int active;           // 1 while the current cycle executes (driver loop)
  int parity;           // which NodeSet holds the current x..zd
  int nodal_done;       // k_nodal workgroups of this instance finished in the
                        // running launch (the last one flips parity, resets it)
};

// node positions and velocities; two sets, ping-ponged by the fused nodal
// kernel (reads set `parity`, writes the other, then parity flips)
struct NodeSet {
  double *x, *y, *z, *xd, *yd, *zd;
};

// structure-of-arrays FP64 fields, one flat device buffer each; instance i
// owns elements [e0, e0+E) and nodes [n0, n0+N) of every buffer
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

// one instance of a (ragged) batch
struct InstDesc {
  int s;                 // elements per edge
  int etx, ety;          // k_nodal tile: etx x ety element lanes (nodal_span)
  int tx, ty;            // k_nodal tiles along x and y
  int64_t e0, n0;        // offsets into the flat arrays
  int64_t E, N;          // sizes
  int eb0, ebn;          // element-kernel blocks [eb0, eb0+ebn)
  int nb0, nbn;          // node-kernel blocks
  int gb0, gbn;          // k_nodal blocks
  int cz;                // k_nodal node planes per chunk (the z-march depth)
  int cz0, nch;          // node planes of the first chunk, chunks per column
  double einit_scale;    // deposited energy = scale * the reference's einit
  double hgcoef;         // initial hgcoef
};

struct Batch {
  int n;                    // instances
  const InstDesc *inst;     // device table [n]
  const int *eblock_inst;   // instance of each element-kernel block
  const int *nblock_inst;   // instance of each node-kernel block
  const int *gblock_inst;   // instance of each k_nodal block
  const int *gblock_order;  // k_nodal block run by each launched workgroup, -1: none
  const KinWork *kin_work;  // k_kin chunk run by each launched workgroup (kin_table)
  const int *qeos_order;    // k_qeos block run by each launched workgroup, -1: none
  int eblocks, nblocks, gblocks;
  int glaunch;              // k_nodal workgroups launched (a multiple of 8)
  int nodal_lead;           // CUs per XCD: k_nodal workgroups dealt to an XCD
                            // before its CUs hold one each
  int kin_launch;           // k_kin workgroups launched
  int qeos_launch;          // k_qeos workgroups launched
  DevScalars *sc;           // [n]
  DevFields f;
  double *dtpart;           // 2 per element-kernel block
  int quirks;
  int gated;                // kernels skip instances whose active flag is 0
  int keep_aux;             // also write fx.., xdd.., dxx.., ql, qq
  int loop_mode;            // k_dt: 1 driver-loop test and next dt, 2 next dt always
  int flip;                 // k_nodal makes the node set it wrote current
  double *dt_record;        // k_time_increment (k_dt in loop_mode 2) stores each instance-0 dt here
  int64_t dt_base;          // ... at index cycle - 1 - dt_base (if non-null)
  int64_t num_iters;
  int cost_mode;            // k_qeos repeats each element's energy update
  const int *elem_rep;      // ... elem_rep[element] times (its region's regReps)
};

enum KernelId {
  KID_NODAL = 0,
  KID_KIN = 1,
  KID_QEOS = 2,
  KID_DT = 3,
  KID_DTPART = 4,
  KID_TIME_INCREMENT = 5,
  KID_RUN_BEGIN = 6,
  KID_CLEAR_ERR = 7,
};

hipError_t launch_init(hipStream_t st, const Batch &b);
// workgroups of k_nodal one CU holds at once (register and LDS limits)
int nodal_blocks_per_cu();
hipError_t launch_kernel(int id, hipStream_t st, const Batch &b);
hipError_t launch_math(hipStream_t st, int op, const double *a, const double *b, double *out,
                       int64_t n);

}  // namespace lulesh
