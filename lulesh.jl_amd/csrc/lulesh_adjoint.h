// Arguments and launchers of the adjoint kernels (single-domain contexts).
#pragma once

#include "lulesh_kernels.h"

namespace lulesh {

struct AdjArgs {
  DevScalars *sc;
  DevFields f;       // fields of cycle c: e p q v ss of S_c, its temporaries (vnew delv
                     // vdov delx_* delv_*), and volo elemMass nodalMass elemBC
  NodeSet cur, old;  // node set holding x', xd' and the one holding S_c
  double dt;         // the cycle's time step
  int s;
  int64_t E, N;
  bool q1, q2, q3;
  // adjoint of the cycle state (S_{c+1} on entry, S_c on exit)
  double *lx, *ly, *lz, *lxd, *lyd, *lzd;  // [N]
  double *le, *lp, *lq, *lv, *lss;         // [E]
  double *lF;                              // [3][N] force adjoint per node
  double *nbr;                             // [6][E] delv adjoint sent to the read neighbour
  double *ldvs, *ldx;                      // [3][E] own delv / delx adjoints
  double *lvnew, *lvnew_delv;              // [E]
  double *corner;                          // [6][8][E] per-corner adjoints from a_kin
  double *corner_f;                        // [6][8][E] per-corner adjoints from a_force
  int force_pending;                       // corner_f holds the later cycle's a_force
                                           // output, which a_node1 gathers first
  double *pacc;                            // [3][E] hgcoef, qlc_monoq, qqc_monoq partials
};

enum AdjKernel { AK_QEOS = 0, AK_KIN, AK_NODE1, AK_FORCE, AK_NODE2, AK_SEED };

hipError_t launch_adjoint(int which, hipStream_t st, const AdjArgs &a);
hipError_t launch_state_copy(hipStream_t st, DevScalars *sc, const DevFields &f, int64_t E,
                             int64_t N, double *buf, int to_buf);
// Between two cycles of a segment replay, in one launch: the temporaries of
// the cycle just run (vnew delv vdov delx_xi delx_eta delx_zeta delv_xi
// delv_eta delv_zeta, 9E) into `temps` (if non-null), the current state
// (x y z xd yd zd of the current node set, 6N, then e p q v ss, 5E) into
// `state` (if non-null), and, with `dt` non-null, the next cycle's time step
// dt[cycle - base] with timeIncrement!'s time/cycle bookkeeping.
hipError_t launch_slot_step(hipStream_t st, DevScalars *sc, const DevFields &f, int64_t E,
                            int64_t N, double *temps, double *state, const double *dt,
                            int64_t base);
hipError_t launch_set_cycle(hipStream_t st, DevScalars *sc, long long cycle);

}  // namespace lulesh
