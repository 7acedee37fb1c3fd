// Host build of the adjoint element routines (lulesh.jl_amd/csrc/lulesh_adj.h)
// for the finite-difference tests.  Built by tests/test_adjoint_cpu.py with
// g++; exports plain C wrappers.
#include <cmath>
#include <cstring>

#include "lulesh_adj.h"

using namespace lulesh_adj;

extern "C" {

double adj_stress_rev(const double *X, const double *Y, const double *Z, int q1, double ms,
                      const double *lF, double *lX, double *lY, double *lZ) {
  double L[8][3];
  std::memcpy(L, lF, sizeof L);
  return stress_rev(X, Y, Z, q1 != 0, ms, L, lX, lY, lZ);
}

void adj_hourglass_rev(const double *X, const double *Y, const double *Z, const double *VX,
                       const double *VY, const double *VZ, double determ, double ss, double mass,
                       double hourg, int q2, const double *lF, double *lX, double *lY, double *lZ,
                       double *lVX, double *lVY, double *lVZ, double *lscal /* determ, ss, hg */) {
  double L[8][3];
  std::memcpy(L, lF, sizeof L);
  HgRev r = hourglass_rev(X, Y, Z, VX, VY, VZ, determ, ss, mass, hourg, std::cbrt(determ), q2 != 0,
                          L, lX, lY, lZ, lVX, lVY, lVZ);
  lscal[0] = r.ldeterm;
  lscal[1] = r.lss;
  lscal[2] = r.lhgcoef;
}

void adj_volume_grad(const double *X, const double *Y, const double *Z, double *gX, double *gY,
                     double *gZ) {
  volume_grad_rev(X, Y, Z, 1.0, gX, gY, gZ);
}

double adj_monoq_grad_rev(const double *X, const double *Y, const double *Z, const double *VX,
                          const double *VY, const double *VZ, double volo, double vnew,
                          const double *l, double *lX, double *lY, double *lZ, double *lVX,
                          double *lVY, double *lVZ) {
  return monoq_grad_rev(X, Y, Z, VX, VY, VZ, volo, vnew, l, lX, lY, lZ, lVX, lVY, lVZ);
}

// in = {dv[3], dm[3], dp[3], dx[3], vdov, elemMass, volo, vnew, qlc, qqc}
// out = {qlin, qquad}; lin = same layout as in (added), lq = {lqlin, lqquad}
void adj_monoq_region_rev(const double *in, const double *lq, double *out, double *lin) {
  QIn q;
  for (int d = 0; d < 3; ++d) {
    q.dv[d] = in[d];
    q.dm[d] = in[3 + d];
    q.dp[d] = in[6 + d];
    q.dx[d] = in[9 + d];
  }
  q.vdov = in[12];
  q.elemMass = in[13];
  q.volo = in[14];
  q.vnew = in[15];
  q.qlc = in[16];
  q.qqc = in[17];
  q.mult = 2.0;
  q.maxslope = 1.0;
  double lqlc = 0.0, lqqc = 0.0, lvnew = 0.0;
  QOut o = monoq_region_rev(q, lq[0], lq[1], lin, lin + 3, lin + 6, lin + 9, &lvnew, &lqlc, &lqqc);
  lin[15] += lvnew;
  lin[16] += lqlc;
  lin[17] += lqqc;
  out[0] = o.qlin;
  out[1] = o.qquad;
}

// in = {vnew, delv, e_old, p_old, q_old, ql, qq}; lo = adjoints of {p, e, q, ss}
void adj_eos_rev(const double *in, const double *lo, double *out, double *li) {
  EosParams P;
  P.e_cut = 1.0e-7; P.p_cut = 1.0e-7; P.q_cut = 1.0e-7;
  P.eosvmax = 1.0e+9; P.eosvmin = 1.0e-9; P.pmin = 0.0; P.emin = -1.0e+15; P.rho0 = 1.0;
  EosOut o = eos_rev(in[0], in[1], in[2], in[3], in[4], in[5], in[6], P, lo, li);
  out[0] = o.p; out[1] = o.e; out[2] = o.q; out[3] = o.ss;
}

}  // extern "C"
