// Reverse-mode (adjoint) arithmetic of one LULESH cycle, element by element.
//
// Hand-written vector-Jacobian products of the forward routines in
// lulesh_elem.h / lulesh_kernels.hip, which restate src/domain.jl.  They
// replace the Enzyme VJP of examples/benchmark.jl:48-51,85-86.  Compiled for
// gfx950 and, for the finite-difference checks of the tests, for the host.
//
// Conventions: every *_rev function recomputes the forward intermediates it
// needs from the forward inputs, then ADDS the adjoints of its inputs into
// the l* outputs.  The time step is an input leaf (the dt schedule of the
// forward run is held fixed), so the Courant/hydro constraints and the
// characteristic length, which only feed dt, carry no adjoint.  Cut-offs,
// clamps and the limiter's min/max are differentiated piecewise: the branch
// the forward arithmetic takes is the one differentiated.
#pragma once

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define LHD __host__ __device__ inline
// every loop here has constant bounds: unrolled, all array indices become
// constants and the arrays live in registers
#define LUNROLL _Pragma("unroll")
#else
#include <cmath>
#define LHD inline
#define LUNROLL
#endif
#include <stdint.h>

namespace lulesh_adj {

LHD double sq(double a) { return a * a; }

// ---------------------------------------------------------------------------
// geometry helpers (forward, no rounding contract: adjoints only)
//
// From here to the end of the hourglass adjoint the GPU build lets the
// compiler fuse multiply-adds: this arithmetic has no branches, no rounding
// contract and no bitwise partner on the forward side, and fusing removes a
// third of a_force's floating-point instructions.  The limiter and EOS
// adjoints below keep the unfused arithmetic of the forward kernels, whose
// branch decisions they recompute.  Coordinates are read through X[n] (an
// array, or on the GPU a view of values parked in LDS).
// ---------------------------------------------------------------------------
#ifdef __HIPCC__
#pragma clang fp contract(fast)
#endif

// the volume derivative's 1/12 as a product (a division is ten dependent
// instructions)
constexpr double TWELFTH = 1.0 / 12.0;

LHD double gamma_of(int n, int m) { return ((0xE0D35B68u >> (4 * n + m)) & 1u) ? -1.0 : 1.0; }

LHD void cross(const double a[3], const double b[3], double r[3]) {
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}

// P(u,w) of voluDer (src/domain.jl:1185-1188): dvdx = P(y,z)/12,
// dvdy = -P(x,z)/12, dvdz = -P(y,x)/12 over the six local nodes
LHD double vd_P(const double u[6], const double w[6]) {
  return (u[1] + u[2]) * (w[0] + w[1]) - (u[0] + u[1]) * (w[1] + w[2]) +
         (u[0] + u[4]) * (w[3] + w[4]) - (u[3] + u[4]) * (w[0] + w[4]) -
         (u[2] + u[5]) * (w[3] + w[5]) + (u[3] + u[5]) * (w[2] + w[5]);
}
// adds s * dP/du to lu and s * dP/dw to lw
LHD void vd_P_rev(const double u[6], const double w[6], double s, double lu[6], double lw[6]) {
  const double A1 = w[0] + w[1], A2 = w[1] + w[2], A3 = w[3] + w[4], A4 = w[0] + w[4],
               A5 = w[3] + w[5], A6 = w[2] + w[5];
  lu[1] += s * A1; lu[2] += s * A1;
  lu[0] -= s * A2; lu[1] -= s * A2;
  lu[0] += s * A3; lu[4] += s * A3;
  lu[3] -= s * A4; lu[4] -= s * A4;
  lu[2] -= s * A5; lu[5] -= s * A5;
  lu[3] += s * A6; lu[5] += s * A6;
  const double U1 = u[1] + u[2], U2 = u[0] + u[1], U3 = u[0] + u[4], U4 = u[3] + u[4],
               U5 = u[2] + u[5], U6 = u[3] + u[5];
  lw[0] += s * U1; lw[1] += s * U1;
  lw[1] -= s * U2; lw[2] -= s * U2;
  lw[3] += s * U3; lw[4] += s * U3;
  lw[0] -= s * U4; lw[4] -= s * U4;
  lw[3] -= s * U5; lw[5] -= s * U5;
  lw[2] += s * U6; lw[5] += s * U6;
}

// node lists of the eight voluDer calls (src/domain.jl:1211-1249) and the
// slot each result lands in (quirk Q2: calls 2 and 4 land in slots 2 and 4
// although they are the derivatives of nodes 4 and 2)
LHD int vd_node(int call, int k) {
  const int P[8][6] = {{1, 2, 3, 4, 5, 7}, {0, 1, 2, 7, 4, 6}, {3, 0, 1, 6, 7, 5},
                       {2, 3, 0, 5, 6, 4}, {7, 6, 5, 0, 3, 1}, {4, 7, 6, 1, 0, 2},
                       {5, 4, 7, 2, 1, 3}, {6, 5, 4, 3, 2, 0}};
  return P[call][k];
}
LHD int vd_slot(int call, bool q2) {
  if (!q2 && call == 1) return 3;
  if (!q2 && call == 3) return 1;
  return call;
}

// the volume derivative landing in slot `slot` (voluDer call `slot`,
// except calls 1 and 3 swap slots without Q2)
template <bool Q2, class XA>
LHD void volume_derivative_slot(const XA &X, const XA &Y, const XA &Z, int slot,
                                double &dx, double &dy, double &dz) {
  const int c = (!Q2 && slot == 1) ? 3 : ((!Q2 && slot == 3) ? 1 : slot);
  double x[6], y[6], z[6];
  LUNROLL
  for (int k = 0; k < 6; ++k) {
    x[k] = X[vd_node(c, k)];
    y[k] = Y[vd_node(c, k)];
    z[k] = Z[vd_node(c, k)];
  }
  dx = vd_P(y, z) * TWELFTH;
  dy = -vd_P(x, z) * TWELFTH;
  dz = -vd_P(y, x) * TWELFTH;
}

template <bool Q2>
LHD void volume_derivative_t(const double X[8], const double Y[8], const double Z[8],
                           double dvdx[8], double dvdy[8], double dvdz[8]) {
  LUNROLL
  for (int c = 0; c < 8; ++c) {
    double x[6], y[6], z[6];
    LUNROLL
    for (int k = 0; k < 6; ++k) {
      x[k] = X[vd_node(c, k)];
      y[k] = Y[vd_node(c, k)];
      z[k] = Z[vd_node(c, k)];
    }
    const int sl = vd_slot(c, Q2);
    dvdx[sl] = vd_P(y, z) * TWELFTH;
    dvdy[sl] = -vd_P(x, z) * TWELFTH;
    dvdz[sl] = -vd_P(y, x) * TWELFTH;
  }
}

// adjoint of volume_derivative: adds to lX, lY, lZ
template <bool Q2>
LHD void volume_derivative_rev_t(const double X[8], const double Y[8], const double Z[8],
                               const double ldx[8], const double ldy[8], const double ldz[8],
                               double lX[8], double lY[8], double lZ[8]) {
  LUNROLL
  for (int c = 0; c < 8; ++c) {
    double x[6], y[6], z[6], lx[6] = {0, 0, 0, 0, 0, 0}, ly[6] = {0, 0, 0, 0, 0, 0},
                                 lz[6] = {0, 0, 0, 0, 0, 0};
    LUNROLL
    for (int k = 0; k < 6; ++k) {
      x[k] = X[vd_node(c, k)];
      y[k] = Y[vd_node(c, k)];
      z[k] = Z[vd_node(c, k)];
    }
    const int sl = vd_slot(c, Q2);
    vd_P_rev(y, z, ldx[sl] * TWELFTH, ly, lz);
    vd_P_rev(x, z, -ldy[sl] * TWELFTH, lx, lz);
    vd_P_rev(y, x, -ldz[sl] * TWELFTH, ly, lx);
    LUNROLL
    for (int k = 0; k < 6; ++k) {
      lX[vd_node(c, k)] += lx[k];
      lY[vd_node(c, k)] += ly[k];
      lZ[vd_node(c, k)] += lz[k];
    }
  }
}

// the Q2 slot swap is a template parameter so that every array index is a
// compile-time constant (a run-time slot would put the arrays in scratch)
LHD void volume_derivative(const double X[8], const double Y[8], const double Z[8], bool q2,
                           double dvdx[8], double dvdy[8], double dvdz[8]) {
  if (q2) volume_derivative_t<true>(X, Y, Z, dvdx, dvdy, dvdz);
  else volume_derivative_t<false>(X, Y, Z, dvdx, dvdy, dvdz);
}

LHD void volume_derivative_rev(const double X[8], const double Y[8], const double Z[8], bool q2,
                               const double ldx[8], const double ldy[8], const double ldz[8],
                               double lX[8], double lY[8], double lZ[8]) {
  if (q2) volume_derivative_rev_t<true>(X, Y, Z, ldx, ldy, ldz, lX, lY, lZ);
  else volume_derivative_rev_t<false>(X, Y, Z, ldx, ldy, ldz, lX, lY, lZ);
}

// exact gradient of calcElemVolume (= the quirk-free volume derivative)
LHD void volume_grad_rev(const double X[8], const double Y[8], const double Z[8], double lvol,
                         double lX[8], double lY[8], double lZ[8]) {
  double dx[8], dy[8], dz[8];
  volume_derivative(X, Y, Z, false, dx, dy, dz);
  LUNROLL
  for (int n = 0; n < 8; ++n) {
    lX[n] += lvol * dx[n];
    lY[n] += lvol * dy[n];
    lZ[n] += lvol * dz[n];
  }
}

// ---------------------------------------------------------------------------
// stress corner forces: f[n][c] = ms * B[n][c], B from the face normals of
// sumElemFaceNormal/calcElemNodeNormals (src/domain.jl:971-1067, quirk Q1)
// ---------------------------------------------------------------------------
LHD int face_node(int f, int k) {
  const int F[6][4] = {{0, 1, 2, 3}, {0, 4, 5, 1}, {1, 5, 6, 2},
                       {2, 6, 7, 3}, {3, 7, 4, 0}, {4, 7, 6, 5}};
  return F[f][k];
}

LHD void face_normal(const double p[4][3], double a[3]) {
  double b0[3], b1[3];
  LUNROLL
  for (int c = 0; c < 3; ++c) {
    b0[c] = 0.5 * (((p[3][c] + p[2][c]) - p[1][c]) - p[0][c]);
    b1[c] = 0.5 * (((p[2][c] + p[1][c]) - p[3][c]) - p[0][c]);
  }
  cross(b0, b1, a);
  LUNROLL
  for (int c = 0; c < 3; ++c) a[c] *= 0.25;
}

LHD void face_normal_rev(const double p[4][3], const double la[3], double lp[4][3]) {
  double b0[3], b1[3], l0[3], l1[3];
  LUNROLL
  for (int c = 0; c < 3; ++c) {
    b0[c] = 0.5 * (((p[3][c] + p[2][c]) - p[1][c]) - p[0][c]);
    b1[c] = 0.5 * (((p[2][c] + p[1][c]) - p[3][c]) - p[0][c]);
  }
  cross(b1, la, l0);  // d(la.(b0 x b1))/db0 = b1 x la
  cross(la, b0, l1);  // d/db1 = la x b0
  LUNROLL
  for (int c = 0; c < 3; ++c) {
    const double g0 = 0.25 * l0[c], g1 = 0.25 * l1[c];
    lp[0][c] += 0.5 * (-g0 - g1);
    lp[1][c] += 0.5 * (-g0 + g1);
    lp[2][c] += 0.5 * (g0 + g1);
    lp[3][c] += 0.5 * (g0 - g1);
  }
}

LHD void node_normals(const double X[8], const double Y[8], const double Z[8], bool q1,
                      double B[8][3]) {
  LUNROLL
  for (int n = 0; n < 8; ++n) B[n][0] = B[n][1] = B[n][2] = 0.0;
  LUNROLL
  for (int f = 0; f < 6; ++f) {
    double p[4][3], a[3];
    LUNROLL
    for (int k = 0; k < 4; ++k) {
      const int nn = face_node(f, k);
      p[k][0] = X[nn];
      p[k][1] = Y[nn];
      p[k][2] = Z[nn];
    }
    face_normal(p, a);
    LUNROLL
    for (int k = 0; k < 4; ++k) {
      const int nn = face_node(f, k);
      LUNROLL
      for (int c = 0; c < 3; ++c) B[nn][c] += a[c];
    }
    if (f == 4 && q1) {  // node 1: areaY once more, areaZ not at all
      B[0][1] += a[1];
      B[0][2] -= a[2];
    }
  }
}

struct HgArraySrcF {
  const double (*lF)[3];
  LHD double lf(int n, int c) const { return lF[n][c]; }
};

// lF[n][c]: adjoint of the stress corner force f[n][c] = ms * B[n][c].
// Face f adds its normal a_f to B of its four nodes (and, quirk Q1, face 5
// adds a_y once more and subtracts a_z at node 1), so with S_f the sum of
// lF over the nodes a_f reaches, lms = sum_f a_f . S_f and the adjoint of
// a_f is ms * S_f.  Reads lF through src.lf(n, c); adds to lX..; returns lms.
template <class Src, class XA>
LHD double stress_rev_s(const XA &X, const XA &Y, const XA &Z, bool q1, double ms,
                        const Src &src, double lX[8], double lY[8], double lZ[8]) {
  double lms = 0.0;
  LUNROLL
  for (int f = 0; f < 6; ++f) {
    double S[3] = {0.0, 0.0, 0.0}, p[4][3];
    LUNROLL
    for (int k = 0; k < 4; ++k) {
      const int nn = face_node(f, k);
      LUNROLL
      for (int c = 0; c < 3; ++c) S[c] += src.lf(nn, c);
      p[k][0] = X[nn];
      p[k][1] = Y[nn];
      p[k][2] = Z[nn];
    }
    if (f == 4 && q1) {
      S[1] += src.lf(0, 1);
      S[2] -= src.lf(0, 2);
    }
    double a[3];
    face_normal(p, a);
    lms += a[0] * S[0] + a[1] * S[1] + a[2] * S[2];
    const double la[3] = {ms * S[0], ms * S[1], ms * S[2]};
    double lp[4][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    face_normal_rev(p, la, lp);
    LUNROLL
    for (int k = 0; k < 4; ++k) {
      const int nn = face_node(f, k);
      lX[nn] += lp[k][0];
      lY[nn] += lp[k][1];
      lZ[nn] += lp[k][2];
    }
  }
  return lms;
}

LHD double stress_rev(const double X[8], const double Y[8], const double Z[8], bool q1, double ms,
                      const double lF[8][3], double lX[8], double lY[8], double lZ[8]) {
  HgArraySrcF src{lF};
  return stress_rev_s(X, Y, Z, q1, ms, src, lX, lY, lZ);
}

// ---------------------------------------------------------------------------
// FB hourglass corner forces (src/domain.jl:1505-1639), adjoint with respect
// to coordinates, velocities, determ (= volo*v), ss and hgcoef
// ---------------------------------------------------------------------------
struct HgRev {
  double ldeterm, lss, lhgcoef;
};

// reverse of voluDer call c (the derivative landing in slot vd_slot(c)),
// scaled by the slot's adjoint (lx, ly, lz); adds to lX, lY, lZ
template <class XA>
LHD void volume_derivative_call_rev(const XA &X, const XA &Y, const XA &Z, int c,
                                    double lx, double ly, double lz, double lX[8], double lY[8],
                                    double lZ[8]) {
  double x[6], y[6], z[6], gx[6] = {0, 0, 0, 0, 0, 0}, gy[6] = {0, 0, 0, 0, 0, 0},
                           gz[6] = {0, 0, 0, 0, 0, 0};
  LUNROLL
  for (int k = 0; k < 6; ++k) {
    x[k] = X[vd_node(c, k)];
    y[k] = Y[vd_node(c, k)];
    z[k] = Z[vd_node(c, k)];
  }
  vd_P_rev(y, z, lx * TWELFTH, gy, gz);
  vd_P_rev(x, z, -ly * TWELFTH, gx, gz);
  vd_P_rev(y, x, -lz * TWELFTH, gy, gx);
  LUNROLL
  for (int k = 0; k < 6; ++k) {
    lX[vd_node(c, k)] += gx[k];
    lY[vd_node(c, k)] += gy[k];
    lZ[vd_node(c, k)] += gz[k];
  }
}

// Streaming form.  With P_n the node positions, d_n the volume derivative,
// h_m = sum_n gamma[n][m] P_n and g[n][m] = gamma[n][m] - (d_n . h_m)/determ,
// the forward force is F[n][c] = coef * sum_m g[n][m] H[c][m] with
// H[c][m] = sum_n g[n][m] V[c][n].  With A[m][c] = sum_n g[n][m] lF[n][c]:
//   lcoef = sum A[m][c] H[c][m],  lV[c][n] = coef sum_m A[m][c] g[n][m],
//   lg[n][m] = coef sum_c (lF[n][c] H[c][m] + A[m][c] V[c][n]),
// and lg flows to determ, d_n and h_m.  Two node-major passes (the first
// forms H and A, the second lg and everything downstream) keep few values
// live: velocities and force adjoints are read node by node from `src`
// (src.v(c, n), src.lf(n, c)) and each lV[c][n] is final when formed, so it
// goes straight to `sink.lv(c, n, value)`.  Adds the position adjoints to
// lX, lY, lZ.
template <bool Q2, class Src, class Sink, class XA>
LHD HgRev hourglass_rev_s(const XA &X, const XA &Y, const XA &Z, const Src &src,
                          Sink &sink, double determ, double ss, double mass, double hourg,
                          double cbrt_determ, double lX[8], double lY[8], double lZ[8]) {
  const double volinv = 1.0 / determ;
  double hx[4], hy[4], hz[4];
  LUNROLL
  for (int m = 0; m < 4; ++m) {
    hx[m] = hy[m] = hz[m] = 0.0;
    LUNROLL
    for (int n = 0; n < 8; ++n) {
      hx[m] += X[n] * gamma_of(n, m);
      hy[m] += Y[n] * gamma_of(n, m);
      hz[m] += Z[n] * gamma_of(n, m);
    }
  }
  // pass 1: H and A
  double H[3][4], A[4][3];
  LUNROLL
  for (int m = 0; m < 4; ++m)
    LUNROLL
    for (int c = 0; c < 3; ++c) H[c][m] = A[m][c] = 0.0;
  LUNROLL
  for (int n = 0; n < 8; ++n) {
    const double v0 = src.v(0, n), v1 = src.v(1, n), v2 = src.v(2, n);
    const double f0 = src.lf(n, 0), f1 = src.lf(n, 1), f2 = src.lf(n, 2);
    // the node's volume derivative is recomputed in each pass: holding all
    // eight would keep 24 more values live through both passes
    double dvx, dvy, dvz;
    volume_derivative_slot<Q2>(X, Y, Z, n, dvx, dvy, dvz);
    LUNROLL
    for (int m = 0; m < 4; ++m) {
      const double g = gamma_of(n, m) - volinv * (dvx * hx[m] + dvy * hy[m] + dvz * hz[m]);
      H[0][m] += g * v0; H[1][m] += g * v1; H[2][m] += g * v2;
      A[m][0] += g * f0; A[m][1] += g * f1; A[m][2] += g * f2;
    }
  }
  const double K = -hourg * 0.01 * mass / cbrt_determ;  // coef = K * ss
  const double coef = K * ss;
  double lcoef = 0.0;
  LUNROLL
  for (int m = 0; m < 4; ++m)
    LUNROLL
    for (int c = 0; c < 3; ++c) lcoef += A[m][c] * H[c][m];
  // pass 2: lg node by node -> lV, determ, d_n (reversed at once), h_m
  double lvolinv = 0.0, lhx[4] = {0, 0, 0, 0}, lhy[4] = {0, 0, 0, 0}, lhz[4] = {0, 0, 0, 0};
  LUNROLL
  for (int n = 0; n < 8; ++n) {
    const double v0 = src.v(0, n), v1 = src.v(1, n), v2 = src.v(2, n);
    const double f0 = src.lf(n, 0), f1 = src.lf(n, 1), f2 = src.lf(n, 2);
    double lv0 = 0.0, lv1 = 0.0, lv2 = 0.0, ldx = 0.0, ldy = 0.0, ldz = 0.0;
    double dvx, dvy, dvz;
    volume_derivative_slot<Q2>(X, Y, Z, n, dvx, dvy, dvz);
    LUNROLL
    for (int m = 0; m < 4; ++m) {
      const double dh = dvx * hx[m] + dvy * hy[m] + dvz * hz[m];
      const double g = gamma_of(n, m) - volinv * dh;
      lv0 += A[m][0] * g; lv1 += A[m][1] * g; lv2 += A[m][2] * g;
      const double lg = coef * (f0 * H[0][m] + f1 * H[1][m] + f2 * H[2][m] +
                                A[m][0] * v0 + A[m][1] * v1 + A[m][2] * v2);
      lvolinv -= lg * dh;
      const double s = volinv * lg;
      ldx -= s * hx[m]; ldy -= s * hy[m]; ldz -= s * hz[m];
      lhx[m] -= s * dvx; lhy[m] -= s * dvy; lhz[m] -= s * dvz;
    }
    sink.lv(0, n, coef * lv0);
    sink.lv(1, n, coef * lv1);
    sink.lv(2, n, coef * lv2);
    // slot n holds voluDer call n, except calls 1 and 3 swap slots without Q2
    const int call = (!Q2 && n == 1) ? 3 : ((!Q2 && n == 3) ? 1 : n);
    volume_derivative_call_rev(X, Y, Z, call, ldx, ldy, ldz, lX, lY, lZ);
  }
  LUNROLL
  for (int n = 0; n < 8; ++n)
    LUNROLL
    for (int m = 0; m < 4; ++m) {
      lX[n] += lhx[m] * gamma_of(n, m);
      lY[n] += lhy[m] * gamma_of(n, m);
      lZ[n] += lhz[m] * gamma_of(n, m);
    }
  HgRev r;
  // coef = -hourg*0.01*ss*mass * determ^(-1/3)
  r.ldeterm = -lvolinv * volinv * volinv - lcoef * coef / (3.0 * determ);
  r.lss = lcoef * K;
  r.lhgcoef = lcoef * (-0.01 * ss * mass / cbrt_determ);
  return r;
}

struct HgArraySrc {
  const double *V[3];
  const double (*lF)[3];
  LHD double v(int c, int n) const { return V[c][n]; }
  LHD double lf(int n, int c) const { return lF[n][c]; }
};
struct HgArraySink {
  double *lV[3];
  LHD void lv(int c, int n, double x) { lV[c][n] += x; }
};

// array form (host tests): adds to lX.. and lVX..
LHD HgRev hourglass_rev(const double X[8], const double Y[8], const double Z[8],
                        const double VX[8], const double VY[8], const double VZ[8], double determ,
                        double ss, double mass, double hourg, double cbrt_determ, bool q2,
                        const double lF[8][3], double lX[8], double lY[8], double lZ[8],
                        double lVX[8], double lVY[8], double lVZ[8]) {
  HgArraySrc src{{VX, VY, VZ}, lF};
  HgArraySink sink{{lVX, lVY, lVZ}};
  if (q2)
    return hourglass_rev_s<true>(X, Y, Z, src, sink, determ, ss, mass, hourg, cbrt_determ, lX, lY, lZ);
  return hourglass_rev_s<false>(X, Y, Z, src, sink, determ, ss, mass, hourg, cbrt_determ, lX, lY, lZ);
}

#ifdef __HIPCC__
#pragma clang fp contract(off)
#endif

// ---------------------------------------------------------------------------
// calcMonotonicQGradientsForElems (src/domain.jl:2246-2317) adjoint.
// Inputs: new coordinates and velocities of the 8 nodes, volo, vnew.
// l[0..2] = adjoints of delx_xi, delx_eta, delx_zeta; l[3..5] of delv_*.
// Adds to lX.., lV.. and returns the adjoint of vnew.
// ---------------------------------------------------------------------------
// coefficient patterns of the three logical-direction differences
LHD double dir_coef(int d, int n) {
  // d = 0: i (xi), 1: j (eta), 2: k (zeta)
  const double Ci[8] = {-.25, .25, .25, -.25, -.25, .25, .25, -.25};
  const double Cj[8] = {-.25, -.25, .25, .25, -.25, -.25, .25, .25};
  const double Ck[8] = {-.25, -.25, -.25, -.25, .25, .25, .25, .25};
  return d == 0 ? Ci[n] : (d == 1 ? Cj[n] : Ck[n]);
}

LHD double monoq_grad_rev(const double X[8], const double Y[8], const double Z[8],
                          const double VX[8], const double VY[8], const double VZ[8], double volo,
                          double vnew, const double l[6], double lX[8], double lY[8], double lZ[8],
                          double lVX[8], double lVY[8], double lVZ[8]) {
  const double ptiny = 1.e-36;
  const double vol = volo * vnew;
  const double norm = 1.0 / (vol + ptiny);
  double D[3][3];  // D[d] = difference vector of direction d (i, j, k)
  double DV[3][3];
  LUNROLL
  for (int d = 0; d < 3; ++d) {
    D[d][0] = D[d][1] = D[d][2] = 0.0;
    DV[d][0] = DV[d][1] = DV[d][2] = 0.0;
    LUNROLL
    for (int n = 0; n < 8; ++n) {
      const double w = dir_coef(d, n);
      D[d][0] += w * X[n]; D[d][1] += w * Y[n]; D[d][2] += w * Z[n];
      DV[d][0] += w * VX[n]; DV[d][1] += w * VY[n]; DV[d][2] += w * VZ[n];
    }
  }
  // outputs xi = (j x k, dv_i), eta = (k x i, dv_j), zeta = (i x j, dv_k)
  const int U[3] = {1, 2, 0}, W[3] = {2, 0, 1}, DVd[3] = {0, 1, 2};
  double lD[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, lDV[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  double lvol = 0.0, lnorm = 0.0;
  LUNROLL
  for (int o = 0; o < 3; ++o) {
    double a[3];
    cross(D[U[o]], D[W[o]], a);
    const double r2 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2] + ptiny;
    const double r = sqrt(r2);
    const double ldelx = l[o], ldelv = l[3 + o];
    const double *dv = DV[DVd[o]];
    const double adv = a[0] * dv[0] + a[1] * dv[1] + a[2] * dv[2];
    double la[3];
    lvol += ldelx / r;
    LUNROLL
    for (int c = 0; c < 3; ++c) {
      la[c] = -ldelx * vol / (r2 * r) * a[c] + ldelv * norm * dv[c];
      lDV[DVd[o]][c] += ldelv * norm * a[c];
    }
    lnorm += ldelv * adv;
    double lu[3], lw[3];
    cross(D[W[o]], la, lu);  // a = u x w
    cross(la, D[U[o]], lw);
    LUNROLL
    for (int c = 0; c < 3; ++c) {
      lD[U[o]][c] += lu[c];
      lD[W[o]][c] += lw[c];
    }
  }
  LUNROLL
  for (int d = 0; d < 3; ++d)
    LUNROLL
    for (int n = 0; n < 8; ++n) {
      const double w = dir_coef(d, n);
      lX[n] += w * lD[d][0]; lY[n] += w * lD[d][1]; lZ[n] += w * lD[d][2];
      lVX[n] += w * lDV[d][0]; lVY[n] += w * lDV[d][1]; lVZ[n] += w * lDV[d][2];
    }
  lvol += -lnorm * norm * norm;
  return lvol * volo;
}

// ---------------------------------------------------------------------------
// calcMonotonicQRegionForElems for one element (src/domain.jl:2332-2511).
// Forward and adjoint.  dv[d] = own delv_{xi,eta,zeta}, dm[d]/dp[d] = the
// minus/plus values the BC case selected (neighbour, self or 0), dx[d] =
// own delx_*.  vdov only selects the branch.
// ---------------------------------------------------------------------------
struct QIn {
  double dv[3], dm[3], dp[3], dx[3];
  double vdov, elemMass, volo, vnew, qlc, qqc, mult, maxslope;
};

// phi and its partials w.r.t. (self delv, delvm, delvp)
LHD double phi_eval(double self, double delvm, double delvp, double mult, double maxslope,
                    double g[3]) {
  const double ptiny = 1e-36;
  const double norm = 1.0 / (self + ptiny);
  const double m = delvm * norm, p = delvp * norm;
  double phi = 0.5 * (m + p);
  double gm = 0.5, gp = 0.5;  // dphi/dm, dphi/dp
  const double mm = m * mult, pm = p * mult;
  if (mm < phi) { phi = mm; gm = mult; gp = 0.0; }
  if (pm < phi) { phi = pm; gm = 0.0; gp = mult; }
  if (phi < 0.0) { phi = 0.0; gm = gp = 0.0; }
  if (phi > maxslope) { phi = maxslope; gm = gp = 0.0; }
  // m = delvm * norm, p = delvp * norm, norm = 1/(self + ptiny)
  g[1] = gm * norm;
  g[2] = gp * norm;
  g[0] = -(gm * delvm + gp * delvp) * norm * norm;
  return phi;
}

struct QOut {
  double qlin, qquad;
};

// forward (smooth form, for checks) and adjoint; lq = {lqlin, lqquad}
// outputs: ldv, ldm, ldp, ldx (added), lvnew, lqlc, lqqc (added)
LHD QOut monoq_region_rev(const QIn &in, double lqlin, double lqquad, double ldv[3], double ldm[3],
                          double ldp[3], double ldx[3], double *lvnew, double *lqlc,
                          double *lqqc) {
  QOut o = {0.0, 0.0};
  if (in.vdov > 0.0) return o;
  double phi[3], g[3][3], dvx[3];
  bool act[3];
  LUNROLL
  for (int d = 0; d < 3; ++d) {
    phi[d] = phi_eval(in.dv[d], in.dm[d], in.dp[d], in.mult, in.maxslope, g[d]);
    dvx[d] = in.dv[d] * in.dx[d];
    act[d] = !(dvx[d] > 0.0);
    if (!act[d]) dvx[d] = 0.0;
  }
  const double rho = in.elemMass / (in.volo * in.vnew);
  double S1 = 0.0, S2 = 0.0;
  LUNROLL
  for (int d = 0; d < 3; ++d) {
    S1 += dvx[d] * (1.0 - phi[d]);
    S2 += dvx[d] * dvx[d] * (1.0 - phi[d] * phi[d]);
  }
  o.qlin = -in.qlc * rho * S1;
  o.qquad = in.qqc * rho * S2;
  const double lrho = lqlin * (-in.qlc * S1) + lqquad * (in.qqc * S2);
  *lqlc += lqlin * (-rho * S1);
  *lqqc += lqquad * (rho * S2);
  LUNROLL
  for (int d = 0; d < 3; ++d) {
    const double ldvx = lqlin * (-in.qlc * rho) * (1.0 - phi[d]) +
                        lqquad * (in.qqc * rho) * 2.0 * dvx[d] * (1.0 - phi[d] * phi[d]);
    const double lphi = lqlin * (-in.qlc * rho) * (-dvx[d]) +
                        lqquad * (in.qqc * rho) * dvx[d] * dvx[d] * (-2.0 * phi[d]);
    if (act[d]) {
      ldv[d] += ldvx * in.dx[d];
      ldx[d] += ldvx * in.dv[d];
    }
    ldv[d] += lphi * g[d][0];
    ldm[d] += lphi * g[d][1];
    ldp[d] += lphi * g[d][2];
  }
  // rho = elemMass / (volo * vnew)
  *lvnew += -lrho * rho / in.vnew;
  return o;
}

// ---------------------------------------------------------------------------
// evalEOSForElems + calcEnergyForElems + calcPressureForElems +
// calcSoundSpeedForElems for one element (src/domain.jl:2585-2857, work = 0)
// forward with the branch record, and its adjoint.
// ---------------------------------------------------------------------------
struct EosParams {
  double e_cut, p_cut, q_cut, eosvmax, eosvmin, pmin, emin, rho0;
};

struct EosOut {
  double p, e, q, ss;
};

// one calcPressureForElems: p = bvc*e unless cut; returns dp/de, dp/dcomp
LHD double eos_pressure(double e, double comp, double vnewc, const EosParams &P, double *dpde,
                        double *dpdc) {
  const double c1s = 2.0 / 3.0;
  const double bvc = c1s * (comp + 1.0);
  double p = bvc * e;
  *dpde = bvc;
  *dpdc = c1s * e;
  if (fabs(p) < P.p_cut) { p = 0.0; *dpde = *dpdc = 0.0; }
  if (vnewc >= P.eosvmax) { p = 0.0; *dpde = *dpdc = 0.0; }
  if (p < P.pmin) { p = P.pmin; *dpde = *dpdc = 0.0; }
  return p;
}

// adjoint of the element EOS.  Inputs vnewc (already clamped), delvc,
// e_old, p_old, q_old, ql, qq; lo = adjoints of (p, e, q, ss) outputs.
// li = adjoints of the inputs in the same order (added).
LHD EosOut eos_rev(double vnewc, double delvc, double e_old, double p_old, double q_old, double ql,
                   double qq, const EosParams &P, const double lo[4], double li[7]) {
  const double TINY1 = 0.111111e-36, TINY3 = 0.333333e-18, SIXTH = 1.0 / 6.0;
  const double c1s = 2.0 / 3.0, pbvc = c1s;
  // ---- forward, keeping what the adjoint needs ----
  double compression = 1.0 / vnewc - 1.0;
  double dcomp_dv = -1.0 / (vnewc * vnewc);
  const double vchalf = vnewc - delvc * 0.5;
  double compH = (1.0 / vchalf) - 1.0;
  double dcH_dv = -1.0 / (vchalf * vchalf), dcH_dd = 0.5 / (vchalf * vchalf);
  if (P.eosvmin != 0.0 && vnewc <= P.eosvmin) {
    compH = compression;
    dcH_dv = dcomp_dv;
    dcH_dd = 0.0;
  }
  double pold = p_old;
  double dpold = 1.0;
  if (P.eosvmax != 0.0 && vnewc >= P.eosvmax) {
    pold = 0.0; dpold = 0.0;
    compression = 0.0; dcomp_dv = 0.0;
    compH = 0.0; dcH_dv = 0.0; dcH_dd = 0.0;
  }
  const double PQ = pold + q_old;
  // stage 1
  double e1 = e_old - 0.5 * delvc * PQ;
  const bool e1cut = e1 < P.emin;
  if (e1cut) e1 = P.emin;
  double dpH_de, dpH_dc;
  const double pH = eos_pressure(e1, compH, vnewc, P, &dpH_de, &dpH_dc);
  const double bvcH = c1s * (compH + 1.0);
  const double vhalf = 1.0 / (1.0 + compH);
  double qH = 0.0, sH = 0.0;
  bool sHcut = true;
  if (!(delvc > 0.0)) {
    const double ssc = (pbvc * e1 + vhalf * vhalf * bvcH * pH) / P.rho0;
    sHcut = ssc <= TINY1;
    sH = sHcut ? TINY3 : sqrt(ssc);
    qH = sH * ql + qq;
  }
  // stage 2
  double e2 = e1 + 0.5 * delvc * (3.0 * PQ - 4.0 * (pH + qH));
  const bool e2cut0 = fabs(e2) < P.e_cut;
  if (e2cut0) e2 = 0.0;
  const bool e2cut1 = e2 < P.emin;
  if (e2cut1) e2 = P.emin;
  double dp2_de, dp2_dc;
  const double p2 = eos_pressure(e2, compression, vnewc, P, &dp2_de, &dp2_dc);
  const double bvc = c1s * (compression + 1.0);
  double qt = 0.0, st = 0.0;
  bool stcut = true;
  if (!(delvc > 0.0)) {
    const double ssc = (pbvc * e2 + vnewc * vnewc * bvc * p2) / P.rho0;
    stcut = ssc <= TINY1;
    st = stcut ? TINY3 : sqrt(ssc);
    qt = st * ql + qq;
  }
  // stage 3
  const double R = 7.0 * PQ - 8.0 * (pH + qH) + (p2 + qt);
  double e3 = e2 - R * delvc * SIXTH;
  const bool e3cut0 = fabs(e3) < P.e_cut;
  if (e3cut0) e3 = 0.0;
  const bool e3cut1 = e3 < P.emin;
  if (e3cut1) e3 = P.emin;
  double dp3_de, dp3_dc;
  const double p3 = eos_pressure(e3, compression, vnewc, P, &dp3_de, &dp3_dc);
  const double ssarg = (pbvc * e3 + vnewc * vnewc * bvc * p3) / P.rho0;
  const bool s3cut = ssarg <= TINY1;
  const double s3 = s3cut ? TINY3 : sqrt(ssarg);
  double q3 = qH;
  bool q3cut = false;
  if (!(delvc > 0.0)) {
    q3 = s3 * ql + qq;
    q3cut = fabs(q3) < P.q_cut;
    if (q3cut) q3 = 0.0;
  }
  EosOut out = {p3, e3, q3, s3};

  // ---- adjoint ----
  double lp3 = lo[0], le3 = lo[1], lq3 = lo[2], ls3 = lo[3];
  double lvn = 0.0, ldel = 0.0, lql = 0.0, lqq = 0.0, lPQ = 0.0, lcomp = 0.0, lcH = 0.0;
  // q3 (and the final sound speed share ssarg)
  double lss_arg = 0.0;
  if (!(delvc > 0.0)) {
    if (!q3cut) {
      ls3 += lq3 * ql;
      lql += lq3 * s3;
      lqq += lq3;
    }
  } else {
    // q3 = qH = 0: no dependence
  }
  if (!s3cut) lss_arg += ls3 * 0.5 / s3;
  // ssarg = (pbvc*e3 + vnewc^2*bvc*p3)/rho0
  le3 += lss_arg * pbvc / P.rho0;
  lp3 += lss_arg * vnewc * vnewc * bvc / P.rho0;
  lvn += lss_arg * 2.0 * vnewc * bvc * p3 / P.rho0;
  lcomp += lss_arg * vnewc * vnewc * c1s * p3 / P.rho0;
  // p3 = pressure(e3, compression)
  le3 += lp3 * dp3_de;
  lcomp += lp3 * dp3_dc;
  // e3 cuts
  if (e3cut0 || e3cut1) le3 = 0.0;
  // e3 = e2 - R*delvc/6
  double le2 = le3;
  const double lR = -le3 * delvc * SIXTH;
  ldel += -le3 * R * SIXTH;
  // R = 7PQ - 8(pH+qH) + (p2+qt)
  lPQ += 7.0 * lR;
  double lpH = -8.0 * lR, lqH = -8.0 * lR;
  double lp2 = lR, lqt = lR;
  // qt
  double lst = 0.0;
  if (!(delvc > 0.0)) {
    lst += lqt * ql;
    lql += lqt * st;
    lqq += lqt;
    if (!stcut) {
      const double la = lst * 0.5 / st;
      le2 += la * pbvc / P.rho0;
      lp2 += la * vnewc * vnewc * bvc / P.rho0;
      lvn += la * 2.0 * vnewc * bvc * p2 / P.rho0;
      lcomp += la * vnewc * vnewc * c1s * p2 / P.rho0;
    }
  }
  // p2 = pressure(e2, compression)
  le2 += lp2 * dp2_de;
  lcomp += lp2 * dp2_dc;
  if (e2cut0 || e2cut1) le2 = 0.0;
  // e2 = e1 + 0.5*delvc*(3PQ - 4(pH+qH))
  double le1 = le2;
  ldel += le2 * 0.5 * (3.0 * PQ - 4.0 * (pH + qH));
  lPQ += le2 * 0.5 * delvc * 3.0;
  lpH += le2 * 0.5 * delvc * (-4.0);
  lqH += le2 * 0.5 * delvc * (-4.0);
  // qH
  if (!(delvc > 0.0)) {
    const double lsH = lqH * ql;
    lql += lqH * sH;
    lqq += lqH;
    if (!sHcut) {
      const double la = lsH * 0.5 / sH;
      // ssc = (pbvc*e1 + vhalf^2*bvcH*pH)/rho0, vhalf = 1/(1+compH), bvcH = c1s*(compH+1)
      le1 += la * pbvc / P.rho0;
      lpH += la * vhalf * vhalf * bvcH / P.rho0;
      // d(vhalf^2*bvcH)/dcompH = d(c1s/(1+compH))/dcompH = -c1s/(1+compH)^2
      lcH += la * pH / P.rho0 * (-c1s * vhalf * vhalf);
    }
  }
  // pH = pressure(e1, compH)
  le1 += lpH * dpH_de;
  lcH += lpH * dpH_dc;
  if (e1cut) le1 = 0.0;
  // e1 = e_old - 0.5*delvc*PQ
  const double le_old = le1;
  ldel += -le1 * 0.5 * PQ;
  lPQ += -le1 * 0.5 * delvc;
  // compH(vnewc, delvc), compression(vnewc)
  lvn += lcH * dcH_dv + lcomp * dcomp_dv;
  ldel += lcH * dcH_dd;
  li[0] += lvn;
  li[1] += ldel;
  li[2] += le_old;
  li[3] += lPQ * dpold;
  li[4] += lPQ;
  li[5] += lql;
  li[6] += lqq;
  return out;
}

}  // namespace lulesh_adj
