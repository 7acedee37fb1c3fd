// Element-local LULESH arithmetic for gfx950, FP64.
//
// Every function restates one reference function and keeps its exact
// operation order: Julia folds n-ary a+b+c / a*b*c left to right and never
// contracts to FMA, so each sum below is written with explicit left
// association and the library is built with -ffp-contract=off.
//
// Node numbering inside an element is the reference's (nodes 1..8 of
// src/domain.jl:158-165 are indices 0..7 here).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lulesh_cbrt.h"

#define LDEV __device__ __forceinline__

namespace lulesh {

// bc masks, src/bc.jl:3-31
constexpr uint32_t XI_M = 0x00007, XI_M_SYMM = 0x00001, XI_M_FREE = 0x00002, XI_M_COMM = 0x00004;
constexpr uint32_t XI_P = 0x00038, XI_P_SYMM = 0x00008, XI_P_FREE = 0x00010, XI_P_COMM = 0x00020;
constexpr uint32_t ETA_M = 0x001c0, ETA_M_SYMM = 0x00040, ETA_M_FREE = 0x00080, ETA_M_COMM = 0x00100;
constexpr uint32_t ETA_P = 0x00e00, ETA_P_SYMM = 0x00200, ETA_P_FREE = 0x00400, ETA_P_COMM = 0x00800;
constexpr uint32_t ZETA_M = 0x07000, ZETA_M_SYMM = 0x01000, ZETA_M_FREE = 0x02000, ZETA_M_COMM = 0x04000;
constexpr uint32_t ZETA_P = 0x38000, ZETA_P_SYMM = 0x08000, ZETA_P_FREE = 0x10000, ZETA_P_COMM = 0x20000;

constexpr int Q1 = 1, Q2 = 2, Q3 = 4;

// Julia 1.7 Base.min for Float64 (NaN-propagating, -0 < +0)
LDEV double jmin(double x, double y) {
  bool takey = (y < x) || (signbit(y) > signbit(x));
  return takey ? (isnan(x) ? x : y) : (isnan(y) ? y : x);
}
LDEV double jmax(double x, double y) {
  bool takey = (y > x) || (signbit(y) < signbit(x));
  return takey ? (isnan(x) ? x : y) : (isnan(y) ? y : x);
}

// correctly rounded cube root: cbrt_cr (lulesh_cbrt.h)

// ---------------------------------------------------------------------------
// calcElemVolume, src/utils.jl:36-122
// ---------------------------------------------------------------------------
LDEV double triple_product(double x1, double y1, double z1, double x2, double y2, double z2,
                           double x3, double y3, double z3) {
  return ((x1 * ((y2 * z3) - (z2 * y3))) + (x2 * ((z1 * y3) - (y1 * z3)))) +
         (x3 * ((y1 * z2) - (z1 * y2)));
}
LDEV double elem_volume(const double (&x)[8], const double (&y)[8], const double (&z)[8]) {
  const double dx61 = x[6] - x[1], dy61 = y[6] - y[1], dz61 = z[6] - z[1];
  const double dx70 = x[7] - x[0], dy70 = y[7] - y[0], dz70 = z[7] - z[0];
  const double dx63 = x[6] - x[3], dy63 = y[6] - y[3], dz63 = z[6] - z[3];
  const double dx20 = x[2] - x[0], dy20 = y[2] - y[0], dz20 = z[2] - z[0];
  const double dx50 = x[5] - x[0], dy50 = y[5] - y[0], dz50 = z[5] - z[0];
  const double dx64 = x[6] - x[4], dy64 = y[6] - y[4], dz64 = z[6] - z[4];
  const double dx31 = x[3] - x[1], dy31 = y[3] - y[1], dz31 = z[3] - z[1];
  const double dx72 = x[7] - x[2], dy72 = y[7] - y[2], dz72 = z[7] - z[2];
  const double dx43 = x[4] - x[3], dy43 = y[4] - y[3], dz43 = z[4] - z[3];
  const double dx57 = x[5] - x[7], dy57 = y[5] - y[7], dz57 = z[5] - z[7];
  const double dx14 = x[1] - x[4], dy14 = y[1] - y[4], dz14 = z[1] - z[4];
  const double dx25 = x[2] - x[5], dy25 = y[2] - y[5], dz25 = z[2] - z[5];
  double v = (triple_product(dx31 + dx72, dx63, dx20, dy31 + dy72, dy63, dy20, dz31 + dz72, dz63,
                             dz20) +
              triple_product(dx43 + dx57, dx64, dx70, dy43 + dy57, dy64, dy70, dz43 + dz57, dz64,
                             dz70)) +
             triple_product(dx14 + dx25, dx61, dx50, dy14 + dy25, dy61, dy50, dz14 + dz25, dz61,
                            dz50);
  return v * (1.0 / 12.0);
}

// ---------------------------------------------------------------------------
// calcElemShapeFunctionDerivatives, src/domain.jl:1932-2021.  Only the four
// independent rows of b are produced (rows 5..8 are their negations and no
// caller needs them); returns the Jacobian volume.
// ---------------------------------------------------------------------------
struct Jac {
  double fjxxi, fjxet, fjxze, fjyxi, fjyet, fjyze, fjzxi, fjzet, fjzze;
};
LDEV Jac jacobian(const double (&x)[8], const double (&y)[8], const double (&z)[8]) {
  Jac j;
  j.fjxxi = .125 * ((((x[6] - x[0]) + (x[5] - x[3])) - (x[7] - x[1])) - (x[4] - x[2]));
  j.fjxet = .125 * ((((x[6] - x[0]) - (x[5] - x[3])) + (x[7] - x[1])) - (x[4] - x[2]));
  j.fjxze = .125 * ((((x[6] - x[0]) + (x[5] - x[3])) + (x[7] - x[1])) + (x[4] - x[2]));
  j.fjyxi = .125 * ((((y[6] - y[0]) + (y[5] - y[3])) - (y[7] - y[1])) - (y[4] - y[2]));
  j.fjyet = .125 * ((((y[6] - y[0]) - (y[5] - y[3])) + (y[7] - y[1])) - (y[4] - y[2]));
  j.fjyze = .125 * ((((y[6] - y[0]) + (y[5] - y[3])) + (y[7] - y[1])) + (y[4] - y[2]));
  j.fjzxi = .125 * ((((z[6] - z[0]) + (z[5] - z[3])) - (z[7] - z[1])) - (z[4] - z[2]));
  j.fjzet = .125 * ((((z[6] - z[0]) - (z[5] - z[3])) + (z[7] - z[1])) - (z[4] - z[2]));
  j.fjzze = .125 * ((((z[6] - z[0]) + (z[5] - z[3])) + (z[7] - z[1])) + (z[4] - z[2]));
  return j;
}
// Jacobian volume only (the stress integration keeps just detJ)
LDEV double jacobian_volume(const double (&x)[8], const double (&y)[8], const double (&z)[8]) {
  const Jac j = jacobian(x, y, z);
  const double cjxet = -(j.fjyxi * j.fjzze) + (j.fjzxi * j.fjyze);
  const double cjyet = (j.fjxxi * j.fjzze) - (j.fjzxi * j.fjxze);
  const double cjzet = -(j.fjxxi * j.fjyze) + (j.fjyxi * j.fjxze);
  return 8.0 * (((j.fjxet * cjxet) + (j.fjyet * cjyet)) + (j.fjzet * cjzet));
}
// b[n][c] for n = 0..3 and the volume
LDEV double shape_derivatives4(const double (&x)[8], const double (&y)[8], const double (&z)[8],
                               double (&b)[4][3]) {
  const Jac j = jacobian(x, y, z);
  const double cjxxi = (j.fjyet * j.fjzze) - (j.fjzet * j.fjyze);
  const double cjxet = -(j.fjyxi * j.fjzze) + (j.fjzxi * j.fjyze);
  const double cjxze = (j.fjyxi * j.fjzet) - (j.fjzxi * j.fjyet);
  const double cjyxi = -(j.fjxet * j.fjzze) + (j.fjzet * j.fjxze);
  const double cjyet = (j.fjxxi * j.fjzze) - (j.fjzxi * j.fjxze);
  const double cjyze = -(j.fjxxi * j.fjzet) + (j.fjzxi * j.fjxet);
  const double cjzxi = (j.fjxet * j.fjyze) - (j.fjyet * j.fjxze);
  const double cjzet = -(j.fjxxi * j.fjyze) + (j.fjyxi * j.fjxze);
  const double cjzze = (j.fjxxi * j.fjyet) - (j.fjyxi * j.fjxet);
  b[0][0] = ((-cjxxi) - cjxet) - cjxze;
  b[1][0] = (cjxxi - cjxet) - cjxze;
  b[2][0] = (cjxxi + cjxet) - cjxze;
  b[3][0] = ((-cjxxi) + cjxet) - cjxze;
  b[0][1] = ((-cjyxi) - cjyet) - cjyze;
  b[1][1] = (cjyxi - cjyet) - cjyze;
  b[2][1] = (cjyxi + cjyet) - cjyze;
  b[3][1] = ((-cjyxi) + cjyet) - cjyze;
  b[0][2] = ((-cjzxi) - cjzet) - cjzze;
  b[1][2] = (cjzxi - cjzet) - cjzze;
  b[2][2] = (cjzxi + cjzet) - cjzze;
  b[3][2] = ((-cjzxi) + cjzet) - cjzze;
  return 8.0 * (((j.fjxet * cjxet) + (j.fjyet * cjyet)) + (j.fjzet * cjzet));
}

// ---------------------------------------------------------------------------
// sumElemFaceNormal (src/domain.jl:971-989) and calcElemNodeNormals
// (src/domain.jl:991-1067, quirk Q1 on face 5).  B starts at zero and each
// node receives its faces' contributions in face order 1..6.
// ---------------------------------------------------------------------------
LDEV void face_normal(double x0, double y0, double z0, double x1, double y1, double z1, double x2,
                      double y2, double z2, double x3, double y3, double z3, double &ax, double &ay,
                      double &az) {
  const double bx0 = 0.5 * (((x3 + x2) - x1) - x0);
  const double by0 = 0.5 * (((y3 + y2) - y1) - y0);
  const double bz0 = 0.5 * (((z3 + z2) - z1) - z0);
  const double bx1 = 0.5 * (((x2 + x1) - x3) - x0);
  const double by1 = 0.5 * (((y2 + y1) - y3) - y0);
  const double bz1 = 0.5 * (((z2 + z1) - z3) - z0);
  ax = 0.25 * ((by0 * bz1) - (bz0 * by1));
  ay = 0.25 * ((bz0 * bx1) - (bx0 * bz1));
  az = 0.25 * ((bx0 * by1) - (by0 * bx1));
}
LDEV void node_normals(const double (&x)[8], const double (&y)[8], const double (&z)[8], int quirks,
                       double (&B)[8][3]) {
#pragma unroll
  for (int n = 0; n < 8; ++n) B[n][0] = B[n][1] = B[n][2] = 0.0;
  double ax, ay, az;
#define LULESH_ADD(n) B[n][0] += ax, B[n][1] += ay, B[n][2] += az
  face_normal(x[0], y[0], z[0], x[1], y[1], z[1], x[2], y[2], z[2], x[3], y[3], z[3], ax, ay, az);
  LULESH_ADD(0); LULESH_ADD(1); LULESH_ADD(2); LULESH_ADD(3);
  face_normal(x[0], y[0], z[0], x[4], y[4], z[4], x[5], y[5], z[5], x[1], y[1], z[1], ax, ay, az);
  LULESH_ADD(0); LULESH_ADD(1); LULESH_ADD(4); LULESH_ADD(5);
  face_normal(x[1], y[1], z[1], x[5], y[5], z[5], x[6], y[6], z[6], x[2], y[2], z[2], ax, ay, az);
  LULESH_ADD(1); LULESH_ADD(2); LULESH_ADD(5); LULESH_ADD(6);
  face_normal(x[2], y[2], z[2], x[6], y[6], z[6], x[7], y[7], z[7], x[3], y[3], z[3], ax, ay, az);
  LULESH_ADD(2); LULESH_ADD(3); LULESH_ADD(6); LULESH_ADD(7);
  face_normal(x[3], y[3], z[3], x[7], y[7], z[7], x[4], y[4], z[4], x[0], y[0], z[0], ax, ay, az);
  if (quirks & Q1) {
    B[0][0] += ax; B[3][0] += ax; B[4][0] += ax; B[7][0] += ax;
    B[0][1] += ay; B[3][1] += ay; B[4][1] += ay; B[7][1] += ay;
    B[0][1] += ay;
    B[3][2] += az; B[4][2] += az; B[7][2] += az;
  } else {
    LULESH_ADD(0); LULESH_ADD(3); LULESH_ADD(4); LULESH_ADD(7);
  }
  face_normal(x[4], y[4], z[4], x[7], y[7], z[7], x[6], y[6], z[6], x[5], y[5], z[5], ax, ay, az);
  LULESH_ADD(4); LULESH_ADD(5); LULESH_ADD(6); LULESH_ADD(7);
#undef LULESH_ADD
}

// ---------------------------------------------------------------------------
// voluDer (src/domain.jl:1178-1207) and calcElemVolumeDerivative
// (src/domain.jl:1209-1256, quirk Q2 = slots 2 and 4 exchanged)
// ---------------------------------------------------------------------------
LDEV void volu_der(double x0, double x1, double x2, double x3, double x4, double x5, double y0,
                   double y1, double y2, double y3, double y4, double y5, double z0, double z1,
                   double z2, double z3, double z4, double z5, double &dx, double &dy, double &dz) {
  double vx = ((((((y1 + y2) * (z0 + z1)) - ((y0 + y1) * (z1 + z2))) + ((y0 + y4) * (z3 + z4))) -
                ((y3 + y4) * (z0 + z4))) -
               ((y2 + y5) * (z3 + z5))) +
              ((y3 + y5) * (z2 + z5));
  double vy = ((((((-(x1 + x2)) * (z0 + z1)) + ((x0 + x1) * (z1 + z2))) - ((x0 + x4) * (z3 + z4))) +
                ((x3 + x4) * (z0 + z4))) +
               ((x2 + x5) * (z3 + z5))) -
              ((x3 + x5) * (z2 + z5));
  double vz = ((((((-(y1 + y2)) * (x0 + x1)) + ((y0 + y1) * (x1 + x2))) - ((y0 + y4) * (x3 + x4))) +
                ((y3 + y4) * (x0 + x4))) +
               ((y2 + y5) * (x3 + x5))) -
              ((y3 + y5) * (x2 + x5));
  const double twelfth = 1.0 / 12.0;
  dx = vx * twelfth;
  dy = vy * twelfth;
  dz = vz * twelfth;
}
LDEV void volume_derivative(const double (&x)[8], const double (&y)[8], const double (&z)[8],
                            int quirks, double (&dvdx)[8], double (&dvdy)[8], double (&dvdz)[8]) {
#define LULESH_VD(slot, a, b, c, d, e, f)                                                      \
  volu_der(x[a], x[b], x[c], x[d], x[e], x[f], y[a], y[b], y[c], y[d], y[e], y[f], z[a], z[b], \
           z[c], z[d], z[e], z[f], dvdx[slot], dvdy[slot], dvdz[slot])
  const bool q2 = (quirks & Q2) != 0;
  LULESH_VD(0, 1, 2, 3, 4, 5, 7);
  if (q2) {
    LULESH_VD(1, 0, 1, 2, 7, 4, 6);
    LULESH_VD(3, 2, 3, 0, 5, 6, 4);
  } else {
    LULESH_VD(3, 0, 1, 2, 7, 4, 6);
    LULESH_VD(1, 2, 3, 0, 5, 6, 4);
  }
  LULESH_VD(2, 3, 0, 1, 6, 7, 5);
  LULESH_VD(4, 7, 6, 5, 0, 3, 1);
  LULESH_VD(5, 4, 7, 6, 1, 0, 2);
  LULESH_VD(6, 5, 4, 7, 2, 1, 3);
  LULESH_VD(7, 6, 5, 4, 3, 2, 0);
#undef LULESH_VD
}

// Gamma of src/domain.jl:1490-1499 as sign bits: bit 4n+m set means
// Gamma[n][m] = -1 (row nibbles 8,6,b,5,3,d,0,e for n = 0..7).
LDEV constexpr double gam(int n, int m) {
  return ((0xE0D35B68u >> (4 * n + m)) & 1u) ? -1.0 : 1.0;
}

// ---------------------------------------------------------------------------
// areaFace (src/domain.jl:1874-1891), calcElemCharacteristicLength
// (src/domain.jl:1893-1930)
// ---------------------------------------------------------------------------
LDEV double area_face(double x0, double x1, double x2, double x3, double y0, double y1, double y2,
                      double y3, double z0, double z1, double z2, double z3) {
  const double fx = (x2 - x0) - (x3 - x1);
  const double fy = (y2 - y0) - (y3 - y1);
  const double fz = (z2 - z0) - (z3 - z1);
  const double gx = (x2 - x0) + (x3 - x1);
  const double gy = (y2 - y0) + (y3 - y1);
  const double gz = (z2 - z0) + (z3 - z1);
  const double fg = ((fx * gx) + (fy * gy)) + (fz * gz);
  return ((((fx * fx) + (fy * fy)) + (fz * fz)) * (((gx * gx) + (gy * gy)) + (gz * gz))) -
         (fg * fg);
}
LDEV double char_length(const double (&x)[8], const double (&y)[8], const double (&z)[8],
                        double volume) {
  double cl = 0.0;
#define LULESH_AF(a, b, c, d) \
  cl = jmax(area_face(x[a], x[b], x[c], x[d], y[a], y[b], y[c], y[d], z[a], z[b], z[c], z[d]), cl)
  LULESH_AF(0, 1, 2, 3);
  LULESH_AF(4, 5, 6, 7);
  LULESH_AF(0, 1, 5, 4);
  LULESH_AF(1, 2, 6, 5);
  LULESH_AF(2, 3, 7, 6);
  LULESH_AF(3, 0, 4, 7);
#undef LULESH_AF
  return (4.0 * volume) / sqrt(cl);
}

// diagonal of calcElemVelocityGradient (src/domain.jl:2023-2043); the
// off-diagonal terms d4..d6 are never read by the reference
LDEV double vgrad_term(const double (&b)[4][3], int c, const double (&v)[8]) {
  return (((b[0][c] * (v[0] - v[6])) + (b[1][c] * (v[1] - v[7]))) + (b[2][c] * (v[2] - v[4]))) +
         (b[3][c] * (v[3] - v[5]));
}

LDEV double sum4(double a, double b, double c, double d) { return ((a + b) + c) + d; }

// limiter of src/domain.jl:2359-2378 (identical for xi, eta, zeta)
LDEV double limit_phi(double delvm, double delvp, double norm, double mult, double maxslope) {
  delvm = delvm * norm;
  delvp = delvp * norm;
  double phi = 0.5 * (delvm + delvp);
  delvm = delvm * mult;
  delvp = delvp * mult;
  if (delvm < phi) phi = delvm;
  if (delvp < phi) phi = delvp;
  if (phi < 0.0) phi = 0.0;
  if (phi > maxslope) phi = maxslope;
  return phi;
}

// calcPressureForElems for one element, src/domain.jl:2585-2614
LDEV double pressure(double &bvc, double &pbvc, double e_old, double compression, double vnewc,
                     double pmin, double p_cut, double eosvmax) {
  const double c1s = 2.0 / 3.0;
  bvc = c1s * (compression + 1.0);
  pbvc = c1s;
  double p_new = bvc * e_old;
  if (fabs(p_new) < p_cut) p_new = 0.0;
  if (vnewc >= eosvmax) p_new = 0.0;
  if (p_new < pmin) p_new = pmin;
  return p_new;
}

}  // namespace lulesh
