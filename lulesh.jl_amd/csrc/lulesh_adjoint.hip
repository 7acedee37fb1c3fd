// Adjoint kernels of one LULESH cycle on gfx950 (reverse of k_nodal ->
// k_kin -> k_qeos with the cycle's dt held fixed).
//
// The kernels read cycle c from a stored slot (AdjArgs): its start state
// S_c = (x y z xd yd zd, e p q v ss), the node set after it (x', xd') and its
// temporaries (vnew delv vdov delx_* delv_*), all kept when the forward
// cycle ran.  On entry the adjoint buffers hold the adjoint of S_{c+1}; on
// exit that of S_c:
//   a_qeos  EOS + volume update + MonoQ region (limiter) adjoints
//   a_kin   MonoQ gradients + relative volume adjoints -> corners (x', xd')
//   a_node1 corner gather + node-update adjoint -> force adjoint per node
//   a_force stress + hourglass corner-force adjoints -> corners (x, xd)
//   a_node2 corner gather -> adjoint of the node state
// Gathers run over fixed lists, so the gradient is deterministic.
#include "lulesh_adj.h"
#include "lulesh_elem.h"
#include "lulesh_adjoint.h"

namespace lulesh {

using lulesh_adj::EosParams;
using lulesh_adj::EosOut;
using lulesh_adj::QIn;
using lulesh_adj::QOut;

namespace {

LDEV void elem_nodes_a(int64_t e, int s, int64_t (&nid)[8]) {
  const int64_t s1 = s + 1, P = s1 * s1;
  const int64_t col = e % s, rest = e / s;
  const int64_t row = rest % s, plane = rest / s;
  const int64_t n0 = plane * P + row * s1 + col;
  nid[0] = n0;
  nid[1] = n0 + 1;
  nid[2] = n0 + s1 + 1;
  nid[3] = n0 + s1;
  nid[4] = n0 + P;
  nid[5] = n0 + P + 1;
  nid[6] = n0 + P + s1 + 1;
  nid[7] = n0 + P + s1;
}

LDEV int64_t clampi(int64_t i, int64_t E) { return i < 0 ? 0 : (i >= E ? E - 1 : i); }

// element whose delv a (direction d, side sd) read of element i addresses,
// as k_qeos computes it (setupElementConnectivities ids, quirk Q3)
LDEV int64_t read_target(int d, int sd, int64_t i, int64_t s, int64_t E, bool q3) {
  const int64_t s2 = s * s;
  if (d == 0) {
    if (sd == 0) return clampi(i == 0 ? 1 : i - 1, E);
    const int64_t nb = i < E - 1 ? i + 1 : E - 1;
    return q3 ? clampi(nb - 1, E) : clampi(nb, E);
  }
  if (d == 1) {
    if (sd == 0) return clampi(i < s ? i : i - s, E);
    return clampi(i < E - s ? i + s : i - (E - s), E);
  }
  if (sd == 0) return clampi(i < s2 ? i : i - s2, E);
  return clampi(i < E - s2 ? i + s2 : i, E);
}

constexpr uint32_t MASK_M[3] = {XI_M, ETA_M, ZETA_M};
constexpr uint32_t MASK_P[3] = {XI_P, ETA_P, ZETA_P};
constexpr uint32_t SYMM_M[3] = {XI_M_SYMM, ETA_M_SYMM, ZETA_M_SYMM};
constexpr uint32_t SYMM_P[3] = {XI_P_SYMM, ETA_P_SYMM, ZETA_P_SYMM};
constexpr uint32_t COMM_M[3] = {XI_M_COMM, ETA_M_COMM, ZETA_M_COMM};
constexpr uint32_t COMM_P[3] = {XI_P_COMM, ETA_P_COMM, ZETA_P_COMM};

// 0 = neighbour read, 1 = self, 2 = zero (FREE or invalid)
LDEV int bc_source(uint32_t bc, int d, int sd) {
  const uint32_t c = bc & (sd == 0 ? MASK_M[d] : MASK_P[d]);
  if (c == 0 || c == (sd == 0 ? COMM_M[d] : COMM_P[d])) return 0;
  if (c == (sd == 0 ? SYMM_M[d] : SYMM_P[d])) return 1;
  return 2;
}

}  // namespace

// ---- reverse of k_qeos ------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void a_qeos(AdjArgs A) {
  const int64_t e = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (e >= A.E) return;
  const DevScalars *sc = A.sc;
  const DevFields &f = A.f;
  const double *dvs[3] = {f.delv_xi, f.delv_eta, f.delv_zeta};
  const double *dxs[3] = {f.delx_xi, f.delx_eta, f.delx_zeta};
  const uint32_t bc = f.elemBC[e];
  QIn qi;
  int src[3][2];
  int64_t tgt[3][2];
  for (int d = 0; d < 3; ++d) {
    qi.dv[d] = dvs[d][e];
    qi.dx[d] = dxs[d][e];
    for (int sd = 0; sd < 2; ++sd) {
      src[d][sd] = bc_source(bc, d, sd);
      tgt[d][sd] = read_target(d, sd, e, A.s, A.E, A.q3);
      const double val = src[d][sd] == 0 ? dvs[d][tgt[d][sd]] : (src[d][sd] == 1 ? qi.dv[d] : 0.0);
      if (sd == 0) qi.dm[d] = val;
      else qi.dp[d] = val;
    }
  }
  qi.vdov = f.vdov[e];
  qi.elemMass = f.elemMass[e];
  qi.volo = f.volo[e];
  qi.vnew = f.vnew[e];
  qi.qlc = sc->qlc_monoq;
  qi.qqc = sc->qqc_monoq;
  qi.mult = sc->monoq_limiter_mult;
  qi.maxslope = sc->monoq_max_slope;
  // forward q terms (a zero adjoint gives just the forward values)
  double z3[3] = {0, 0, 0}, z3b[3] = {0, 0, 0}, z3c[3] = {0, 0, 0}, z3d[3] = {0, 0, 0};
  double zv = 0.0, zl = 0.0, zq = 0.0;
  const QOut q = lulesh_adj::monoq_region_rev(qi, 0.0, 0.0, z3, z3b, z3c, z3d, &zv, &zl, &zq);

  // EOS adjoint
  EosParams P;
  P.e_cut = sc->e_cut; P.p_cut = sc->p_cut; P.q_cut = sc->q_cut; P.eosvmax = sc->eosvmax;
  P.eosvmin = sc->eosvmin; P.pmin = sc->pmin; P.emin = sc->emin; P.rho0 = sc->refdens;
  const double vnew = qi.vnew;
  double vnewc = vnew, dclamp = 1.0;
  if (P.eosvmin != 0.0 && vnewc < P.eosvmin) { vnewc = P.eosvmin; dclamp = 0.0; }
  if (P.eosvmax != 0.0 && vnewc > P.eosvmax) { vnewc = P.eosvmax; dclamp = 0.0; }
  const double lo[4] = {A.lp[e], A.le[e], A.lq[e], A.lss[e]};
  double li[7] = {0, 0, 0, 0, 0, 0, 0};
  lulesh_adj::eos_rev(vnewc, f.delv[e], f.e[e], f.p[e], f.q[e], q.qlin, q.qquad, P, lo, li);
  double lvnew = li[0] * dclamp;
  // v' = vnew unless snapped to 1
  if (!(fabs(vnew - 1.0) < sc->v_cut)) lvnew += A.lv[e];

  // MonoQ region adjoint with the EOS's (lql, lqq)
  double ldv[3] = {0, 0, 0}, ldm[3] = {0, 0, 0}, ldp[3] = {0, 0, 0}, ldx[3] = {0, 0, 0};
  double lqlc = 0.0, lqqc = 0.0;
  lulesh_adj::monoq_region_rev(qi, li[5], li[6], ldv, ldm, ldp, ldx, &lvnew, &lqlc, &lqqc);
  for (int d = 0; d < 3; ++d) {
    double own = ldv[d];
    const double side[2] = {ldm[d], ldp[d]};
    for (int sd = 0; sd < 2; ++sd) {
      double out = 0.0;
      if (src[d][sd] == 0) {
        if (tgt[d][sd] == e) own += side[sd];
        else out = side[sd];
      } else if (src[d][sd] == 1) {
        own += side[sd];
      }
      A.nbr[(int64_t)(d * 2 + sd) * A.E + e] = out;
    }
    A.ldvs[(int64_t)d * A.E + e] = own;
    A.ldx[(int64_t)d * A.E + e] = ldx[d];
  }
  A.lvnew[e] = lvnew;
  // adjoint of S_c so far: e final; p, q from the EOS (stress adds later);
  // v through delv = vnew - v (hourglass adds later); ss only via hourglass
  A.le[e] = li[2];
  A.lp[e] = li[3];
  A.lq[e] = li[4];
  A.lv[e] = -li[1];
  A.lss[e] = 0.0;
  A.lvnew_delv[e] = li[1];
  A.pacc[(int64_t)1 * A.E + e] += lqlc;
  A.pacc[(int64_t)2 * A.E + e] += lqqc;
}

// ---- reverse of k_kin (MonoQ gradients + relative volume) -------------------
__global__ __launch_bounds__(BLOCK) void a_kin(AdjArgs A) {
  const int64_t e = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (e >= A.E) return;
  const DevFields &f = A.f;
  const NodeSet S = A.cur;
  // own delv adjoint plus what the elements that read this delv sent back
  double l6[6];
  for (int d = 0; d < 3; ++d) {
    double acc = A.ldvs[(int64_t)d * A.E + e];
    for (int sd = 0; sd < 2; ++sd) {
      const double *nb = A.nbr + (int64_t)(d * 2 + sd) * A.E;
      // readers of e are among these candidates; each is checked exactly
      const int64_t s = A.s, s2 = (int64_t)A.s * A.s, E = A.E;
      const int64_t step = d == 0 ? 1 : (d == 1 ? s : s2);
      const int64_t cand[5] = {e - step, e + step, d == 1 ? e + (E - s) : -1,
                               d == 0 && e == 1 ? 0 : -1, d == 0 && e == E - 2 ? E - 1 : -1};
      for (int k = 0; k < 5; ++k) {
        const int64_t i = cand[k];
        if (i < 0 || i >= E || i == e) continue;
        bool dup = false;
        for (int k2 = 0; k2 < k; ++k2) dup = dup || cand[k2] == i;
        if (dup) continue;
        if (read_target(d, sd, i, s, E, A.q3) == e) acc += nb[i];
      }
    }
    l6[3 + d] = acc;
    l6[d] = A.ldx[(int64_t)d * A.E + e];
  }
  int64_t nid[8];
  elem_nodes_a(e, A.s, nid);
  double X[8], Y[8], Z[8], VX[8], VY[8], VZ[8];
  for (int n = 0; n < 8; ++n) {
    X[n] = S.x[nid[n]]; Y[n] = S.y[nid[n]]; Z[n] = S.z[nid[n]];
    VX[n] = S.xd[nid[n]]; VY[n] = S.yd[nid[n]]; VZ[n] = S.zd[nid[n]];
  }
  double lX[8] = {0}, lY[8] = {0}, lZ[8] = {0}, lVX[8] = {0}, lVY[8] = {0}, lVZ[8] = {0};
  const double volo = f.volo[e];
  double lvnew = A.lvnew[e] + A.lvnew_delv[e];  // delv = vnew - v
  lvnew += lulesh_adj::monoq_grad_rev(X, Y, Z, VX, VY, VZ, volo, f.vnew[e], l6, lX, lY, lZ, lVX,
                                      lVY, lVZ);
  lulesh_adj::volume_grad_rev(X, Y, Z, lvnew / volo, lX, lY, lZ);
  double *cr = A.corner;
  for (int n = 0; n < 8; ++n) {
    cr[(int64_t)(0 * 8 + n) * A.E + e] = lX[n];
    cr[(int64_t)(1 * 8 + n) * A.E + e] = lY[n];
    cr[(int64_t)(2 * 8 + n) * A.E + e] = lZ[n];
    cr[(int64_t)(3 * 8 + n) * A.E + e] = lVX[n];
    cr[(int64_t)(4 * 8 + n) * A.E + e] = lVY[n];
    cr[(int64_t)(5 * 8 + n) * A.E + e] = lVZ[n];
  }
}

// corner of element (ei,ej,ek) that node (i,j,k) occupies
LDEV int corner_of_a(int ox, int oy, int oz) { return 4 * oz + (oy ? (ox ? 2 : 3) : (ox ? 1 : 0)); }

// sum over the node's corners of comp c (0..5) of the corner buffer
LDEV void gather_corners(const AdjArgs &A, const double *corner, int64_t n, double out[6]) {
  const int s = A.s;
  const int64_t s1 = s + 1;
  const int i = (int)(n % s1), j = (int)((n / s1) % s1), k = (int)(n / (s1 * s1));
  for (int c = 0; c < 6; ++c) out[c] = 0.0;
  for (int dk = -1; dk <= 0; ++dk) {
    const int ek = k + dk;
    if (ek < 0 || ek >= s) continue;
    for (int dj = -1; dj <= 0; ++dj) {
      const int ej = j + dj;
      if (ej < 0 || ej >= s) continue;
      for (int di = -1; di <= 0; ++di) {
        const int ei = i + di;
        if (ei < 0 || ei >= s) continue;
        const int64_t e = ((int64_t)ek * s + ej) * s + ei;
        const int cn = corner_of_a(-di, -dj, -dk);
        for (int c = 0; c < 6; ++c) out[c] += corner[(int64_t)(c * 8 + cn) * A.E + e];
      }
    }
  }
}

// ---- reverse of the node update (calcAcceleration .. calcPosition) ----------
__global__ __launch_bounds__(BLOCK) void a_node1(AdjArgs A) {
  const int64_t n = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (n >= A.N) return;
  double g[6];
  if (A.force_pending) {
    // what a_node2 of the later cycle would add: corner_f into the node state
    gather_corners(A, A.corner_f, n, g);
    A.lx[n] += g[0]; A.ly[n] += g[1]; A.lz[n] += g[2];
    A.lxd[n] += g[3]; A.lyd[n] += g[4]; A.lzd[n] += g[5];
  }
  gather_corners(A, A.corner, n, g);
  const double dt = A.dt;
  const int64_t s1 = A.s + 1;
  const int i = (int)(n % s1), j = (int)((n / s1) % s1), k = (int)(n / (s1 * s1));
  const double mass = A.f.nodalMass[n];
  double *lxs[3] = {A.lx, A.ly, A.lz};
  double *lvs[3] = {A.lxd, A.lyd, A.lzd};
  const double *vnew[3] = {A.cur.xd, A.cur.yd, A.cur.zd};
  const bool bcz[3] = {i == 0, j == 0, k == 0};
  for (int c = 0; c < 3; ++c) {
    const double lxn = lxs[c][n] + g[c];        // adjoint of x'
    double lvn = lvs[c][n] + g[3 + c];          // adjoint of xd'
    lvn += lxn * dt;                             // x' = x + xd' dt
    double lv_old = 0.0, lF = 0.0;
    if (vnew[c][n] != 0.0) {                     // not cut by u_cut
      lv_old = lvn;
      if (!bcz[c]) lF = lvn * dt / mass;          // xd' = xd + (F/m) dt
    }
    lxs[c][n] = lxn;
    lvs[c][n] = lv_old;
    A.lF[(int64_t)c * A.N + n] = lF;
  }
}

// ---- reverse of the corner forces (stress + FB hourglass) -------------------
// velocities of S_c and force adjoints read from memory node by node, so
// they do not occupy registers across the whole element
// The element's node values are addressed by 32-bit byte offsets from
// uniform array bases (the host keeps 8 N and 8 E below 2^32), so each
// node costs one offset register shared by every array read there instead
// of a 64-bit address per array.
LDEV const double &at_off(const double *base, uint32_t off) {
  return *(const double *)((const char *)base + off);
}
struct NodeSrc {
  const double *vx, *vy, *vz, *lF;
  int64_t N;
  uint32_t off[8];  // byte offsets of the element's nodes
  LDEV double v(int c, int n) const { return at_off(c == 0 ? vx : (c == 1 ? vy : vz), off[n]); }
  LDEV double lf(int n, int c) const { return at_off(lF + (int64_t)c * N, off[n]); }
};
// per-corner adjoints straight to the corner buffer ([6][8][E])
struct CornerSink {
  double *cr;
  int64_t E;
  uint32_t off;  // byte offset of the element
  LDEV void put(int k, int n, double x) {
    *(double *)((char *)(cr + (int64_t)(k * 8 + n) * E) + off) = x;
  }
  LDEV void lv(int c, int n, double x) { put(3 + c, n, x); }
};
// One coordinate axis of the element's eight nodes, parked in LDS (stride
// BLOCK, so a wave's reads are conflict-free) and re-read at every use: a
// volatile LDS read cannot be merged with an earlier one, so the 24
// coordinates hold no registers between uses.  That, with the 32-bit node
// offsets, keeps a_force within 256 VGPRs: two waves per SIMD instead of one.
struct LdsAxis {
  const double *p;
  __device__ double operator[](int n) const {
    typedef __attribute__((address_space(3))) const volatile double lds_double;
    return *(lds_double *)(p + n * BLOCK);
  }
};

template <bool Q1, bool Q2>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(2))) void a_force(AdjArgs A) {
  const int64_t e = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (e >= A.E) return;
  const DevFields &f = A.f;
  const NodeSet S = A.old;
  int64_t nid[8];
  elem_nodes_a(e, A.s, nid);
  NodeSrc src{S.xd, S.yd, S.zd, A.lF, A.N, {}};
  for (int n = 0; n < 8; ++n) src.off[n] = (uint32_t)nid[n] * 8u;
  __shared__ double pos[3][8][BLOCK];
  for (int n = 0; n < 8; ++n) {
    pos[0][n][threadIdx.x] = S.x[nid[n]];
    pos[1][n][threadIdx.x] = S.y[nid[n]];
    pos[2][n][threadIdx.x] = S.z[nid[n]];
  }
  const LdsAxis X{&pos[0][0][threadIdx.x]}, Y{&pos[1][0][threadIdx.x]}, Z{&pos[2][0][threadIdx.x]};
  CornerSink sink{A.corner_f, A.E, (uint32_t)e * 8u};
  double lX[8] = {0}, lY[8] = {0}, lZ[8] = {0};
  const double hourg = A.sc->hgcoef;
  if (hourg > 0.0) {
    const double volo = f.volo[e];
    const double determ = volo * f.v[e];
    const double ss = f.ss[e], mass = f.elemMass[e], c13 = cbrt_cr(determ);
    const lulesh_adj::HgRev r = lulesh_adj::hourglass_rev_s<Q2>(X, Y, Z, src, sink, determ, ss,
                                                                mass, hourg, c13, lX, lY, lZ);
    A.lv[e] += r.ldeterm * volo;
    A.lss[e] += r.lss;
    A.pacc[e] += r.lhgcoef;
  } else {
    for (int c = 0; c < 3; ++c)
      for (int n = 0; n < 8; ++n) sink.lv(c, n, 0.0);
  }
  const double ms = f.p[e] + f.q[e];
  const double lms = lulesh_adj::stress_rev_s(X, Y, Z, Q1, ms, src, lX, lY, lZ);
  A.lp[e] += lms;
  A.lq[e] += lms;
  for (int n = 0; n < 8; ++n) {
    sink.put(0, n, lX[n]);
    sink.put(1, n, lY[n]);
    sink.put(2, n, lZ[n]);
  }
}

__global__ __launch_bounds__(BLOCK) void a_node2(AdjArgs A) {
  const int64_t n = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (n >= A.N) return;
  double g[6];
  gather_corners(A, A.corner_f, n, g);
  A.lx[n] += g[0];
  A.ly[n] += g[1];
  A.lz[n] += g[2];
  A.lxd[n] += g[3];
  A.lyd[n] += g[4];
  A.lzd[n] += g[5];
}

// loss seed: dL/de = elemMass (L = sum e * elemMass), everything else 0
__global__ __launch_bounds__(BLOCK) void a_seed(AdjArgs A) {
  const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i < A.E) {
    A.le[i] = A.f.elemMass[i];
    A.lp[i] = 0.0; A.lq[i] = 0.0; A.lv[i] = 0.0; A.lss[i] = 0.0;
    A.pacc[i] = 0.0; A.pacc[A.E + i] = 0.0; A.pacc[2 * A.E + i] = 0.0;
  }
  if (i < A.N) {
    A.lx[i] = 0.0; A.ly[i] = 0.0; A.lz[i] = 0.0;
    A.lxd[i] = 0.0; A.lyd[i] = 0.0; A.lzd[i] = 0.0;
  }
}

// copy the cycle state (current node set + e p q v ss) to/from a buffer
// laid out as [x y z xd yd zd][N] then [e p q v ss][E]
__global__ __launch_bounds__(BLOCK) void a_state_copy(DevScalars *sc, DevFields f, int64_t E,
                                                      int64_t N, double *buf, int to_buf) {
  const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  const int par = sc->parity;
  double *nd[6] = {par ? f.ns[1].x : f.ns[0].x, par ? f.ns[1].y : f.ns[0].y,
                   par ? f.ns[1].z : f.ns[0].z, par ? f.ns[1].xd : f.ns[0].xd,
                   par ? f.ns[1].yd : f.ns[0].yd, par ? f.ns[1].zd : f.ns[0].zd};
  double *el[5] = {f.e, f.p, f.q, f.v, f.ss};
  if (i < N)
    for (int k = 0; k < 6; ++k) {
      if (to_buf) buf[k * N + i] = nd[k][i];
      else nd[k][i] = buf[k * N + i];
    }
  if (i < E)
    for (int k = 0; k < 5; ++k) {
      if (to_buf) buf[6 * N + k * E + i] = el[k][i];
      else el[k][i] = buf[6 * N + k * E + i];
    }
}

__global__ __launch_bounds__(BLOCK) void a_slot_step(DevScalars *sc, DevFields f, int64_t E,
                                                     int64_t N, double *temps, double *state,
                                                     const double *dt, int64_t base) {
  const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  const int par = sc->parity;
  if (temps && i < E) {
    const double *el[9] = {f.vnew, f.delv, f.vdov, f.delx_xi, f.delx_eta, f.delx_zeta,
                           f.delv_xi, f.delv_eta, f.delv_zeta};
    for (int k = 0; k < 9; ++k) temps[k * E + i] = el[k][i];
  }
  if (state) {
    if (i < N) {
      const double *nd[6] = {par ? f.ns[1].x : f.ns[0].x, par ? f.ns[1].y : f.ns[0].y,
                             par ? f.ns[1].z : f.ns[0].z, par ? f.ns[1].xd : f.ns[0].xd,
                             par ? f.ns[1].yd : f.ns[0].yd, par ? f.ns[1].zd : f.ns[0].zd};
      for (int k = 0; k < 6; ++k) state[k * N + i] = nd[k][i];
    }
    if (i < E) {
      const double *el[5] = {f.e, f.p, f.q, f.v, f.ss};
      for (int k = 0; k < 5; ++k) state[6 * N + k * E + i] = el[k][i];
    }
  }
  if (dt && i == 0) {
    const double d = dt[sc->cycle - base];
    sc->deltatime = d;
    sc->time = sc->time + d;
    sc->cycle = sc->cycle + 1;
  }
}

__global__ void a_set_cycle(DevScalars *sc, long long cycle) {
  if (threadIdx.x == 0 && blockIdx.x == 0) sc->cycle = cycle;
}

static inline unsigned nblk_a(int64_t n) { return (unsigned)((n + BLOCK - 1) / BLOCK); }

hipError_t launch_adjoint(int which, hipStream_t st, const AdjArgs &a) {
  switch (which) {
    case AK_QEOS:
      hipLaunchKernelGGL(a_qeos, dim3(nblk_a(a.E)), dim3(BLOCK), 0, st, a);
      break;
    case AK_KIN:
      hipLaunchKernelGGL(a_kin, dim3(nblk_a(a.E)), dim3(BLOCK), 0, st, a);
      break;
    case AK_NODE1:
      hipLaunchKernelGGL(a_node1, dim3(nblk_a(a.N)), dim3(BLOCK), 0, st, a);
      break;
    case AK_FORCE:
      if (a.q1 && a.q2)
        hipLaunchKernelGGL((a_force<true, true>), dim3(nblk_a(a.E)), dim3(BLOCK), 0, st, a);
      else if (a.q1)
        hipLaunchKernelGGL((a_force<true, false>), dim3(nblk_a(a.E)), dim3(BLOCK), 0, st, a);
      else if (a.q2)
        hipLaunchKernelGGL((a_force<false, true>), dim3(nblk_a(a.E)), dim3(BLOCK), 0, st, a);
      else
        hipLaunchKernelGGL((a_force<false, false>), dim3(nblk_a(a.E)), dim3(BLOCK), 0, st, a);
      break;
    case AK_NODE2:
      hipLaunchKernelGGL(a_node2, dim3(nblk_a(a.N)), dim3(BLOCK), 0, st, a);
      break;
    case AK_SEED:
      hipLaunchKernelGGL(a_seed, dim3(nblk_a(a.N > a.E ? a.N : a.E)), dim3(BLOCK), 0, st, a);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_state_copy(hipStream_t st, DevScalars *sc, const DevFields &f, int64_t E,
                             int64_t N, double *buf, int to_buf) {
  hipLaunchKernelGGL(a_state_copy, dim3(nblk_a(N > E ? N : E)), dim3(BLOCK), 0, st, sc, f, E, N,
                     buf, to_buf);
  return hipGetLastError();
}

hipError_t launch_slot_step(hipStream_t st, DevScalars *sc, const DevFields &f, int64_t E,
                            int64_t N, double *temps, double *state, const double *dt,
                            int64_t base) {
  hipLaunchKernelGGL(a_slot_step, dim3(nblk_a(N > E ? N : E)), dim3(BLOCK), 0, st, sc, f, E, N,
                     temps, state, dt, base);
  return hipGetLastError();
}

hipError_t launch_set_cycle(hipStream_t st, DevScalars *sc, long long cycle) {
  hipLaunchKernelGGL(a_set_cycle, dim3(1), dim3(64), 0, st, sc, cycle);
  return hipGetLastError();
}

}  // namespace lulesh
