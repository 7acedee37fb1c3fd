// MI355X (gfx950) kernels of one LULESH Lagrange leap-frog cycle, FP64.
//
// One cycle = five launches on the context stream:
//   k_force  element-parallel: stress corner forces (integrateStressForElems)
//            and Flanagan-Belytschko hourglass corner forces
//            (calcHourglassControlForElems + calcFBHourglassForceForElems)
//   k_node   node-parallel: fixed-order corner gather (stress sum, then
//            hourglass sum, as the reference sums them), acceleration,
//            symmetry-plane BCs, velocity, position
//   k_kin    element-parallel: calcKinematicsForElems + the deviatoric loop
//            + calcMonotonicQGradientsForElems
//   k_qeos   element-parallel: calcMonotonicQRegionForElems, the QStop check,
//            applyMaterialPropertiesForElems/evalEOSForElems,
//            updateVolumesForElems and the per-element Courant/hydro terms,
//            reduced to one minimum per block
//   k_dt     one block: the final minima (calcTimeConstraintsForElems) and,
//            in driver-loop mode, the loop test and the next timeIncrement!
// No atomics touch field data: corners are written once and gathered in the
// reference's order, so the FP64 result is deterministic run to run.
#include "lulesh_elem.h"
#include "lulesh_kernels.h"

namespace lulesh {

// ---------------------------------------------------------------------------
// structured-mesh indexing (buildMesh!, src/domain.jl:117-174)
// ---------------------------------------------------------------------------
LDEV void elem_nodes(int64_t e, int s, int64_t (&nid)[8]) {
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

template <typename T>
LDEV void gather8(const T *__restrict__ src, const int64_t (&nid)[8], T (&out)[8]) {
#pragma unroll
  for (int n = 0; n < 8; ++n) out[n] = src[nid[n]];
}

// An element's 8 nodes as 4 adjacent pairs (n, n+1) -- nodes (0,1) (3,2)
// (4,5) (7,6) of the hex order of buildMesh! -- each pair one 16-byte load
// at a 32-bit byte offset from a uniform base (node arrays of a context stay
// below 4 GiB).  o[] = byte offsets of nodes 0, 3, 4, 7.
LDEV void hex_pairs(uint32_t n0, uint32_t s1, uint32_t (&o)[4]) {
  const uint32_t P = s1 * s1;
  o[0] = n0 * 8u;
  o[1] = (n0 + s1) * 8u;
  o[2] = (n0 + P) * 8u;
  o[3] = (n0 + P + s1) * 8u;
}

LDEV double2 ld_pair(const double *base, uint32_t boff) {
  const double *p = (const double *)__builtin_assume_aligned((const char *)base + boff, 8);
  double2 v;
  __builtin_memcpy(&v, p, 16);
  return v;
}

LDEV void gather8p(const double *__restrict__ base, const uint32_t (&o)[4], double (&out)[8]) {
  const double2 a = ld_pair(base, o[0]), b = ld_pair(base, o[1]), c = ld_pair(base, o[2]),
                d = ld_pair(base, o[3]);
  out[0] = a.x; out[1] = a.y;
  out[3] = b.x; out[2] = b.y;
  out[4] = c.x; out[5] = c.y;
  out[7] = d.x; out[6] = d.y;
}

// element of a field at a 32-bit byte offset from a uniform base
template <typename T>
LDEV T &at(T *base, uint32_t boff) {
  return *(T *)((char *)base + boff);
}

LDEV void raise(DevScalars *sc, int code) { atomicMin(&sc->err, code); }

// Stores of k_kin's and k_qeos' outputs that do not keep their lines in the
// caches: the next kernel reads them back from memory anyway, and with the
// lines not left behind the following kernel's reads run faster (k_qeos
// 229 -> 214 us at s=192 with k_kin's outputs so written, and k_kin 289 ->
// 285 us with k_qeos' outputs; k_nodal's node stores so written make
// k_nodal slower, 670 -> 685 us, and stay ordinary).  k_qeos reads its own
// element's values the same way (they are used once): k_qeos 213 -> 205 us,
// and k_dt, whose partial minima then stay in cache, 13 -> 10 us.
template <typename T>
LDEV T ld_nt(const T *p) {
  return __builtin_nontemporal_load(p);
}
template <typename T>
LDEV void st_nt(T *base, uint32_t boff, T v) {
  __builtin_nontemporal_store(v, (T *)((char *)base + boff));
}
template <typename T>
LDEV void st_nt(T *p, T v) {
  __builtin_nontemporal_store(v, p);
}

// instance i's view of the flat field arrays
LDEV DevFields slice(const DevFields &f, const InstDesc &I) {
  DevFields g;
  const int64_t a = I.e0, b = I.n0;
  g.e = f.e + a; g.p = f.p + a; g.q = f.q + a; g.ql = f.ql + a; g.qq = f.qq + a;
  g.v = f.v + a; g.volo = f.volo + a; g.delv = f.delv + a; g.vdov = f.vdov + a;
  g.arealg = f.arealg + a; g.ss = f.ss + a; g.elemMass = f.elemMass + a; g.vnew = f.vnew + a;
  g.delv_xi = f.delv_xi + a; g.delv_eta = f.delv_eta + a; g.delv_zeta = f.delv_zeta + a;
  g.delx_xi = f.delx_xi + a; g.delx_eta = f.delx_eta + a; g.delx_zeta = f.delx_zeta + a;
  g.dxx = f.dxx + a; g.dyy = f.dyy + a; g.dzz = f.dzz + a; g.elemBC = f.elemBC + a;
  for (int k = 0; k < 2; ++k) {
    g.ns[k].x = f.ns[k].x + b; g.ns[k].y = f.ns[k].y + b; g.ns[k].z = f.ns[k].z + b;
    g.ns[k].xd = f.ns[k].xd + b; g.ns[k].yd = f.ns[k].yd + b; g.ns[k].zd = f.ns[k].zd + b;
  }
  g.xdd = f.xdd + b; g.ydd = f.ydd + b; g.zdd = f.zdd + b;
  g.fx = f.fx + b; g.fy = f.fy + b; g.fz = f.fz + b; g.nodalMass = f.nodalMass + b;
  return g;
}

// NodeSet k of a sliced field view, chosen without indexing a local array
// by a run-time value (that would put the struct in scratch memory)
LDEV NodeSet pick(const DevFields &f, int k) {
  NodeSet r;
  r.x = k ? f.ns[1].x : f.ns[0].x;
  r.y = k ? f.ns[1].y : f.ns[0].y;
  r.z = k ? f.ns[1].z : f.ns[0].z;
  r.xd = k ? f.ns[1].xd : f.ns[0].xd;
  r.yd = k ? f.ns[1].yd : f.ns[0].yd;
  r.zd = k ? f.ns[1].zd : f.ns[0].zd;
  return r;
}

// Per-block instance lookup for the element-, node- and k_nodal-shaped
// grids: every block belongs to exactly one instance (instances start on a
// block boundary), so the lookup is uniform over the block.  k_kin and
// k_qeos take their blocks from host-built tables (kin_table, qeos_table:
// XCD-aware orders that leave finished instances out).
// k_kin (no barriers, no LDS, three waves per SIMD by its registers) runs
// one-wave workgroups of KIN_CHUNK elements: a finished wave's slot is
// refilled without waiting for a whole 4-wave workgroup to fit (at s=192:
// 0.292 against 0.304 ms per launch)
static_assert(KIN_CHUNK == 64, "k_kin workgroups are one wave");

#define LULESH_ELEM_PREAMBLE(B)                                                  \
  const int inst_ = (B).eblock_inst[blockIdx.x];                               \
  const InstDesc I = (B).inst[inst_];                                          \
  DevScalars *sc = (B).sc + inst_;                                             \
  const int s = I.s;                                                           \
  const int64_t E = I.E;                                                       \
  const int64_t e = (int64_t)(blockIdx.x - I.eb0) * BLOCK + threadIdx.x;       \
  const DevFields f = slice((B).f, I);
#define LULESH_ELEM_PREAMBLE_T(B, TABLE)                                         \
  const int vb_ = (B).TABLE[blockIdx.x];                                       \
  if (vb_ < 0) return;                                                         \
  const int inst_ = (B).eblock_inst[vb_];                                      \
  const InstDesc I = (B).inst[inst_];                                          \
  DevScalars *sc = (B).sc + inst_;                                             \
  const int s = I.s;                                                           \
  const int64_t E = I.E;                                                       \
  const int64_t e = (int64_t)(vb_ - I.eb0) * BLOCK + threadIdx.x;              \
  const DevFields f = slice((B).f, I);
#define LULESH_NODE_PREAMBLE(B)                                                  \
  const int inst_ = (B).nblock_inst[blockIdx.x];                               \
  const InstDesc I = (B).inst[inst_];                                          \
  DevScalars *sc = (B).sc + inst_;                                             \
  const int s = I.s;                                                           \
  const int64_t N = I.N;                                                       \
  const int64_t n = (int64_t)(blockIdx.x - I.nb0) * BLOCK + threadIdx.x;       \
  const DevFields f = slice((B).f, I);

// timeIncrement!, src/domain.jl:920-961
LDEV void time_increment(DevScalars *sc) {
  double targetdt = sc->stoptime - sc->time;
  if (sc->dtfixed <= 0.0 && sc->cycle != 0) {
    const double olddt = sc->deltatime;
    double gnewdt = __builtin_huge_val();
    if (sc->dtcourant < gnewdt) gnewdt = sc->dtcourant / 2.0;
    if (sc->dthydro < gnewdt) gnewdt = (sc->dthydro * 2.0) / 3.0;
    double newdt = gnewdt;
    const double ratio = newdt / olddt;
    if (ratio >= 1.0) {
      if (ratio < sc->deltatimemultlb) newdt = olddt;
      else if (ratio > sc->deltatimemultub) newdt = olddt * sc->deltatimemultub;
    }
    newdt = jmin(newdt, sc->dtmax);
    sc->deltatime = newdt;
  }
  if (sc->deltatime < targetdt && targetdt < (4.0 * sc->deltatime) / 3.0)
    targetdt = (2.0 * sc->deltatime) / 3.0;
  if (targetdt < sc->deltatime) sc->deltatime = targetdt;
  sc->time = sc->time + sc->deltatime;
  sc->cycle = sc->cycle + 1;
}

// ---------------------------------------------------------------------------
// Domain(prob) on the device (src/domain.jl:117-174, 615-771, 822-916)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_init_nodes(Batch B) {
  LULESH_NODE_PREAMBLE(B)
  (void)sc;
  if (n >= N) return;
  const int64_t s1 = s + 1;
  const int64_t col = n % s1, rest = n / s1;
  const int64_t row = rest % s1, plane = rest / s1;
  // tx = 1.125*(colLoc*nx + col)/meshEdgeElems with colLoc = 0, tp = 1
  const NodeSet &S = f.ns[0];
  S.x[n] = (1.125 * (double)col) / (double)s;
  S.y[n] = (1.125 * (double)row) / (double)s;
  S.z[n] = (1.125 * (double)plane) / (double)s;
  S.xd[n] = 0.0; S.yd[n] = 0.0; S.zd[n] = 0.0;
  f.xdd[n] = 0.0; f.ydd[n] = 0.0; f.zdd[n] = 0.0;
  f.fx[n] = 0.0; f.fy[n] = 0.0; f.fz[n] = 0.0;
}

__global__ __launch_bounds__(BLOCK) void k_init_elems(Batch B) {
  LULESH_ELEM_PREAMBLE(B)
  (void)sc;
  if (e >= E) return;
  int64_t nid[8];
  elem_nodes(e, s, nid);
  double X[8], Y[8], Z[8];
  gather8(f.ns[0].x, nid, X); gather8(f.ns[0].y, nid, Y); gather8(f.ns[0].z, nid, Z);
  const double vol = elem_volume(X, Y, Z);
  f.volo[e] = vol;
  f.elemMass[e] = vol;
  f.v[e] = 1.0;
  f.e[e] = 0.0; f.p[e] = 0.0; f.q[e] = 0.0; f.ss[e] = 0.0;
  f.ql[e] = 0.0; f.qq[e] = 0.0; f.delv[e] = 0.0; f.vdov[e] = 0.0; f.arealg[e] = 0.0;
  f.vnew[e] = 0.0; f.dxx[e] = 0.0; f.dyy[e] = 0.0; f.dzz[e] = 0.0;
  f.delv_xi[e] = 0.0; f.delv_eta[e] = 0.0; f.delv_zeta[e] = 0.0;
  f.delx_xi[e] = 0.0; f.delx_eta[e] = 0.0; f.delx_zeta[e] = 0.0;
  // setupBoundaryConditions for the single rank: SYMM on min faces, FREE on max
  const int64_t col = e % s, rest = e / s;
  const int64_t row = rest % s, plane = rest / s;
  uint32_t bc = 0;
  if (plane == 0) bc |= ZETA_M_SYMM;
  if (plane == s - 1) bc |= ZETA_P_FREE;
  if (row == 0) bc |= ETA_M_SYMM;
  if (row == s - 1) bc |= ETA_P_FREE;
  if (col == 0) bc |= XI_M_SYMM;
  if (col == s - 1) bc |= XI_P_FREE;
  f.elemBC[e] = bc;
}

// nodal mass: volume/8 of each adjacent element added in ascending element
// order, the order of the element loop at src/domain.jl:709-727
__global__ __launch_bounds__(BLOCK) void k_init_mass(Batch B) {
  LULESH_NODE_PREAMBLE(B)
  (void)sc;
  if (n >= N) return;
  const int64_t s1 = s + 1;
  const int i = (int)(n % s1), j = (int)((n / s1) % s1), k = (int)(n / (s1 * s1));
  double m = 0.0;
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
        m = m + f.volo[e] / 8.0;
      }
    }
  }
  f.nodalMass[n] = m;
}

// scalars of src/domain.jl:663-701, energy deposit and dt0 (738-749)
// one thread per instance; the deposit is einit_scale * einit (1 for the
// reference problem) and dt0 is computed from it
__global__ void k_init_scalars(Batch B) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= B.n) return;
  const InstDesc I = B.inst[i];
  DevScalars *sc = B.sc + i;
  const int s = I.s;
  double *e0 = B.f.e + I.e0;
  const double *volo = B.f.volo + I.e0;
  sc->dthydro = 1e20;
  sc->dtcourant = 1e20;
  sc->time = 0.;
  sc->dtfixed = -1.0e-6;
  sc->deltatimemultlb = 1.1;
  sc->deltatimemultub = 1.2;
  sc->stoptime = 1.0e-2;
  sc->dtmax = 1.0e-2;
  sc->cycle = 0;
  sc->e_cut = 1.0e-7;
  sc->p_cut = 1.0e-7;
  sc->q_cut = 1.0e-7;
  sc->u_cut = 1.0e-7;
  sc->v_cut = 1.0e-10;
  sc->hgcoef = I.hgcoef;
  sc->ss4o3 = 4.0 / 3.0;
  sc->qstop = 1.0e+12;
  sc->monoq_max_slope = 1.0;
  sc->monoq_limiter_mult = 2.0;
  sc->qlc_monoq = 0.5;
  sc->qqc_monoq = 2.0 / 3.0;
  sc->qqc = 2.0;
  sc->pmin = 0.;
  sc->emin = -1.0e+15;
  sc->dvovmax = 0.1;
  sc->eosvmax = 1.0e+9;
  sc->eosvmin = 1.0e-9;
  sc->refdens = 1.0;
  const double ebase = 3.948746e+7;
  const double scale = (double)s / 45.0;
  const double einit = (((ebase * scale) * scale) * scale) * I.einit_scale;
  e0[0] = einit;
  sc->deltatime = (.5 * cbrt_cr(volo[0])) / sqrt(2.0 * einit);
  sc->err = ERR_NONE;
  sc->active = 0;
  sc->num_iters = 0;
  sc->parity = 0;
  sc->nodal_done = 0;
}

// ---------------------------------------------------------------------------
// the cycle
// ---------------------------------------------------------------------------
__global__ void k_time_increment(Batch B) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= B.n) return;
  DevScalars *sc = B.sc + i;
  time_increment(sc);
  if (i == 0 && B.dt_record) B.dt_record[sc->cycle - 1 - B.dt_base] = sc->deltatime;
}

// driver loop head (examples/benchmark.jl:82-84): the first iteration is
// entered whenever time < stoptime
__global__ void k_run_begin(Batch B) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= B.n) return;
  DevScalars *sc = B.sc + i;
  sc->num_iters = B.num_iters;
  sc->err = ERR_NONE;
  const int go = sc->time < sc->stoptime;
  sc->active = go;
  if (go) time_increment(sc);
}

// ---------------------------------------------------------------------------
// lagrangeNodal (src/domain.jl:1840-1872) as one kernel.
//
// A workgroup owns a column of nodes (up to etx x ety; 16 x 14 for inner
// tiles at s=192, see nodal_tile) and a z-chunk of its node planes (the
// instance's NodalCut: cz0 planes for the first chunk, cz for the others,
// the last one up to node plane s).  It marches
// up the chunk one element plane at a time: its threads compute the corner
// forces of the etx x ety elements around the node tile
// (integrateStressForElems + calcFBHourglassForceForElems arithmetic), stage
// them in LDS, and the node threads fold them into per-node running sums.
// A node's sum takes its elements in ascending element id -- the plane below
// first, then its own plane, each in (row, col) order -- which is exactly
// the nodeElemCornerList order of the reference's gathers
// (src/domain.jl:1112-1127, 1678-1693); stress and hourglass corners are
// summed separately and added at the end, as the reference does.  Finished
// node planes are updated at once (acceleration, symmetry BCs, velocity,
// position; src/domain.jl:1781-1838) into the other NodeSet, so tiles that
// still read the old positions of shared halo nodes never see new values.
// Elements on tile edges are computed by every tile that needs them
// (17x15x(cz+1) per 16x14xcz nodes for an inner 17 x 15 tile): no corner
// force ever goes to HBM.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_nodal(Batch B) {
  // XCD-aware order: the dispatcher deals workgroups round-robin over the 8
  // XCDs, and the host's table gives each XCD a contiguous run of blocks
  // (neighbouring tiles, whose halos then share its L2; -8 % at s=192)
  const int vb = B.gblock_order[blockIdx.x];
  if (vb < 0) return;
  const int inst_ = B.gblock_inst[vb];
  const InstDesc I = B.inst[inst_];
  DevScalars *sc = B.sc + inst_;
  if (B.gated && !sc->active) return;
  const int s = I.s, quirks = B.quirks, keep_aux = B.keep_aux;
  const int etx = I.etx, ety = I.ety;
  const DevFields f = slice(B.f, I);
  const int lb = vb - I.gb0;
  const int tx = lb % I.tx, ty = (lb / I.tx) % I.ty, chunk = lb / (I.tx * I.ty);
  // the tile's first node column (the first tile owns etx columns, inner
  // ones etx-1) and how far its first element lane sits before it (none
  // before node 0)
  const int i0 = tx == 0 ? 0 : etx + (tx - 1) * (etx - 1);
  const int j0 = ty == 0 ? 0 : ety + (ty - 1) * (ety - 1);
  const int oi = tx != 0, oj = ty != 0;
  const int iend = tx == I.tx - 1 ? s + 1 : i0 + etx - oi;
  const int jend = ty == I.ty - 1 ? s + 1 : j0 + ety - oj;
  const NodalCut cut{I.cz0, I.cz, I.nch};
  const int k0 = nodal_chunk_begin(cut, chunk);
  const int k1 = nodal_chunk_end(cut, chunk, s);
  const int t = threadIdx.x, ti = t % etx, tj = t / etx;
  const int s1 = s + 1;
  // this thread's element in each plane, and its node column (threads past
  // the etx x ety tile take part in the barriers only)
  const int ei = i0 - oi + ti, ej = j0 - oj + tj;
  const bool elem_col = tj < ety && ei < s && ej < s;
  const int ni = i0 + ti, nj = j0 + tj;
  const bool node_col = tj < ety && ni < iend && nj < jend;
  // the four elements around the node column, in ascending id order:
  // (ni-1,nj-1) (ni,nj-1) (ni-1,nj) (ni,nj) at lanes (ti+oi-1, tj+oj-1) ..
  // (ti+oi, tj+oj); each is read only where it exists, which keeps every
  // read inside the tile (a node column the tile owns has its elements
  // among the tile's lanes)
  const bool h00 = ni >= 1 && nj >= 1, h10 = ni < s && nj >= 1;
  const bool h01 = ni >= 1 && nj < s, h11 = ni < s && nj < s;
  const int l00 = (tj + oj - 1) * etx + (ti + oi - 1), l10 = l00 + 1, l01 = l00 + etx,
            l11 = l00 + etx + 1;
  // corner of the node in element (ox,oy) = node offset inside it:
  // (1,1)->2, (0,1)->3, (1,0)->1, (0,0)->0 at the bottom; +4 at the top

  const int par = sc->parity;
  const NodeSet in = pick(f, par), out = pick(f, par ^ 1);
  const double hourg = sc->hgcoef;
  const bool hg_on = hourg > 0.0;
  const double dt = sc->deltatime, u_cut = sc->u_cut;

  __shared__ double lds[24 * BLOCK];  // [corner][component][element slot]

  double cs0 = 0.0, cs1 = 0.0, cs2 = 0.0, ch0 = 0.0, ch1 = 0.0, ch2 = 0.0;  // node plane ek
  double ns0 = 0.0, ns1 = 0.0, ns2 = 0.0, nh0 = 0.0, nh1 = 0.0, nh2 = 0.0;  // node plane ek+1

  // Issue order.  A CU holds two k_nodal workgroups, one wave of each per
  // SIMD, and a SIMD issues the older wave first when both are ready, so a
  // workgroup that starts beside a running one falls behind it and marches
  // its last planes alone, at the issue rate of one wave.  Such a workgroup
  // (any after the first nodal_lead dealt to its XCD, one per CU) is issued
  // first for the first 55 % of its planes (50 % leaves it ~20 us behind
  // at s=192).
  const int lead_end = k0 - 1 + (11 * (k1 - k0 + 1)) / 20;
  if ((int)(blockIdx.x >> 3) >= B.nodal_lead) __builtin_amdgcn_s_setprio(1);

  for (int ek = k0 - 1; ek < k1; ++ek) {
    const bool plane = ek >= 0 && ek < s;
    if (ek == lead_end) __builtin_amdgcn_s_setprio(0);
    double hgf[8][3];
    if (plane && elem_col) {
      const int64_t e = ((int64_t)ek * s + ej) * s + ei;
      const uint32_t eo = (uint32_t)e * 8u;
      uint32_t nid[4];
      hex_pairs(((uint32_t)ek * (uint32_t)s1 + (uint32_t)ej) * (uint32_t)s1 + (uint32_t)ei,
                (uint32_t)s1, nid);
      double X[8], Y[8], Z[8];
      gather8p(in.x, nid, X); gather8p(in.y, nid, Y); gather8p(in.z, nid, Z);
      // stress corners, src/domain.jl:1097-1108 (sig from 963-969)
      const double detJ = jacobian_volume(X, Y, Z);
      if (detJ <= 0.0) raise(sc, ERR_EARLY_VOLUME);
      {
        double B[8][3];
        node_normals(X, Y, Z, quirks, B);
        const double ms = -((-at(f.p, eo)) - at(f.q, eo));
#pragma unroll
        for (int n = 0; n < 8; ++n)
#pragma unroll
          for (int c = 0; c < 3; ++c) lds[(n * 3 + c) * BLOCK + t] = ms * B[n][c];
      }
      // hourglass corners, src/domain.jl:1707-1729 and 1505-1639
      const double v = at(f.v, eo);
      const double determ = at(f.volo, eo) * v;
      if (v <= 0.0) raise(sc, ERR_HG_VOLUME);
      if (hg_on) {
        // Mode-major form of the same arithmetic: for each hourglass mode m
        // the column hourgam[.][m], its three projections h_c[m] and the
        // running sums sum_m hourgam[n][m] h_c[m] (m = 0..3 in order, as the
        // reference's 4-term sums fold) -- fewer values live at once.
        double dvdx[8], dvdy[8], dvdz[8];
        volume_derivative(X, Y, Z, quirks, dvdx, dvdy, dvdz);
        double hmx[4], hmy[4], hmz[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
          double hx = X[0] * gam(0, m), hy = Y[0] * gam(0, m), hz = Z[0] * gam(0, m);
#pragma unroll
          for (int n = 1; n < 8; ++n) {
            hx = hx + X[n] * gam(n, m);
            hy = hy + Y[n] * gam(n, m);
            hz = hz + Z[n] * gam(n, m);
          }
          hmx[m] = hx; hmy[m] = hy; hmz[m] = hz;
        }
        const double volinv = 1.0 / determ;
        double VX[8], VY[8], VZ[8];
        gather8p(in.xd, nid, VX); gather8p(in.yd, nid, VY); gather8p(in.zd, nid, VZ);
#pragma unroll
        for (int m = 0; m < 4; ++m) {
          double g[8];
#pragma unroll
          for (int n = 0; n < 8; ++n)
            g[n] = gam(n, m) - volinv * (((dvdx[n] * hmx[m]) + (dvdy[n] * hmy[m])) + (dvdz[n] * hmz[m]));
          double h0 = g[0] * VX[0], h1 = g[0] * VY[0], h2 = g[0] * VZ[0];
#pragma unroll
          for (int n = 1; n < 8; ++n) {
            h0 = h0 + g[n] * VX[n];
            h1 = h1 + g[n] * VY[n];
            h2 = h2 + g[n] * VZ[n];
          }
#pragma unroll
          for (int n = 0; n < 8; ++n) {
            if (m == 0) {
              hgf[n][0] = g[n] * h0;
              hgf[n][1] = g[n] * h1;
              hgf[n][2] = g[n] * h2;
            } else {
              hgf[n][0] = hgf[n][0] + g[n] * h0;
              hgf[n][1] = hgf[n][1] + g[n] * h1;
              hgf[n][2] = hgf[n][2] + g[n] * h2;
            }
          }
        }
        const double volume13 = cbrt_cr(determ);
        const double coefficient = ((((-hourg) * 0.01) * at(f.ss, eo)) * at(f.elemMass, eo)) / volume13;
#pragma unroll
        for (int n = 0; n < 8; ++n) {
          hgf[n][0] = coefficient * hgf[n][0];
          hgf[n][1] = coefficient * hgf[n][1];
          hgf[n][2] = coefficient * hgf[n][2];
        }
      }
    }
    __syncthreads();
    // stress: finish node plane ek with the bottom corners of plane ek, start
    // node plane ek+1 with its top corners (plane ek precedes plane ek+1)
    if (node_col && plane) {
      if (ek >= k0) {
#define LULESH_ACC(flag, slot, corner)                                  \
  if (flag) {                                                           \
    cs0 = cs0 + lds[((corner) * 3 + 0) * BLOCK + (slot)];               \
    cs1 = cs1 + lds[((corner) * 3 + 1) * BLOCK + (slot)];               \
    cs2 = cs2 + lds[((corner) * 3 + 2) * BLOCK + (slot)];               \
  }
        LULESH_ACC(h00, l00, 2) LULESH_ACC(h10, l10, 3) LULESH_ACC(h01, l01, 1)
        LULESH_ACC(h11, l11, 0)
#undef LULESH_ACC
      }
      if (ek + 1 < k1) {
#define LULESH_ACC(flag, slot, corner)                                  \
  if (flag) {                                                           \
    ns0 = ns0 + lds[((corner) * 3 + 0) * BLOCK + (slot)];               \
    ns1 = ns1 + lds[((corner) * 3 + 1) * BLOCK + (slot)];               \
    ns2 = ns2 + lds[((corner) * 3 + 2) * BLOCK + (slot)];               \
  }
        LULESH_ACC(h00, l00, 6) LULESH_ACC(h10, l10, 7) LULESH_ACC(h01, l01, 5)
        LULESH_ACC(h11, l11, 4)
#undef LULESH_ACC
      }
    }
    __syncthreads();
    if (hg_on) {
      if (plane && elem_col) {
#pragma unroll
        for (int n = 0; n < 8; ++n)
#pragma unroll
          for (int c = 0; c < 3; ++c) lds[(n * 3 + c) * BLOCK + t] = hgf[n][c];
      }
      __syncthreads();
      if (node_col && plane) {
        if (ek >= k0) {
#define LULESH_ACC(flag, slot, corner)                                  \
  if (flag) {                                                           \
    ch0 = ch0 + lds[((corner) * 3 + 0) * BLOCK + (slot)];               \
    ch1 = ch1 + lds[((corner) * 3 + 1) * BLOCK + (slot)];               \
    ch2 = ch2 + lds[((corner) * 3 + 2) * BLOCK + (slot)];               \
  }
          LULESH_ACC(h00, l00, 2) LULESH_ACC(h10, l10, 3) LULESH_ACC(h01, l01, 1)
          LULESH_ACC(h11, l11, 0)
#undef LULESH_ACC
        }
        if (ek + 1 < k1) {
#define LULESH_ACC(flag, slot, corner)                                  \
  if (flag) {                                                           \
    nh0 = nh0 + lds[((corner) * 3 + 0) * BLOCK + (slot)];               \
    nh1 = nh1 + lds[((corner) * 3 + 1) * BLOCK + (slot)];               \
    nh2 = nh2 + lds[((corner) * 3 + 2) * BLOCK + (slot)];               \
  }
          LULESH_ACC(h00, l00, 6) LULESH_ACC(h10, l10, 7) LULESH_ACC(h01, l01, 5)
          LULESH_ACC(h11, l11, 4)
#undef LULESH_ACC
        }
      }
      __syncthreads();
    }
    // node plane ek is complete: calcAccelerationForNodes ..
    // calcPositionForNodes for it
    if (node_col && ek >= k0) {
      const uint32_t no = (((uint32_t)ek * (uint32_t)s1 + (uint32_t)nj) * (uint32_t)s1 + (uint32_t)ni) * 8u;
      double fx = cs0, fy = cs1, fz = cs2;
      if (hg_on) {
        fx = fx + ch0;
        fy = fy + ch1;
        fz = fz + ch2;
      }
      const double mass = at(f.nodalMass, no);
      double xdd = fx / mass, ydd = fy / mass, zdd = fz / mass;
      // symmetry planes: the col 0, row 0 and plane 0 node sets
      // (src/domain.jl:773-791, 1787-1806)
      if (ni == 0) xdd = 0.0;
      if (nj == 0) ydd = 0.0;
      if (ek == 0) zdd = 0.0;
      double xd = at(in.xd, no) + xdd * dt;
      if (fabs(xd) < u_cut) xd = 0.0;
      double yd = at(in.yd, no) + ydd * dt;
      if (fabs(yd) < u_cut) yd = 0.0;
      double zd = at(in.zd, no) + zdd * dt;
      if (fabs(zd) < u_cut) zd = 0.0;
      at(out.xd, no) = xd;
      at(out.yd, no) = yd;
      at(out.zd, no) = zd;
      at(out.x, no) = at(in.x, no) + xd * dt;
      at(out.y, no) = at(in.y, no) + yd * dt;
      at(out.z, no) = at(in.z, no) + zd * dt;
      if (keep_aux) {
        at(f.fx, no) = fx; at(f.fy, no) = fy; at(f.fz, no) = fz;
        at(f.xdd, no) = xdd; at(f.ydd, no) = ydd; at(f.zdd, no) = zdd;
      }
    }
    cs0 = ns0; cs1 = ns1; cs2 = ns2;
    ch0 = nh0; ch1 = nh1; ch2 = nh2;
    ns0 = ns1 = ns2 = nh0 = nh1 = nh2 = 0.0;
  }
  // the instance's last workgroup to finish makes the node set just written
  // current: every workgroup of the instance has then started (and read
  // the old parity), and the next kernel sees the new one (a kernel
  // boundary orders the stores; no fence is needed inside this launch)
  if (B.flip) {
    __syncthreads();
    if (t == 0) {
      if (atomicAdd(&sc->nodal_done, 1) == I.gbn - 1) {
        sc->nodal_done = 0;
        sc->parity = par ^ 1;
      }
    }
  }
}

// calcKinematicsForElems (src/domain.jl:2099-2147), the deviatoric loop
// (2156-2171) and calcMonotonicQGradientsForElems (2175-2320)
__global__ __launch_bounds__(KIN_CHUNK) void k_kin(Batch B) {
  // the chunk of one element plane this workgroup runs: kin_table bands
  // the planes so that an XCD's consecutive chunks climb one band of the
  // domain plane by plane and share node rows in its L2
  const KinWork w = B.kin_work[blockIdx.x];
  if (w.inst < 0) return;
  const InstDesc I = B.inst[w.inst];
  DevScalars *sc = B.sc + w.inst;
  if (B.gated && !sc->active) return;
  const int s = I.s;
  const uint32_t plane = (uint32_t)s * (uint32_t)s;
  const uint32_t plane_end = ((uint32_t)w.start / plane + 1u) * plane;
  const int64_t e = (int64_t)w.start + threadIdx.x;
  const DevFields f = slice(B.f, I);
  if (e >= (int64_t)plane_end) return;
  const int keep_aux = B.keep_aux;
  uint32_t nid[4];
  {
    const uint32_t ue = (uint32_t)e, us = (uint32_t)s, s1 = us + 1u;
    const uint32_t col = ue % us, rest = ue / us;
    hex_pairs(((rest / us) * s1 + rest % us) * s1 + col, s1, nid);
  }
  const uint32_t eo = (uint32_t)e * 8u;
  double X[8], Y[8], Z[8], XV[8], YV[8], ZV[8];
  const NodeSet S = pick(f, sc->parity);
  gather8p(S.x, nid, X); gather8p(S.y, nid, Y); gather8p(S.z, nid, Z);
  gather8p(S.xd, nid, XV); gather8p(S.yd, nid, YV); gather8p(S.zd, nid, ZV);
  const double volo = at(f.volo, eo);
  const double volume = elem_volume(X, Y, Z);
  const double vnew = volume / volo;
  if (vnew <= 0.0) raise(sc, ERR_KIN_VOLUME);
  st_nt(f.vnew, eo, vnew);
  st_nt(f.delv, eo, vnew - at(f.v, eo));
  st_nt(f.arealg, eo, char_length(X, Y, Z, volume));
  {
    const double dt2 = 0.5 * sc->deltatime;
    double XS[8], YS[8], ZS[8], b[4][3];
#pragma unroll
    for (int n = 0; n < 8; ++n) {
      XS[n] = X[n] - dt2 * XV[n];
      YS[n] = Y[n] - dt2 * YV[n];
      ZS[n] = Z[n] - dt2 * ZV[n];
    }
    const double detJ = shape_derivatives4(XS, YS, ZS, b);
    const double inv_detJ = 1.0 / detJ;
    const double d1 = inv_detJ * vgrad_term(b, 0, XV);
    const double d2 = inv_detJ * vgrad_term(b, 1, YV);
    const double d3 = inv_detJ * vgrad_term(b, 2, ZV);
    const double vdov = (d1 + d2) + d3;
    st_nt(f.vdov, eo, vdov);
    if (keep_aux) {
      const double vdovthird = vdov / 3.0;
      st_nt(f.dxx, eo, d1 - vdovthird);
      st_nt(f.dyy, eo, d2 - vdovthird);
      st_nt(f.dzz, eo, d3 - vdovthird);
    }
    if (vnew <= 0.0) raise(sc, ERR_LAG_VOLUME);
  }
  // MonoQ gradients
  const double ptiny = 1.e-36;
  const double vol = volo * vnew;
  const double norm = 1.0 / (vol + ptiny);
  const double dxj = -0.25 * (sum4(X[0], X[1], X[5], X[4]) - sum4(X[3], X[2], X[6], X[7]));
  const double dyj = -0.25 * (sum4(Y[0], Y[1], Y[5], Y[4]) - sum4(Y[3], Y[2], Y[6], Y[7]));
  const double dzj = -0.25 * (sum4(Z[0], Z[1], Z[5], Z[4]) - sum4(Z[3], Z[2], Z[6], Z[7]));
  const double dxi = 0.25 * (sum4(X[1], X[2], X[6], X[5]) - sum4(X[0], X[3], X[7], X[4]));
  const double dyi = 0.25 * (sum4(Y[1], Y[2], Y[6], Y[5]) - sum4(Y[0], Y[3], Y[7], Y[4]));
  const double dzi = 0.25 * (sum4(Z[1], Z[2], Z[6], Z[5]) - sum4(Z[0], Z[3], Z[7], Z[4]));
  const double dxk = 0.25 * (sum4(X[4], X[5], X[6], X[7]) - sum4(X[0], X[1], X[2], X[3]));
  const double dyk = 0.25 * (sum4(Y[4], Y[5], Y[6], Y[7]) - sum4(Y[0], Y[1], Y[2], Y[3]));
  const double dzk = 0.25 * (sum4(Z[4], Z[5], Z[6], Z[7]) - sum4(Z[0], Z[1], Z[2], Z[3]));
  double ax, ay, az, dxv, dyv, dzv;
  // zeta: i cross j
  ax = (dyi * dzj) - (dzi * dyj);
  ay = (dzi * dxj) - (dxi * dzj);
  az = (dxi * dyj) - (dyi * dxj);
  st_nt(f.delx_zeta, eo, vol / sqrt((((ax * ax) + (ay * ay)) + (az * az)) + ptiny));
  ax = ax * norm; ay = ay * norm; az = az * norm;
  dxv = 0.25 * (sum4(XV[4], XV[5], XV[6], XV[7]) - sum4(XV[0], XV[1], XV[2], XV[3]));
  dyv = 0.25 * (sum4(YV[4], YV[5], YV[6], YV[7]) - sum4(YV[0], YV[1], YV[2], YV[3]));
  dzv = 0.25 * (sum4(ZV[4], ZV[5], ZV[6], ZV[7]) - sum4(ZV[0], ZV[1], ZV[2], ZV[3]));
  st_nt(f.delv_zeta, eo, ((ax * dxv) + (ay * dyv)) + (az * dzv));
  // xi: j cross k
  ax = (dyj * dzk) - (dzj * dyk);
  ay = (dzj * dxk) - (dxj * dzk);
  az = (dxj * dyk) - (dyj * dxk);
  st_nt(f.delx_xi, eo, vol / sqrt((((ax * ax) + (ay * ay)) + (az * az)) + ptiny));
  ax = ax * norm; ay = ay * norm; az = az * norm;
  dxv = 0.25 * (sum4(XV[1], XV[2], XV[6], XV[5]) - sum4(XV[0], XV[3], XV[7], XV[4]));
  dyv = 0.25 * (sum4(YV[1], YV[2], YV[6], YV[5]) - sum4(YV[0], YV[3], YV[7], YV[4]));
  dzv = 0.25 * (sum4(ZV[1], ZV[2], ZV[6], ZV[5]) - sum4(ZV[0], ZV[3], ZV[7], ZV[4]));
  st_nt(f.delv_xi, eo, ((ax * dxv) + (ay * dyv)) + (az * dzv));
  // eta: k cross i
  ax = (dyk * dzi) - (dzk * dyi);
  ay = (dzk * dxi) - (dxk * dzi);
  az = (dxk * dyi) - (dyk * dxi);
  st_nt(f.delx_eta, eo, vol / sqrt((((ax * ax) + (ay * ay)) + (az * az)) + ptiny));
  ax = ax * norm; ay = ay * norm; az = az * norm;
  dxv = -0.25 * (sum4(XV[0], XV[1], XV[5], XV[4]) - sum4(XV[3], XV[2], XV[6], XV[7]));
  dyv = -0.25 * (sum4(YV[0], YV[1], YV[5], YV[4]) - sum4(YV[3], YV[2], YV[6], YV[7]));
  dzv = -0.25 * (sum4(ZV[0], ZV[1], ZV[5], ZV[4]) - sum4(ZV[3], ZV[2], ZV[6], ZV[7]));
  st_nt(f.delv_eta, eo, ((ax * dxv) + (ay * dyv)) + (az * dzv));
}

// neighbour ids exactly as setupElementConnectivities stores them
// (src/domain.jl:793-820), 0-based
LDEV int64_t nb_xim(int64_t e) { return e == 0 ? 1 : e - 1; }
LDEV int64_t nb_xip(int64_t e, int64_t E) { return e < E - 1 ? e + 1 : E - 1; }
LDEV int64_t nb_etam(int64_t e, int64_t s) { return e < s ? e : e - s; }
LDEV int64_t nb_etap(int64_t e, int64_t s, int64_t E) { return e < E - s ? e + s : e - (E - s); }
LDEV int64_t nb_zetam(int64_t e, int64_t s2) { return e < s2 ? e : e - s2; }
LDEV int64_t nb_zetap(int64_t e, int64_t s2, int64_t E) { return e < E - s2 ? e + s2 : e; }
// a read the reference would make out of bounds (only reachable with an
// uploaded elemBC that marks a boundary face interior, or s = 1) is clamped
LDEV int64_t clampi(int64_t i, int64_t E) { return i < 0 ? 0 : (i >= E ? E - 1 : i); }

// block-wide minimum of two values (64-wide wavefronts, BLOCK threads)
LDEV void block_min2(double &a, double &b) {
  __shared__ double sa[BLOCK / 64], sb[BLOCK / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double oa = __shfl_xor(a, off, 64);
    const double ob = __shfl_xor(b, off, 64);
    a = oa < a ? oa : a;
    b = ob < b ? ob : b;
  }
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) { sa[w] = a; sb[w] = b; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < BLOCK / 64; ++i) {
      a = sa[i] < a ? sa[i] : a;
      b = sb[i] < b ? sb[i] : b;
    }
  }
}

// calcMonotonicQRegionForElems (src/domain.jl:2322-2513, quirk Q3), the
// QStop check (2567-2582), applyMaterialPropertiesForElems + evalEOSForElems
// + calcEnergyForElems + calcSoundSpeedForElems (2585-2910, one element at a
// time; the reference's element-wise temporaries never mix elements),
// updateVolumesForElems (2912-2930) and the per-element terms of
// calcCourantConstraintForElems / calcHydroConstraintForElems (2964-3054)
template <bool COST>
__global__ __launch_bounds__(BLOCK) void k_qeos(Batch B) {
  LULESH_ELEM_PREAMBLE_T(B, qeos_order)
  if (B.gated && !sc->active) return;  // uniform over the block
  const int quirks = B.quirks, keep_aux = B.keep_aux;
  double *__restrict__ dtpart = B.dtpart;
  double cand_c = 1.0e+20, cand_h = 1.0e+20;
  if (e < E) {
    const double ptiny = 1e-36;
    const double mult = sc->monoq_limiter_mult, maxslope = sc->monoq_max_slope;
    const uint32_t bc = ld_nt(&f.elemBC[e]);
    const double dvxi = f.delv_xi[e], dveta = f.delv_eta[e], dvzeta = f.delv_zeta[e];
    bool bad = false;
    double delvm = 0.0, delvp = 0.0;
    uint32_t c;
    // xi
    c = bc & XI_M;
    if (c == 0 || c == XI_M_COMM) delvm = f.delv_xi[clampi(nb_xim(e), E)];
    else if (c == XI_M_SYMM) delvm = dvxi;
    else if (c == XI_M_FREE) delvm = 0.0;
    else { delvm = 0.0; bad = true; }
    c = bc & XI_P;
    if (c == 0 || c == XI_P_COMM) {
      const int64_t nb = nb_xip(e, E);  // stored 0-based id
      delvp = (quirks & Q3) ? f.delv_xi[clampi(nb - 1, E)] : f.delv_xi[clampi(nb, E)];
    } else if (c == XI_P_SYMM) delvp = dvxi;
    else if (c == XI_P_FREE) delvp = 0.0;
    else { delvp = 0.0; bad = true; }
    const double phixi = limit_phi(delvm, delvp, 1.0 / (dvxi + ptiny), mult, maxslope);
    // eta
    c = bc & ETA_M;
    if (c == 0 || c == ETA_M_COMM) delvm = f.delv_eta[clampi(nb_etam(e, s), E)];
    else if (c == ETA_M_SYMM) delvm = dveta;
    else if (c == ETA_M_FREE) delvm = 0.0;
    else { delvm = 0.0; bad = true; }
    c = bc & ETA_P;
    if (c == 0 || c == ETA_P_COMM) delvp = f.delv_eta[clampi(nb_etap(e, s, E), E)];
    else if (c == ETA_P_SYMM) delvp = dveta;
    else if (c == ETA_P_FREE) delvp = 0.0;
    else { delvp = 0.0; bad = true; }
    const double phieta = limit_phi(delvm, delvp, 1.0 / (dveta + ptiny), mult, maxslope);
    // zeta
    const int64_t s2 = (int64_t)s * s;
    c = bc & ZETA_M;
    if (c == 0 || c == ZETA_M_COMM) delvm = f.delv_zeta[clampi(nb_zetam(e, s2), E)];
    else if (c == ZETA_M_SYMM) delvm = dvzeta;
    else if (c == ZETA_M_FREE) delvm = 0.0;
    else { delvm = 0.0; bad = true; }
    c = bc & ZETA_P;
    if (c == 0 || c == ZETA_P_COMM) delvp = f.delv_zeta[clampi(nb_zetap(e, s2, E), E)];
    else if (c == ZETA_P_SYMM) delvp = dvzeta;
    else if (c == ZETA_P_FREE) delvp = 0.0;
    else { delvp = 0.0; bad = true; }
    const double phizeta = limit_phi(delvm, delvp, 1.0 / (dvzeta + ptiny), mult, maxslope);
    if (bad) raise(sc, ERR_BC);

    const double vdov = ld_nt(&f.vdov[e]), volo = ld_nt(&f.volo[e]), vnew = ld_nt(&f.vnew[e]);
    double qlin, qquad;
    if (vdov > 0.0) {
      qlin = 0.0;
      qquad = 0.0;
    } else {
      double delvxxi = dvxi * ld_nt(&f.delx_xi[e]);
      double delvxeta = dveta * ld_nt(&f.delx_eta[e]);
      double delvxzeta = dvzeta * ld_nt(&f.delx_zeta[e]);
      if (delvxxi > 0.0) delvxxi = 0.0;
      if (delvxeta > 0.0) delvxeta = 0.0;
      if (delvxzeta > 0.0) delvxzeta = 0.0;
      const double rho = ld_nt(&f.elemMass[e]) / (volo * vnew);
      qlin = ((-sc->qlc_monoq) * rho) *
             (((delvxxi * (1.0 - phixi)) + (delvxeta * (1.0 - phieta))) + (delvxzeta * (1.0 - phizeta)));
      qquad = (sc->qqc_monoq * rho) *
              ((((delvxxi * delvxxi) * (1.0 - (phixi * phixi))) +
                ((delvxeta * delvxeta) * (1.0 - (phieta * phieta)))) +
               ((delvxzeta * delvxzeta) * (1.0 - (phizeta * phizeta))));
    }
    if (keep_aux) {
      st_nt(&f.qq[e], qquad);
      st_nt(&f.ql[e], qlin);
    }

    // QStop on the q of the previous cycle
    double q_old = ld_nt(&f.q[e]);
    if (q_old > sc->qstop) raise(sc, ERR_QSTOP);

    // material properties
    const double eosvmin = sc->eosvmin, eosvmax = sc->eosvmax;
    double vc = ld_nt(&f.v[e]);
    if (eosvmin != 0.0 && vc < eosvmin) vc = eosvmin;
    if (eosvmax != 0.0 && vc > eosvmax) vc = eosvmax;
    if (vc <= 0.0) raise(sc, ERR_EOS_VOLUME);
    double vnewc = vnew;
    if (eosvmin != 0.0 && vnewc < eosvmin) vnewc = eosvmin;
    if (eosvmax != 0.0 && vnewc > eosvmax) vnewc = eosvmax;

    const double delvc_in = ld_nt(&f.delv[e]), e_old_in = ld_nt(&f.e[e]), p_old_in = ld_nt(&f.p[e]);
    const double TINY1 = 0.111111e-36, TINY3 = 0.333333e-18, SIXTH = 1.0 / 6.0;
    const double pmin = sc->pmin, p_cut = sc->p_cut, e_cut = sc->e_cut, q_cut = sc->q_cut;
    const double emin = sc->emin, rho0 = sc->refdens;
    // LULESH-2.0 cost mode: the energy update of an element is evaluated as
    // many times as its region's regReps says (the same inputs each time, so
    // the same result); the empty asm statements make every evaluation real
    int rep = 1;
    if constexpr (COST) rep = B.elem_rep[I.e0 + e];
    double p_new = 0.0, e_new = 0.0, q_new = 0.0, bvc = 0.0, pbvc = 0.0;
    for (int rr = 0; rr < rep; ++rr) {
    double delvc = delvc_in, e_old = e_old_in, p_old = p_old_in, q_o = q_old, vn = vnewc,
           ql_r = qlin, qq_r = qquad;
    if constexpr (COST) {
      __asm__ volatile("" : "+v"(delvc), "+v"(e_old), "+v"(p_old), "+v"(q_o));
      __asm__ volatile("" : "+v"(vn), "+v"(ql_r), "+v"(qq_r));
    }
    {
    const double q_old = q_o, vnewc = vn, qlin = ql_r, qquad = qq_r;
    double compression = 1.0 / vnewc - 1.0;
    const double vchalf = vnewc - delvc * 0.5;
    double compHalfStep = (1.0 / vchalf) - 1.0;
    if (eosvmin != 0.0 && vnewc <= eosvmin) compHalfStep = compression;
    if (eosvmax != 0.0 && vnewc >= eosvmax) {
      p_old = 0.0;
      compression = 0.0;
      compHalfStep = 0.0;
    }
    const double work = 0.0;
    e_new = (e_old - ((0.5 * delvc) * (p_old + q_old))) + (0.5 * work);
    if (e_new < emin) e_new = emin;
    const double pHalfStep = pressure(bvc, pbvc, e_new, compHalfStep, vnewc, pmin, p_cut, eosvmax);
    const double vhalf = 1.0 / (1.0 + compHalfStep);
    if (delvc > 0.0) {
      q_new = 0.0;
    } else {
      double ssc = ((pbvc * e_new) + (((vhalf * vhalf) * bvc) * pHalfStep)) / rho0;
      ssc = (ssc <= TINY1) ? TINY3 : sqrt(ssc);
      q_new = (ssc * qlin) + qquad;
    }
    e_new = e_new + ((0.5 * delvc) * ((3.0 * (p_old + q_old)) - (4.0 * (pHalfStep + q_new))));
    e_new = e_new + 0.5 * work;
    if (fabs(e_new) < e_cut) e_new = 0.0;
    if (e_new < emin) e_new = emin;
    p_new = pressure(bvc, pbvc, e_new, compression, vnewc, pmin, p_cut, eosvmax);
    double q_tilde;
    if (delvc > 0.0) {
      q_tilde = 0.0;
    } else {
      double ssc = ((pbvc * e_new) + (((vnewc * vnewc) * bvc) * p_new)) / rho0;
      ssc = (ssc <= TINY1) ? TINY3 : sqrt(ssc);
      q_tilde = (ssc * qlin) + qquad;
    }
    e_new = e_new - (((((7.0 * (p_old + q_old)) - (8.0 * (pHalfStep + q_new))) + (p_new + q_tilde)) *
                      delvc) *
                     SIXTH);
    if (fabs(e_new) < e_cut) e_new = 0.0;
    if (e_new < emin) e_new = emin;
    p_new = pressure(bvc, pbvc, e_new, compression, vnewc, pmin, p_cut, eosvmax);
    if (delvc <= 0.0) {
      double ssc = ((pbvc * e_new) + (((vnewc * vnewc) * bvc) * p_new)) / rho0;
      ssc = (ssc <= TINY1) ? TINY3 : sqrt(ssc);
      q_new = (ssc * qlin) + qquad;
      if (fabs(q_new) < q_cut) q_new = 0.0;
    }
    }
    if constexpr (COST) {
      __asm__ volatile("" ::"v"(p_new), "v"(e_new), "v"(q_new));
      __asm__ volatile("" ::"v"(bvc), "v"(pbvc));
    }
    }
    double ss = ((pbvc * e_new) + (((vnewc * vnewc) * bvc) * p_new)) / rho0;
    ss = (ss <= TINY1) ? TINY3 : sqrt(ss);
    st_nt(&f.p[e], p_new);
    st_nt(&f.e[e], e_new);
    st_nt(&f.q[e], q_new);
    st_nt(&f.ss[e], ss);

    // updateVolumesForElems
    double tmpV = vnew;
    if (fabs(tmpV - 1.0) < sc->v_cut) tmpV = 1.0;
    st_nt(&f.v[e], tmpV);
    if (tmpV <= 0.0) raise(sc, ERR_UPDATE_VOLUME);

    // time-constraint terms; an element only counts where vdov != 0 and its
    // value beats the 1e20 start value (strict <, as in the reference)
    if (vdov != 0.0) {
      const double qqc = sc->qqc;
      const double qqc2 = (64.0 * qqc) * qqc;
      const double arealg = ld_nt(&f.arealg[e]);
      double dtf = ss * ss;
      if (vdov < 0.0) dtf = dtf + ((((qqc2 * arealg) * arealg) * vdov) * vdov);
      dtf = sqrt(dtf);
      dtf = arealg / dtf;
      if (dtf < cand_c) cand_c = dtf;
      const double dtdvov = sc->dvovmax / (fabs(vdov) + 1.e-20);
      if (dtdvov < cand_h) cand_h = dtdvov;
    }
  }
  block_min2(cand_c, cand_h);
  if (threadIdx.x == 0) {
    dtpart[2 * vb_] = cand_c;
    dtpart[2 * vb_ + 1] = cand_h;
  }
}

// per-element time-constraint terms from the fields alone, for a
// calcTimeConstraintsForElems call that does not follow k_qeos
__global__ __launch_bounds__(BLOCK) void k_dtpart(Batch B) {
  LULESH_ELEM_PREAMBLE(B)
  (void)s;
  double *__restrict__ dtpart = B.dtpart;
  double cand_c = 1.0e+20, cand_h = 1.0e+20;
  if (e < E) {
    const double vdov = f.vdov[e];
    if (vdov != 0.0) {
      const double qqc2 = (64.0 * sc->qqc) * sc->qqc;
      const double arealg = f.arealg[e], ss = f.ss[e];
      double dtf = ss * ss;
      if (vdov < 0.0) dtf = dtf + ((((qqc2 * arealg) * arealg) * vdov) * vdov);
      dtf = sqrt(dtf);
      dtf = arealg / dtf;
      if (dtf < cand_c) cand_c = dtf;
      const double dtdvov = sc->dvovmax / (fabs(vdov) + 1.e-20);
      if (dtdvov < cand_h) cand_h = dtdvov;
    }
  }
  block_min2(cand_c, cand_h);
  if (threadIdx.x == 0) {
    dtpart[2 * blockIdx.x] = cand_c;
    dtpart[2 * blockIdx.x + 1] = cand_h;
  }
}

// final minima; in driver-loop mode (loop_mode 1) also the loop test of
// examples/benchmark.jl:82,95 and the timeIncrement! of the next cycle;
// with loop_mode 2 the next cycle's timeIncrement! unconditionally, its dt
// recorded as k_time_increment records it (a fixed number of cycles)
__global__ __launch_bounds__(1024) void k_dt(Batch B) {
  const InstDesc I = B.inst[blockIdx.x];
  DevScalars *sc = B.sc + blockIdx.x;
  if (B.gated && !sc->active) return;
  const int loop_mode = B.loop_mode;
  // (courant, hydro) pairs, 16-byte aligned; four independent loads in
  // flight per thread (a minimum does not depend on the order it is taken in)
  const double2 *__restrict__ dtpart = (const double2 *)(B.dtpart + 2 * (int64_t)I.eb0);
  const int64_t nb = I.ebn, bd = blockDim.x;
  double a = 1.0e+20, b = 1.0e+20, a1 = 1.0e+20, b1 = 1.0e+20;
  int64_t i = threadIdx.x;
  for (; i + 3 * bd < nb; i += 4 * bd) {
    const double2 p0 = dtpart[i], p1 = dtpart[i + bd], p2 = dtpart[i + 2 * bd],
                  p3 = dtpart[i + 3 * bd];
    a = p0.x < a ? p0.x : a;
    b = p0.y < b ? p0.y : b;
    a1 = p1.x < a1 ? p1.x : a1;
    b1 = p1.y < b1 ? p1.y : b1;
    a = p2.x < a ? p2.x : a;
    b = p2.y < b ? p2.y : b;
    a1 = p3.x < a1 ? p3.x : a1;
    b1 = p3.y < b1 ? p3.y : b1;
  }
  for (; i < nb; i += bd) {
    const double2 p0 = dtpart[i];
    a = p0.x < a ? p0.x : a;
    b = p0.y < b ? p0.y : b;
  }
  a = a1 < a ? a1 : a;
  b = b1 < b ? b1 : b;
  __shared__ double sa[16], sb[16];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double oa = __shfl_xor(a, off, 64);
    const double ob = __shfl_xor(b, off, 64);
    a = oa < a ? oa : a;
    b = ob < b ? ob : b;
  }
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) { sa[w] = a; sb[w] = b; }
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int i = 1; i < (int)(blockDim.x >> 6); ++i) {
    a = sa[i] < a ? sa[i] : a;
    b = sb[i] < b ? sb[i] : b;
  }
  if (a < 1.0e+20) sc->dtcourant = a;
  if (b < 1.0e+20) sc->dthydro = b;
  if (loop_mode == 1) {
    const int go = (sc->err == ERR_NONE) && (sc->cycle < sc->num_iters) && (sc->time < sc->stoptime);
    sc->active = go;
    if (go) time_increment(sc);
  } else if (loop_mode == 2) {
    time_increment(sc);
    if (blockIdx.x == 0 && B.dt_record) B.dt_record[sc->cycle - 1 - B.dt_base] = sc->deltatime;
  }
}

__global__ void k_math(int op, const double *__restrict__ a, const double *__restrict__ b,
                       double *__restrict__ out, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double x = a[i];
  out[i] = op == 0 ? cbrt_cr(x) : (op == 1 ? sqrt(x) : x / b[i]);
}

__global__ void k_clear_err(Batch B) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < B.n) B.sc[i].err = ERR_NONE;
}

}  // namespace lulesh

// ---------------------------------------------------------------------------
// launchers (called from the C ABI)
// ---------------------------------------------------------------------------
namespace lulesh {

static inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }
static inline unsigned per_inst_grid(int n) { return (unsigned)((n + 63) / 64); }

int nodal_blocks_per_cu() {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_nodal, BLOCK, 0) != hipSuccess || n < 1)
    n = 1;
  return n;
}

hipError_t launch_init(hipStream_t st, const Batch &b) {
  hipLaunchKernelGGL(k_init_nodes, dim3(b.nblocks), dim3(BLOCK), 0, st, b);
  hipLaunchKernelGGL(k_init_elems, dim3(b.eblocks), dim3(BLOCK), 0, st, b);
  hipLaunchKernelGGL(k_init_mass, dim3(b.nblocks), dim3(BLOCK), 0, st, b);
  hipLaunchKernelGGL(k_init_scalars, dim3(per_inst_grid(b.n)), dim3(64), 0, st, b);
  return hipGetLastError();
}

hipError_t launch_math(hipStream_t st, int op, const double *a, const double *b, double *out,
                       int64_t n) {
  hipLaunchKernelGGL(k_math, dim3(nblk(n, BLOCK)), dim3(BLOCK), 0, st, op, a, b, out, n);
  return hipGetLastError();
}

hipError_t launch_kernel(int id, hipStream_t st, const Batch &b) {
  switch (id) {
    case KID_NODAL:
      // glaunch workgroups, dealt to blocks by the XCD-aware order table
      hipLaunchKernelGGL(k_nodal, dim3(b.glaunch), dim3(BLOCK), 0, st, b);
      break;
    case KID_KIN:
      hipLaunchKernelGGL(k_kin, dim3(b.kin_launch), dim3(KIN_CHUNK), 0, st, b);
      break;
    case KID_QEOS:
      if (b.cost_mode)
        hipLaunchKernelGGL(k_qeos<true>, dim3(b.qeos_launch), dim3(BLOCK), 0, st, b);
      else
        hipLaunchKernelGGL(k_qeos<false>, dim3(b.qeos_launch), dim3(BLOCK), 0, st, b);
      break;
    case KID_DT:
      hipLaunchKernelGGL(k_dt, dim3(b.n), dim3(1024), 0, st, b);
      break;
    case KID_DTPART:
      hipLaunchKernelGGL(k_dtpart, dim3(b.eblocks), dim3(BLOCK), 0, st, b);
      break;
    case KID_TIME_INCREMENT:
      hipLaunchKernelGGL(k_time_increment, dim3(per_inst_grid(b.n)), dim3(64), 0, st, b);
      break;
    case KID_RUN_BEGIN:
      hipLaunchKernelGGL(k_run_begin, dim3(per_inst_grid(b.n)), dim3(64), 0, st, b);
      break;
    case KID_CLEAR_ERR:
      hipLaunchKernelGGL(k_clear_err, dim3(per_inst_grid(b.n)), dim3(64), 0, st, b);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace lulesh
