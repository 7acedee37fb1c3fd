// lulesh_lagrange_leapfrog_noalloc: one lagrangeLeapFrog on a Domain whose
// fields the caller holds in device memory, with no device or host
// allocation by the library.  The cycle's scratch -- the second node set
// k_nodal writes, the kinematics temporaries k_kin hands to k_qeos, the dt
// partial minima, the instance descriptor, the block tables and the
// device-side scalars -- lives in `scratch`, a fixed array in each GPU's
// memory (LULESH_NOALLOC_SCRATCH_BYTES, held until the process ends).  The
// kernels are the ones lulesh_lagrange_leapfrog launches.
//
// Every call on one GPU uses that GPU's single copy of the array and leaves
// its contents behind: calls that overlap on one GPU race.  Each call sets
// up every table and counter it reads before its kernels run.
#include <hip/hip_runtime.h>

#include <mutex>

#include "lulesh_hip.h"
#include "lulesh_kernels.h"

using namespace lulesh;

static __device__ __align__(256) double scratch[LULESH_NOALLOC_SCRATCH_BYTES / sizeof(double)];

namespace {

constexpr int kMaxDevices = 64;

// per device: the scratch array's device address (hipGetSymbolAddress; the
// array's name is a host-side stand-in, never a device address) and the
// k_nodal planning constants
struct DeviceScratch {
  std::once_flag once;
  hipError_t status = hipErrorNotInitialized;
  char *ptr = nullptr;
  int cus = 256, nodal_per_cu = 2;
};
DeviceScratch g_dev[kMaxDevices];

void setup_device(DeviceScratch &d, int device) {
  hipError_t e = hipSetDevice(device);
  void *p = nullptr;
  if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e == hipSuccess && p == nullptr) e = hipErrorInvalidSymbol;
  int cus = 0;
  if (e == hipSuccess &&
      (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
       cus < 1))
    cus = 256;
  if (e == hipSuccess) {
    d.ptr = (char *)p;
    d.cus = cus;
    d.nodal_per_cu = nodal_blocks_per_cu();
  }
  d.status = e;
}

constexpr size_t kAlign = 256;
size_t round_up(size_t n) { return (n + kAlign - 1) / kAlign * kAlign; }

// the cycle's scratch for one domain of nx elements per edge, carved out of
// the scratch array (offsets in bytes, each 256-byte aligned)
struct Carve {
  int64_t E, N;
  int eb, gb, glaunch, kin_launch, qeos_launch;
  NodalTile tile;
  NodalCut cut;
  size_t off_nodes, off_temps, off_dtpart, off_sc, off_inst, off_eblk, off_gblk, off_gord,
      off_kin, off_qeos, bytes;
};

Carve carve(int nx, const DeviceScratch &d) {
  Carve c{};
  c.E = (int64_t)nx * nx * nx;
  c.N = (int64_t)(nx + 1) * (nx + 1) * (nx + 1);
  c.eb = (int)((c.E + BLOCK - 1) / BLOCK);
  c.tile = nodal_tile(nx);
  const NodalPlanInst pi{nx, c.tile};
  c.cut = depth_cut(nx, auto_depth(&pi, 1, (int64_t)d.cus * d.nodal_per_cu, true), true);
  c.gb = nodal_blocks(c.tile, c.cut);
  // one instance, its blocks taken as equally heavy (under the balanced
  // cut only the last chunk marches fewer planes): nodal_order's XCD runs
  // are then cut at ceil(gb * x / 8) and its longest run has ceil(gb / 8)
  // blocks
  const int runlen = (c.gb + 7) / 8;
  c.glaunch = runlen * 8 > 8 ? runlen * 8 : 8;
  // kin_table for one instance (band rotation 0): XCD x runs band x of
  // every plane, so the launch is 8 x nx x the widest band
  const int nch = kin_chunks(nx);
  int widest = 1;
  for (int x = 0; x < 8; ++x) {
    const int w = kin_band_lo(nch, x + 1) - kin_band_lo(nch, x);
    widest = w > widest ? w : widest;
  }
  c.kin_launch = 8 * nx * widest;
  c.qeos_launch = qeos_launch(c.eb);
  size_t o = 0;
  c.off_nodes = o;  o += round_up(6 * (size_t)c.N * sizeof(double));
  c.off_temps = o;  o += round_up(7 * (size_t)c.E * sizeof(double));
  c.off_dtpart = o; o += round_up(2 * (size_t)c.eb * sizeof(double));
  c.off_sc = o;     o += round_up(sizeof(DevScalars));
  c.off_inst = o;   o += round_up(sizeof(InstDesc));
  c.off_eblk = o;   o += round_up((size_t)c.eb * sizeof(int));
  c.off_gblk = o;   o += round_up((size_t)c.gb * sizeof(int));
  c.off_gord = o;   o += round_up((size_t)c.glaunch * sizeof(int));
  c.off_kin = o;    o += round_up((size_t)c.kin_launch * sizeof(KinWork));
  c.off_qeos = o;   o += round_up((size_t)c.qeos_launch * sizeof(int));
  c.bytes = o;
  return c;
}

// Writes everything the cycle's kernels read from the scratch array: the
// instance descriptor, the scalars (error code clear, node-set parity 0,
// k_nodal's finish count 0), the element-block table (all instance 0) and
// the three dispatch tables, each as the host's planner builds it for a
// single instance (nodal_order_one, kin_work_one, qeos_order_one).
__global__ void k_noalloc_setup(InstDesc I, DevScalars S, InstDesc *inst, DevScalars *sc,
                                int *eblock_inst, int *gblock_inst, int *gblock_order,
                                KinWork *kin_work, int *qeos_order, int eb, int gb, int glaunch,
                                int kin_launch, int qeos_launch) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  if (t == 0) {
    *inst = I;
    *sc = S;
  }
  for (int64_t i = t; i < eb; i += stride) eblock_inst[i] = 0;
  for (int64_t i = t; i < gb; i += stride) gblock_inst[i] = 0;
  for (int64_t g = t; g < glaunch; g += stride) gblock_order[g] = nodal_order_one(gb, g);
  for (int64_t g = t; g < kin_launch; g += stride) kin_work[g] = kin_work_one(I.s, g);
  for (int64_t g = t; g < qeos_launch; g += stride) qeos_order[g] = qeos_order_one(eb, g);
}

bool set_scalars(DevScalars &S, const double *v) {
  // the order of the scalar names of lulesh_get_scalar
  S.time = v[0];
  S.deltatime = v[1];
  if (!(v[2] >= 0.0 && v[2] < 9.0e15)) return false;
  S.cycle = (long long)v[2];
  S.dtcourant = v[3];
  S.dthydro = v[4];
  S.stoptime = v[5];
  S.dtmax = v[6];
  S.dtfixed = v[7];
  S.deltatimemultlb = v[8];
  S.deltatimemultub = v[9];
  S.u_cut = v[10];
  S.hgcoef = v[11];
  S.qstop = v[12];
  S.monoq_max_slope = v[13];
  S.monoq_limiter_mult = v[14];
  S.e_cut = v[15];
  S.p_cut = v[16];
  S.ss4o3 = v[17];
  S.q_cut = v[18];
  S.v_cut = v[19];
  S.qlc_monoq = v[20];
  S.qqc_monoq = v[21];
  S.qqc = v[22];
  S.eosvmax = v[23];
  S.eosvmin = v[24];
  S.pmin = v[25];
  S.emin = v[26];
  S.dvovmax = v[27];
  S.refdens = v[28];
  S.num_iters = 0;
  S.err = ERR_NONE;
  S.active = 0;
  S.parity = 0;
  S.nodal_done = 0;
  return true;
}

int run(const lulesh_device_domain *dom, double *scalars, hipStream_t st, int device) {
  if (!dom || !scalars) return LULESH_ERR_INVALID_ARG;
  if (device < 0 || device >= kMaxDevices) return LULESH_ERR_INVALID_ARG;
  const int nx = dom->nx;
  if (nx < 1) return LULESH_ERR_INVALID_ARG;
  if (nx > LULESH_NOALLOC_MAX_NX) return LULESH_ERR_SCRATCH;
  if (dom->quirks < 0 || dom->quirks > LULESH_QUIRKS_REFERENCE) return LULESH_ERR_INVALID_ARG;
  const void *fields[] = {dom->e,  dom->p,  dom->q,  dom->v,    dom->volo,     dom->delv,
                          dom->vdov, dom->arealg, dom->ss, dom->elemMass, dom->elemBC, dom->x,
                          dom->y,  dom->z,  dom->xd, dom->yd,   dom->zd,       dom->nodalMass};
  for (const void *f : fields)
    if (!f) return LULESH_ERR_INVALID_ARG;
  DevScalars S{};
  if (!set_scalars(S, scalars)) return LULESH_ERR_INVALID_ARG;

  DeviceScratch &d = g_dev[device];
  std::call_once(d.once, setup_device, std::ref(d), device);
  if (d.status != hipSuccess) return LULESH_ERR_HIP;
  if (hipSetDevice(device) != hipSuccess) return LULESH_ERR_HIP;
  const Carve c = carve(nx, d);
  // every nx <= LULESH_NOALLOC_MAX_NX fits whatever depth the device's CU
  // count leads to; the check keeps it so
  if (c.bytes > LULESH_NOALLOC_SCRATCH_BYTES) return LULESH_ERR_SCRATCH;

  char *base = d.ptr;
  InstDesc I{};
  I.s = nx;
  I.etx = c.tile.etx;
  I.ety = c.tile.ety;
  I.tx = c.tile.tx;
  I.ty = c.tile.ty;
  I.e0 = 0;
  I.n0 = 0;
  I.E = c.E;
  I.N = c.N;
  I.eb0 = 0;
  I.ebn = c.eb;
  I.nb0 = 0;
  I.nbn = (int)((c.N + BLOCK - 1) / BLOCK);
  I.gb0 = 0;
  I.gbn = c.gb;
  I.cz = c.cut.cz;
  I.cz0 = c.cut.first;
  I.nch = c.cut.chunks;
  I.einit_scale = 1.0;
  I.hgcoef = S.hgcoef;

  Batch B{};
  B.n = 1;
  B.inst = (const InstDesc *)(base + c.off_inst);
  B.eblock_inst = (const int *)(base + c.off_eblk);
  B.nblock_inst = nullptr;  // no node-shaped kernel runs in the cycle
  B.gblock_inst = (const int *)(base + c.off_gblk);
  B.gblock_order = (const int *)(base + c.off_gord);
  B.kin_work = (const KinWork *)(base + c.off_kin);
  B.qeos_order = (const int *)(base + c.off_qeos);
  B.kin_launch = c.kin_launch;
  B.qeos_launch = c.qeos_launch;
  B.eblocks = c.eb;
  B.nblocks = I.nbn;
  B.gblocks = c.gb;
  B.glaunch = c.glaunch;
  B.nodal_lead = d.cus / 8 > 1 ? d.cus / 8 : 1;
  B.sc = (DevScalars *)(base + c.off_sc);
  DevFields &f = B.f;
  f.e = dom->e; f.p = dom->p; f.q = dom->q; f.v = dom->v; f.volo = dom->volo;
  f.delv = dom->delv; f.vdov = dom->vdov; f.arealg = dom->arealg; f.ss = dom->ss;
  f.elemMass = dom->elemMass; f.elemBC = dom->elemBC;
  double *t = (double *)(base + c.off_temps);
  const int64_t E = c.E, N = c.N;
  f.vnew = t;
  f.delv_xi = t + E; f.delv_eta = t + 2 * E; f.delv_zeta = t + 3 * E;
  f.delx_xi = t + 4 * E; f.delx_eta = t + 5 * E; f.delx_zeta = t + 6 * E;
  // written only under keep_aux, which this entry point leaves off
  f.ql = f.qq = f.dxx = f.dyy = f.dzz = nullptr;
  f.xdd = f.ydd = f.zdd = f.fx = f.fy = f.fz = nullptr;
  f.nodalMass = dom->nodalMass;
  // k_nodal reads set 0 (the caller's) and writes set 1 (scratch), then
  // makes set 1 current for k_kin and k_qeos; set 1 is copied back below
  f.ns[0] = NodeSet{dom->x, dom->y, dom->z, dom->xd, dom->yd, dom->zd};
  double *ns = (double *)(base + c.off_nodes);
  f.ns[1] = NodeSet{ns, ns + N, ns + 2 * N, ns + 3 * N, ns + 4 * N, ns + 5 * N};
  B.dtpart = (double *)(base + c.off_dtpart);
  B.quirks = dom->quirks;
  B.gated = 0;
  B.keep_aux = 0;
  B.loop_mode = 0;
  B.flip = 1;
  B.dt_record = nullptr;
  B.dt_base = 0;
  B.num_iters = 0;

  int work = c.eb > c.glaunch ? c.eb : c.glaunch;
  work = work > c.kin_launch ? work : c.kin_launch;
  work = work > c.qeos_launch ? work : c.qeos_launch;
  const unsigned grid = (unsigned)((work + BLOCK - 1) / BLOCK);
  hipLaunchKernelGGL(k_noalloc_setup, dim3(grid), dim3(BLOCK), 0, st, I, S,
                     (InstDesc *)B.inst, B.sc, (int *)B.eblock_inst, (int *)B.gblock_inst,
                     (int *)B.gblock_order, (KinWork *)B.kin_work, (int *)B.qeos_order, c.eb,
                     c.gb, c.glaunch, c.kin_launch, c.qeos_launch);
  if (hipGetLastError() != hipSuccess) return LULESH_ERR_HIP;
  const int ids[] = {KID_NODAL, KID_KIN, KID_QEOS, KID_DT};
  for (int id : ids)
    if (launch_kernel(id, st, B) != hipSuccess) return LULESH_ERR_HIP;
  double *const dst[] = {dom->x, dom->y, dom->z, dom->xd, dom->yd, dom->zd};
  for (int k = 0; k < 6; ++k)
    if (hipMemcpyAsync(dst[k], ns + k * N, sizeof(double) * (size_t)N, hipMemcpyDeviceToDevice,
                       st) != hipSuccess)
      return LULESH_ERR_HIP;
  DevScalars out{};
  if (hipMemcpyAsync(&out, B.sc, sizeof(DevScalars), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return LULESH_ERR_HIP;
  scalars[3] = out.dtcourant;
  scalars[4] = out.dthydro;
  if (out.err != ERR_NONE) return (out.err >= 1 && out.err <= 9) ? out.err : LULESH_ERR_HIP;
  return 0;
}

}  // namespace

extern "C" int lulesh_lagrange_leapfrog_noalloc(const lulesh_device_domain *dom, double *scalars,
                                                 void *stream, int32_t device) {
  try {
    return run(dom, scalars, (hipStream_t)stream, device);
  } catch (...) {
    // std::call_once can throw std::system_error; no exception leaves the ABI
    return LULESH_ERR_HIP;
  }
}
