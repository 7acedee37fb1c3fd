// C ABI of liblulesh_hip (declared in include/lulesh_hip.h): context
// lifetime, Domain field access, the hot-path entry points, the device-side
// driver loop (graph replay) and per-kernel event timing.
#include "lulesh_hip.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "lulesh_kernels.h"

using namespace lulesh;

struct lulesh_ctx {
  int nx = 0, nr = 1, balance = 1, cost = 1, quirks = LULESH_QUIRKS_REFERENCE, device = 0;
  int64_t E = 0, N = 0;
  hipStream_t stream = nullptr;
  DevFields f{};
  DevScalars *sc = nullptr;   // device
  DevScalars *hsc = nullptr;  // pinned host mirror
  double *dtpart = nullptr;
  int keep_aux = 0;
  bool partials_valid = false;
  // timing
  bool timing = false;
  struct Pending {
    int id;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> event_pool;
  double ktime[LULESH_K_COUNT] = {0};
  int64_t kcount[LULESH_K_COUNT] = {0};
  // driver-loop graph (one cycle)
  hipGraphExec_t graph = nullptr;
  int graph_keep_aux = -1;
  std::string last_error;
  std::vector<void *> allocations;
};

namespace {

const char *kErrText[] = {"",
                          "Early Volume Error",
                          "Mid Volume Error",
                          "Volume Error: Volume is negative",
                          "Volume Error: calcKinematicsForElems",
                          "Volume Error :2157",
                          "Error",
                          "QStopError",
                          "Volume Error :2887",
                          "Volume Error :2908"};

int fail(lulesh_ctx *c, int code, const std::string &msg) {
  if (c) c->last_error = msg;
  return code;
}

int hip_fail(lulesh_ctx *c, hipError_t e, const char *what) {
  std::string m = std::string(what) + ": " + hipGetErrorString(e);
  return fail(c, LULESH_ERR_HIP, m);
}

#define HIP_TRY(c, expr)                                 \
  do {                                                   \
    hipError_t _e = (expr);                              \
    if (_e != hipSuccess) return hip_fail((c), _e, #expr); \
  } while (0)

LaunchArgs make_args(lulesh_ctx *c, int gated, int loop_mode) {
  LaunchArgs a;
  a.sc = c->sc;
  a.f = c->f;
  a.dtpart = c->dtpart;
  a.s = c->nx;
  a.E = c->E;
  a.N = c->N;
  a.quirks = c->quirks;
  a.gated = gated;
  a.keep_aux = c->keep_aux;
  a.loop_mode = loop_mode;
  a.num_iters = 0;
  return a;
}

hipEvent_t get_event(lulesh_ctx *c) {
  if (!c->event_pool.empty()) {
    hipEvent_t e = c->event_pool.back();
    c->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// one launch on the context stream, bracketed by events when timing is on
int launch(lulesh_ctx *c, int id, const LaunchArgs &a) {
  hipEvent_t ea = nullptr, eb = nullptr;
  const bool timed = c->timing && id < LULESH_K_COUNT;
  if (timed) {
    ea = get_event(c);
    eb = get_event(c);
    if (!ea || !eb) return fail(c, LULESH_ERR_HIP, "hipEventCreate failed");
    HIP_TRY(c, hipEventRecord(ea, c->stream));
  }
  hipError_t e = launch_kernel(id, c->stream, a);
  if (e != hipSuccess) return hip_fail(c, e, "kernel launch");
  if (timed) {
    HIP_TRY(c, hipEventRecord(eb, c->stream));
    c->pending.push_back({id, ea, eb});
  }
  return 0;
}

// synchronise, fold pending event pairs into the timing totals, then report
// the lowest physics error raised on the device (if any)
int finish(lulesh_ctx *c, bool check_err) {
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  for (auto &p : c->pending) {
    float ms = 0.f;
    HIP_TRY(c, hipEventElapsedTime(&ms, p.a, p.b));
    c->ktime[p.id] += ms;
    c->kcount[p.id] += 1;
    c->event_pool.push_back(p.a);
    c->event_pool.push_back(p.b);
  }
  c->pending.clear();
  if (!check_err) return 0;
  HIP_TRY(c, hipMemcpy(c->hsc, c->sc, sizeof(DevScalars), hipMemcpyDeviceToHost));
  const int err = c->hsc->err;
  if (err != ERR_NONE) {
    if (err >= 1 && err <= 9) return fail(c, err, kErrText[err]);
    return fail(c, LULESH_ERR_HIP, "unknown device error code");
  }
  return 0;
}

int clear_err(lulesh_ctx *c) { return launch(c, KID_CLEAR_ERR, make_args(c, 0, 0)); }

struct FieldRef {
  double *ptr;
  int64_t n;
};

// node positions/velocities live in the NodeSet the device parity selects
FieldRef find_field(lulesh_ctx *c, const char *name, int parity) {
  if (!name) return {nullptr, 0};
  const std::string s(name);
  DevFields &f = c->f;
  const NodeSet &S = f.ns[parity & 1];
#define EF(nm) \
  if (s == #nm) return {f.nm, c->E};
#define NF(nm) \
  if (s == #nm) return {f.nm, c->N};
#define SF(nm) \
  if (s == #nm) return {S.nm, c->N};
  EF(e) EF(p) EF(q) EF(ql) EF(qq) EF(v) EF(volo) EF(delv) EF(vdov) EF(arealg) EF(ss)
  EF(elemMass) EF(vnew) EF(delv_xi) EF(delv_eta) EF(delv_zeta) EF(delx_xi) EF(delx_eta)
  EF(delx_zeta) EF(dxx) EF(dyy) EF(dzz)
  SF(x) SF(y) SF(z) SF(xd) SF(yd) SF(zd)
  NF(xdd) NF(ydd) NF(zdd) NF(fx) NF(fy) NF(fz) NF(nodalMass)
#undef EF
#undef NF
#undef SF
  return {nullptr, 0};
}

// current parity (syncs the scalar mirror)
int current_parity(lulesh_ctx *c, int *parity) {
  HIP_TRY(c, hipMemcpyAsync(c->hsc, c->sc, sizeof(DevScalars), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  *parity = c->hsc->parity;
  return 0;
}

double *scalar_slot(DevScalars *h, const std::string &s) {
#define SF(nm) \
  if (s == #nm) return &h->nm;
  SF(time) SF(deltatime) SF(dtcourant) SF(dthydro) SF(stoptime) SF(dtmax) SF(dtfixed)
  SF(deltatimemultlb) SF(deltatimemultub) SF(u_cut) SF(hgcoef) SF(qstop) SF(monoq_max_slope)
  SF(monoq_limiter_mult) SF(e_cut) SF(p_cut) SF(ss4o3) SF(q_cut) SF(v_cut) SF(qlc_monoq)
  SF(qqc_monoq) SF(qqc) SF(eosvmax) SF(eosvmin) SF(pmin) SF(emin) SF(dvovmax) SF(refdens)
#undef SF
  return nullptr;
}

template <typename T>
int dev_alloc(lulesh_ctx *c, T **p, int64_t n) {
  void *q = nullptr;
  hipError_t e = hipMalloc(&q, (size_t)n * sizeof(T));
  if (e != hipSuccess) return hip_fail(c, e, "hipMalloc");
  c->allocations.push_back(q);
  *p = (T *)q;
  return 0;
}

void release(lulesh_ctx *c) {
  if (c->graph) (void)hipGraphExecDestroy(c->graph);
  for (auto &p : c->pending) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (auto e : c->event_pool) (void)hipEventDestroy(e);
  for (void *p : c->allocations) (void)hipFree(p);
  if (c->hsc) (void)hipHostFree(c->hsc);
  if (c->stream) (void)hipStreamDestroy(c->stream);
}

int build_graph(lulesh_ctx *c) {
  if (c->graph && c->graph_keep_aux == c->keep_aux) return 0;
  if (c->graph) {
    (void)hipGraphExecDestroy(c->graph);
    c->graph = nullptr;
  }
  const LaunchArgs a = make_args(c, 1, 1);
  hipGraph_t g = nullptr;
  HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  const int ids[] = {KID_NODAL, KID_FLIP, KID_KIN, KID_QEOS, KID_DT};
  for (int id : ids) {
    hipError_t e = launch_kernel(id, c->stream, a);
    if (e != hipSuccess) {
      (void)hipStreamEndCapture(c->stream, &g);
      if (g) (void)hipGraphDestroy(g);
      return hip_fail(c, e, "graph capture");
    }
  }
  HIP_TRY(c, hipStreamEndCapture(c->stream, &g));
  hipError_t e = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    c->graph = nullptr;
    return hip_fail(c, e, "hipGraphInstantiate");
  }
  c->graph_keep_aux = c->keep_aux;
  return 0;
}

}  // namespace

extern "C" {

const char *lulesh_error_string(int code) {
  if (code == LULESH_OK) return "ok";
  if (code >= 1 && code <= 9) return kErrText[code];
  if (code == LULESH_ERR_INVALID_ARG) return "invalid argument";
  if (code == LULESH_ERR_HIP) return "HIP runtime error";
  if (code == LULESH_ERR_UNSUPPORTED) return "unsupported";
  return "unknown error";
}

int lulesh_create(int32_t nx, int32_t nr, int32_t balance, int32_t cost, int32_t quirks,
                  int32_t device, lulesh_ctx **out) {
  if (!out) return LULESH_ERR_INVALID_ARG;
  *out = nullptr;
  if (nx < 1 || nx > 1200 || quirks < 0 || quirks > LULESH_QUIRKS_REFERENCE)
    return LULESH_ERR_INVALID_ARG;
  lulesh_ctx *c = new lulesh_ctx();
  c->nx = nx;
  c->nr = nr;
  c->balance = balance;
  c->cost = cost;
  c->quirks = quirks;
  c->device = device;
  c->E = (int64_t)nx * nx * nx;
  c->N = (int64_t)(nx + 1) * (nx + 1) * (nx + 1);
  int rc = 0;
  hipError_t he = hipSetDevice(device);
  if (he != hipSuccess) rc = hip_fail(c, he, "hipSetDevice");
  if (!rc) {
    he = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (he != hipSuccess) rc = hip_fail(c, he, "hipStreamCreate");
  }
  if (!rc) {
    he = hipHostMalloc((void **)&c->hsc, sizeof(DevScalars), hipHostMallocDefault);
    if (he != hipSuccess) rc = hip_fail(c, he, "hipHostMalloc");
  }
  DevFields &f = c->f;
  const int64_t E = c->E, N = c->N;
  double **ef[] = {&f.e, &f.p, &f.q, &f.ql, &f.qq, &f.v, &f.volo, &f.delv, &f.vdov, &f.arealg,
                   &f.ss, &f.elemMass, &f.vnew, &f.delv_xi, &f.delv_eta, &f.delv_zeta,
                   &f.delx_xi, &f.delx_eta, &f.delx_zeta, &f.dxx, &f.dyy, &f.dzz};
  double **nf[] = {&f.ns[0].x, &f.ns[0].y, &f.ns[0].z, &f.ns[0].xd, &f.ns[0].yd, &f.ns[0].zd,
                   &f.ns[1].x, &f.ns[1].y, &f.ns[1].z, &f.ns[1].xd, &f.ns[1].yd, &f.ns[1].zd,
                   &f.xdd, &f.ydd, &f.zdd, &f.fx, &f.fy, &f.fz, &f.nodalMass};
  for (double **p : ef)
    if (!rc) rc = dev_alloc(c, p, E);
  for (double **p : nf)
    if (!rc) rc = dev_alloc(c, p, N);
  if (!rc) rc = dev_alloc(c, &f.elemBC, E);
  if (!rc) rc = dev_alloc(c, &c->dtpart, 2 * qeos_blocks(E));
  if (!rc) rc = dev_alloc(c, &c->sc, 1);
  if (!rc) rc = lulesh_reset(c);
  if (rc) {
    release(c);
    delete c;
    return rc;
  }
  *out = c;
  return 0;
}

int lulesh_destroy(lulesh_ctx *c) {
  if (!c) return LULESH_ERR_INVALID_ARG;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  release(c);
  delete c;
  return 0;
}

int lulesh_reset(lulesh_ctx *c) {
  if (!c) return LULESH_ERR_INVALID_ARG;
  HIP_TRY(c, hipSetDevice(c->device));
  hipError_t e = launch_init(c->stream, c->nx, c->E, c->N, c->f, c->sc);
  if (e != hipSuccess) return hip_fail(c, e, "init kernels");
  c->partials_valid = false;
  return finish(c, false);
}

int64_t lulesh_num_elem(const lulesh_ctx *c) { return c ? c->E : -1; }
int64_t lulesh_num_node(const lulesh_ctx *c) { return c ? c->N : -1; }
const char *lulesh_last_error(const lulesh_ctx *c) { return c ? c->last_error.c_str() : "null context"; }

int lulesh_download(lulesh_ctx *c, const char *field, double *dst, int64_t n) {
  if (!c || !dst) return LULESH_ERR_INVALID_ARG;
  HIP_TRY(c, hipSetDevice(c->device));
  int par = 0;
  int rc0 = current_parity(c, &par);
  if (rc0) return rc0;
  FieldRef r = find_field(c, field, par);
  if (!r.ptr) return fail(c, LULESH_ERR_INVALID_ARG, std::string("unknown field ") + (field ? field : "(null)"));
  if (n != r.n) return fail(c, LULESH_ERR_INVALID_ARG, "field length mismatch");
  HIP_TRY(c, hipSetDevice(c->device));
  HIP_TRY(c, hipMemcpyAsync(dst, r.ptr, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  return finish(c, false);
}

int lulesh_upload(lulesh_ctx *c, const char *field, const double *src, int64_t n) {
  if (!c || !src) return LULESH_ERR_INVALID_ARG;
  HIP_TRY(c, hipSetDevice(c->device));
  int par = 0;
  int rc0 = current_parity(c, &par);
  if (rc0) return rc0;
  FieldRef r = find_field(c, field, par);
  if (!r.ptr) return fail(c, LULESH_ERR_INVALID_ARG, std::string("unknown field ") + (field ? field : "(null)"));
  if (n != r.n) return fail(c, LULESH_ERR_INVALID_ARG, "field length mismatch");
  HIP_TRY(c, hipSetDevice(c->device));
  HIP_TRY(c, hipMemcpyAsync(r.ptr, src, (size_t)n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  c->partials_valid = false;
  return finish(c, false);
}

int lulesh_download_elem_bc(lulesh_ctx *c, uint32_t *dst, int64_t n) {
  if (!c || !dst || n != c->E) return LULESH_ERR_INVALID_ARG;
  HIP_TRY(c, hipSetDevice(c->device));
  HIP_TRY(c, hipMemcpyAsync(dst, c->f.elemBC, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
  return finish(c, false);
}

int lulesh_upload_elem_bc(lulesh_ctx *c, const uint32_t *src, int64_t n) {
  if (!c || !src || n != c->E) return LULESH_ERR_INVALID_ARG;
  HIP_TRY(c, hipSetDevice(c->device));
  HIP_TRY(c, hipMemcpyAsync(c->f.elemBC, src, (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
  return finish(c, false);
}

int lulesh_get_scalar(lulesh_ctx *c, const char *name, double *value) {
  if (!c || !name || !value) return LULESH_ERR_INVALID_ARG;
  HIP_TRY(c, hipSetDevice(c->device));
  HIP_TRY(c, hipMemcpyAsync(c->hsc, c->sc, sizeof(DevScalars), hipMemcpyDeviceToHost, c->stream));
  int rc = finish(c, false);
  if (rc) return rc;
  const std::string s(name);
  if (s == "cycle") {
    *value = (double)c->hsc->cycle;
    return 0;
  }
  double *slot = scalar_slot(c->hsc, s);
  if (!slot) return fail(c, LULESH_ERR_INVALID_ARG, "unknown scalar " + s);
  *value = *slot;
  return 0;
}

int lulesh_set_scalar(lulesh_ctx *c, const char *name, double value) {
  if (!c || !name) return LULESH_ERR_INVALID_ARG;
  HIP_TRY(c, hipSetDevice(c->device));
  HIP_TRY(c, hipMemcpyAsync(c->hsc, c->sc, sizeof(DevScalars), hipMemcpyDeviceToHost, c->stream));
  int rc = finish(c, false);
  if (rc) return rc;
  const std::string s(name);
  if (s == "cycle") {
    if (!(value >= 0.0) || value != std::floor(value)) return fail(c, LULESH_ERR_INVALID_ARG, "cycle must be a non-negative integer");
    c->hsc->cycle = (long long)value;
  } else {
    double *slot = scalar_slot(c->hsc, s);
    if (!slot) return fail(c, LULESH_ERR_INVALID_ARG, "unknown scalar " + s);
    *slot = value;
  }
  HIP_TRY(c, hipMemcpyAsync(c->sc, c->hsc, sizeof(DevScalars), hipMemcpyHostToDevice, c->stream));
  c->partials_valid = false;
  return finish(c, false);
}

int lulesh_set_keep_aux(lulesh_ctx *c, int32_t on) {
  if (!c) return LULESH_ERR_INVALID_ARG;
  c->keep_aux = on ? 1 : 0;
  return 0;
}

int lulesh_time_increment(lulesh_ctx *c) {
  if (!c) return LULESH_ERR_INVALID_ARG;
  HIP_TRY(c, hipSetDevice(c->device));
  int rc = launch(c, KID_TIME_INCREMENT, make_args(c, 0, 0));
  if (rc) return rc;
  return finish(c, false);
}

int lulesh_lagrange_nodal(lulesh_ctx *c) {
  if (!c) return LULESH_ERR_INVALID_ARG;
  HIP_TRY(c, hipSetDevice(c->device));
  const LaunchArgs a = make_args(c, 0, 0);
  int rc = clear_err(c);
  if (!rc) rc = launch(c, KID_NODAL, a);
  if (!rc) rc = launch(c, KID_FLIP, a);
  if (rc) return rc;
  c->partials_valid = false;
  return finish(c, true);
}

int lulesh_lagrange_elements(lulesh_ctx *c) {
  if (!c) return LULESH_ERR_INVALID_ARG;
  HIP_TRY(c, hipSetDevice(c->device));
  const LaunchArgs a = make_args(c, 0, 0);
  int rc = clear_err(c);
  if (!rc) rc = launch(c, KID_KIN, a);
  if (!rc) rc = launch(c, KID_QEOS, a);
  if (rc) return rc;
  rc = finish(c, true);
  c->partials_valid = (rc == 0);
  return rc;
}

int lulesh_calc_time_constraints(lulesh_ctx *c) {
  if (!c) return LULESH_ERR_INVALID_ARG;
  HIP_TRY(c, hipSetDevice(c->device));
  const LaunchArgs a = make_args(c, 0, 0);
  int rc = 0;
  if (!c->partials_valid) rc = launch(c, KID_DTPART, a);
  if (!rc) rc = launch(c, KID_DT, a);
  if (rc) return rc;
  c->partials_valid = false;
  return finish(c, false);
}

int lulesh_lagrange_leapfrog(lulesh_ctx *c) {
  if (!c) return LULESH_ERR_INVALID_ARG;
  HIP_TRY(c, hipSetDevice(c->device));
  const LaunchArgs a = make_args(c, 0, 0);
  int rc = clear_err(c);
  const int ids[] = {KID_NODAL, KID_FLIP, KID_KIN, KID_QEOS, KID_DT};
  for (int id : ids)
    if (!rc) rc = launch(c, id, a);
  if (rc) return rc;
  c->partials_valid = false;
  return finish(c, true);
}

int lulesh_run(lulesh_ctx *c, int64_t num_iters, int64_t *cycles_run) {
  if (!c) return LULESH_ERR_INVALID_ARG;
  HIP_TRY(c, hipSetDevice(c->device));
  HIP_TRY(c, hipMemcpyAsync(c->hsc, c->sc, sizeof(DevScalars), hipMemcpyDeviceToHost, c->stream));
  int rc = finish(c, false);
  if (rc) return rc;
  const long long cycle0 = c->hsc->cycle;
  LaunchArgs a = make_args(c, 1, 1);
  a.num_iters = num_iters;
  rc = launch(c, KID_RUN_BEGIN, a);
  if (rc) return rc;
  // with timing on, kernels are launched one by one so each can carry its
  // events; otherwise one captured cycle is replayed
  const bool use_graph = !c->timing;
  if (use_graph) {
    rc = build_graph(c);
    if (rc) return rc;
  }
  // replay exactly the cycles num_iters still allows (the first one always
  // runs); only a stop-time exit leaves gated no-op replays behind
  int64_t cap = 8;
  long long next_cycle = cycle0 + 1;
  for (;;) {
    const long long left = (long long)num_iters - next_cycle + 1;
    const int64_t batch = left < 1 ? 1 : (left < cap ? left : cap);
    for (int64_t i = 0; i < batch; ++i) {
      if (use_graph) {
        HIP_TRY(c, hipGraphLaunch(c->graph, c->stream));
      } else {
        const int ids[] = {KID_NODAL, KID_FLIP, KID_KIN, KID_QEOS, KID_DT};
        for (int id : ids) {
          rc = launch(c, id, a);
          if (rc) return rc;
        }
      }
    }
    HIP_TRY(c, hipMemcpyAsync(c->hsc, c->sc, sizeof(DevScalars), hipMemcpyDeviceToHost, c->stream));
    rc = finish(c, false);
    if (rc) return rc;
    if (!c->hsc->active) break;
    next_cycle = c->hsc->cycle;
    if (cap < 256) cap *= 2;
  }
  c->partials_valid = false;
  if (cycles_run) *cycles_run = (int64_t)(c->hsc->cycle - cycle0);
  const int err = c->hsc->err;
  if (err != ERR_NONE) {
    if (err >= 1 && err <= 9) return fail(c, err, kErrText[err]);
    return fail(c, LULESH_ERR_HIP, "unknown device error code");
  }
  return 0;
}

int lulesh_verify(lulesh_ctx *c, double out[4]) {
  if (!c || !out) return LULESH_ERR_INVALID_ARG;
  const int64_t nx = c->nx;
  std::vector<double> e(nx * nx);
  HIP_TRY(c, hipSetDevice(c->device));
  HIP_TRY(c, hipMemcpyAsync(e.data(), c->f.e, e.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  int rc = finish(c, false);
  if (rc) return rc;
  // src/utils.jl:134-152
  double MaxAbsDiff = 0.0, MaxRelDiff = 0.0, TotalAbsDiff = 0.0;
  for (int64_t j = 1; j <= nx; ++j)
    for (int64_t k = j + 1; k <= nx; ++k) {
      const double a = e[(j - 1) * nx + k - 1], b = e[(k - 1) * nx + j - 1];
      const double AbsDiff = std::fabs(a - b);
      TotalAbsDiff = TotalAbsDiff + AbsDiff;
      if (MaxAbsDiff < AbsDiff) MaxAbsDiff = AbsDiff;
      const double RelDiff = AbsDiff / b;
      if (MaxRelDiff < RelDiff) MaxRelDiff = RelDiff;
    }
  out[0] = e[0];
  out[1] = MaxAbsDiff;
  out[2] = TotalAbsDiff;
  out[3] = MaxRelDiff;
  return 0;
}

int lulesh_set_timing(lulesh_ctx *c, int32_t on) {
  if (!c) return LULESH_ERR_INVALID_ARG;
  c->timing = on != 0;
  for (int i = 0; i < LULESH_K_COUNT; ++i) {
    c->ktime[i] = 0.0;
    c->kcount[i] = 0;
  }
  return 0;
}

int lulesh_kernel_time(lulesh_ctx *c, int32_t id, double *total_ms, int64_t *launches) {
  if (!c || id < 0 || id >= LULESH_K_COUNT) return LULESH_ERR_INVALID_ARG;
  if (total_ms) *total_ms = c->ktime[id];
  if (launches) *launches = c->kcount[id];
  return 0;
}

int lulesh_math_check(int32_t op, const double *a, const double *b, double *out, int64_t n,
                      int32_t device) {
  if (op < 0 || op > 2 || !a || !out || n < 0 || (op == 2 && !b)) return LULESH_ERR_INVALID_ARG;
  if (n == 0) return 0;
  if (hipSetDevice(device) != hipSuccess) return LULESH_ERR_HIP;
  double *d = nullptr;
  const size_t bytes = (size_t)n * sizeof(double);
  if (hipMalloc(&d, 3 * bytes) != hipSuccess) return LULESH_ERR_HIP;
  int rc = 0;
  if (hipMemcpy(d, a, bytes, hipMemcpyHostToDevice) != hipSuccess) rc = LULESH_ERR_HIP;
  if (!rc && op == 2 && hipMemcpy(d + n, b, bytes, hipMemcpyHostToDevice) != hipSuccess)
    rc = LULESH_ERR_HIP;
  if (!rc && launch_math(nullptr, op, d, d + n, d + 2 * n, n) != hipSuccess) rc = LULESH_ERR_HIP;
  if (!rc && hipDeviceSynchronize() != hipSuccess) rc = LULESH_ERR_HIP;
  if (!rc && hipMemcpy(out, d + 2 * n, bytes, hipMemcpyDeviceToHost) != hipSuccess)
    rc = LULESH_ERR_HIP;
  (void)hipFree(d);
  return rc;
}

const char *lulesh_kernel_name(int32_t id) {
  static const char *names[] = {"k_nodal", "k_kin", "k_qeos", "k_dt", "k_flip"};
  if (id < 0 || id >= LULESH_K_COUNT) return "";
  return names[id];
}

}  // extern "C"
