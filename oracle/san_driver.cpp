// Drives the CPU oracle through every exported entry point, built with the
// address and undefined-behaviour sanitizers (make -C oracle sanitize) --
// TEST INFRASTRUCTURE ONLY.  Prints one "key value" line per checked
// result; tests/test_sanitize.py compares them with the golden fixtures.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>

#include "lulesh_oracle.h"

static int failures = 0;
#define CHECK(cond)                                                  \
  do {                                                               \
    if (!(cond)) {                                                   \
      std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #cond);     \
      ++failures;                                                    \
    }                                                                \
  } while (0)

static void hexline(const char *key, double v) { std::printf("%s %a\n", key, v); }

int main() {
  // Domain(prob) with every region configuration the tests use, fields read
  const int regs[][3] = {{1, 1, 1}, {2, 1, 1}, {5, 0, 2}, {11, 1, 1}, {16, 2, 1}, {32, 1, 0}};
  for (int nx : {1, 2, 3, 8})
    for (const auto &r : regs) {
      oracle_domain *d = oracle_create_regions(nx, 7, 1, r[0], r[1], r[2]);
      CHECK(d != nullptr);
      const char *inames[] = {"matElemlist", "nodelist", "lxim", "lxip", "letam", "letap",
                              "lzetam", "lzetap", "symmX", "symmY", "symmZ", "nodeElemCount",
                              "nodeElemStart", "nodeElemCornerList", "regNumList",
                              "regElemlist", "regCSR", "regReps", "regSorted"};
      int64_t sum = 0;
      for (const char *nm : inames) {
        int64_t n = 0;
        const int64_t *p = oracle_ifield(d, nm, &n);
        CHECK(p != nullptr && n > 0);
        for (int64_t i = 0; i < n; ++i) sum += p[i];
      }
      CHECK(sum > 0);
      oracle_destroy(d);
    }
  CHECK(oracle_create_regions(4, 7, 1, 0, 1, 1) == nullptr);
  CHECK(oracle_create_regions(4, 7, 1, 16, 40, 1) == nullptr);

  // run to stop time at s=8, reference and quirk-free arithmetic, serial and
  // with OpenMP threads (the golden s8_to_stoptime holds the reference run)
  for (int quirks : {7, 0})
    for (int threads : {1, 4}) {
      oracle_domain *d = oracle_create(8, quirks, threads);
      double el = 0.0;
      CHECK(oracle_run(d, 1000000000, &el) == 0);
      double v[4];
      oracle_verify(d, v);
      double cyc = 0.0, t = 0.0;
      oracle_get_scalar(d, "cycle", &cyc);
      oracle_get_scalar(d, "time", &t);
      char key[64];
      std::snprintf(key, sizeof key, "s8_q%d_t%d_cycle", quirks, threads);
      std::printf("%s %.0f\n", key, cyc);
      std::snprintf(key, sizeof key, "s8_q%d_t%d_time", quirks, threads);
      hexline(key, t);
      for (int k = 0; k < 4; ++k) {
        std::snprintf(key, sizeof key, "s8_q%d_t%d_verify%d", quirks, threads, k);
        hexline(key, v[k]);
      }
      oracle_destroy(d);
    }

  // phases one by one, scalar and field access
  {
    oracle_domain *d = oracle_create(5, 7, 2);
    for (int c = 0; c < 20; ++c) {
      CHECK(oracle_time_increment(d) == 0);
      CHECK(oracle_lagrange_nodal(d) == 0);
      CHECK(oracle_lagrange_elements(d) == 0);
      CHECK(oracle_calc_time_constraints(d) == 0);
    }
    const char *dnames[] = {"e", "p", "q", "ql", "qq", "v", "volo", "delv", "vdov", "arealg",
                            "ss", "elemMass", "vnew", "delv_xi", "delv_eta", "delv_zeta",
                            "delx_xi", "delx_eta", "delx_zeta", "dxx", "dyy", "dzz", "x", "y",
                            "z", "xd", "yd", "zd", "xdd", "ydd", "zdd", "fx", "fy", "fz",
                            "nodalMass"};
    for (const char *nm : dnames) {
      int64_t n = 0;
      double *p = oracle_dfield(d, nm, &n);
      CHECK(p != nullptr && n > 0);
      double s = 0.0;
      for (int64_t i = 0; i < n; ++i) s += p[i];
      CHECK(std::isfinite(s));
    }
    int64_t n = 0;
    CHECK(oracle_dfield(d, "nosuchfield", &n) == nullptr);
    double v = 0.0;
    CHECK(oracle_set_scalar(d, "hgcoef", 2.5) == 0 && oracle_get_scalar(d, "hgcoef", &v) == 0 &&
          v == 2.5);
    CHECK(oracle_get_scalar(d, "nosuchscalar", &v) != 0);
    // QStop (code 7) and an invalid elemBC mask (code 6)
    CHECK(oracle_set_scalar(d, "qstop", -1.0) == 0);
    oracle_time_increment(d);
    CHECK(oracle_lagrange_leapfrog(d) == 7);
    oracle_set_scalar(d, "qstop", 1.0e12);
    uint32_t *bc = oracle_elemBC(d, &n);
    bc[0] = 0x3;  // XI_M_SYMM | XI_M_FREE
    oracle_time_increment(d);
    CHECK(oracle_lagrange_leapfrog(d) == 6);
    oracle_destroy(d);
  }

  // element functions on random hexes
  std::mt19937_64 rng(12345);
  std::normal_distribution<double> nd(0.0, 0.1);
  const double U[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                          {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  for (int t = 0; t < 200; ++t) {
    double x[8], y[8], z[8], xd[8], yd[8], zd[8], B[24], b[24], d6[6], hg[24], dx[8], dy[8],
        dz[8], g[6], q[2], eos[4];
    for (int k = 0; k < 8; ++k) {
      x[k] = U[k][0] + nd(rng);
      y[k] = U[k][1] + nd(rng);
      z[k] = U[k][2] + nd(rng);
      xd[k] = nd(rng);
      yd[k] = nd(rng);
      zd[k] = nd(rng);
    }
    const double vol = oracle_calc_elem_volume(x, y, z);
    CHECK(vol > 0.0);
    oracle_calc_elem_node_normals(x, y, z, t & 1, B);
    const double detJ = oracle_calc_elem_shape_function_derivatives(x, y, z, b);
    oracle_calc_elem_velocity_gradient(xd, yd, zd, b, detJ, d6);
    oracle_calc_elem_volume_derivative(x, y, z, t & 2, dx, dy, dz);
    CHECK(std::isfinite(oracle_calc_elem_characteristic_length(x, y, z, vol)));
    oracle_calc_elem_fb_hourglass(x, y, z, xd, yd, zd, vol, 1.0, vol, 3.0, t & 2, hg);
    oracle_monoq_gradients_elem(x, y, z, xd, yd, zd, vol, 1.0, g);
    const double dv[3] = {g[3], g[4], g[5]}, dm[3] = {nd(rng), nd(rng), nd(rng)},
                 dp[3] = {nd(rng), nd(rng), nd(rng)}, dxs[3] = {g[0], g[1], g[2]};
    oracle_monoq_region_elem(dv, dm, dp, dxs, -0.1, vol, vol, 1.0, 0.5, 2.0 / 3.0, q);
    oracle_eval_eos_elem(1.0 + nd(rng) * 0.1, 1.0, nd(rng) * 0.01, 1.0e3, 1.0e3, 0.0, q[0], q[1],
                         eos);
    CHECK(std::isfinite(eos[0] + eos[1] + eos[2] + eos[3]));
    CHECK(std::fabs(std::cbrt(vol) - oracle_cbrt(vol)) <= 1e-15 * std::cbrt(vol));
  }
  double st[10] = {0.0, 1e-6, 1.0, 1e20, 1e20, 1e-2, 1e-2, -1e-6, 1.1, 1.2};
  oracle_time_increment_scalars(st);
  CHECK(st[2] == 2.0);
  uint64_t s[4];
  oracle_julia_rng_state(s);
  std::printf("rng_state0 %llx\n", (unsigned long long)s[0]);
  std::printf("failures %d\n", failures);
  return failures == 0 ? 0 : 1;
}
