// Host check of the fast correctly rounded cube root (lulesh.jl_amd/csrc/
// lulesh_cbrt.h, the code the kernels run), built with the sanitizers by
// tests/host/Makefile.  Its results must equal, bit for bit, both its own
// exact 192-bit path and the oracle's cube root (an independent exact
// implementation, itself pinned against exact integer arithmetic):
//   * random doubles over the whole exponent range, subnormals included;
//   * inputs next to a rounding midpoint: for random K, the two doubles
//     around ((2K+1) 2^-53)^3, whose cube roots lie within about 2^-52 ulp
//     of the midpoint between K 2^-52 and (K+1) 2^-52 -- the cases the fast
//     decision must hand to the exact path;
//   * exact cubes k^3 and powers of two.
// It also reports how often the fast decision defers to the exact path.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>

#include "lulesh_cbrt.h"
#include "lulesh_oracle.h"

using namespace lulesh;

static int failures = 0;
static uint64_t bitsof(double v) {
  uint64_t u;
  std::memcpy(&u, &v, 8);
  return u;
}
static void check(double x, int64_t *slow) {
  const double a = cbrt_cr(x, false, slow), b = cbrt_cr(x, true), c = oracle_cbrt(x);
  if (bitsof(a) != bitsof(b) || bitsof(a) != bitsof(c)) {
    if (failures < 20) std::printf("FAIL x=%a fast=%a exact=%a oracle=%a\n", x, a, b, c);
    ++failures;
  }
}

int main() {
  std::mt19937_64 rng(7);
  int64_t slow_random = 0, slow_hard = 0, n_random = 0, n_hard = 0;
  // random doubles: uniform mantissas, every exponent, both signs
  for (int i = 0; i < 3000000; ++i) {
    const uint64_t u = rng();
    uint64_t b = (u & 0x800fffffffffffffull) | ((uint64_t)(rng() % 2046 + 1) << 52);
    if (i % 97 == 0) b &= 0x800fffffffffffffull;  // subnormal
    double x;
    std::memcpy(&x, &b, 8);
    check(x, &slow_random);
    ++n_random;
  }
  // next to rounding midpoints
  for (int i = 0; i < 300000; ++i) {
    const uint64_t K = (1ull << 52) + (rng() & ((1ull << 52) - 1));
    uint64_t c[3];
    cube54(2 * K + 1, c);  // (2K+1)^3; t = (2K+1)^3 2^-159 = X 2^-52 with X = c >> 107
    const uint64_t X = (c[1] >> 43) | (c[2] << 21);
    for (uint64_t Xd : {X, X + 1}) {
      const double t = (double)Xd * 2.220446049250313080847e-16;
      const int sh = (int)(rng() % 600) - 300;
      check(std::ldexp(t, 3 * sh), &slow_hard);
      ++n_hard;
    }
  }
  // hard cases: K whose midpoint cube ((2K+1) 2^-53)^3 = (2K+1)^3 2^-159
  // sits within 2^-18 of a multiple of 2^-52, so the double t there has a
  // cube root within about 2^-20 units of the rounding midpoint -- where the
  // fast decision must defer to the exact path
  int64_t n_tight = 0, slow_tight = 0;
  for (int i = 0; i < (1 << 23); ++i) {
    const uint64_t K = (1ull << 52) + (rng() & ((1ull << 52) - 1));
    uint64_t c[3];
    cube54(2 * K + 1, c);
    // fraction of (2K+1)^3 / 2^107: the low 107 bits; keep its top 64 bits
    const uint64_t frac_top = (c[1] << 21) | (c[0] >> 43);
    const uint64_t tol = 1ull << 46;  // 2^-18 of 2^64
    if (frac_top < tol || frac_top > ~tol) {
      const uint64_t X = (c[1] >> 43) | (c[2] << 21);
      for (uint64_t Xd : {X, X + 1}) {
        check((double)Xd * 2.220446049250313080847e-16, &slow_tight);
        ++n_tight;
      }
    }
  }
  std::printf("tight-midpoint %lld exact-path %lld\n", (long long)n_tight, (long long)slow_tight);
  for (int64_t k = 1; k < 200000; ++k) {
    const double kk = (double)k;
    check(kk * kk * kk, nullptr);
    check(-kk, nullptr);
  }
  for (int e = -1074; e <= 1023; ++e) check(std::ldexp(1.0, e), nullptr);
  check(0.0, nullptr);
  check(-0.0, nullptr);
  check(INFINITY, nullptr);
  check(-INFINITY, nullptr);
  std::printf("random %lld exact-path %lld\n", (long long)n_random, (long long)slow_random);
  std::printf("near-midpoint %lld exact-path %lld\n", (long long)n_hard, (long long)slow_hard);
  std::printf("failures %d\n", failures);
  return failures == 0 ? 0 : 1;
}
