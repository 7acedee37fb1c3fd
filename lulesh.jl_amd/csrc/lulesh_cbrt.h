// Correctly rounded FP64 cube root for the GPU kernels (and, compiled for
// the host, for its test harness).  Any correctly rounded cube root returns
// the same bits, so the method only decides the speed:
//
//   1. x = t * 2^(3q) with t in [1, 8); a single-precision seed refined by
//      two Newton steps y -= (y^3 - t) c, c ~ 1/(3 y^2) from a single-
//      precision reciprocal, lands within two units of 2^-52 of t^(1/3).
//   2. K = y rounded to a multiple of 2^-52 (K < 2^53, so yK = K 2^-52 is a
//      double).  The residual r = t - yK^3 is formed in double-double with
//      fused multiply-adds (yK^2 = y2 + e2 and y2 yK = y3 + e3 exactly,
//      t - y3 exact by Sterbenz), and d = r / (3 yK^2) in units of 2^-52 is
//      the distance from yK to t^(1/3) with an error below 2^-40.
//   3. When d is more than 2^-20 away from every half-integer, the nearest
//      integer n to d is the rounding of t^(1/3) and the result is
//      (K + n) 2^(q-52).  Otherwise (about one input in 10^6) the last bit is
//      decided exactly: 8 t 2^104 against the cubes (2K +- 1)^3 of the two
//      rounding midpoints, in 192-bit integer arithmetic; a cube root is
//      never exactly a midpoint, so that comparison never ties.
//
// Subnormal arguments are scaled by 2^54 first.  No other part of the
// build depends on how the result is obtained.
#pragma once

#include <math.h>
#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#define LULESH_CBRT_FN __host__ __device__ __forceinline__
#else
#define LULESH_CBRT_FN inline
#endif

namespace lulesh {

LULESH_CBRT_FN uint64_t cbrt_mul64hi(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umul64hi(a, b);
#else
  return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}
// a single-precision reciprocal, good to a few units of 2^-24
LULESH_CBRT_FN float cbrt_rcpf(float a) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcpf(a);
#else
  return 1.0f / a;
#endif
}

// a^3 for a < 2^54 as a 192-bit integer (w[0] lowest)
LULESH_CBRT_FN void cube54(uint64_t a, uint64_t w[3]) {
  const uint64_t sl = a * a, sh = cbrt_mul64hi(a, a);
  const uint64_t p0l = sl * a, p0h = cbrt_mul64hi(sl, a);
  const uint64_t p1l = sh * a, p1h = cbrt_mul64hi(sh, a);
  const uint64_t mid = p0h + p1l;
  const uint64_t carry = mid < p0h ? 1u : 0u;
  w[0] = p0l;
  w[1] = mid;
  w[2] = p1h + carry;
}
LULESH_CBRT_FN int cmp192(const uint64_t a[3], const uint64_t b[3]) {
  if (a[2] != b[2]) return a[2] < b[2] ? -1 : 1;
  if (a[1] != b[1]) return a[1] < b[1] ? -1 : 1;
  if (a[0] != b[0]) return a[0] < b[0] ? -1 : 1;
  return 0;
}
// the correctly rounded K for t = X 2^-52 (X in [2^52, 2^55)), starting
// from an estimate within a few units
LULESH_CBRT_FN int64_t cbrt_exact_k(uint64_t X, int64_t K) {
  const uint64_t T8[3] = {0ull, X << 43, X >> 21};  // X * 2^107
  for (int it = 0; it < 8; ++it) {
    uint64_t c[3];
    cube54((uint64_t)(2 * K + 1), c);
    if (cmp192(c, T8) < 0) {
      ++K;
      continue;
    }
    cube54((uint64_t)(2 * K - 1), c);
    if (cmp192(c, T8) > 0) {
      --K;
      continue;
    }
    break;
  }
  return K;
}

// the cube root, correctly rounded; for tests, `exact_only` skips the fast
// decision and `*slow` counts the calls the exact path decided
LULESH_CBRT_FN double cbrt_cr(double x, bool exact_only = false, int64_t *slow = nullptr) {
  if (x == 0.0 || !isfinite(x)) return x + x;
  double ax = fabs(x);
  int sub = 0;
  if (ax < 2.2250738585072014e-308) {  // subnormal: scale by 2^54, undo 2^18 later
    ax = ax * 18014398509481984.0;
    sub = 18;
  }
  uint64_t bits;
  memcpy(&bits, &ax, sizeof bits);
  const int E = (int)((bits >> 52) & 0x7ff) - 1023;
  const uint64_t M = (bits & 0xfffffffffffffull) | 0x10000000000000ull;
  const int q = E >= 0 ? E / 3 : -((-E + 2) / 3);
  const int r = E - 3 * q;
  const uint64_t X = M << r;  // t = X * 2^-52 in [1, 8)
  const double t = (double)X * 2.220446049250313080847e-16;
  double y = (double)cbrtf((float)t);
  for (int k = 0; k < 2; ++k) {
    const double y2 = y * y;
    const double f = y2 * y - t;
    y = y - f * (double)cbrt_rcpf((float)(3.0 * y2));
  }
  int64_t K = (int64_t)rint(y * 4503599627370496.0);
  bool decided = false;
  if (!exact_only) {
    const double yk = (double)K * 2.220446049250313080847e-16;
    const double y2 = yk * yk, e2 = fma(yk, yk, -y2);
    const double y3 = y2 * yk, e3 = fma(y2, yk, -y3);
    const double res = ((t - y3) - e3) - e2 * yk;
    const double a = 3.0 * y2;
    double c = (double)cbrt_rcpf((float)a);
    c = c * (2.0 - a * c);
    const double d = (res * c) * 4503599627370496.0;
    const double n = rint(d);
    if (fabs(d - n) < 0.5 - 9.5367431640625e-07 && fabs(n) <= 4.0) {  // 2^-20 margin
      K += (int64_t)n;
      decided = true;
    }
  }
  if (!decided) {
    K = cbrt_exact_k(X, K);
    if (slow) ++*slow;
  }
  const double out = ldexp((double)K, q - 52 - sub);
  return signbit(x) ? -out : out;
}

}  // namespace lulesh
