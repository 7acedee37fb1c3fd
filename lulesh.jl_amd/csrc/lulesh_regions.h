// Region index sets of Domain(prob): createRegionIndexSets! and sortRegions
// (src/domain.jl:313-465), built on the host (as the reference builds them)
// the first time a context needs them.  No step of the reference's cycle
// reads them (its material loops walk the identity matElemlist); here they
// feed the LULESH-2.0 cost mode of k_qeos (lulesh_set_cost_mode) and are
// downloadable by name.  Values are stored as the reference stores them:
// region numbers and element ids 1-based, regCSR the running end offsets
// the fill loop leaves behind (src/domain.jl:451-456), regReps/regSorted as
// sortRegions' partial bubble sort leaves them (its inner loop stops one
// pair short, src/domain.jl:317).
//
// Random draws: the reference calls Random.seed!(myRank) and rand(Int) on
// Julia 1.7's default task-local generator, xoshiro256++ (Blackman & Vigna,
// "Scrambled linear pseudorandom number generators", ACM TOMS 47, 2021)
// seeded with the SHA-256 digest of the seed's 32-bit words (Julia's
// make_seed: seed 0 -> UInt32[0]).  The generator below is that
// construction; every domain here is rank 0, so its state starts from the
// SHA-256 of four zero bytes.  Only a Julia run could confirm the draw
// stream (nr = 1 draws nothing and is exact).
#pragma once

#include <stdint.h>

#include <vector>

namespace lulesh {

struct JuliaXoshiro {
  // SHA-256(00 00 00 00) = df3f6198 04a92fdb 4057192d c43dd748 ea778adc
  // 52bc498c e80524c0 14b81119, read as four little-endian 64-bit words
  uint64_t s0 = 0xdb2fa90498613fdfull, s1 = 0x48d73dc42d195740ull,
           s2 = 0x8c49bc52dc8a77eaull, s3 = 0x1911b814c02405e8ull;
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next_u64() {
    const uint64_t res = rotl(s0 + s3, 23) + s0;
    const uint64_t t = s1 << 17;
    s2 ^= s0;
    s3 ^= s1;
    s1 ^= s2;
    s0 ^= s3;
    s2 ^= t;
    s3 = rotl(s3, 45);
    return res;
  }
  // rand(Int): the 64 bits read as a signed Int64
  int64_t rand_int() { return (int64_t)next_u64(); }
};

struct RegionSets {
  std::vector<int32_t> numList;   // regNumList[e], 1-based region number
  std::vector<int32_t> elemList;  // regElemlist, 1-based element ids grouped by region
  std::vector<int64_t> csr;       // regCSR (end offsets after the fill)
  std::vector<int64_t> reps;      // regReps (partially sorted, descending)
  std::vector<int64_t> sorted;    // regSorted[region] = its slot in the sort
};

// base^e for base >= 1, e >= 0 (Julia's Int ^ Int), false when the result
// does not fit a signed 64-bit integer
inline bool ipow_checked(int64_t base, int64_t e, int64_t *out) {
  if (base == 1 || e == 0) {
    *out = 1;
    return true;
  }
  __int128 r = 1;
  for (int64_t k = 0; k < e; ++k) {  // base >= 2: at most 63 steps
    r *= base;
    if (r > INT64_MAX) return false;
  }
  *out = (int64_t)r;
  return true;
}

// createRegionIndexSets!(domain, nr, balance) for numElem elements at rank
// 0 with the given cost.  Returns false for arguments the reference cannot
// run with (nr < 1, negative balance, a cost denominator beyond Int64).
inline bool build_regions(int64_t numElem, int nr, int balance, int cost, RegionSets &R) {
  if (nr < 1 || balance < 0 || numElem < 1) return false;
  const int myRank = 0;
  R.numList.assign((size_t)numElem, 0);
  std::vector<int64_t> regElemSize((size_t)nr, 0);
  int64_t nextIndex = 0;
  if (nr == 1) {
    while (nextIndex < numElem) R.numList[(size_t)nextIndex++] = 1;
  } else {
    JuliaXoshiro rng;
    int64_t lastReg = -1, costDenominator = 0;
    std::vector<int64_t> regBinEnd((size_t)nr);
    for (int i = 1; i <= nr; ++i) {
      int64_t w = 0;
      if (!ipow_checked(i, balance, &w) || costDenominator > INT64_MAX - w) return false;
      costDenominator += w;
      regBinEnd[(size_t)i - 1] = costDenominator;
    }
    // the bin of a draw; a negative draw (rand(Int) is signed and % keeps
    // the dividend's sign) stays in the first bin
    auto pick = [&]() {
      const int64_t regionVar = rng.rand_int() % costDenominator;
      int64_t i = 0;
      while (regionVar >= regBinEnd[(size_t)i]) ++i;
      return ((i + myRank) % nr) + 1;
    };
    while (nextIndex < numElem) {
      int64_t regionNum = pick();
      while (regionNum == lastReg) regionNum = pick();
      const int64_t binSize = rng.rand_int() % 1000;
      int64_t elements;
      if (binSize < 773) elements = rng.rand_int() % 15 + 1;
      else if (binSize < 937) elements = rng.rand_int() % 16 + 16;
      else if (binSize < 970) elements = rng.rand_int() % 32 + 32;
      else if (binSize < 974) elements = rng.rand_int() % 64 + 64;
      else if (binSize < 978) elements = rng.rand_int() % 128 + 128;
      else if (binSize < 981) elements = rng.rand_int() % 256 + 256;
      else elements = rng.rand_int() % 1537 + 512;
      const int64_t runto = elements + nextIndex;
      while (nextIndex < runto && nextIndex < numElem) R.numList[(size_t)nextIndex++] = (int32_t)regionNum;
      lastReg = regionNum;
    }
  }
  for (int64_t i = 0; i < numElem; ++i) regElemSize[(size_t)R.numList[(size_t)i] - 1] += 1;
  R.reps.assign((size_t)nr, 0);
  for (int r = 1; r <= nr; ++r) {
    int64_t rep;
    if (r < nr / 2) rep = 1;
    else if (r < nr - (nr + 15) / 20) rep = 1 + cost;
    else rep = 10 * (1 + (int64_t)cost);
    R.reps[(size_t)r - 1] = rep;
  }
  // sortRegions, src/domain.jl:313-336
  std::vector<int64_t> regIndex((size_t)nr);
  for (int v = 1; v <= nr; ++v) regIndex[(size_t)v - 1] = v;
  for (int i = 1; i <= nr - 1; ++i)
    for (int j = 1; j <= nr - i - 1; ++j)
      if (R.reps[(size_t)j - 1] < R.reps[(size_t)j]) {
        std::swap(R.reps[(size_t)j - 1], R.reps[(size_t)j]);
        std::swap(regElemSize[(size_t)j - 1], regElemSize[(size_t)j]);
        std::swap(regIndex[(size_t)j - 1], regIndex[(size_t)j]);
      }
  R.sorted.assign((size_t)nr, 0);
  for (int i = 1; i <= nr; ++i) R.sorted[(size_t)regIndex[(size_t)i - 1] - 1] = i;
  R.csr.assign((size_t)nr, 0);
  for (int i = 2; i <= nr; ++i) R.csr[(size_t)i - 1] = R.csr[(size_t)i - 2] + regElemSize[(size_t)i - 2];
  R.elemList.assign((size_t)numElem, 0);
  for (int64_t i = 1; i <= numElem; ++i) {
    const int64_t r = R.sorted[(size_t)R.numList[(size_t)i - 1] - 1];
    R.elemList[(size_t)R.csr[(size_t)r - 1]] = (int32_t)i;
    R.csr[(size_t)r - 1] += 1;
  }
  return true;
}

}  // namespace lulesh
