// Host checks of liblulesh_hip's planning code (lulesh.jl_amd/csrc/
// lulesh_plan.h) and region builder (lulesh_regions.h), compiled with the
// address and undefined-behaviour sanitizers by tests/host/Makefile and run
// by tests/test_sanitize.py.  Every check is a property the kernels or the
// adjoint driver rely on, evaluated over many more cases than the GPU
// tests can run:
//   * k_nodal tiles: along each axis the tiles own every node column once,
//     and an owned column's elements are inside the tile's lanes;
//   * z-chunks: for every depth and both cuts the chunks own every node
//     plane of a column once and each marches at most depth + 1 element
//     planes (the balanced cut's all but the last exactly that); for a
//     requested depth the balanced cut is never longer or deeper than the
//     plain one;
//   * auto_depth: its choice minimises the launch-round cost model (brute
//     force over every depth);
//   * nodal_order: every live block is dispatched exactly once, no finished
//     instance's block is, and each XCD's run holds at most 1/8 of the work
//     plus one block;
//   * kin_table: every element of a live instance is in exactly one chunk,
//     no chunk crosses a plane, and each band of an instance's planes is
//     climbed by one XCD;
//   * the single-instance closed forms of the three dispatch tables (what
//     the no-allocation entry point writes on the device) equal the tables
//     (k_nodal's for blocks of equal weight);
//   * revolve_reverse: a recorded schedule reverses steps n-1..0 once each
//     in that order, never needs more saved states than allowed, only
//     advances from a saved state, and runs n + revolve_cost(n, c-1) steps;
//   * build_regions: the product's arrays equal the oracle's.
#include <cstdio>
#include <algorithm>
#include <cstdlib>
#include <random>
#include <set>
#include <vector>

#include "lulesh_oracle.h"
#include "lulesh_plan.h"
#include "lulesh_regions.h"

using namespace lulesh;

static int failures = 0;
#define CHECK(cond)                                                  \
  do {                                                               \
    if (!(cond)) {                                                   \
      if (failures < 20) std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #cond); \
      ++failures;                                                    \
    }                                                                \
  } while (0)

// node columns owned by the tiles along one axis of s elements, a lanes a tile
static void check_axis(int s, int a) {
  const int n = nodal_span(s, a);
  std::vector<int> owner(s + 1, 0);
  for (int k = 0; k < n; ++k) {
    const int i0 = k == 0 ? 0 : a + (k - 1) * (a - 1);
    const int oi = k != 0;
    const int iend = k == n - 1 ? s + 1 : i0 + a - oi;
    CHECK(iend - i0 <= a);
    for (int i = i0; i < iend; ++i) {
      CHECK(i >= 0 && i <= s);
      if (i >= 0 && i <= s) owner[i] += 1;
      // the elements around node column i (i-1 and i) sit on lanes
      // i-1-(i0-oi) and i-(i0-oi), both in [0, a)
      const int first_lane = i0 - oi;
      if (i >= 1) CHECK(i - 1 - first_lane >= 0 && i - 1 - first_lane < a);
      if (i < s) CHECK(i - first_lane >= 0 && i - first_lane < a);
    }
  }
  for (int i = 0; i <= s; ++i) CHECK(owner[i] == 1);
}

static void check_tiles() {
  const int fam[][2] = {{16, 32}, {16, 16}, {12, 21}, {8, 32}, {24, 32}, {2, 128}};
  for (const auto &f : fam)
    for (int s = 1; s <= 800; ++s) {
      const NodalTile t = nodal_tile(s, f[0], f[1]);
      CHECK(t.etx >= f[0] && t.etx <= f[1] && t.etx * t.ety <= BLOCK);
      CHECK(t.tx == nodal_span(s, t.etx) && t.ty == nodal_span(s, t.ety));
      if (s <= 200 || s % 37 == 0) {
        check_axis(s, t.etx);
        check_axis(s, t.ety);
      }
    }
}

// the node planes k_nodal's chunk c owns: [nodal_chunk_begin(k, c),
// nodal_chunk_end(k, c, s)); it marches element planes
// max(k0 - 1, 0) .. min(k1, s) - 1.  Returns the most any chunk marches.
static int check_cut(int s, const NodalCut &k, bool balanced) {
  CHECK(k.chunks >= 1 && k.cz >= 1);
  std::vector<int> owner((size_t)s + 1, 0);
  int widest = 0;
  for (int c = 0; c < k.chunks; ++c) {
    const int k0 = nodal_chunk_begin(k, c), k1 = nodal_chunk_end(k, c, s);
    CHECK(k0 < k1);
    for (int n = k0; n < k1 && n >= 0 && n <= s; ++n) owner[(size_t)n] += 1;
    const int planes = std::min(k1, s) - std::max(k0 - 1, 0);
    CHECK(planes >= 1 && planes <= k.cz + 1);
    if (balanced && c + 1 < k.chunks) CHECK(planes == k.cz + 1);
    widest = std::max(widest, planes);
  }
  for (int n = 0; n <= s; ++n) CHECK(owner[(size_t)n] == 1);
  return widest;
}

static void check_chunks() {
  for (int s = 1; s <= 260; ++s)
    for (int cz = 1; cz <= s + 2; ++cz) {
      const NodalCut p = nodal_cut(s, cz, false), b = nodal_cut(s, cz, true);
      CHECK(p.cz == cz && b.cz == cz);
      const int n = p.chunks;
      CHECK(check_cut(s, p, false) == (n == 1 ? s : n == 2 ? cz : cz + 1));
      check_cut(s, b, true);
      // a requested depth: the plain cut keeps its chunk count and lowers
      // the depth; the balanced cut is never longer or deeper than it
      const NodalCut dp = depth_cut(s, cz, false), db = depth_cut(s, cz, true);
      CHECK(dp.cz >= 1 && dp.cz <= cz && dp.chunks == n);
      CHECK(db.chunks <= dp.chunks);
      CHECK(check_cut(s, db, true) <= check_cut(s, dp, false));
    }
}

static void check_depth(std::mt19937 &rng) {
  for (int trial = 0; trial < 300; ++trial) {
    const int n = 1 + (int)(rng() % 8);
    std::vector<NodalPlanInst> in;
    int maxm = 2;
    for (int i = 0; i < n; ++i) {
      const int s = 1 + (int)(rng() % (trial % 3 == 0 ? 200 : 64));
      in.push_back({s, nodal_tile(s)});
      maxm = s + 1 > maxm ? s + 1 : maxm;
    }
    const int64_t slots = 1 + (int64_t)(rng() % 1024);
    const bool balanced = trial % 2 == 1;
    const int d = auto_depth(in.data(), in.size(), slots, balanced);
    auto cost = [&](int dd) {
      int64_t total = 0;
      int deepest = 0;
      for (const auto &I : in) {
        const NodalCut k = depth_cut(I.s, dd, balanced);
        CHECK(k.cz >= 1 && k.cz <= I.s + 1);
        total += nodal_blocks(I.t, k);
        deepest = k.cz > deepest ? k.cz : deepest;
      }
      return ((total + slots - 1) / slots) * (deepest + 1);
    };
    CHECK(d >= 1 && d <= maxm);
    const int64_t cd = cost(d);
    for (int dd = 1; dd <= maxm; ++dd) CHECK(cd <= cost(dd));
  }
}

// the single-instance closed forms the no-allocation entry point writes on
// the device equal the planner's tables
static void check_single_instance_tables() {
  for (int s = 1; s <= 120; ++s) {
    const NodalTile t = nodal_tile(s);
    for (int cz : {1, 2, 5, s + 1}) {
      const std::vector<int> gi((size_t)nodal_blocks(t, nodal_cut(s, cz, true)), 0);
      const std::vector<int> pl(gi.size(), cz + 1);
      int glaunch = 0;
      const std::vector<int> ord = nodal_order(gi, pl, nullptr, glaunch);
      const int gb = (int)gi.size(), runlen = (gb + 7) / 8;
      CHECK(glaunch == (runlen * 8 > 8 ? runlen * 8 : 8));
      for (int g = 0; g < glaunch; ++g) CHECK(ord[(size_t)g] == nodal_order_one(gb, g));
    }
    std::vector<int> base;
    int launch = 0;
    const std::vector<KinWork> kin = kin_table({s}, nullptr, base, launch);
    CHECK(base.size() == 1 && base[0] == 0);
    for (int g = 0; g < launch; ++g) {
      const KinWork a = kin[(size_t)g], b = kin_work_one(s, g);
      CHECK(a.inst == b.inst && (a.inst < 0 || a.start == b.start));
    }
    const int eb = (s * s * s + BLOCK - 1) / BLOCK;
    const std::vector<int> q = qeos_table(std::vector<int>((size_t)eb, 0), nullptr);
    CHECK((int)q.size() == qeos_launch(eb));
    for (int g = 0; g < (int)q.size(); ++g) CHECK(q[(size_t)g] == qeos_order_one(eb, g));
  }
}

// k_kin's table: every element of every live instance is in exactly one
// chunk, a chunk never crosses a plane, and XCD x's chunks of an instance
// all lie in one band of the plane (one XCD climbs each band)
static void check_kin(std::mt19937 &rng) {
  for (int trial = 0; trial < 60; ++trial) {
    const int n = 1 + (int)(rng() % 12);
    std::vector<int> s((size_t)n);
    for (int i = 0; i < n; ++i) s[(size_t)i] = 1 + (int)(rng() % (trial < 5 ? 100 : 40));
    std::vector<int> base;
    int launch = 0;
    for (int pass = 0; pass < 2; ++pass) {
      std::vector<char> live((size_t)n, 1);
      if (pass)
        for (int i = 0; i < n; ++i) live[(size_t)i] = (char)(rng() % 2);
      const std::vector<KinWork> t = kin_table(s, pass ? &live : nullptr, base, launch);
      CHECK((int)t.size() == launch && launch % 8 == 0);
      std::vector<std::vector<int>> cover((size_t)n);
      std::vector<std::vector<int>> xcd_of((size_t)n);
      for (int i = 0; i < n; ++i) {
        cover[(size_t)i].assign((size_t)s[(size_t)i] * s[(size_t)i] * s[(size_t)i], 0);
        xcd_of[(size_t)i].assign((size_t)kin_chunks(s[(size_t)i]), -1);
      }
      for (int g = 0; g < launch; ++g) {
        const KinWork w = t[(size_t)g];
        if (w.inst < 0) continue;
        CHECK(w.inst < n && live[(size_t)w.inst]);
        const int64_t si = s[(size_t)w.inst], plane = si * si;
        const int64_t p = w.start / plane, in = w.start % plane;
        CHECK(in % KIN_CHUNK == 0 && p < si);
        const int64_t end = std::min<int64_t>(w.start + KIN_CHUNK, (p + 1) * plane);
        for (int64_t e = w.start; e < end; ++e) cover[(size_t)w.inst][(size_t)e] += 1;
        int &x = xcd_of[(size_t)w.inst][(size_t)(in / KIN_CHUNK)];
        CHECK(x < 0 || x == g % 8);
        x = g % 8;
      }
      for (int i = 0; i < n; ++i)
        for (int c : cover[(size_t)i]) CHECK(c == (live[(size_t)i] ? 1 : 0));
    }
  }
}

static void check_order(std::mt19937 &rng) {
  for (int trial = 0; trial < 300; ++trial) {
    const int n = 1 + (int)(rng() % 40);
    std::vector<int> gi, pl;
    for (int i = 0; i < n; ++i) {
      const int s = 1 + (int)(rng() % 64);
      const NodalTile t = nodal_tile(s);
      const NodalCut k = depth_cut(s, 1 + (int)(rng() % 20), trial % 2 == 1);
      gi.insert(gi.end(), nodal_blocks(t, k), i);
      for (int ch = 0; ch < k.chunks; ++ch)
        pl.insert(pl.end(), t.tx * t.ty, nodal_chunk_planes(k, ch, s));
    }
    CHECK(pl.size() == gi.size());
    int glaunch = 0;
    for (int pass = 0; pass < 3; ++pass) {
      std::vector<char> live(n, 1);
      if (pass > 0)
        for (int i = 0; i < n; ++i) live[i] = (char)(rng() % 3 != 0);
      // with the slot count known and one round enough, equal block counts
      const int64_t slots = trial % 3 == 2 ? (int64_t)gi.size() + (int64_t)(rng() % 64) : 0;
      const std::vector<int> ord = nodal_order(gi, pl, pass ? &live : nullptr, glaunch, slots);
      CHECK(glaunch >= 8 && glaunch % 8 == 0 && (int)ord.size() == glaunch);
      std::vector<int> seen(gi.size(), 0);
      int64_t total = 0, maxw = 0;
      for (size_t v = 0; v < gi.size(); ++v)
        if (live[gi[v]]) {
          total += pl[v];
          maxw = pl[v] > maxw ? pl[v] : maxw;
        }
      std::vector<int64_t> xcd(8, 0);
      for (int g = 0; g < glaunch; ++g) {
        const int v = ord[g];
        if (v < 0) continue;
        CHECK(v < (int)gi.size());
        if (v >= (int)gi.size()) continue;
        seen[v] += 1;
        CHECK(live[gi[v]]);
        xcd[g % 8] += pl[v];
      }
      for (size_t v = 0; v < gi.size(); ++v) CHECK(seen[v] == (live[gi[v]] ? 1 : 0));
      // the first plan cuts the block order into eight equal-work runs, or
      // equal-count ones when the slots hold every block
      if (pass == 0 && slots == 0)
        for (int x = 0; x < 8; ++x) CHECK(xcd[x] * 8 <= total + 8 * maxw);
      if (pass == 0 && slots > 0) {
        int64_t nl = 0;
        std::vector<int64_t> cnt(8, 0);
        for (int g = 0; g < glaunch; ++g)
          if (ord[g] >= 0) {
            cnt[g % 8] += 1;
            nl += 1;
          }
        for (int x = 0; x < 8; ++x) CHECK(cnt[x] * 8 >= nl - 7 && cnt[x] * 8 <= nl + 7);
      }
    }
  }
}

// records a revolve schedule
struct Recorder {
  int64_t n, snapshots;
  std::set<int64_t> saved;  // states held (the start one included)
  int64_t next_rev;         // the step the next youturn must be
  int64_t forward = 0, peak = 0;
  bool ok = true;
  int advance(int64_t from, int64_t to) {
    if (!saved.count(from) || to < from) ok = false;
    forward += to - from;
    return 0;
  }
  int save(int64_t m) {
    if (saved.count(m)) ok = false;
    saved.insert(m);
    peak = (int64_t)saved.size() > peak ? (int64_t)saved.size() : peak;
    return 0;
  }
  int drop(int64_t m) {
    if (!saved.erase(m)) ok = false;
    return 0;
  }
  int youturn(int64_t k) {
    if (k != next_rev) ok = false;
    next_rev = k - 1;
    forward += 1;  // the step run again with its temporaries kept
    return 0;
  }
};

static void check_revolve() {
  for (int64_t n = 1; n <= 150; ++n)
    for (int64_t snaps = 1; snaps <= 12; ++snaps) {
      const int64_t c = snaps < n ? snaps : n;
      Recorder R{n, c, {0}, n - 1};
      R.peak = 1;
      revolve_reverse(R, 0, n, c - 1);
      CHECK(R.ok);
      CHECK(R.next_rev == -1);
      CHECK(R.peak <= c);
      CHECK(R.saved.size() == 1 && *R.saved.begin() == 0);
      // the first descent of the schedule is the recording forward pass, so
      // advancing steps + youturns = revolve_cost + n
      CHECK(R.forward == n + revolve_cost(n, c - 1));
    }
  CHECK(revolve_cost(10000, 19) < 5 * 10000);
  CHECK(beta_sat(1000, 1000) == kRevolveSat);
}

static void check_regions() {
  const int cases[][4] = {{1, 1, 1, 1}, {2, 11, 1, 1}, {8, 11, 1, 1}, {8, 2, 1, 3},
                          {12, 5, 0, 2}, {16, 16, 2, 1}, {20, 32, 1, 0}, {9, 7, 3, 4}};
  for (const auto &k : cases) {
    RegionSets R;
    CHECK(build_regions((int64_t)k[0] * k[0] * k[0], k[1], k[2], k[3], R));
    oracle_domain *d = oracle_create_regions(k[0], 7, 1, k[1], k[2], k[3]);
    CHECK(d != nullptr);
    if (!d) continue;
    int64_t n = 0;
    const int64_t *num = oracle_ifield(d, "regNumList", &n);
    CHECK(n == (int64_t)R.numList.size());
    for (int64_t i = 0; i < n; ++i) CHECK(num[i] == R.numList[i]);
    const int64_t *el = oracle_ifield(d, "regElemlist", &n);
    for (int64_t i = 0; i < n; ++i) CHECK(el[i] == R.elemList[i]);
    const std::vector<int64_t> *small[] = {&R.csr, &R.reps, &R.sorted};
    const char *names[] = {"regCSR", "regReps", "regSorted"};
    for (int j = 0; j < 3; ++j) {
      const int64_t *p = oracle_ifield(d, names[j], &n);
      CHECK(n == (int64_t)small[j]->size());
      for (int64_t i = 0; i < n; ++i) CHECK(p[i] == (*small[j])[i]);
    }
    oracle_destroy(d);
  }
  RegionSets R;
  CHECK(!build_regions(8, 0, 1, 1, R));
  CHECK(!build_regions(8, 16, 40, 1, R));
  int64_t w = 0;
  CHECK(ipow_checked(1, 1 << 30, &w) && w == 1);
  CHECK(!ipow_checked(2, 63, &w) && ipow_checked(2, 62, &w));
}

int main() {
  std::mt19937 rng(2024);
  check_tiles();
  check_chunks();
  check_depth(rng);
  check_order(rng);
  check_single_instance_tables();
  check_kin(rng);
  check_revolve();
  check_regions();
  std::printf("failures %d\n", failures);
  return failures == 0 ? 0 : 1;
}
