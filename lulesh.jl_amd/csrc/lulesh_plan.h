// Host-side planning arithmetic of liblulesh_hip, free of HIP so that a
// CPU-only harness (built with the address and undefined-behaviour
// sanitizers) compiles exactly this code:
//   * the k_nodal tile of an instance (nodal_span, nodal_tile),
//   * the z-chunk cuts and the depth model (nodal_cut, depth_cut, auto_depth),
//   * the XCD-aware k_nodal dispatch table (nodal_order),
//   * the binomial checkpoint schedule of the adjoint (revolve_cost,
//     revolve_split, revolve_reverse).
#pragma once

#include <stdint.h>

#include <algorithm>
#include <vector>

namespace lulesh {

constexpr int BLOCK = 256;

// k_nodal tiling: a workgroup marches a column of etx x ety element lanes
// (etx * ety <= BLOCK) up one z-chunk (below) and owns the node columns whose
// elements all lie in its column.  Along an axis of s elements (s+1 nodes),
// with a lanes per tile: the first tile's elements start at node 0 and it
// owns a nodes (no element precedes node 0); an inner tile's first element
// is the halo shared with the tile before and it owns a-1 nodes; the last
// tile owns the rest, at most a (no element follows node s); one tile covers
// the axis when s+1 <= a.
inline int nodal_span(int s, int a) {
  const int m = s + 1;
  if (m <= a) return 1;
  const int rest = m - 2 * a;
  return 2 + (rest > 0 ? (rest + a - 2) / (a - 1) : 0);
}
struct NodalTile {
  int etx, ety, tx, ty;
};
// Per instance: the tile, etx in lo..hi lanes wide (by default 16..32: a
// row's node loads span whole 128-byte lines), ety = BLOCK / etx, that
// covers the (s+1)^2 node columns with the fewest tiles (the squarest among
// equals).  At s=192 that is 17 x 15 (12 x 14 tiles); small domains avoid a
// mostly idle last tile.  2 <= lo <= hi <= BLOCK / 2.
inline NodalTile nodal_tile(int s, int lo = 16, int hi = 32) {
  NodalTile best{0, 0, 0, 0};
  for (int a = lo; a <= hi; ++a) {
    const int b = BLOCK / a;
    const int tx = nodal_span(s, a), ty = nodal_span(s, b);
    const int dcand = a > b ? a - b : b - a;
    const int dcur = best.etx > best.ety ? best.etx - best.ety : best.ety - best.etx;
    if (best.etx == 0 || tx * ty < best.tx * best.ty || (tx * ty == best.tx * best.ty && dcand < dcur))
      best = NodalTile{a, b, tx, ty};
  }
  return best;
}
// z-chunks of a column.  Node plane k takes the corners of element planes
// k-1 and k, so a chunk marches one element plane more than it owns node
// planes, except where there is none below (node plane 0) or above (node
// plane s).  A cut gives the first chunk `first` node planes and every later
// one `cz` (the z-march depth), the last one what is left up to node plane
// s; chunk c owns node planes [nodal_chunk_begin(c), nodal_chunk_end(c)).
// Two cuts:
//   plain     first = cz and ceil((s + 1) / cz) chunks: the first and the
//             last chunk march one element plane fewer than the others;
//   balanced  first = cz + 1 and ceil((s - 1) / cz) chunks (one if
//             cz >= s - 1): every chunk but the last marches cz + 1 element
//             planes, the last at most that.
struct NodalCut {
  int first, cz, chunks;
};
constexpr int nodal_chunk_begin(const NodalCut &k, int c) {
  return c == 0 ? 0 : k.first + (c - 1) * k.cz;
}
constexpr int nodal_chunk_end(const NodalCut &k, int c, int s) {
  return c + 1 >= k.chunks ? s + 1 : k.first + c * k.cz;
}
inline NodalCut nodal_cut(int s, int cz, bool balanced) {
  if (!balanced) return NodalCut{cz, cz, (s + cz) / cz};
  return NodalCut{cz + 1, cz, s > 1 && cz < s - 1 ? (s + cz - 2) / cz : 1};
}
// The cut for a requested `depth`: the plain cut's chunk count for it,
// ceil((s + 1) / depth), with the depth lowered as far as that count allows
// (nearly equal chunks).  The balanced cut of a depth never has more chunks
// or a deeper chunk than the plain one.
inline NodalCut depth_cut(int s, int depth, bool balanced) {
  const int chunks = (s + depth) / depth;
  if (!balanced) return nodal_cut(s, (s + chunks) / chunks, false);
  const int cz = (s + 2 * chunks - 2) / chunks - 1;
  return nodal_cut(s, cz > 1 ? cz : 1, true);
}
inline int nodal_blocks(const NodalTile &t, const NodalCut &k) { return t.tx * t.ty * k.chunks; }
// the element planes chunk c marches: max(k0 - 1, 0) .. min(k1, s) - 1
inline int nodal_chunk_planes(const NodalCut &k, int c, int s) {
  const int k0 = nodal_chunk_begin(k, c), k1 = nodal_chunk_end(k, c, s);
  return (k1 < s ? k1 : s) - (k0 > 1 ? k0 - 1 : 0);
}

// one instance as the k_nodal planner sees it
struct NodalPlanInst {
  int s;
  NodalTile t;
};

// Default z-march depth.  A k_nodal workgroup's time grows with the planes
// it marches (depth + 1 element planes), and the launch runs in rounds of
// `slots` (CUs x resident workgroups per CU): the estimate
//   ceil(workgroups / slots) * (depth + 1)
// is minimised over all depths (the smaller depth on ties, for balance),
// each instance cut by depth_cut.  At s=192 on 256 CUs x 2 this picks 3
// chunks (504 workgroups, one round; plain cut 65, 66 and 63 element planes,
// balanced 65, 65 and 64) over 13 chunks of depth 15 (2184 workgroups, 4.3
// rounds).
inline int auto_depth(const NodalPlanInst *insts, size_t n, int64_t slots, bool balanced) {
  int maxm = 2;
  for (size_t i = 0; i < n; ++i) maxm = insts[i].s + 1 > maxm ? insts[i].s + 1 : maxm;
  if (slots < 1) slots = 1;
  int best = 2;
  int64_t best_cost = -1;
  for (int d = 1; d <= maxm; ++d) {
    int64_t total = 0;
    int deepest = 0;
    for (size_t i = 0; i < n; ++i) {
      const NodalPlanInst &I = insts[i];
      const NodalCut k = depth_cut(I.s, d, balanced);
      total += nodal_blocks(I.t, k);
      deepest = k.cz > deepest ? k.cz : deepest;
    }
    const int64_t cost = ((total + slots - 1) / slots) * (deepest + 1);
    if (best_cost < 0 || cost < best_cost) {
      best = d;
      best_cost = cost;
    }
  }
  return best;
}

// k_nodal dispatch order.  The dispatcher deals a launch's workgroups
// round-robin over the eight XCDs (workgroup g to XCD g % 8), each with its
// own L2.  XCD x gets the x-th of eight contiguous runs of the block order
// (instance, chunk, tile row, tile), cut at equal sums of marched element
// planes: neighbouring tiles, which read each other's halo nodes, share an
// L2, and every XCD gets the same work.  Within a run the blocks that march
// the most planes go first (stable, so an instance's blocks stay together):
// the shallow ones -- short chunks included -- fill the last round.
//   gblock_inst[v]  instance of k_nodal block v,
//   planes[v]       element planes block v marches (nodal_chunk_planes),
//   live            instances to list (null: all),
//   glaunch         workgroups launched: 0 sizes it (a multiple of 8, at
//                   least 8); a later call keeps it, and lists the live
//                   blocks in plain order when they need more rounds,
//   slots           workgroups the GPU holds at once (0: unknown).  When
//                   the live blocks fit in one round the runs are cut at
//                   equal block counts instead: every block then has a
//                   slot of its own from the start, and a run longer than
//                   its XCD's share of the slots would need a second round.
// Returns the table: block run by each launched workgroup, -1 for none.
inline std::vector<int> nodal_order(const std::vector<int> &gblock_inst,
                                    const std::vector<int> &planes,
                                    const std::vector<char> *live, int &glaunch,
                                    int64_t slots = 0) {
  std::vector<int> lv;
  lv.reserve(gblock_inst.size());
  int64_t total = 0;
  for (int v = 0; v < (int)gblock_inst.size(); ++v)
    if (!live || (*live)[gblock_inst[v]]) {
      lv.push_back(v);
      total += planes[v];
    }
  const int nl = (int)lv.size();
  int cut[9];
  cut[0] = 0;
  int x = 1;
  int64_t acc = 0;
  if (slots > 0 && nl <= slots) {
    for (; x < 8; ++x) cut[x] = (int)(((int64_t)nl * x + 7) / 8);
  } else {
    for (int k = 0; k < nl; ++k) {
      acc += planes[lv[k]];
      while (x < 8 && acc * 8 >= total * x) cut[x++] = k + 1;
    }
  }
  while (x <= 8) cut[x++] = nl;
  int runlen = 0;
  for (int r = 0; r < 8; ++r) runlen = cut[r + 1] - cut[r] > runlen ? cut[r + 1] - cut[r] : runlen;
  if (glaunch == 0) glaunch = runlen * 8 > 8 ? runlen * 8 : 8;
  std::vector<int> order((size_t)glaunch, -1);
  if (runlen * 8 <= glaunch) {
    for (int r = 0; r < 8; ++r) {
      std::stable_sort(lv.begin() + cut[r], lv.begin() + cut[r + 1],
                       [&](int a, int b) { return planes[a] > planes[b]; });
      for (int k = 0; k < cut[r + 1] - cut[r]; ++k) order[(size_t)k * 8 + r] = lv[cut[r] + k];
    }
  } else {
    for (int k = 0; k < nl; ++k) order[k] = lv[k];
  }
  return order;
}

// ---------------------------------------------------------------------------
// Work tables of the element kernels.  The dispatcher deals workgroup g to
// XCD g % 8, and each XCD has its own L2.
//
// k_kin gathers the nodes of two node planes per element plane, so node
// plane p+1 is read by element planes p and p+1.  Each element plane is cut
// into KIN_CHUNK-element chunks and its chunks into eight contiguous bands;
// XCD x runs band (x - base) mod 8 of every plane of an instance, plane
// after plane, so the node rows an element plane shares with the next one
// are read again by the same XCD while its L2 still holds them.  `base`
// rotates the bands per instance to even out the XCDs' work in a batch.
// ---------------------------------------------------------------------------
constexpr int KIN_CHUNK = 64;
struct KinWork {
  int32_t inst;   // instance, -1: no work
  int32_t start;  // first element of the chunk (instance-local); the chunk
                  // ends at the next multiple of KIN_CHUNK or the plane end
};
inline int kin_chunks(int s) { return (int)(((int64_t)s * s + KIN_CHUNK - 1) / KIN_CHUNK); }
// band b of a plane of nch chunks: chunks [kin_band_lo(nch, b), kin_band_lo(nch, b + 1))
inline int kin_band_lo(int nch, int b) { return (int)(((int64_t)b * nch + 7) / 8); }

//   s[i]      elements per edge of instance i,
//   live      instances to list (null: all),
//   base      band rotation per instance; an empty vector is filled in
//             (greedy: each instance in turn takes the rotation that keeps
//             the busiest XCD least busy, the first on ties),
//   launch    workgroups launched: 0 sizes it (8 x the longest XCD
//             sequence); a later call for fewer instances keeps it.
// Returns the table: the chunk run by each launched workgroup.
inline std::vector<KinWork> kin_table(const std::vector<int> &s, const std::vector<char> *live,
                                      std::vector<int> &base, int &launch) {
  const int n = (int)s.size();
  if ((int)base.size() != n) {
    base.assign((size_t)n, 0);
    int64_t load[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = 0; i < n; ++i) {
      const int nch = kin_chunks(s[i]);
      int best = 0;
      int64_t best_max = -1;
      for (int b = 0; b < 8; ++b) {
        int64_t mx = 0;
        for (int x = 0; x < 8; ++x) {
          const int band = (x - b + 8) % 8;
          const int64_t w = load[x] + (int64_t)s[i] * (kin_band_lo(nch, band + 1) - kin_band_lo(nch, band));
          mx = w > mx ? w : mx;
        }
        if (best_max < 0 || mx < best_max) {
          best_max = mx;
          best = b;
        }
      }
      base[(size_t)i] = best;
      for (int x = 0; x < 8; ++x) {
        const int band = (x - best + 8) % 8;
        load[x] += (int64_t)s[i] * (kin_band_lo(nch, band + 1) - kin_band_lo(nch, band));
      }
    }
  }
  // XCD x's sequence: the live instances in order, each band x's chunks
  // plane by plane; entry j of it is workgroup 8 j + x.  Counted first, then
  // written in place (the driver loop rebuilds the table whenever an
  // instance finishes, while the GPU waits).
  auto band_of = [&](int i, int x, int &lo, int &hi) {
    const int nch = kin_chunks(s[i]);
    const int band = (x - base[(size_t)i] + 8) % 8;
    lo = kin_band_lo(nch, band);
    hi = kin_band_lo(nch, band + 1);
  };
  int64_t len = 1;
  for (int x = 0; x < 8; ++x) {
    int64_t m = 0;
    for (int i = 0; i < n; ++i) {
      if (live && !(*live)[(size_t)i]) continue;
      int lo, hi;
      band_of(i, x, lo, hi);
      m += (int64_t)s[i] * (hi - lo);
    }
    len = m > len ? m : len;
  }
  if (launch == 0) launch = (int)(8 * len);
  std::vector<KinWork> table((size_t)launch, KinWork{-1, 0});
  for (int x = 0; x < 8; ++x) {
    size_t g = (size_t)x;
    for (int i = 0; i < n && g < table.size(); ++i) {
      if (live && !(*live)[(size_t)i]) continue;
      const int64_t plane = (int64_t)s[i] * s[i];
      int lo, hi;
      band_of(i, x, lo, hi);
      for (int p = 0; p < s[i] && g < table.size(); ++p)
        for (int ch = lo; ch < hi && g < table.size(); ++ch, g += 8)
          table[g] = KinWork{i, (int32_t)(p * plane + (int64_t)ch * KIN_CHUNK)};
    }
  }
  return table;
}

// k_qeos runs 256-element blocks (each writes one pair of dt partial
// minima); XCD x takes runs of QEOS_GROUP consecutive blocks, the runs
// dealt round-robin, so the eight XCDs sweep neighbouring addresses
// together.  Blocks of instances not in `live` (null: all) are left out
// (-1), as are the launch's padding workgroups.  launch = the block count
// rounded up to a multiple of 8 x QEOS_GROUP.
constexpr int QEOS_GROUP = 4;
inline int qeos_launch(int eblocks) {
  const int g = 8 * QEOS_GROUP;
  return (eblocks + g - 1) / g * g;
}
inline std::vector<int> qeos_table(const std::vector<int> &eblock_inst, const std::vector<char> *live) {
  const int nb = (int)eblock_inst.size();
  std::vector<int> t((size_t)qeos_launch(nb), -1);
  for (int g = 0; g < (int)t.size(); ++g) {
    const int x = g % 8, k = g / 8;
    const int v = ((k / QEOS_GROUP) * 8 + x) * QEOS_GROUP + k % QEOS_GROUP;
    if (v < nb && (!live || (*live)[(size_t)eblock_inst[(size_t)v]])) t[(size_t)g] = v;
  }
  return t;
}

// The three tables of a single instance (all live, band rotation 0) in
// closed form, entry g at a time: what the no-allocation entry point
// writes on the device, equal to nodal_order / kin_table / qeos_table.
#if defined(__HIPCC__)
#define LULESH_HD __host__ __device__
#else
#define LULESH_HD
#endif
// nodal_order: every block weighs the same, so the XCD runs are cut at
// ceil(gb r / 8)
LULESH_HD inline int nodal_order_one(int64_t gb, int64_t g) {
  const int64_t r = g % 8, k = g / 8;
  const int64_t lo = (gb * r + 7) / 8, hi = (gb * (r + 1) + 7) / 8;
  return k < hi - lo ? (int)(lo + k) : -1;
}
LULESH_HD inline KinWork kin_work_one(int64_t s, int64_t g) {
  const int64_t nch = (s * s + KIN_CHUNK - 1) / KIN_CHUNK;
  const int64_t x = g % 8, j = g / 8;
  const int64_t lo = (x * nch + 7) / 8, hi = ((x + 1) * nch + 7) / 8, nb = hi - lo;
  if (nb > 0 && j < s * nb) return KinWork{0, (int32_t)((j / nb) * s * s + (lo + j % nb) * KIN_CHUNK)};
  return KinWork{-1, 0};
}
LULESH_HD inline int qeos_order_one(int64_t eb, int64_t g) {
  const int64_t x = g % 8, k = g / 8;
  const int64_t v = ((k / QEOS_GROUP) * 8 + x) * QEOS_GROUP + k % QEOS_GROUP;
  return v < eb ? (int)v : -1;
}

// ---------------------------------------------------------------------------
// binomial checkpointing: Griewank, "Achieving logarithmic growth of temporal
// and spatial complexity in reverse automatic differentiation" (Optimization
// Methods and Software 1, 1992) and Griewank & Walther, "Algorithm 799:
// revolve" (ACM TOMS 26, 2000).  With c saved states (the start one
// included), reversing n steps needs at least
//   t(n, c) = r n - beta(c + 1, r - 1)   forward steps besides the n
// recording ones, r the least integer with beta(c, r) >= n,
// beta(c, r) = C(c + r, c).
// ---------------------------------------------------------------------------
constexpr int64_t kRevolveSat = (int64_t)1 << 62;

inline int64_t beta_sat(int64_t c, int64_t r) {
  if (r < 0) return 0;
  unsigned __int128 v = 1;  // C(c + i, i) for i = 0..r, each step exact
  for (int64_t i = 1; i <= r; ++i) {
    v = v * (unsigned __int128)(c + i) / (unsigned __int128)i;
    if (v >= (unsigned __int128)kRevolveSat) return kRevolveSat;
  }
  return (int64_t)v;
}

// fewest advancing forward steps that reverse n steps from a saved start
// state with `spare` further states free
inline int64_t revolve_cost(int64_t n, int64_t spare) {
  if (n <= 1) return 0;
  if (spare == 0) return n * (n - 1) / 2;
  const int64_t c = spare + 1;
  int64_t r = 0;
  while (beta_sat(c, r) < n) ++r;
  return r * n - beta_sat(c + 1, r - 1);
}

// where to save the next state: advance l steps (1 <= l < n).  The total
// l + cost(n - l, spare - 1) + cost(l, spare) is convex in l (the cost is
// piecewise linear in n with slopes r, non-decreasing), so the first l whose
// forward difference is >= 0 minimises it.
inline int64_t revolve_split(int64_t n, int64_t spare) {
  auto g = [&](int64_t l) { return l + revolve_cost(n - l, spare - 1) + revolve_cost(l, spare); };
  int64_t lo = 1, hi = n - 1;
  while (lo < hi) {
    const int64_t mid = lo + (hi - lo) / 2;
    if (g(mid + 1) - g(mid) >= 0) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

// Reverse steps b-1 .. a0 with the start state of step a0 saved and `spare`
// state buffers free.  `act` carries out the schedule:
//   advance(from, to)  run steps from..to-1 starting at the saved state of `from`,
//   save(m)            keep the current state (that of step m) in a free buffer,
//   drop(m)            free the buffer holding step m's state,
//   youturn(k)         run step k once more keeping its temporaries, then its adjoint;
// each returns 0 or an error code, which ends the recursion.
template <typename Actions>
int revolve_reverse(Actions &act, int64_t a0, int64_t b, int64_t spare) {
  if (b - a0 == 1) return act.youturn(a0);
  if (spare == 0) {
    for (int64_t k = b - 1; k >= a0; --k) {
      int rc = act.advance(a0, k);
      if (!rc) rc = act.youturn(k);
      if (rc) return rc;
    }
    return 0;
  }
  const int64_t m = a0 + revolve_split(b - a0, spare);
  int rc = act.advance(a0, m);
  if (!rc) rc = act.save(m);
  if (!rc) rc = revolve_reverse(act, m, b, spare - 1);
  if (!rc) rc = act.drop(m);
  if (rc) return rc;
  return revolve_reverse(act, a0, m, spare);
}

}  // namespace lulesh
