// Host-side plan of the inverse side tasks that bta_potrf_inverse runs inside
// the blocked Cholesky's step launches (chol.hip).  Plain C++ (no HIP types)
// so that tests can build it alone and replay a plan on the CPU.
//
// X = L^-1 is built row block by row block (64-row leaves):
//   X_rc = -X_rr S_rc,  S_rc = sum_{k=c}^{r-1} L_rk X_kc   (c < r)
// with X_rr = L_rr^-1 from step r's diagonal workgroup.  Partial tasks
// accumulate the terms k < 2q - 1 of the row pair (2q, 2q + 1) in place in
// X's storage as 128 x 64 tile products over K chunks of leaves [k0, k1); a
// chunk may run in step j >= k1 + 2 (L_rk is final after step k, X row block
// k after step k + 2) and every chunk of pair q by step 2q + 1.  The finish
// task of X_rc runs in step r + 2: it adds the terms k in [max(c, 2q - 1), r)
// and multiplies by -X_rr.  Steps NB and NB + 1 carry only side tasks.
#pragma once
#include <algorithm>
#include <cstdint>
#include <vector>

namespace bta {
namespace invplan {

// Chunks span at most KCMAX leaves (a task of about one step's length on one
// workgroup) and at least KCMIN unless they end the pair's range; a step takes
// optional chunks while their units stay within CAP_PER_CU / 100 units per CU
// less the trailing update's tiles.  One unit is one 128 x 64 x 64 product.
constexpr int KCMAX = 2, KCMIN = 1, CAP_PER_CU = 240;
constexpr int FINISH_UNITS = 2;

// x = kind (0 partial, 1 finish) | first << 4 (partial: the chunk that starts
// the tile's sum stores instead of adding); partial: y = q, z = c,
// w = k0 | k1 << 16; finish: y = r, z = c, w = 0.  Same layout as HIP's int4.
struct Task {
  int32_t x, y, z, w;
};

struct Plan {
  std::vector<Task> task;
  std::vector<int> off;    // step j's tasks: task[off[j] .. off[j + 1]); empty on failure
  std::vector<int> units;  // step j's side work in units
  std::vector<int> npart;  // step j's partial tasks (they come before its finish tasks)
};

// 128 x 64 tiles of the rank-64 trailing update in step j of NB
inline int bulk_tiles(int NB, int j) {
  if (j <= 0 || NB - j - 1 <= 0) return 0;
  const int MT = (NB - j) / 2;
  return MT * (MT + 1);
}

inline Plan build(int NB, int ncu) {
  Plan P;
  const int steps = NB + 2;
  const int cap = (int)((int64_t)CAP_PER_CU * ncu / 100);
  struct Tile {
    int q, c, next, end;
  };
  std::vector<Tile> tiles;
  for (int q = 1; 2 * q < NB; ++q)
    for (int c = 0; c <= 2 * q - 2; ++c) tiles.push_back({q, c, c, 2 * q - 1});
  struct Cand {
    int urgent, dl, len, idx;
  };
  std::vector<Cand> cand;
  std::vector<Task> chunk;
  P.off.push_back(0);
  for (int j = 0; j < steps; ++j) {
    int budget = std::max(0, cap - bulk_tiles(NB, j));
    cand.clear();
    for (int i = 0; i < (int)tiles.size(); ++i) {
      const Tile& t = tiles[i];
      if (t.next >= t.end) continue;
      const int ae = std::min(j - 2, t.end);  // leaves k < ae are usable in step j
      if (ae <= t.next) continue;
      const int dl = 2 * t.q + 1;
      // a pair that could not finish in KCMAX-leaf chunks from step j + 1 on
      const int urgent = (t.end - t.next) > KCMAX * (dl - j) ? 1 : 0;
      cand.push_back({urgent, dl, ae - t.next, i});
    }
    std::stable_sort(cand.begin(), cand.end(), [](const Cand& a, const Cand& b) {
      if (a.urgent != b.urgent) return a.urgent > b.urgent;
      if (a.dl != b.dl) return a.dl < b.dl;
      return a.len > b.len;
    });
    chunk.clear();
    int units = 0;
    for (const Cand& cd : cand) {
      Tile& t = tiles[cd.idx];
      const int len = std::min(cd.len, KCMAX);
      if (!cd.urgent) {
        if (len < KCMIN && t.next + len < t.end) continue;
        if (budget < len) continue;
      }
      budget -= len;
      units += len;
      const int first = t.next == t.c ? 1 : 0;
      chunk.push_back({first << 4, t.q, t.c, t.next | ((t.next + len) << 16)});
      t.next += len;
    }
    // longest chunks first (the workers take tasks in a plain stride)
    std::stable_sort(chunk.begin(), chunk.end(), [](const Task& a, const Task& b) {
      return ((a.w >> 16) - (a.w & 0xffff)) > ((b.w >> 16) - (b.w & 0xffff));
    });
    P.task.insert(P.task.end(), chunk.begin(), chunk.end());
    P.npart.push_back((int)chunk.size());
    const int r = j - 2;
    if (r >= 1 && r < NB)
      for (int c = 0; c < r; ++c) {
        P.task.push_back({1, r, c, 0});
        units += FINISH_UNITS;
      }
    P.off.push_back((int)P.task.size());
    P.units.push_back(units);
  }
  for (const Tile& t : tiles)
    if (t.next != t.end) {  // the urgency rule meets every deadline; kept as a guard
      P.off.clear();
      break;
    }
  return P;
}

}  // namespace invplan
}  // namespace bta
