// Host-side atom orders of the device arrays (plain C++: the C-ABI layer
// includes it, and host tests build it with a C++ compiler).
#pragma once
#include <math.h>

#include <algorithm>
#include <vector>

#include "pst_types.hpp"

namespace pst {

// Spatial order inside a span of atom indices.  The pose-mode force loops
// hand atom i of a run to lane i mod G (G = 16..64), so the atoms one wave
// gathers together are an aligned run of 64 or 32 entries.  Recursive
// bisection along the longest edge of the bounding box, at multiples of 64
// while a span holds more than 64 atoms and at 32, then 16, below that (ties
// broken by atom index), makes every such run a compact cluster: its stencils
// share grid bricks and cache lines.  Only the order of each sum changes.
inline void bisect_runs(const double* pos, int* idx, int n) {
  struct Span { int lo, hi; };
  std::vector<Span> todo{{0, n}};
  while (!todo.empty()) {
    const Span s = todo.back();
    todo.pop_back();
    const int cnt = s.hi - s.lo;
    if (cnt <= 16) continue;
    double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
    for (int j = s.lo; j < s.hi; ++j)
      for (int d = 0; d < 3; ++d) {
        mn[d] = fmin(mn[d], pos[3 * idx[j] + d]);
        mx[d] = fmax(mx[d], pos[3 * idx[j] + d]);
      }
    int ax = 0;
    for (int d = 1; d < 3; ++d) if (mx[d] - mn[d] > mx[ax] - mn[ax]) ax = d;
    const int run = cnt > 64 ? 64 : cnt > 32 ? 32 : 16;
    const int mid = s.lo + (((cnt + run - 1) / run) / 2) * run;
    std::nth_element(idx + s.lo, idx + mid, idx + s.hi, [&](int a, int b) {
      const double pa = pos[3 * a + ax], pb = pos[3 * b + ax];
      return pa < pb || (pa == pb && a < b);
    });
    todo.push_back({mid, s.hi});
    todo.push_back({s.lo, mid});
  }
}

// Recursive bisection of the receptor atoms into runs of PST_REC_CHUNK
// (the last run takes the remainder): split along the longest edge of the
// bounding box at a multiple of PST_REC_CHUNK atoms, ties broken by atom
// index; when there is more than one run, each is then put in the nested
// spatial order of bisect_runs (its aligned halves of 32 and quarters of 16
// atoms are compact as well), else kept in input order.  spheres gets (centre,
// radius) per run, the radius rounded up so that |r_j - centre| <= radius
// holds for every atom of the run in floating point.
inline void rec_chunk_order(int n, const double* pos, std::vector<int>& order,
                            std::vector<double>& spheres) {
  order.resize(n);
  for (int j = 0; j < n; ++j) order[j] = j;
  struct Span { int lo, hi; };
  std::vector<Span> todo{{0, n}}, leaves;
  while (!todo.empty()) {
    const Span s = todo.back();
    todo.pop_back();
    const int cnt = s.hi - s.lo;
    if (cnt <= PST_REC_CHUNK) { leaves.push_back(s); continue; }
    double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
    for (int j = s.lo; j < s.hi; ++j)
      for (int d = 0; d < 3; ++d) {
        mn[d] = fmin(mn[d], pos[3 * order[j] + d]);
        mx[d] = fmax(mx[d], pos[3 * order[j] + d]);
      }
    int ax = 0;
    for (int d = 1; d < 3; ++d) if (mx[d] - mn[d] > mx[ax] - mn[ax]) ax = d;
    const int chunks = (cnt + PST_REC_CHUNK - 1) / PST_REC_CHUNK;
    const int mid = s.lo + (chunks / 2) * PST_REC_CHUNK;
    std::nth_element(order.begin() + s.lo, order.begin() + mid, order.begin() + s.hi,
                     [&](int a, int b) {
                       const double pa = pos[3 * a + ax], pb = pos[3 * b + ax];
                       return pa < pb || (pa == pb && a < b);
                     });
    todo.push_back({mid, s.hi});
    todo.push_back({s.lo, mid});
  }
  std::sort(leaves.begin(), leaves.end(), [](const Span& a, const Span& b) { return a.lo < b.lo; });
  spheres.assign(4 * leaves.size(), 0.0);
  for (size_t k = 0; k < leaves.size(); ++k) {
    const Span s = leaves[k];
    std::sort(order.begin() + s.lo, order.begin() + s.hi);
    if (n > PST_REC_CHUNK) bisect_runs(pos, order.data() + s.lo, s.hi - s.lo);
    double c[3] = {0, 0, 0};
    for (int j = s.lo; j < s.hi; ++j)
      for (int d = 0; d < 3; ++d) c[d] += pos[3 * order[j] + d];
    for (int d = 0; d < 3; ++d) c[d] /= (double)(s.hi - s.lo);
    double r = 0;
    for (int j = s.lo; j < s.hi; ++j) {
      const double* p = pos + 3 * order[j];
      r = fmax(r, sqrt((p[0] - c[0]) * (p[0] - c[0]) + (p[1] - c[1]) * (p[1] - c[1]) +
                       (p[2] - c[2]) * (p[2] - c[2])));
    }
    spheres[4 * k] = c[0]; spheres[4 * k + 1] = c[1]; spheres[4 * k + 2] = c[2];
    spheres[4 * k + 3] = r * (1.0 + 1e-12) + 1e-9;
  }
}

// Device order of a ligand's atoms: input order up to 64 atoms (one pass of
// the force loop covers them all), else the nested spatial order above.
inline void ligand_order(int m, const double* body, std::vector<int>& order) {
  order.resize(m);
  for (int i = 0; i < m; ++i) order[i] = i;
  if (m > 64) bisect_runs(body, order.data(), m);
}

}  // namespace pst
