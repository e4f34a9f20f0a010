// pmg_cellgroups.cpp -- see pmg_cellgroups.hpp.
#include "pmg_cellgroups.hpp"

#include <algorithm>

namespace pmg {

namespace {
constexpr uint32_t kNoDof = 0xFFFFFFFFu;

// DoF -> cells incidence in CSR form
void dof_cells(int64_t n_cells, int64_t nm3, const uint32_t *map, int64_t n_dofs,
               std::vector<int64_t> &off, std::vector<uint32_t> &adj) {
  off.assign(static_cast<size_t>(n_dofs + 1), 0);
  for (int64_t e = 0; e < n_cells * nm3; ++e)
    if (map[e] != kNoDof) ++off[map[e] + 1];
  for (int64_t d = 0; d < n_dofs; ++d) off[d + 1] += off[d];
  adj.resize(static_cast<size_t>(off[n_dofs]));
  std::vector<int64_t> fill(off.begin(), off.end() - 1);
  for (int64_t c = 0; c < n_cells; ++c)
    for (int64_t i = 0; i < nm3; ++i) {
      const uint32_t d = map[c * nm3 + i];
      if (d != kNoDof) adj[fill[d]++] = static_cast<uint32_t>(c);
    }
}
}  // namespace

std::vector<uint32_t> group_cells(int64_t n_cells, int64_t nm3, const uint32_t *map,
                                  int64_t n_dofs, int E) {
  std::vector<int64_t> off;
  std::vector<uint32_t> adj;
  dof_cells(n_cells, nm3, map, n_dofs, off, adj);
  std::vector<uint32_t> order;
  order.reserve(static_cast<size_t>(n_cells));
  std::vector<char> assigned(static_cast<size_t>(n_cells), 0);
  std::vector<int32_t> score(static_cast<size_t>(n_cells), 0);
  std::vector<uint32_t> touched;
  int64_t seed = 0, scan = 0;
  auto add = [&](uint32_t c) {
    assigned[c] = 1;
    order.push_back(c);
    for (int64_t i = 0; i < nm3; ++i) {
      const uint32_t d = map[int64_t(c) * nm3 + i];
      if (d == kNoDof) continue;
      for (int64_t k = off[d]; k < off[d + 1]; ++k) {
        const uint32_t o = adj[k];
        if (assigned[o]) continue;
        if (score[o] == 0) touched.push_back(o);
        ++score[o];
      }
    }
  };
  while (static_cast<int64_t>(order.size()) < n_cells) {
    while (assigned[seed]) ++seed;
    const size_t start = order.size();
    add(static_cast<uint32_t>(seed));
    while (static_cast<int>(order.size() - start) < E) {
      int64_t best = -1;
      int32_t best_score = 0;
      for (uint32_t o : touched)
        if (!assigned[o] && (score[o] > best_score || (score[o] == best_score && o < best))) {
          best = o;
          best_score = score[o];
        }
      if (best < 0) break;
      add(static_cast<uint32_t>(best));
    }
    // no connected candidate left: complete the batch with the next cells
    scan = std::max(scan, seed);
    while (static_cast<int>(order.size() - start) < E &&
           static_cast<int64_t>(order.size()) < n_cells) {
      while (assigned[scan]) ++scan;
      add(static_cast<uint32_t>(scan));
    }
    for (uint32_t o : touched) score[o] = 0;
    touched.clear();
  }
  return order;
}

}  // namespace pmg
