// pmg_cellgroups.cpp -- see pmg_cellgroups.hpp.
#include "pmg_cellgroups.hpp"

#include <algorithm>
#include <stdexcept>
#include <exception>
#include <thread>
#include <utility>

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

CellPatches build_patches(int64_t n_cells, int64_t nm3, const uint32_t *map, int64_t n_dofs,
                          int E) {
  if (int64_t(E) * nm3 >= 0xFFFF) throw std::length_error("patch index exceeds 16 bits");
  CellPatches P;
  const int64_t nb = (n_cells + E - 1) / E, pmax = int64_t(E) * nm3;
  RawVector<int32_t> mult(static_cast<size_t>(n_dofs));  // cells per DoF
  P.pdof.resize(static_cast<size_t>(nb * pmax));
  P.pcount.resize(static_cast<size_t>(nb));
  P.pidx.resize(static_cast<size_t>(n_cells * nm3));
  // Worker threads over contiguous cell (batch) ranges: first the DoF
  // multiplicities (relaxed atomic increments), then the batches, which are
  // independent given those; each writes only its own slices.  Per batch, the
  // sorted (DoF, entry) keys give the patch (ascending DoFs, so the
  // write-back coalesces), each DoF's count in the batch and every entry's slot.
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const int64_t n_threads = std::min<int64_t>(std::min<unsigned>(hw, 16u), std::max<int64_t>(nb / 64, 1));
  // Exception-safe: a worker's exception is caught in the worker and
  // rethrown here after every started thread has been joined; if a thread
  // cannot be started, its share runs on the calling thread.
  auto run = [n_threads](auto &&fn) {
    std::vector<std::exception_ptr> err(static_cast<size_t>(n_threads));
    auto guarded = [&fn, &err](int64_t t) {
      try {
        fn(t);
      } catch (...) {
        err[static_cast<size_t>(t)] = std::current_exception();
      }
    };
    std::vector<std::thread> pool;
    std::vector<int64_t> inline_shares;
    pool.reserve(static_cast<size_t>(n_threads));
    for (int64_t t = 1; t < n_threads; ++t) {
      try {
        pool.emplace_back(guarded, t);
      } catch (...) {
        inline_shares.push_back(t);
      }
    }
    guarded(int64_t(0));
    for (int64_t t : inline_shares) guarded(t);
    for (auto &th : pool) th.join();
    for (auto &e : err)
      if (e) std::rethrow_exception(e);
  };
  run([&](int64_t t) {
    const int64_t d0 = n_dofs * t / n_threads, d1 = n_dofs * (t + 1) / n_threads;
    std::fill(mult.begin() + d0, mult.begin() + d1, 0);
  });
  run([&](int64_t t) {
    const int64_t e0 = n_cells * nm3 * t / n_threads, e1 = n_cells * nm3 * (t + 1) / n_threads;
    for (int64_t e = e0; e < e1; ++e)
      if (map[e] != kNoDof) __atomic_fetch_add(&mult[map[e]], 1, __ATOMIC_RELAXED);
  });
  std::vector<int64_t> shared(static_cast<size_t>(n_threads), 0), owned(shared);
  run([&](int64_t t) {
    const int64_t b0 = nb * t / n_threads, b1 = nb * (t + 1) / n_threads;
    std::vector<uint64_t> ent;  // DoF << 16 | entry offset in the batch (< 2^16)
    ent.reserve(static_cast<size_t>(pmax));
    for (int64_t b = b0; b < b1; ++b) {
      const int64_t c0 = b * E, c1 = std::min(n_cells, c0 + E);
      std::fill(P.pidx.begin() + c0 * nm3, P.pidx.begin() + c1 * nm3, uint16_t(0xFFFF));
      ent.clear();
      for (int64_t e = c0 * nm3; e < c1 * nm3; ++e)
        if (map[e] != kNoDof) ent.push_back(uint64_t(map[e]) << 16 | uint64_t(e - c0 * nm3));
      std::sort(ent.begin(), ent.end());
      uint32_t n = 0;
      for (size_t k = 0; k < ent.size();) {
        const uint32_t d = static_cast<uint32_t>(ent[k] >> 16);
        size_t r = k;
        while (r < ent.size() && (ent[r] >> 16) == d) {
          P.pidx[c0 * nm3 + static_cast<int64_t>(ent[r] & 0xFFFFu)] = static_cast<uint16_t>(n);
          ++r;
        }
        const bool sh = static_cast<int64_t>(r - k) < mult[d];
        P.pdof[b * pmax + n] = d | (sh ? 0x80000000u : 0u);
        ++(sh ? shared[t] : owned[t]);
        ++n;
        k = r;
      }
      P.pcount[b] = n;
      std::fill(P.pdof.begin() + b * pmax + n, P.pdof.begin() + (b + 1) * pmax, 0u);
    }
  });
  for (int64_t t = 0; t < n_threads; ++t) {
    P.n_shared += shared[t];
    P.n_owned += owned[t];
  }
  return P;
}

}  // namespace pmg
