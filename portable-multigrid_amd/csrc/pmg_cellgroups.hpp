// pmg_cellgroups.hpp -- host-side set-up of the batched cell schedules:
// grouping cells into compact batches and the per-batch DoF patches.
#pragma once

#include <cstdint>
#include <memory>
#include <new>
#include <utility>
#include <vector>

namespace pmg {

// Order of the cells such that every run of E consecutive cells is a compact
// group (a 2x2x2 block on a structured mesh for E = 8): starting from the
// first unassigned cell in caller order, repeatedly add the unassigned cell
// with the largest number of (DoF, group cell) incidences, ties to the lowest
// index; a group that runs out of neighbours is filled with the next
// unassigned cells, so groups and batches stay aligned.  map: [C][nm3],
// 0xFFFFFFFF entries ignored.  Returns order[internal] = caller cell.
std::vector<uint32_t> group_cells(int64_t n_cells, int64_t nm3, const uint32_t *map,
                                  int64_t n_dofs, int E);

// Patches of consecutive batches of E cells (internal order, map already
// permuted): pdof[b * E nm3 + i] = i-th smallest DoF of batch b's cells, bit
// 31 set when a cell outside the batch also has it; pcount[b]; pidx[c][i] =
// patch index of cell-local DoF i (0xFFFF for constrained entries).
// Allocator whose vectors leave new elements uninitialised, so that the
// threads that fill them first-touch their own pages.
template <typename T>
struct NoInitAllocator : std::allocator<T> {
  template <typename U> struct rebind { using other = NoInitAllocator<U>; };
  NoInitAllocator() = default;
  template <typename U> NoInitAllocator(const NoInitAllocator<U> &) {}
  template <typename U> void construct(U *) {}
  template <typename U, typename... A> void construct(U *p, A &&...a) {
    ::new (static_cast<void *>(p)) U(std::forward<A>(a)...);
  }
};
template <typename T> using RawVector = std::vector<T, NoInitAllocator<T>>;

struct CellPatches {
  RawVector<uint32_t> pdof, pcount;
  RawVector<uint16_t> pidx;
  int64_t n_shared = 0, n_owned = 0;  // over all batches
};
CellPatches build_patches(int64_t n_cells, int64_t nm3, const uint32_t *map, int64_t n_dofs,
                          int E);

}  // namespace pmg
