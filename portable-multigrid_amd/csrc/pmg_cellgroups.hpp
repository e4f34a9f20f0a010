// pmg_cellgroups.hpp -- host-side grouping of cells into compact batches
// (the optional grouped cell order).
#pragma once

#include <cstdint>
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

}  // namespace pmg
