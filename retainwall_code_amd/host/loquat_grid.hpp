// loquat_grid.hpp -- LoquatGrid surface (loquat_grid.h:29-67) over the C ABI.
//
// The cell arrays live inside the lq_ctx; these methods keep the reference's
// names and call the matching entry points.  Any failure prints the library
// message and exits, like checkCudaErrors in the reference.
#pragma once
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "retainwall.h"

inline void lq_check(lq_ctx* ctx, int st, const char* what) {
  if (st != LQ_OK) {
    std::fprintf(stderr, "%s failed (%d): %s\n", what, st, ctx ? lq_last_error(ctx) : "");
    std::exit(EXIT_FAILURE);
  }
}

class LoquatGrid {
 public:
  explicit LoquatGrid(lq_ctx* ctx) : ctx_(ctx) {}
  // GridConfig (loquat_grid.cu:59-68): cell arrays are allocated by lq_create.
  void GridConfig(const lq_params*) {}
  // NeighborSearch (loquat_grid.cu:147-191).
  void NeighborSearch() { lq_check(ctx_, lq_neighbor_search(ctx_), "NeighborSearch"); }
  // GetCellBeg / GetCellEnd (loquat_grid.h:54-55), copied to the host.
  std::vector<int32_t> GetCellBeg(int ncell) { return get(LQ_F_CELL_BEG, ncell); }
  std::vector<int32_t> GetCellEnd(int ncell) { return get(LQ_F_CELL_END, ncell); }

 private:
  std::vector<int32_t> get(int f, int n) {
    std::vector<int32_t> v(static_cast<size_t>(n), 0);
    lq_check(ctx_, lq_get_field(ctx_, f, v.data(), v.size() * 4), "lq_get_field");
    return v;
  }
  lq_ctx* ctx_;
};
