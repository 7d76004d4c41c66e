// aph_poisson_tree.hpp -- the composite Poisson operator of a multilevel mesh, assembled on
// the host from the mesh tree (aph_amr_tree.hip) for the device solver (aph_poisson.hip).
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

#include "../../include/apophis_hydro.h"

namespace aph {

// Leaf cells are numbered in block order, cell (i, j, k) of block b at
// b * nx0 nx1 nx2 + (k nx1 + j) nx0 + i.  Operator rows: one per cell.  Neighbour rows:
// (2 d + side) * n + c, the value beyond cell c's face on that side of axis d as a linear
// map of the cell values (empty on a physical domain face).
struct PoissonTreeCsr {
  int64_t n = 0;
  std::vector<int64_t> rowptr, nb_rowptr;
  std::vector<int32_t> col, nb_col;
  std::vector<double> val, nb_val;
  std::vector<int32_t> wall;  // per block: bit 2 d + side set on a physical domain face
};

int poisson_tree_csr(const aph_block *blocks, int32_t nblocks, const int32_t nbase[3],
                     const int32_t *level, const int64_t *loc, int32_t flux_correct,
                     int32_t prolongation, double alpha, PoissonTreeCsr &out, char *err,
                     size_t err_len);

// One coarser level of the multigrid preconditioner: the same problem re-assembled on a
// coarser mesh (the tree cut one level lower, or below the base level each block with half
// its cells per axis), and the map from the next finer level's cells to its own.
struct PoissonMgLevel {
  int64_t n = 0;
  std::vector<int64_t> rowptr;
  std::vector<int32_t> col;
  std::vector<double> val;
  std::vector<int32_t> cmap;     // per cell of the finer level: the cell of this level
  std::vector<int64_t> rptr;     // per cell of this level: its finer cells (CSR) ...
  std::vector<int32_t> rcol;
  std::vector<double> rval;      // ... weighted 1 / their count (the mean)
};

// The coarser levels below the mesh of poisson_tree_csr (the same arguments), finest first,
// down to a level of at most max_coarse cells or blocks of 2 cells per axis.
int poisson_tree_levels(const aph_block *blocks, int32_t nblocks, const int32_t nbase[3],
                        const int32_t *level, const int64_t *loc, int32_t flux_correct,
                        int32_t prolongation, double alpha, int64_t max_coarse,
                        std::vector<PoissonMgLevel> &out, char *err, size_t err_len);

}  // namespace aph
