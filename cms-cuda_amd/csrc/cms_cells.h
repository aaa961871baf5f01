#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cms_internal.h"

namespace cms {

struct GridInfo {
  double x0, y0, z0, h;  // grid origin and cell edge
  int nx, ny, nz;        // cells per axis
  int cell0;             // first cell of this complex in the batch-wide cell array
};

__global__ void k_grid_setup(int B, const int* aoff, const double4* atom, Settings s, GridInfo* grid,
                             double4* frame, int* ncell);
__global__ void k_grid_offsets(int B, const int* cscan, GridInfo* grid);
__global__ void k_cell_count(BatchView b, const GridInfo* grid, int* acell, int* ccount);
__global__ void k_cell_scatter(int n, const int* acell, const int* cstart, int* cursor, int* catom,
                               const double4* atom, double4* cxyz);
__global__ void k_attention_cells(BatchView b, Settings s, int all_atoms, int full, const GridInfo* grid,
                                  const int* cstart, const int* catom, const double4* cxyz, int8_t* atten,
                                  int8_t* access, int* need, unsigned long long* coinc, int* status);
__global__ void k_region_cells(BatchView b, Settings s, const GridInfo* grid, const int* cstart, const int* catom,
                               const double4* cxyz, const int8_t* atten, int* need);
__global__ void k_attention_cells_t(BatchView b, Settings s, int all_atoms, int full, const GridInfo* grid,
                                    const int* cstart, const int* catom, const double4* cxyz, int8_t* atten,
                                    int8_t* access, int* need, unsigned long long* coinc, int* status);
__global__ void k_region_cells_t(BatchView b, Settings s, const GridInfo* grid, const int* cstart,
                                 const int* catom, const double4* cxyz, const int8_t* atten, int* need);
__global__ void k_neighbors_cells(BatchView b, Settings s, Work w, const GridInfo* grid, const int* cstart,
                                  const int* catom, const double4* cxyz);

}  // namespace cms
