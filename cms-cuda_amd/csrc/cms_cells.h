#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cms_internal.h"

namespace cms {

struct GridInfo {
  double x0, y0, z0, h;  // grid origin and cell edge
  int nx, ny, nz;        // cells per axis
  int cell0;             // first cell of this complex in the batch-wide cell array
  double rmax;           // largest atom radius of the complex
};

// Cells of the 3x3x3 block around (ix, iy, iz) (bit k = 9 (dx + 1) + 3 (dy + 1) + dz + 1) whose
// box comes within lim of the point (x, y, z) of cell (ix, iy, iz): atoms
// farther than lim lie only in the other cells.  The point's offsets from its
// cell's faces are clamped at 0 and the test carries a 1e-6 A margin, far
// above the rounding of the offsets and of the cell index.
__device__ __forceinline__ uint32_t cells_within(const GridInfo& gi, double x, double y, double z, int ix, int iy,
                                                 int iz, double lim) {
  const double fx = x - (gi.x0 + ix * gi.h), fy = y - (gi.y0 + iy * gi.h), fz = z - (gi.z0 + iz * gi.h);
  const double gx[3] = {fmax(fx, 0.0), 0.0, fmax(gi.h - fx, 0.0)};
  const double gy[3] = {fmax(fy, 0.0), 0.0, fmax(gi.h - fy, 0.0)};
  const double gz[3] = {fmax(fz, 0.0), 0.0, fmax(gi.h - fz, 0.0)};
  const double l2 = (lim + 1e-6) * (lim + 1e-6);
  uint32_t mask = 0;
#pragma unroll
  for (int k = 0; k < 27; ++k) {
    const double a = gx[k / 9], b = gy[(k / 3) % 3], c = gz[k % 3];
    if ((a * a + b * b) + c * c <= l2) mask |= 1u << k;
  }
  return mask;
}

// The same test column by column: the run of cells z0 .. z1 (grid-clipped)
// of column (kx, ky) (0..2 each) of the block whose boxes come within lim
// (l2 = (lim + 1e-6)^2) of the point at offsets (fx, fy, fz) from the low
// faces of its cell (ix, iy, iz), with the arithmetic of cells_within, so
// the same cells; false when none does or the column is outside the grid.
// The column's middle cell is its nearest, so the run holds it.
__device__ __forceinline__ bool column_within(const GridInfo& gi, double fx, double fy, double fz, int ix, int iy,
                                              int iz, int kx, int ky, double l2, int* z0, int* z1) {
  const int cx = ix + kx - 1, cy = iy + ky - 1;
  if (cx < 0 || cx >= gi.nx || cy < 0 || cy >= gi.ny) return false;
  const double a = kx == 0 ? fmax(fx, 0.0) : (kx == 1 ? 0.0 : fmax(gi.h - fx, 0.0));
  const double b = ky == 0 ? fmax(fy, 0.0) : (ky == 1 ? 0.0 : fmax(gi.h - fy, 0.0));
  const double gxy = a * a + b * b;
  if (!(gxy + 0.0 * 0.0 <= l2)) return false;
  const double cl = fmax(fz, 0.0), ch = fmax(gi.h - fz, 0.0);
  *z0 = max(iz - ((gxy + cl * cl <= l2) ? 1 : 0), 0);
  *z1 = min(iz + ((gxy + ch * ch <= l2) ? 1 : 0), gi.nz - 1);
  return true;
}

// The same test for the one cell k of the block (one cell per lane of a
// wavefront, the mask then taken by ballot).
__device__ __forceinline__ bool cell_within(const GridInfo& gi, double x, double y, double z, int ix, int iy, int iz,
                                            double lim, int k) {
  const double fx = x - (gi.x0 + ix * gi.h), fy = y - (gi.y0 + iy * gi.h), fz = z - (gi.z0 + iz * gi.h);
  const int kx = k / 9, ky = (k / 3) % 3, kz = k % 3;
  const double a = kx == 0 ? fmax(fx, 0.0) : (kx == 1 ? 0.0 : fmax(gi.h - fx, 0.0));
  const double b = ky == 0 ? fmax(fy, 0.0) : (ky == 1 ? 0.0 : fmax(gi.h - fy, 0.0));
  const double c = kz == 0 ? fmax(fz, 0.0) : (kz == 1 ? 0.0 : fmax(gi.h - fz, 0.0));
  const double l2 = (lim + 1e-6) * (lim + 1e-6);
  return (a * a + b * b) + c * c <= l2;
}

__global__ void k_grid_setup(int B, const int* aoff, const double4* atom, Settings s, GridInfo* grid,
                             double4* frame, int* ncell);
__global__ void k_grid_offsets(int B, const int* cscan, GridInfo* grid);
__global__ void k_cell_count_atoms(BatchView b, const GridInfo* grid, int* acell, int* ccount);
__global__ void k_cell_scatter_atoms(int n, const int* acell, const int* cstart, int* cursor, int* catom,
                                     const double4* atom, double4* cxyz);
__global__ void k_cell_build(BatchView b, const GridInfo* grid, int* ccount, int* acell, int* arank, int* cstart,
                             int* catom, double4* cxyz);
__global__ void k_attention_cells(BatchView b, Settings s, int all_atoms, int full, const GridInfo* grid,
                                  const int* cstart, const int* catom, const double4* cxyz, int8_t* atten,
                                  int8_t* catt, int8_t* access, int* need, unsigned long long* coinc, int* status);
__global__ void k_region_cells(BatchView b, Settings s, const GridInfo* grid, const int* cstart, const int* catom,
                               const double4* cxyz, const int8_t* atten, const int* list, const int* list_n,
                               int* need);
__global__ void k_attention_cells_t(BatchView b, Settings s, int all_atoms, int full, const GridInfo* grid,
                                    const int* cstart, const int* catom, const double4* cxyz, int8_t* atten,
                                    int8_t* catt, int8_t* access, int* need, unsigned long long* coinc, int* status,
                                    int* list, int* list_n);
// catt: the atoms' attention values in cell-slot order (the attention pass
// writes them beside atten)
__global__ void k_neighbors_cells(BatchView b, Settings s, Work w, const GridInfo* grid, const int* cstart,
                                  const int* catom, const double4* cxyz, const int8_t* catt);

}  // namespace cms
