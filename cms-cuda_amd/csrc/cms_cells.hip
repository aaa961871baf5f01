// Cell lists: per-complex uniform grids used by the attention sweep, the
// coincident-atom check, the interface region and the neighbour rows.
//
// Cell edge h >= max(sep, 2 r_max + 2 rp) of the complex (slightly padded so
// that rounding of the cell index cannot push two atoms closer than h into
// non-adjacent cells), so every partner an atom can have -- other-molecule
// atoms closer than sep, same-molecule atoms inside the bridge distance
// r_i + r_j + 2rp, coincident atoms -- lies in the 27 cells around it.
// Each cell holds two atom ranges, molecule 0 then molecule 1 (slot
// 2 * cell + molecule of the count / start arrays), so a sweep reads only
// the molecule it needs.  Results do not depend on the order atoms are
// visited in: minima, "any" tests, atomic minima of (i, j) keys, idempotent
// flags, and rows sorted afterwards.
#include "cms_cells.h"
#include "cms_device.h"
#include "cms_internal.h"

namespace cms {

// Per complex: bounding box, cell edge, grid dims (capped at 8 N + 64 cells
// by growing the edge), and the FP32-search frame (centre, R).
__global__ __launch_bounds__(256) void k_grid_setup(int B, const int* __restrict__ aoff,
                                                    const double4* __restrict__ atom, Settings s,
                                                    GridInfo* __restrict__ grid, double4* __restrict__ frame,
                                                    int* __restrict__ ncell) {
  __shared__ double red[7][256];
  for (int c = blockIdx.x; c < B; c += gridDim.x) {
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    double rmax = 0.0;
    for (int i = aoff[c] + threadIdx.x; i < aoff[c + 1]; i += blockDim.x) {
      const double4 a = atom[i];
      mn[0] = fmin(mn[0], a.x); mn[1] = fmin(mn[1], a.y); mn[2] = fmin(mn[2], a.z);
      mx[0] = fmax(mx[0], a.x); mx[1] = fmax(mx[1], a.y); mx[2] = fmax(mx[2], a.z);
      rmax = fmax(rmax, a.w);
    }
    for (int k = 0; k < 3; ++k) { red[k][threadIdx.x] = mn[k]; red[3 + k][threadIdx.x] = mx[k]; }
    red[6][threadIdx.x] = rmax;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if (threadIdx.x < o) {
        for (int k = 0; k < 3; ++k) red[k][threadIdx.x] = fmin(red[k][threadIdx.x], red[k][threadIdx.x + o]);
        for (int k = 3; k < 7; ++k) red[k][threadIdx.x] = fmax(red[k][threadIdx.x], red[k][threadIdx.x + o]);
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      const int n = aoff[c + 1] - aoff[c];
      GridInfo gi;
      double lo[3] = {red[0][0], red[1][0], red[2][0]};
      double hi[3] = {red[3][0], red[4][0], red[5][0]};
      if (n == 0 || !(hi[0] >= lo[0]) || !(hi[1] >= lo[1]) || !(hi[2] >= lo[2])) {
        lo[0] = lo[1] = lo[2] = 0.0;
        hi[0] = hi[1] = hi[2] = 0.0;
      }
      const double rm = red[6][0];
      double h = fmax(s.sep, 2.0 * rm + 2.0 * s.rp);
      h = h * (1.0 + 1e-9) + 1e-9;
      if (!(h > 0.0)) h = 1.0;
      int d[3];
      const int64_t cap = 8ll * n + 64;
      for (int it = 0; it < 64; ++it) {
        int64_t tot = 1;
        for (int k = 0; k < 3; ++k) {
          const double e = (hi[k] - lo[k]) / h;
          d[k] = e < 1e6 ? (int)floor(e) + 1 : 1000001;
          tot *= d[k];
        }
        if (tot <= cap) break;
        h *= 1.5;
      }
      gi.x0 = lo[0]; gi.y0 = lo[1]; gi.z0 = lo[2]; gi.h = h;
      gi.nx = d[0]; gi.ny = d[1]; gi.nz = d[2];
      gi.cell0 = 0;
      gi.rmax = rm;
      grid[c] = gi;
      ncell[c] = d[0] * d[1] * d[2];
      const double cx = 0.5 * (lo[0] + hi[0]), cy = 0.5 * (lo[1] + hi[1]), cz = 0.5 * (lo[2] + hi[2]);
      const double half = fmax(fmax(hi[0] - cx, hi[1] - cy), hi[2] - cz);
      frame[c] = make_double4(cx, cy, cz, half + rm + 2.0 * s.rp + 1.0);
    }
    __syncthreads();
  }
}

__global__ void k_grid_offsets(int B, const int* __restrict__ cscan, GridInfo* __restrict__ grid) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < B; c += gridDim.x * blockDim.x)
    grid[c].cell0 = cscan[c];
}

__device__ __forceinline__ void cell_coords(const GridInfo& gi, double x, double y, double z, int* ix, int* iy,
                                            int* iz) {
  *ix = min(max((int)floor((x - gi.x0) / gi.h), 0), gi.nx - 1);
  *iy = min(max((int)floor((y - gi.y0) / gi.h), 0), gi.ny - 1);
  *iz = min(max((int)floor((z - gi.z0) / gi.h), 0), gi.nz - 1);
}

// Cell counts and the cell-sorted atom lists for small batches, one thread
// per atom (wavefront-aggregated device atomics).
__global__ void k_cell_count_atoms(BatchView b, const GridInfo* __restrict__ grid, int* __restrict__ acell,
                                   int* __restrict__ ccount) {
  CMS_WAVE_LOOP(g, b.ntot) {
    int cell = 0;
    if (g < b.ntot) {
      const GridInfo gi = grid[b.acplx[g]];
      const double4 a = b.atom[g];
      int ix, iy, iz;
      cell_coords(gi, a.x, a.y, a.z, &ix, &iy, &iz);
      cell = 2 * (gi.cell0 + (ix * gi.ny + iy) * gi.nz + iz) + b.amol[g];
      acell[g] = cell;
    }
    wave_counter_inc(ccount, cell, g < b.ntot);
  }
}

__global__ void k_cell_scatter_atoms(int n, const int* __restrict__ acell, const int* __restrict__ cstart,
                                     int* __restrict__ cursor, int* __restrict__ catom,
                                     const double4* __restrict__ atom, double4* __restrict__ cxyz) {
  CMS_WAVE_LOOP(g, n) {
    const int cell = g < n ? acell[g] : 0;
    const int slot = wave_counter_add(cursor, cell, g < n);
    if (g < n) {
      catom[cstart[cell] + slot] = (int)g;
      cxyz[cstart[cell] + slot] = atom[g];
    }
  }
}

constexpr int CELL_LDS = 4096;
constexpr int CELL_NT = 512;    // threads per complex
constexpr int CELL_U = 4;       // atoms per thread per round (loads in flight)

// Count, scan and scatter of the large-batch cell lists in one pass, one
// workgroup per complex.  A complex's atoms occupy the atom slots
// aoff[c] .. aoff[c + 1] and its cell slots are consecutive, so the cell
// starts are the complex's atom offset plus a scan over its own cells: no
// batch-wide scan.  Each atom's rank within its cell comes from the
// counting atomic (LDS counters when the grid has at most CELL_LDS cell
// slots, otherwise device atomics on the zero-filled ccount); within a cell
// the order is whatever the atomics give, which nothing reads.
__global__ __launch_bounds__(CELL_NT) void k_cell_build(BatchView b, const GridInfo* __restrict__ grid,
                                                        int* __restrict__ ccount, int* __restrict__ acell,
                                                        int* __restrict__ arank, int* __restrict__ cstart,
                                                        int* __restrict__ catom, double4* __restrict__ cxyz) {
  __shared__ int cnt[CELL_LDS];
  __shared__ int wsum[CELL_NT / WAVE];
  for (int c = blockIdx.x; c < b.B; c += gridDim.x) {
    const GridInfo gi = grid[c];
    const int ns = 2 * gi.nx * gi.ny * gi.nz;
    const int base = 2 * gi.cell0;
    const bool lds = ns <= CELL_LDS;
    if (lds)
      for (int k = threadIdx.x; k < ns; k += CELL_NT) cnt[k] = 0;
    __syncthreads();
    const int a0 = b.aoff[c], a1 = b.aoff[c + 1];
    // ranks
    for (int g0 = a0 + threadIdx.x; g0 < a1; g0 += CELL_U * CELL_NT) {
      double4 a[CELL_U];
      int8_t mol[CELL_U];
#pragma unroll
      for (int u = 0; u < CELL_U; ++u) {
        const int g = g0 + u * CELL_NT;
        if (g < a1) { a[u] = b.atom[g]; mol[u] = b.amol[g]; }
      }
#pragma unroll
      for (int u = 0; u < CELL_U; ++u) {
        const int g = g0 + u * CELL_NT;
        if (g < a1) {
          int ix, iy, iz;
          cell_coords(gi, a[u].x, a[u].y, a[u].z, &ix, &iy, &iz);
          const int k = 2 * ((ix * gi.ny + iy) * gi.nz + iz) + mol[u];
          acell[g] = k;
          arank[g] = lds ? atomicAdd(&cnt[k], 1) : atomicAdd(&ccount[base + k], 1);
        }
      }
    }
    __syncthreads();
    // cell starts: the complex's atom offset + an exclusive scan of its
    // counts (kept in cnt / ccount for the placement)
    int carry = a0;
    for (int k0 = 0; k0 < ns; k0 += CELL_NT) {
      const int k = k0 + threadIdx.x;
      const int v = k < ns ? (lds ? cnt[k] : __hip_atomic_load(&ccount[base + k], __ATOMIC_RELAXED,
                                                                  __HIP_MEMORY_SCOPE_AGENT))
                           : 0;
      const int inc = wave_incl_scan(v);
      if (lane() == WAVE - 1) wsum[threadIdx.x / WAVE] = inc;
      __syncthreads();
      int off = 0, tot = 0;
      for (int w = 0; w < CELL_NT / WAVE; ++w) {
        off += w < (int)(threadIdx.x / WAVE) ? wsum[w] : 0;
        tot += wsum[w];
      }
      if (k < ns) {
        const int start = carry + off + inc - v;
        cstart[base + k] = start;
        if (lds)
          cnt[k] = start;
        else
          ccount[base + k] = start;
      }
      carry += tot;
      __syncthreads();
    }
    // the slot after the complex's last one: the next complex's first
    // start, and after the last complex the batch total
    if (threadIdx.x == 0) cstart[base + ns] = a1;
    __syncthreads();
    for (int g = a0 + threadIdx.x; g < a1; g += CELL_NT) {
      const int k = acell[g];
      catom[(lds ? cnt[k] : ccount[base + k]) + arank[g]] = g;
    }
    // coordinates in slot order (coalesced stores) once this workgroup's
    // index stores are visible to it
    __syncthreads();
    for (int t0 = a0 + threadIdx.x; t0 < a1; t0 += CELL_U * CELL_NT) {
      int src[CELL_U];
#pragma unroll
      for (int u = 0; u < CELL_U; ++u) {
        const int t = t0 + u * CELL_NT;
        src[u] = t < a1 ? catom[t] : 0;
      }
#pragma unroll
      for (int u = 0; u < CELL_U; ++u) {
        const int t = t0 + u * CELL_NT;
        if (t < a1) cxyz[t] = b.atom[src[u]];
      }
    }
    __syncthreads();
  }
}

// Molecule selector of a sweep: 0 / 1 one molecule's range, 2 both.
__device__ __forceinline__ int sel_lo(int sel) { return sel == 2 ? 0 : sel; }
__device__ __forceinline__ int sel_hi(int sel) { return sel == 2 ? 2 : sel + 1; }

// Cell k of the 3x3x3 block around (ix, iy, iz) (k = 9 dx + 3 dy + dz with
// offsets -1..1), or -1 outside the grid.
__device__ __forceinline__ int cell_of(const GridInfo& gi, int ix, int iy, int iz, int k) {
  const int cx = ix + k / 9 - 1, cy = iy + (k / 3) % 3 - 1, cz = iz + k % 3 - 1;
  if (cx < 0 || cx >= gi.nx || cy < 0 || cy >= gi.ny || cz < 0 || cz >= gi.nz) return -1;
  return gi.cell0 + (cx * gi.ny + cy) * gi.nz + cz;
}

// Cells that can hold an atom within COINC_R of x: the atom's own cell,
// and a neighbour only across a face, edge or corner closer than COINC_R
// (bit k of the mask = cell k of the 3x3x3 block).
constexpr double COINC_R = 0.0101;  // > 0.01, the coincident-atom distance
__device__ __forceinline__ uint32_t coincident_cells(const GridInfo& gi, double x, double y, double z, int ix,
                                                     int iy, int iz) {
  const double fx = x - (gi.x0 + ix * gi.h), fy = y - (gi.y0 + iy * gi.h), fz = z - (gi.z0 + iz * gi.h);
  const bool lx = fx <= COINC_R, hx = gi.h - fx <= COINC_R;
  const bool ly = fy <= COINC_R, hy = gi.h - fy <= COINC_R;
  const bool lz = fz <= COINC_R, hz = gi.h - fz <= COINC_R;
  uint32_t mask = 0;
  for (int k = 0; k < 27; ++k) {
    const int dx = k / 9 - 1, dy = (k / 3) % 3 - 1, dz = k % 3 - 1;
    const bool okx = dx == 0 || (dx < 0 ? lx : hx);
    const bool oky = dy == 0 || (dy < 0 ? ly : hy);
    const bool okz = dz == 0 || (dz < 0 ? lz : hz);
    if (okx && oky && okz) mask |= 1u << k;
  }
  return mask;
}

// Visits the cells of the 3x3x3 block around (ix, iy, iz) whose bit is set
// in mask.
#define FOR_NEIGHBOR_CELLS(gi, ix, iy, iz, mask, cellvar)                             \
  for (int k_ = 0; k_ < 27; ++k_)                                                    \
    if (((mask) >> k_) & 1u)                                                         \
      for (int cellvar = cell_of((gi), (ix), (iy), (iz), k_), once_ = 1; once_ && cellvar >= 0; once_ = 0)

// Atom ranges of molecule slot ms (0 or 1) in the cells z0 .. z1 (at most
// three, inside the grid) of column (cx, cy): consecutive cells, so their
// bounds come from one run of cstart and are loaded together; col_index
// maps a flattened position over the ranges to a cell-list slot.
struct ColRanges {
  int b[3], n[3], total;
};
__device__ __forceinline__ ColRanges column_ranges(const GridInfo& gi, const int* __restrict__ cstart, int cx,
                                                   int cy, int z0, int z1, int ms) {
  ColRanges r;
  r.total = 0;
  const int base = 2 * (gi.cell0 + (cx * gi.ny + cy) * gi.nz + z0) + ms;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const bool ok = z0 + k <= z1;
    const int bb = ok ? cstart[base + 2 * k] : 0, ee = ok ? cstart[base + 2 * k + 1] : 0;
    r.b[k] = bb;
    r.n[k] = ee - bb;
    r.total += ee - bb;
  }
  return r;
}
__device__ __forceinline__ int col_index(const ColRanges& r, int t) {
  return t < r.n[0] ? r.b[0] + t : (t < r.n[0] + r.n[1] ? r.b[1] + (t - r.n[0]) : r.b[2] + (t - r.n[0] - r.n[1]));
}

// Flattened candidate list of the cells around (ix, iy, iz) selected by
// mask, molecule range sel: cbeg[k] / cpre[k] are the start in catom / cxyz and the
// prefix count of cell k; cpre[27] is the total, cpre[28 .. 31] INT_MAX
// (sentinels for cell_slot's search).  Wave-cooperative (lanes 0..26 each
// own a cell).
constexpr int CPRE_N = 32;   // cpre entries: 27 prefixes, the total, three sentinels
__device__ __forceinline__ void cells_around(const GridInfo& gi, int ix, int iy, int iz,
                                             const int* __restrict__ cstart, int sel, uint32_t mask, int* cbeg,
                                             int* cpre) {
  const int k = lane();
  int cnt = 0, beg = 0;
  if (k < 27 && ((mask >> k) & 1u)) {
    const int cell = cell_of(gi, ix, iy, iz, k);
    if (cell >= 0) {
      beg = cstart[2 * cell + sel_lo(sel)];
      cnt = cstart[2 * cell + sel_hi(sel)] - beg;
    }
  }
  const int inc = wave_incl_scan(cnt);
  if (k < 27) { cbeg[k] = beg; cpre[k] = inc - cnt; }
  if (k == 26) cpre[27] = inc;
  if (k >= 28 && k < CPRE_N) cpre[k] = INT_MAX;
  __syncthreads();
}

// The cell of flattened index idx < cpre[27] is the last k with cpre[k] <=
// idx (it is non-empty: cpre[k + 1] > idx); a fixed five-step search over
// the 32 entries (the sentinels are never <= idx), fully unrolled and
// branch-free, so a wave issues no loop control and no exec-mask updates.
__device__ __forceinline__ int cell_slot(const int* cbeg, const int* cpre, int idx) {
  int lo = 0;
#pragma unroll
  for (int step = 16; step > 0; step >>= 1) lo = cpre[lo + step] <= idx ? lo + step : lo;
  return cbeg[lo] + idx - cpre[lo];
}

// Attention (_core.py:1150-1195) and coincident atoms (_core.py:1285-1293):
// one wavefront per atom.  The minimum distance to the other molecule reads
// only that molecule's atoms in the 27 cells; coincident partners (same
// molecule, within 0.01 A) are looked for in the atom's own cell and in the
// neighbour cells within 0.01 A of it.  Also seeds the interface-region
// flag (non-blockers; every atom when full != 0).  Atoms are visited in
// cell order so that consecutive workgroups read the same cells.
// Returns whether the atom is a non-blocker.
__device__ __forceinline__ bool attention_store(int g, int sidx, int c, double mn, unsigned long long ckey,
                                                int all_atoms, int full, const Settings& s, int8_t* atten,
                                                int8_t* catt, int8_t* access, int* need,
                                                unsigned long long* coinc, int* status) {
  // atten, access and need come pre-filled in atom order (atten: blocker,
  // or buried in all-atoms mode; access 0; need 1 in full mode, else 0), so
  // only the non-blockers of a normal run are stored here: few scattered
  // stores instead of three per atom
  const double dmin = (mn == INFINITY) ? 99999.0 : sqrt(mn);
  const bool nonblocker = !all_atoms && !(dmin >= s.sep);
  if (nonblocker) {
    atten[g] = ATTEN_BURIED;
    catt[sidx] = ATTEN_BURIED;   // the same in cell-slot order
    if (!full) need[g] = 1;
  }
  if (ckey != ~0ull) {
    atomicMin(&coinc[c], ckey);
    atomicOr(&status[c], ST_COINCIDENT);
  }
  return nonblocker;
}

__global__ __launch_bounds__(64) void k_attention_cells(BatchView b, Settings s, int all_atoms, int full,
                                                        const GridInfo* __restrict__ grid,
                                                        const int* __restrict__ cstart,
                                                        const int* __restrict__ catom,
                                                        const double4* __restrict__ cxyz,
                                                        int8_t* __restrict__ atten, int8_t* __restrict__ catt,
                                                        int8_t* __restrict__ access,
                                                        int* __restrict__ need,
                                                        unsigned long long* __restrict__ coinc,
                                                        int* __restrict__ status) {
  __shared__ int cbeg[27], cpre[CPRE_N];
  CMS_UNIT_LOOP(sidx, b.ntot) {
    const int g = catom[sidx];
    const int c = b.acplx[g];
    const int a0 = b.aoff[c];
    const int m = b.amol[g];
    const GridInfo gi = grid[c];
    const double4 ai = cxyz[sidx];
    const V3 xi = ld3(ai);
    int ix, iy, iz;
    cell_coords(gi, ai.x, ai.y, ai.z, &ix, &iy, &iz);
    double mn = INFINITY;
    if (!all_atoms) {
      cells_around(gi, ix, iy, iz, cstart, 1 - m, 0x7ffffffu, cbeg, cpre);
      const int T = cpre[27];
      for (int idx = lane(); idx < T; idx += WAVE) {
        const V3 d = xi - ld3(cxyz[cell_slot(cbeg, cpre, idx)]);
        mn = fmin(mn, dot_seq(d, d));
      }
      mn = wave_min_double(mn);
      __syncthreads();
    }
    unsigned long long ckey = ~0ull;
    cells_around(gi, ix, iy, iz, cstart, m, coincident_cells(gi, ai.x, ai.y, ai.z, ix, iy, iz), cbeg, cpre);
    const int T = cpre[27];
    for (int idx = lane(); idx < T; idx += WAVE) {
      const int t = cell_slot(cbeg, cpre, idx);
      const int j = catom[t];
      const V3 d = xi - ld3(cxyz[t]);
      if (j != g && dot_seq(d, d) <= 0.0001) {
        const unsigned long long key = ((unsigned long long)(g - a0) << 32) | (unsigned)(j - a0);
        ckey = ckey < key ? ckey : key;
      }
    }
    for (int o = 32; o > 0; o >>= 1) {
      const unsigned long long t = __shfl_xor(ckey, o, WAVE);
      ckey = ckey < t ? ckey : t;
    }
    if (lane() == 0)
      attention_store(g, (int)sidx, c, mn, ckey, all_atoms, full, s, atten, catt, access, need, coinc, status);
    __syncthreads();
  }
}

// Interface region: every non-blocker marks the same-molecule atoms within
// the bridge distance r_i + r_j + 2 rp of it (the relation is symmetric, so
// this is the set of atoms within the bridge distance of a non-blocker).
// One wavefront per cell slot, the candidates over the lanes; with a list
// (the non-blockers' cell slots, *list_n of them, in any order) only its
// slots.
__global__ __launch_bounds__(64) void k_region_cells(BatchView b, Settings s, const GridInfo* __restrict__ grid,
                                                     const int* __restrict__ cstart,
                                                     const int* __restrict__ catom,
                                                     const double4* __restrict__ cxyz,
                                                     const int8_t* __restrict__ atten,
                                                     const int* __restrict__ list,
                                                     const int* __restrict__ list_n, int* __restrict__ need) {
  __shared__ int cbeg[27], cpre[CPRE_N];
  CMS_UNIT_LOOP(u, list ? *list_n : b.ntot) {
    const int sidx = list ? list[u] : (int)u;
    const int g = catom[sidx];
    if (atten[g] <= ATTEN_BLOCKER) continue;
    const int c = b.acplx[g];
    const int m = b.amol[g];
    const GridInfo gi = grid[c];
    const double4 ai = cxyz[sidx];
    const V3 xi = ld3(ai);
    int ix, iy, iz;
    cell_coords(gi, ai.x, ai.y, ai.z, &ix, &iy, &iz);
    // the bridge distance is at most r_i + r_max + 2 rp
    const double reach = (ai.w + gi.rmax) + 2.0 * s.rp;
    cells_around(gi, ix, iy, iz, cstart, m, cells_within(gi, ai.x, ai.y, ai.z, ix, iy, iz, reach), cbeg, cpre);
    const int T = cpre[27];
    for (int idx = lane(); idx < T; idx += WAVE) {
      const int t = cell_slot(cbeg, cpre, idx);
      const int j = catom[t];
      const double4 aj = cxyz[t];
      const V3 d = xi - ld3(aj);
      const double br = (ai.w + aj.w) + 2.0 * s.rp;
      if (dot_seq(d, d) < br * br) need[j] = 1;
    }
    __syncthreads();
  }
}

// Thread-per-atom variants for large batches: consecutive threads take
// consecutive cell-sorted atoms, so a wavefront reads a handful of cells
// that stay in L1/L2 instead of every lane's own neighbourhood.  The
// attention sweep waits on its loads: it is held to 64 VGPRs for eight
// wavefronts per SIMD.
__global__ __launch_bounds__(256, 8) void k_attention_cells_t(BatchView b, Settings s, int all_atoms, int full,
                                                           const GridInfo* __restrict__ grid,
                                                           const int* __restrict__ cstart,
                                                           const int* __restrict__ catom,
                                                           const double4* __restrict__ cxyz,
                                                           int8_t* __restrict__ atten, int8_t* __restrict__ catt,
                                                        int8_t* __restrict__ access,
                                                           int* __restrict__ need,
                                                           unsigned long long* __restrict__ coinc,
                                                           int* __restrict__ status, int* __restrict__ list,
                                                           int* __restrict__ list_n) {
  for (int sidx = blockIdx.x * blockDim.x + threadIdx.x; sidx < b.ntot; sidx += gridDim.x * blockDim.x) {
    const int g = catom[sidx];
    const int c = b.acplx[g];
    const int a0 = b.aoff[c];
    const int m = b.amol[g];
    const GridInfo gi = grid[c];
    const double4 ai = cxyz[sidx];
    const V3 xi = ld3(ai);
    int ix, iy, iz;
    cell_coords(gi, ai.x, ai.y, ai.z, &ix, &iy, &iz);
    double mn = INFINITY;
    if (!all_atoms) {
      // the other molecule's atoms of the 27 cells, by (x, y) column, the
      // atom's own column first.  Only sqrt(min d^2) < sep is used, so the
      // sweep stops once a partial minimum lies below sep^2 (1 - 1e-12):
      // its square root, and so that of the full minimum, rounds below sep
      // (the margin is 1e4 times the roundings of sep^2 and of the root)
      const double stop = (s.sep * s.sep) * (1.0 - 1e-12);
      const int z0 = max(iz - 1, 0), z1 = min(iz + 1, gi.nz - 1);
      for (int k = 0; k < 9 && !(mn < stop); ++k) {
        const int col = k == 0 ? 4 : (k <= 4 ? k - 1 : k);
        const int cx = ix + col / 3 - 1, cy = iy + col % 3 - 1;
        if (cx < 0 || cx >= gi.nx || cy < 0 || cy >= gi.ny) continue;
        const ColRanges cr = column_ranges(gi, cstart, cx, cy, z0, z1, 1 - m);
#pragma unroll 4
        for (int t = 0; t < cr.total; ++t) {
          const V3 d = xi - ld3(cxyz[col_index(cr, t)]);
          mn = fmin(mn, dot_seq(d, d));
        }
      }
    }
    // coincident partners, each pair once with key (lower, higher index):
    // the minimum over pairs is np.argwhere's first entry.  The atom's own
    // cell range is taken cyclically, the floor(L / 2) slots after the
    // atom's (every pair of the range is within that distance one way
    // round, and each lane does the same number of tests); the other cells
    // within 0.01 A of the atom in full
    unsigned long long ckey = ~0ull;
    auto coincident = [&](int t) {
      const V3 d = xi - ld3(cxyz[t]);
      if (dot_seq(d, d) <= 0.0001) {
        const int j = catom[t];
        const int lo = min(g, j) - a0, hi = max(g, j) - a0;
        const unsigned long long key = ((unsigned long long)lo << 32) | (unsigned)hi;
        ckey = ckey < key ? ckey : key;
      }
    };
    {
      const int own = gi.cell0 + (ix * gi.ny + iy) * gi.nz + iz;
      const int lo = cstart[2 * own + m], hi = cstart[2 * own + m + 1];
      const int L = hi - lo;
      for (int k = 1; k <= L / 2; ++k) {
        const int t = sidx + k < hi ? sidx + k : sidx + k - L;
        coincident(t);
      }
    }
    const uint32_t cmask = coincident_cells(gi, ai.x, ai.y, ai.z, ix, iy, iz) & ~(1u << 13);   // bit 13: own cell
    FOR_NEIGHBOR_CELLS(gi, ix, iy, iz, cmask, cell) {
      for (int t = cstart[2 * cell + m]; t < cstart[2 * cell + m + 1]; ++t) coincident(t);
    }
    const bool nonblocker =
        attention_store(g, sidx, c, mn, ckey, all_atoms, full, s, atten, catt, access, need, coinc, status);
    if (list) {
      // the region pass's list of non-blocker cell slots: one atomic per
      // wavefront
      const uint64_t mk = ballot(nonblocker);
      if (mk) {
        const int leader = __builtin_ctzll(mk);
        int base = 0;
        if (lane() == leader) base = atomicAdd(list_n, __popcll(mk));
        base = __shfl(base, leader, WAVE);
        if (nonblocker) list[base + lanes_below(mk)] = sidx;
      }
    }
  }
}

// Neighbour rows (_core.py:1251-1349) from the 27 cells: one wavefront per
// region atom; the cells' atom ranges are flattened across lanes; rows are
// sorted by (d^2, j), buried rows by j.
__global__ __launch_bounds__(64) void k_neighbors_cells(BatchView b, Settings s, Work w,
                                                        const GridInfo* __restrict__ grid,
                                                        const int* __restrict__ cstart,
                                                        const int* __restrict__ catom,
                                                        const double4* __restrict__ cxyz,
                                                        const int8_t* __restrict__ catt) {
  // row entries (d^2, j) in one 16 B word each: one LDS read per entry in
  // the rank loop
  __shared__ double2 ce[KCAP];
  __shared__ int bj[BCAP];
  __shared__ int cbeg[27], cpre[CPRE_N];
  const int nR = *w.nR;
  CMS_UNIT_LOOP(r, nR) {
    const int g = w.ratom[r];
    const int c = b.acplx[g];
    const int m = b.amol[g];
    const GridInfo gi = grid[c];
    const double4 ai = b.atom[g];
    const V3 xi = ld3(ai);
    int ix, iy, iz;
    cell_coords(gi, ai.x, ai.y, ai.z, &ix, &iy, &iz);
    int cnt = 0, cntb = 0, nbmax = -1;
    const unsigned g1 = (unsigned)g + 1u;
    // both rows take atoms within the bridge distance r_i + r_j + 2 rp, at
    // most r_i + r_max + 2 rp: only the cells whose box comes that close
    // can hold them (one cell tested per lane)
    const uint32_t near = (uint32_t)ballot(
        lane() < 27 && cell_within(gi, ai.x, ai.y, ai.z, ix, iy, iz, (ai.w + gi.rmax) + 2.0 * s.rp, lane()));
    // two sweeps: the atom's own molecule (neighbour row), then the other
    // molecule (buried row); the cell slot gives the molecule
    for (int sweep = 0; sweep < 2; ++sweep) {
      const bool own = sweep == 0;
      cells_around(gi, ix, iy, iz, cstart, own ? m : 1 - m, near, cbeg, cpre);
      const int T = cpre[27];
      for (int i0 = 0; i0 < T; i0 += WAVE) {
        const int idx = i0 + lane();
        bool isn = false, isb = false;
        double d2 = 0.0;
        int j = -1;
        if (idx < T) {
          const int t = cell_slot(cbeg, cpre, idx);
          j = catom[t];
          const double4 aj = cxyz[t];
          const V3 d = xi - ld3(aj);
          d2 = dot_seq(d, d);
          const double br = (ai.w + aj.w) + 2.0 * s.rp;
          const bool inb = d2 < br * br;
          if (own) {
            isn = inb && j != g;
            if (isn && catt[t] > ATTEN_BLOCKER) nbmax = max(nbmax, j);
          } else {
            isb = inb && catt[t] >= ATTEN_BURIED;
          }
        }
        const uint64_t mn = ballot(isn), mb = ballot(isb);
        if (isn) {
          // the entry keeps ju = j - (g + 1) as unsigned (see the rank loop)
          const int p = cnt + lanes_below(mn);
          if (p < KCAP) ce[p] = make_double2(d2, __longlong_as_double((long long)((unsigned)j - g1)));
        }
        if (isb) {
          const int p = cntb + lanes_below(mb);
          if (p < BCAP) bj[p] = j;
        }
        cnt += __popcll(mn);
        cntb += __popcll(mb);
      }
      __syncthreads();
    }
    __syncthreads();
    if (cnt > KCAP || cntb > BCAP) {
      if (lane() == 0) atomicOr(&w.status[c], cnt > KCAP ? ST_NBR_CAP : ST_BUR_CAP);
      cnt = min(cnt, KCAP);
      cntb = min(cntb, BCAP);
    }
    int* row = w.nbr + (int64_t)r * KCAP;
    uint8_t* fwd = w.fwd + (int64_t)r * KCAP;
    // rank by d^2 alone (entries with a smaller d^2), and the forward rank:
    // with ju = j - (g + 1) as unsigned (what the entries hold), the forward entries (j > g) are
    // exactly those with ju < 2^31 and keep their index order, the others
    // wrap above them, so one unsigned compare counts the forward entries of
    // smaller index.  Without exact d^2 ties the ranks are a permutation of
    // 0 .. cnt-1 and sum to cnt (cnt - 1) / 2; tied entries share the lower
    // rank, the sum falls short, and the row is ranked again with the
    // (d^2, j) order.
    int rsum = 0;
    for (int p = lane(); p < cnt; p += WAVE) {
      const double dp = ce[p].x;
      const unsigned jup = (unsigned)__double_as_longlong(ce[p].y);
      int rank = 0, frank = 0;
#pragma unroll 8
      for (int t = 0; t < cnt; ++t) {
        const double2 e = ce[t];
        rank += e.x < dp;
        frank += (unsigned)__double_as_longlong(e.y) < jup;
      }
      rsum += rank;
      row[rank] = jup + g1;
      // forward neighbours (index > g) by ascending index: their row slots
      if (jup < 0x80000000u) fwd[frank] = (uint8_t)rank;
    }
    if (wave_sum_int(rsum) != cnt * (cnt - 1) / 2) {
      for (int p = lane(); p < cnt; p += WAVE) {
        const double dp = ce[p].x;
        const int jp = (int)((unsigned)__double_as_longlong(ce[p].y) + g1);
        int rank = 0, frank = 0;
        for (int t = 0; t < cnt; ++t) {
          const double2 e = ce[t];
          const double dt = e.x;
          const int jt = (int)((unsigned)__double_as_longlong(e.y) + g1);
          rank += (dt < dp) || (dt == dp && jt < jp);
          frank += (jt > g) & (jt < jp);
        }
        row[rank] = jp;
        if (jp > g) fwd[frank] = (uint8_t)rank;
      }
    }
    int* brow = w.bur + (int64_t)r * BCAP;
    for (int p = lane(); p < cntb; p += WAVE) {
      const int jp = bj[p];
      int rank = 0;
      for (int t = 0; t < cntb; ++t) rank += bj[t] < jp;
      brow[rank] = jp;
    }
    for (int o = 32; o > 0; o >>= 1) nbmax = max(nbmax, __shfl_xor(nbmax, o, WAVE));
    // the row's forward pairs that matter (what the pair pass numbers,
    // _core.py:1351-1442): valid torus and relevant; an atom whose only
    // neighbour forms a valid forward pair takes the single-neighbour
    // shortcut (both atoms accessible, no pairs)
    int npair = 0;
    bool shortcut = false;
    for (int p0 = 0; p0 < cnt; p0 += WAVE) {
      const int p = p0 + lane();
      bool ok = false, rel = false;
      int jp = -1;
      if (p < cnt) {
        jp = (int)((unsigned)__double_as_longlong(ce[p].y) + g1);
        if (jp > g) {
          const double4 aj = b.atom[jp];
          ok = pair_geom_ok(xi, ai.w, ld3(aj), aj.w, s.rp);
          rel = ok && pair_relevant(w.atten[g], w.atten[jp], nbmax, jp);
        }
      }
      if (cnt == 1 && wave_any(ok)) {
        shortcut = true;
        if (ok) {
          w.access[g] = 1;
          w.access[jp] = 1;
        }
      }
      npair += __popcll(ballot(rel));
    }
    if (lane() == 0) {
      w.nn[r] = cnt;
      w.nb[r] = cntb;
      w.nbmax[r] = nbmax;
      w.pcnt[r] = shortcut ? 0 : npair;
    }
    __syncthreads();
  }
}

}  // namespace cms
