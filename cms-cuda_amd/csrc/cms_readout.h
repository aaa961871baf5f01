#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cms_cells.h"
#include "cms_internal.h"

namespace cms {

// FP32 search tiles: NT32 queries per tile (references staged NT32 per
// round), TPB32 threads per block, two queries per thread.  The host sizes
// tiles and launches from these; k_nearest32 assumes blockDim.x == TPB32.
constexpr int NT32 = 128;
constexpr int TPB32 = NT32 / 2;
// FP64 search: queries per tile = threads per block
constexpr int NT64 = 256;

__global__ void k_dot_flag(Count n, const int8_t* bur, int want, int* flag);
__global__ void k_dot_split(Count cnt, const int8_t* dbur, const int* bscan, const double4* dpos, int* blist,
                            double4* bxyz);
__global__ void k_gather_compl(int G, const int* idx, const int* scan, int* out);
__global__ void k_tile_counts(int G, const int* qgrp, int tq, int* cnt);
template <bool IDX>
__global__ void k_nearest64(const double4* Q, const int* qgrp, const double4* R, const int* rgrp,
                            int xmask, const int* toff, int G, int nsplit, int64_t pstride,
                            double* out_d2, int* out_idx);
__global__ void k_nearest_combine(Count nq, int nsplit, int64_t pstride, const double* pd2, const int* pidx,
                                  double* out_d2, int* out_idx);
__global__ void k_cms_weights(Count n, const double4* bxyz, const double* d2, double weight, double* w);
__global__ void k_owner_count(Count n, const int* list, const int* datom, int* cnt);
__global__ void k_owner_scatter(Count n, const int* list, const int* datom, const int* start, int* cur,
                                int* perm);
__global__ void k_owner_small(int natoms, const int* start, const int* perm, const double* val, int64_t vstride,
                              double* out, int* multi, int* nmulti);
__global__ void k_owner_sums(const int* multi, const int* nmulti, const int* start, int* perm, const double* val,
                             int64_t vstride, double* out);
__global__ void k_group_np_sum(int G, const int* start, const double* vals, int64_t stride, double* out);
__global__ void k_trim_flag(Count n, const double* d2acc, double band2, int* flag);
__global__ void k_trim_compact(Count n, const int* flag, const int* scan, const int* blist,
                               const double4* dpos, const double4* dnml, int* tlist, double4* txyz,
                               double4* tnml);
__global__ void k_sc_terms(Count n, const int* tgrp_of, const int* tgrp, const double* d2, const int* idx,
                           const double4* tnml, double weight, double* dist, double* rr);
__global__ void k_group_of_items(int G, const int* start, int* gof);
__global__ void k_sc_sides(int G, const int* tgrp, const double4* txyz, const double* dist,
                           const double* rr, const int* dgrp, Settings s, double* side);
__global__ void k_sc_combine(int B, const double* side, double* res);

template <bool IDX>
__global__ void k_nearest32(const float4* Q32, const double4* Q64, const int* qgrp, const float4* R32,
                            const double4* R64, const int* rgrp, int xmask, const int* toff, int G, int nsplit,
                            int64_t pstride, const double4* frame, double* out_d2, int* out_idx,
                            unsigned long long* nrescored);
__global__ void k_dot_grid_setup(int G, const int* agrp, const int* bgrp, const double4* bxyz, double band,
                                 GridInfo* grid, int* ncell);
__global__ void k_dot_grid_build(int G, const int* dgrp, const double4* dpos, const int8_t* dbur,
                                 const GridInfo* grid, int* tcount, int* tstart, double4* sorted);
template <int L>
__global__ void k_trim_grid(Count nb, const double4* bxyz, const int* bgrp, int G, const GridInfo* grid,
                            const int* cstart, const double4* sorted, double band2, int* flag);
__global__ void k_to_float(Count n, const double4* in, const int* grp, int G, int percomplex,
                           const double4* frame, float4* out);

}  // namespace cms
