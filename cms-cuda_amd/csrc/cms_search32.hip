// FP32 nearest-dot search with FP64 rescoring.
//
// Coordinates are recentred per complex (centre c = middle of the complex's
// atom bounding box) and rounded to float.  With R bounding |x - c| per
// component for every dot of the complex and u = 2^-24, the float squared
// distance of a pair at true distance d satisfies
//     |d2f - d^2| <= delta(d) = 4 sqrt(3) R u d + 7 u d^2 + 12 R^2 u^2
// (recentring and rounding of each coordinate: |e| <= R u; rounding of the
// difference: |dx| u; three roundings in (dx^2 + dy^2) + dz^2: 3u d^2).
// While streaming over the references, every reference with
//     d2f <= m + 2 * 1.5 * delta(sqrt(m) + 0.05)
// (m = running float minimum) is rescored in FP64 exactly as the FP64 path
// computes it, and the (d^2, index) minimum is kept among rescored ones.  A
// reference left out has d^2 > (true minimum), so the FP64 argmin -- and
// every reference tied with it -- is always rescored: distances and partner
// indices are identical to the FP64 search.
#include "cms_device.h"
#include "cms_internal.h"
#include "cms_readout.h"

namespace cms {

constexpr int NT32 = 256;
constexpr float U32 = 5.9604645e-08f;  // 2^-24

template <bool IDX>
__global__ __launch_bounds__(NT32) void k_nearest32(const float4* __restrict__ Q32, const double4* __restrict__ Q64,
                                                    const int* __restrict__ qgrp, const float4* __restrict__ R32,
                                                    const double4* __restrict__ R64, const int* __restrict__ rgrp,
                                                    int xmask, const int* __restrict__ toff, int G, int nsplit,
                                                    int64_t pstride, const double4* __restrict__ frame,
                                                    double* __restrict__ out_d2, int* __restrict__ out_idx,
                                                    unsigned long long* __restrict__ nrescored) {
  __shared__ float rx[NT32], ry[NT32], rz[NT32];
  const int64_t nwork = (int64_t)toff[G] * nsplit;
  unsigned long long resc = 0;
  for (int64_t wk = blockIdx.x; wk < nwork; wk += gridDim.x) {
    const int tile = (int)(wk / nsplit);
    const int sp = (int)(wk - (int64_t)tile * nsplit);
    int lo = 0, hi = G;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (toff[mid] <= tile) lo = mid; else hi = mid;
    }
    const int g = lo;
    const int q = qgrp[g] + (tile - toff[g]) * NT32 + threadIdx.x;
    const bool live = q < qgrp[g + 1];
    const int h = g ^ xmask;
    const int rg0 = rgrp[h], rn = rgrp[h + 1] - rg0;
    const int r0 = rg0 + (int)(((int64_t)rn * sp) / nsplit);
    const int r1 = rg0 + (int)(((int64_t)rn * (sp + 1)) / nsplit);
    const float Rb = (float)frame[g >> 1].w;
    const float cA = 1.5f * 4.0f * 1.7320508f * Rb * U32;
    const float cB = 1.5f * 7.0f * U32;
    const float cC = 1.5f * 12.0f * Rb * Rb * U32 * U32;
    float qx = 0.f, qy = 0.f, qz = 0.f;
    double qx64 = 0.0, qy64 = 0.0, qz64 = 0.0;
    if (live) {
      const float4 a = Q32[q];
      qx = a.x; qy = a.y; qz = a.z;
      const double4 b = Q64[q];
      qx64 = b.x; qy64 = b.y; qz64 = b.z;
    }
    float m = INFINITY, thr = INFINITY;
    double best = INFINITY;
    int bi = -1;
    for (int rb = r0; rb < r1; rb += NT32) {
      __syncthreads();
      const int rr = rb + threadIdx.x;
      if (rr < r1) {
        const float4 a = R32[rr];
        rx[threadIdx.x] = a.x; ry[threadIdx.x] = a.y; rz[threadIdx.x] = a.z;
      }
      __syncthreads();
      const int n = min(NT32, r1 - rb);
      if (live) {
        for (int t = 0; t < n; ++t) {
          const float dx = rx[t] - qx, dy = ry[t] - qy, dz = rz[t] - qz;
          const float d2f = (dx * dx + dy * dy) + dz * dz;
          if (d2f <= thr) {
            ++resc;
            if (d2f < m) {
              m = d2f;
              const float de = sqrtf(m) + 0.05f;
              thr = m + 2.0f * ((cA * de + cB * de * de) + cC);
            }
            const double4 a = R64[rb + t];
            const double ex = a.x - qx64, ey = a.y - qy64, ez = a.z - qz64;
            const double d2 = (ex * ex + ey * ey) + ez * ez;
            if (d2 < best || (d2 == best && rb - rg0 + t < bi)) { best = d2; bi = rb - rg0 + t; }
          }
        }
      }
    }
    if (live) {
      const int64_t o = (int64_t)sp * pstride + q;
      out_d2[o] = best;
      if (IDX) out_idx[o] = bi;
    }
    __syncthreads();
  }
  if (nrescored && resc) atomicAdd(nrescored, resc);
}

// float copy of a double4 point list, recentred on its complex's frame;
// group of item i found from grp (G+1 starts), complex = group / 2 (or 0
// when percomplex == 0).
__global__ void k_to_float(Count cnt, const double4* __restrict__ in, const int* __restrict__ grp, int G,
                           int percomplex, const double4* __restrict__ frame, float4* __restrict__ out) {
  const int n = cnt.get();
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    int c = 0;
    if (percomplex) {
      int lo = 0, hi = G;
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (grp[mid] <= i) lo = mid; else hi = mid;
      }
      c = lo >> 1;
    }
    const double4 f = frame[c];
    const double4 a = in[i];
    out[i] = make_float4((float)(a.x - f.x), (float)(a.y - f.y), (float)(a.z - f.z), 0.f);
  }
}

template __global__ void k_nearest32<false>(const float4*, const double4*, const int*, const float4*, const double4*,
                                            const int*, int, const int*, int, int, int64_t, const double4*,
                                            double*, int*, unsigned long long*);
template __global__ void k_nearest32<true>(const float4*, const double4*, const int*, const float4*, const double4*,
                                           const int*, int, const int*, int, int, int64_t, const double4*,
                                           double*, int*, unsigned long long*);

}  // namespace cms
