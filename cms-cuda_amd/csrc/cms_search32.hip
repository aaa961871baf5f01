// FP32 nearest-dot search with FP64 rescoring.
//
// Coordinates are recentred per complex (centre c = middle of the complex's
// atom bounding box) and rounded to float; R bounds |x - c| per component
// for every dot of the complex and u = 2^-24.  Each float point carries
// its squared norm rounded to float in .w, and a (query q, reference r)
// pair is filtered on
//     e = |r|^2 - 2 r.q       (three fused multiply-adds)
// which differs from d^2 - |q|^2 (d = distance of the FP64 points, |q|^2 a
// per-query constant) by at most 64 u R^2:
//   - float rounding of the coordinates moves |r - q| by at most 2 sqrt(3) u R,
//     and d <= 2 sqrt(3) R, so |(r - q)_float|^2 is within 24 u R^2 (+ O(u^2)) of d^2;
//   - rounding |r|^2 to float costs 3 u R^2, the three fused steps, whose
//     partial sums stay below 5, 7 and 9 R^2, cost 21 u R^2.
// While streaming over the references, every reference with
//     e <= m + 1.5 * 2 * 64 u R^2
// (m = running minimum of e) is rescored in FP64 exactly as the FP64 path
// computes it, and the (d^2, index) minimum is kept among rescored ones.  A
// reference left out has d^2 above that of the one that set m, so the FP64
// argmin -- and every reference tied with it -- is always rescored:
// distances and partner indices are identical to the FP64 search.  (The 1.5
// covers the float rounding of the threshold, at most 9 u R^2, and the
// FP64 rounding of d^2.)
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
  __shared__ float4 rs[NT32];
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
    const float band = (1.5f * 2.0f * 64.0f * U32) * (Rb * Rb);
    float qx = 0.f, qy = 0.f, qz = 0.f;   // -2 q
    double qx64 = 0.0, qy64 = 0.0, qz64 = 0.0;
    if (live) {
      const float4 a = Q32[q];
      qx = -2.0f * a.x; qy = -2.0f * a.y; qz = -2.0f * a.z;
      const double4 b = Q64[q];
      qx64 = b.x; qy64 = b.y; qz64 = b.z;
    }
    float m = INFINITY, thr = INFINITY;
    double best = INFINITY;
    int bi = -1;
    for (int rb = r0; rb < r1; rb += NT32) {
      __syncthreads();
      const int rr = rb + threadIdx.x;
      if (rr < r1) rs[threadIdx.x] = R32[rr];
      __syncthreads();
      const int n = min(NT32, r1 - rb);
      if (live) {
        auto take = [&](float e, int t) {
          if (e <= thr) {
            ++resc;
            if (e < m) {
              m = e;
              thr = m + band;
            }
            const double4 a = R64[rb + t];
            const double ex = a.x - qx64, ey = a.y - qy64, ez = a.z - qz64;
            const double d2 = (ex * ex + ey * ey) + ez * ez;
            if (d2 < best || (d2 == best && rb - rg0 + t < bi)) { best = d2; bi = rb - rg0 + t; }
          }
        };
        int t = 0;
        // four references per test: the threshold only falls, so a group
        // whose smallest e is above it holds no candidate
        for (; t + 4 <= n; t += 4) {
          float e[4];
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            const float4 r = rs[t + k];
            e[k] = fmaf(r.x, qx, fmaf(r.y, qy, fmaf(r.z, qz, r.w)));
          }
          if (fminf(fminf(e[0], e[1]), fminf(e[2], e[3])) <= thr) {
#pragma unroll
            for (int k = 0; k < 4; ++k) take(e[k], t + k);
          }
        }
        for (; t < n; ++t) {
          const float4 r = rs[t];
          take(fmaf(r.x, qx, fmaf(r.y, qy, fmaf(r.z, qz, r.w))), t);
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

// float copy of a double4 point list, recentred on its complex's frame,
// with the float point's squared norm (rounded once) in .w; group of item i
// found from grp (G+1 starts), complex = group / 2 (or 0 when
// percomplex == 0).
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
    const float x = (float)(a.x - f.x), y = (float)(a.y - f.y), z = (float)(a.z - f.z);
    // float products are exact in double
    const double n2 = ((double)x * x + (double)y * y) + (double)z * z;
    out[i] = make_float4(x, y, z, (float)n2);
  }
}

template __global__ void k_nearest32<false>(const float4*, const double4*, const int*, const float4*, const double4*,
                                            const int*, int, const int*, int, int, int64_t, const double4*,
                                            double*, int*, unsigned long long*);
template __global__ void k_nearest32<true>(const float4*, const double4*, const int*, const float4*, const double4*,
                                           const int*, int, const int*, int, int, int64_t, const double4*,
                                           double*, int*, unsigned long long*);

}  // namespace cms
