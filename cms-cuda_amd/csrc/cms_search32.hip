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

constexpr float U32 = 5.9604645e-08f;  // 2^-24

// One query's running state: -2 q (float), q (FP64), the float minimum of
// e, the rescoring threshold, and the (d^2, index) minimum among rescored
// references.
struct Near32 {
  float qx, qy, qz, m, thr;
  double x, y, z, best;
  int bi;
};

template <bool IDX>
__global__ __launch_bounds__(TPB32) void k_nearest32(const float4* __restrict__ Q32, const double4* __restrict__ Q64,
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
    const int q0 = qgrp[g] + (tile - toff[g]) * NT32 + threadIdx.x;
    const int qe = qgrp[g + 1];
    const int h = g ^ xmask;
    const int rg0 = rgrp[h], rn = rgrp[h + 1] - rg0;
    const int r0 = rg0 + (int)(((int64_t)rn * sp) / nsplit);
    const int r1 = rg0 + (int)(((int64_t)rn * (sp + 1)) / nsplit);
    const float Rb = (float)frame[g >> 1].w;
    const float band = (1.5f * 2.0f * 64.0f * U32) * (Rb * Rb);
    // queries q0 and q0 + TPB32; a query past the group takes nothing
    // (threshold -inf)
    Near32 st[2];
#pragma unroll
    for (int v = 0; v < 2; ++v) {
      const int q = q0 + v * TPB32;
      Near32& z = st[v];
      z.qx = z.qy = z.qz = 0.f;
      z.x = z.y = z.z = 0.0;
      z.m = INFINITY;
      z.thr = q < qe ? INFINITY : -INFINITY;
      z.best = INFINITY;
      z.bi = -1;
      if (q < qe) {
        const float4 a = Q32[q];
        z.qx = -2.0f * a.x; z.qy = -2.0f * a.y; z.qz = -2.0f * a.z;
        const double4 b = Q64[q];
        z.x = b.x; z.y = b.y; z.z = b.z;
      }
    }
    const bool live = q0 < qe;
    for (int rb = r0; rb < r1; rb += NT32) {
      __syncthreads();
      for (int k = threadIdx.x; k < NT32; k += TPB32)
        if (rb + k < r1) rs[k] = R32[rb + k];
      __syncthreads();
      const int n = min(NT32, r1 - rb);
      if (live) {
        auto take = [&](Near32& z, float e, int t) {
          if (e <= z.thr) {
            ++resc;
            if (e < z.m) {
              z.m = e;
              z.thr = z.m + band;
            }
            const double4 a = R64[rb + t];
            const double ex = a.x - z.x, ey = a.y - z.y, ez = a.z - z.z;
            const double d2 = (ex * ex + ey * ey) + ez * ez;
            if (d2 < z.best || (d2 == z.best && rb - rg0 + t < z.bi)) { z.best = d2; z.bi = rb - rg0 + t; }
          }
        };
        int t = 0;
        // four references per test: a threshold only falls, so a group
        // whose smallest e is above it holds no candidate for that query
        for (; t + 4 <= n; t += 4) {
          float e0[4], e1[4];
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            const float4 r = rs[t + k];
            e0[k] = fmaf(r.x, st[0].qx, fmaf(r.y, st[0].qy, fmaf(r.z, st[0].qz, r.w)));
            e1[k] = fmaf(r.x, st[1].qx, fmaf(r.y, st[1].qy, fmaf(r.z, st[1].qz, r.w)));
          }
          const bool h0 = fminf(fminf(e0[0], e0[1]), fminf(e0[2], e0[3])) <= st[0].thr;
          const bool h1 = fminf(fminf(e1[0], e1[1]), fminf(e1[2], e1[3])) <= st[1].thr;
          if (h0 || h1) {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
              take(st[0], e0[k], t + k);
              take(st[1], e1[k], t + k);
            }
          }
        }
        for (; t < n; ++t) {
          const float4 r = rs[t];
          take(st[0], fmaf(r.x, st[0].qx, fmaf(r.y, st[0].qy, fmaf(r.z, st[0].qz, r.w))), t);
          take(st[1], fmaf(r.x, st[1].qx, fmaf(r.y, st[1].qy, fmaf(r.z, st[1].qz, r.w))), t);
        }
      }
    }
#pragma unroll
    for (int v = 0; v < 2; ++v) {
      const int q = q0 + v * TPB32;
      if (q < qe) {
        const int64_t o = (int64_t)sp * pstride + q;
        out_d2[o] = st[v].best;
        if (IDX) out_idx[o] = st[v].bi;
      }
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
