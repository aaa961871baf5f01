// Readouts: buried / accessible / trimmed dot lists, brute-force nearest
// dot search, CMS weights and per-atom sums in dot order, NumPy-order
// pairwise totals, SC trim and side statistics with the binned median.
#include "cms_device.h"
#include "cms_internal.h"
#include "cms_readout.h"

namespace cms {

// ------------------------------------------------------------ lists
__global__ void k_dot_flag(Count cnt, const int8_t* __restrict__ bur, int want, int* __restrict__ flag) {
  const int n = cnt.get();
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    flag[i] = (bur[i] != 0) == (want != 0) ? 1 : 0;
}

// Buried dot list in one pass: index and position at the exclusive scan
// of the burial flags (the trim grid reads the accessible dots in place).
__global__ void k_dot_split(Count cnt, const int8_t* __restrict__ dbur, const int* __restrict__ bscan,
                            const double4* __restrict__ dpos, int* __restrict__ blist, double4* __restrict__ bxyz) {
  const int n = cnt.get();
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    if (dbur[i]) {
      const int b = bscan[i];
      blist[b] = i;
      bxyz[b] = dpos[i];
    }
  }
}
// Accessible-list group starts: idx[g] - scan[idx[g]].
__global__ void k_gather_compl(int G, const int* __restrict__ idx, const int* __restrict__ scan,
                               int* __restrict__ out) {
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g <= G; g += gridDim.x * blockDim.x) out[g] = idx[g] - scan[idx[g]];
}

// ------------------------------------------------------------ search
__global__ void k_tile_counts(int G, const int* __restrict__ qgrp, int tq, int* __restrict__ cnt) {
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < G; g += gridDim.x * blockDim.x)
    cnt[g] = (qgrp[g + 1] - qgrp[g] + tq - 1) / tq;
}

// Nearest reference point of every query, FP64.  Query group g searches the
// reference group g ^ xmask.  d^2 = (dx^2 + dy^2) + dz^2 as in
// scipy cdist(sqeuclidean); ties keep the first (lowest) reference index,
// as np.argmin does.  Work item = (query tile of 256, reference split s of
// nsplit): split s scans the s-th contiguous slice of the group's
// references, staged in LDS 256 at a time.  With nsplit > 1 every split
// writes a partial (d^2, index) and k_nearest_combine takes the minimum in
// split order, so the first index still wins ties.
constexpr int NT = NT64;
template <bool IDX>
__global__ __launch_bounds__(NT) void k_nearest64(const double4* __restrict__ Q, const int* __restrict__ qgrp,
                                                  const double4* __restrict__ R, const int* __restrict__ rgrp,
                                                  int xmask, const int* __restrict__ toff, int G, int nsplit,
                                                  int64_t pstride, double* __restrict__ out_d2,
                                                  int* __restrict__ out_idx) {
  __shared__ double rx[NT], ry[NT], rz[NT];
  const int64_t nwork = (int64_t)toff[G] * nsplit;
  for (int64_t wk = blockIdx.x; wk < nwork; wk += gridDim.x) {
    const int tile = (int)(wk / nsplit);
    const int sp = (int)(wk - (int64_t)tile * nsplit);
    int lo = 0, hi = G;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (toff[mid] <= tile) lo = mid; else hi = mid;
    }
    const int g = lo;
    const int q = qgrp[g] + (tile - toff[g]) * NT + threadIdx.x;
    const bool live = q < qgrp[g + 1];
    const int h = g ^ xmask;
    const int rg0 = rgrp[h], rn = rgrp[h + 1] - rg0;
    const int r0 = rg0 + (int)(((int64_t)rn * sp) / nsplit);
    const int r1 = rg0 + (int)(((int64_t)rn * (sp + 1)) / nsplit);
    double qx = 0.0, qy = 0.0, qz = 0.0;
    if (live) {
      const double4 a = Q[q];
      qx = a.x; qy = a.y; qz = a.z;
    }
    double best = INFINITY;
    int bi = -1;
    for (int rb = r0; rb < r1; rb += NT) {
      __syncthreads();
      const int rr = rb + threadIdx.x;
      if (rr < r1) {
        const double4 a = R[rr];
        rx[threadIdx.x] = a.x; ry[threadIdx.x] = a.y; rz[threadIdx.x] = a.z;
      }
      __syncthreads();
      const int n = min(NT, r1 - rb);
      for (int t = 0; t < n; ++t) {
        const double dx = rx[t] - qx, dy = ry[t] - qy, dz = rz[t] - qz;
        const double d2 = (dx * dx + dy * dy) + dz * dz;
        if (d2 < best) { best = d2; bi = rb - rg0 + t; }
      }
    }
    if (live) {
      const int64_t o = (int64_t)sp * pstride + q;
      out_d2[o] = best;
      if (IDX) out_idx[o] = bi;
    }
    __syncthreads();
  }
}

__global__ void k_nearest_combine(Count cnt, int nsplit, int64_t pstride, const double* __restrict__ pd2,
                                  const int* __restrict__ pidx, double* __restrict__ out_d2,
                                  int* __restrict__ out_idx) {
  const int nq = cnt.get();
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nq; q += gridDim.x * blockDim.x) {
    double best = INFINITY;
    int bi = -1;
    for (int s = 0; s < nsplit; ++s) {
      const double d = pd2[(int64_t)s * pstride + q];
      if (d < best) {
        best = d;
        if (pidx) bi = pidx[(int64_t)s * pstride + q];
      }
    }
    out_d2[q] = best;
    if (out_idx) out_idx[q] = bi;
  }
}

// ------------------------------------------------------------ CMS weights
// per buried dot: area * exp(-d^2 * weight)   (_core.py:995)
__global__ void k_cms_weights(Count cnt, const double4* __restrict__ bxyz, const double* __restrict__ d2,
                              double weight, double* __restrict__ w) {
  const int n = cnt.get();
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    w[i] = bxyz[i].w * exp(-d2[i] * weight);
}

// ------------------------------------------------------------ per-atom sums
// np.bincount(owner, weights) accumulates item values per atom in item
// order starting from 0.0.  Counting sort by owner (integer atomics), then
// each atom sorts its own (short) segment by item index and sums in order,
// so the result is deterministic and bit-identical to bincount.
__global__ void k_owner_count(Count items, const int* __restrict__ list, const int* __restrict__ datom,
                              int* __restrict__ cnt) {
  const int n = items.get();
  CMS_WAVE_LOOP(i, n) {
    const int a = i < n ? datom[list ? list[i] : i] : 0;
    wave_counter_inc(cnt, a, i < n);
  }
}
__global__ void k_owner_scatter(Count cnt, const int* __restrict__ list, const int* __restrict__ datom,
                                const int* __restrict__ start, int* __restrict__ cur,
                                int* __restrict__ perm) {
  const int n = cnt.get();
  CMS_WAVE_LOOP(i, n) {
    const int a = i < n ? datom[list ? list[i] : i] : 0;
    const int slot = wave_counter_add(cur, a, i < n);
    if (i < n) perm[start[a] + slot] = (int)i;
  }
}
// One wavefront per atom: the atom's item indices (scattered in arbitrary
// order by k_owner_scatter) are put in ascending order -- up to 64 of them
// by rank (each lane counts the smaller indices), longer segments with a
// bitonic network in LDS -- their values gathered in that order and summed
// left to right by lane 0.
// Atoms with no item (0.0) or one (0.0 + value), one thread each.
// Atoms with at most one item, one thread each; the others are listed
// (multi, *nmulti of them, in any order) for k_owner_sums, gathered per
// workgroup in LDS and appended with one atomic per workgroup and flush.
constexpr int OWN_LIST = 1024;
__global__ __launch_bounds__(256) void k_owner_small(int natoms, const int* __restrict__ start,
                                                     const int* __restrict__ perm, const double* __restrict__ val,
                                                     int64_t vstride, double* __restrict__ out,
                                                     int* __restrict__ multi, int* __restrict__ nmulti) {
  __shared__ int lst[OWN_LIST];
  __shared__ int nl, base;
  if (threadIdx.x == 0) nl = 0;
  __syncthreads();
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t end = ((int64_t)natoms + blockDim.x - 1) / blockDim.x * blockDim.x;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t - threadIdx.x < end; t += stride) {
    const int a = (int)t;
    bool many = false;
    if (a < natoms) {
      const int lo = start[a], n = start[a + 1] - lo;
      if (n <= 1)
        out[a] = n == 1 ? 0.0 + val[(int64_t)perm[lo] * vstride] : 0.0;
      else
        many = true;
    }
    if (many) lst[atomicAdd(&nl, 1)] = a;
    __syncthreads();
    // flush when another round could overflow the list, and at the end
    const bool last = t - threadIdx.x + stride >= end;
    if (nl > OWN_LIST - (int)blockDim.x || last) {
      if (threadIdx.x == 0) base = nl ? atomicAdd(nmulti, nl) : 0;
      __syncthreads();
      for (int k = threadIdx.x; k < nl; k += blockDim.x) multi[base + k] = lst[k];
      __syncthreads();
      if (threadIdx.x == 0) nl = 0;
      __syncthreads();
    }
  }
}
constexpr int OWN_CAP = 1024;
__global__ __launch_bounds__(64) void k_owner_sums(const int* __restrict__ multi, const int* __restrict__ nmulti,
                                                   const int* __restrict__ start, int* __restrict__ perm,
                                                   const double* __restrict__ val, int64_t vstride,
                                                   double* __restrict__ out) {
  __shared__ int key[OWN_CAP];
  __shared__ double v[OWN_CAP];
  // the atoms with two items or more, listed by k_owner_count (the others
  // are summed by k_owner_small)
  CMS_UNIT_LOOP(u, *nmulti) {
    const int a = multi[u];
    const int lo = start[a], n = start[a + 1] - lo;
    if (n > OWN_CAP) {
      // very long segment: in-place insertion sort by one lane
      if (lane() == 0) {
        for (int i = lo + 1; i < lo + n; ++i) {
          const int x = perm[i];
          int j = i - 1;
          while (j >= lo && perm[j] > x) { perm[j + 1] = perm[j]; --j; }
          perm[j + 1] = x;
        }
        double s = 0.0;
        for (int i = lo; i < lo + n; ++i) s += val[(int64_t)perm[i] * vstride];
        out[a] = s;
      }
      continue;
    }
    if (n <= WAVE) {
      // item indices are distinct, so the ranks are a permutation of 0..n-1
      const int k = lane() < n ? perm[lo + lane()] : 0x7fffffff;
      const double x = lane() < n ? val[(int64_t)k * vstride] : 0.0;
      int rank = 0;
      for (int t = 0; t < n; ++t) rank += __builtin_amdgcn_readlane(k, t) < k;
      if (lane() < n) v[rank] = x;
      __syncthreads();
      if (lane() == 0) {
        double s = 0.0;
        for (int t = 0; t < n; ++t) s += v[t];
        out[a] = s;
      }
      __syncthreads();
      continue;
    }
    int P = 1;
    while (P < n) P <<= 1;
    for (int k = lane(); k < P; k += WAVE) key[k] = k < n ? perm[lo + k] : 0x7fffffff;
    __syncthreads();
    for (int size = 2; size <= P; size <<= 1) {
      for (int stride = size >> 1; stride > 0; stride >>= 1) {
        for (int t = lane(); t < P / 2; t += WAVE) {
          const int i = 2 * t - (t & (stride - 1));
          const int j = i + stride;
          const bool up = (i & size) == 0;
          const int x = key[i], y = key[j];
          if ((x > y) == up) { key[i] = y; key[j] = x; }
        }
        __syncthreads();
      }
    }
    for (int k = lane(); k < n; k += WAVE) v[k] = val[(int64_t)key[k] * vstride];
    __syncthreads();
    if (lane() == 0) {
      double s = 0.0;
      for (int k = 0; k < n; ++k) s += v[k];
      out[a] = s;
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------ pairwise sums
// np.add.reduce on a contiguous float64 array: blocks of <= 128 values are
// summed with 8 interleaved accumulators, longer ranges split at
// n/2 rounded down to a multiple of 8 and the halves added.
__device__ double np_leaf_sum(const double* a, int64_t stride, int n) {
  if (n < 8) {
    double res = 0.0;
    for (int i = 0; i < n; ++i) res += a[i * stride];
    return res;
  }
  double r[8];
  for (int j = 0; j < 8; ++j) r[j] = a[j * stride];
  int i = 8;
  for (; i < n - (n % 8); i += 8)
    for (int j = 0; j < 8; ++j) r[j] += a[(int64_t)(i + j) * stride];
  double res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
  for (; i < n; ++i) res += a[(int64_t)i * stride];
  return res;
}

constexpr int PW_MAXLEV = 32;

// Block-wide np.sum of K arrays of the same length n (a[k][i * stride[k]]),
// each the pairwise tree of np_leaf_sum leaves; the tree depends only on n,
// so it is laid out once for the K arrays.  Wavefront 0 lays it out level by
// level (the children of a level's split nodes, in order, form the next
// level); each leaf is summed by 8 lanes, lane j holding the accumulator of
// the values j, j + 8, ... (at most 16, all loaded before the first add),
// the 8 accumulators combined by a butterfly that forms
// ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7)) (addition commutes
// exactly), the leaf's last m % 8 values added in order by its first lane;
// the internal nodes are added bottom-up one level at a time as left + right
// -- the association of the recursion.  A tree with more than MAXN nodes is
// followed by one thread instead.  Every thread gets the sums.  Expects
// blockDim.x a multiple of WAVE.
template <int K, int MAXN>
__device__ void block_np_sums(const double* const (&a)[K], const int64_t (&stride)[K], int n, double (&out)[K]) {
  __shared__ int nd_lo[MAXN], nd_m[MAXN], nd_ch[MAXN];
  __shared__ double nd_v[K][MAXN];
  __shared__ int lev_start[PW_MAXLEV + 1];
  __shared__ int nlev, overflow;
  __shared__ int stk_n[64], stk_lo[64], stk_v[64];
  __shared__ double vstk[64];
  __shared__ double result[K];
  if (threadIdx.x < WAVE) {
    int s0 = 0, cnt = 1, lev = 0, of = 0;
    if (lane() == 0) {
      nd_lo[0] = 0;
      nd_m[0] = n;
      lev_start[0] = 0;
    }
    for (;;) {
      // the split nodes of level lev get two children each at the next level
      const int next = s0 + cnt;
      int ncnt = 0;
      for (int b0 = 0; b0 < cnt; b0 += WAVE) {
        const int k = b0 + lane();
        const int m = k < cnt ? nd_m[s0 + k] : 0;
        const bool split = k < cnt && m > 128;
        const int c = split ? 2 : 0;
        const int inc = wave_incl_scan(c);
        const int pos = next + ncnt + inc - c;
        if (k < cnt) nd_ch[s0 + k] = split ? pos : -1;
        if (split && pos + 1 < MAXN) {
          const int lo = nd_lo[s0 + k];
          int n2 = m / 2;
          n2 -= n2 % 8;
          nd_lo[pos] = lo;
          nd_m[pos] = n2;
          nd_lo[pos + 1] = lo + n2;
          nd_m[pos + 1] = m - n2;
        }
        ncnt += __shfl(inc, WAVE - 1, WAVE);
      }
      if (ncnt == 0) break;
      if (next + ncnt > MAXN || lev + 1 >= PW_MAXLEV) { of = 1; break; }
      s0 = next;
      cnt = ncnt;
      ++lev;
      if (lane() == 0) lev_start[lev] = s0;
    }
    if (lane() == 0) {
      nlev = lev + 1;
      lev_start[lev + 1] = s0 + cnt;
      overflow = of;
    }
  }
  __syncthreads();
  if (overflow) {
    // very long arrays: one thread follows the same tree
    if (threadIdx.x == 0) {
      for (int q = 0; q < K; ++q) {
        int sp = 0, vp = 0;
        stk_n[0] = n; stk_lo[0] = 0; stk_v[0] = 0;
        while (sp >= 0) {
          const int m = stk_n[sp], lo = stk_lo[sp], v = stk_v[sp];
          --sp;
          if (m <= 128) {
            vstk[vp++] = np_leaf_sum(a[q] + (int64_t)lo * stride[q], stride[q], m);
          } else if (!v) {
            int n2 = m / 2;
            n2 -= n2 % 8;
            ++sp; stk_n[sp] = m; stk_lo[sp] = lo; stk_v[sp] = 1;
            ++sp; stk_n[sp] = m - n2; stk_lo[sp] = lo + n2; stk_v[sp] = 0;
            ++sp; stk_n[sp] = n2; stk_lo[sp] = lo; stk_v[sp] = 0;
          } else {
            const double b2 = vstk[--vp];
            const double b1 = vstk[--vp];
            vstk[vp++] = b1 + b2;
          }
        }
        result[q] = vstk[0];
      }
    }
    __syncthreads();
    for (int q = 0; q < K; ++q) out[q] = result[q];
    __syncthreads();
    return;
  }
  const int nodes = lev_start[nlev];
  const int sub = threadIdx.x & 7;
  const int lbase = lane() & ~7;
  for (int k0 = 0; k0 < nodes; k0 += blockDim.x / 8) {
    // uniform trip count: every lane reaches the shuffles
    const int k = k0 + (int)(threadIdx.x >> 3);
    const bool leaf = k < nodes && nd_ch[k] < 0;
    const int lo = leaf ? nd_lo[k] : 0, m = leaf ? nd_m[k] : 0;
    const int m8 = m < 8 ? 0 : m - m % 8;
    const int cnt = m8 / 8;   // values per lane, at most 16
    const int tcnt = m - m8;  // the leaf's last values, added in order
    double v[K][16], tv[K];
#pragma unroll
    for (int q = 0; q < K; ++q) {
#pragma unroll
      for (int t = 0; t < 16; ++t) v[q][t] = t < cnt ? a[q][(int64_t)(lo + 8 * t + sub) * stride[q]] : 0.0;
      tv[q] = sub < tcnt ? a[q][(int64_t)(lo + m8 + sub) * stride[q]] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < K; ++q) {
      double r = v[q][0];
#pragma unroll
      for (int t = 1; t < 16; ++t)
        if (t < cnt) r += v[q][t];
      r += __shfl_xor(r, 1, WAVE);
      r += __shfl_xor(r, 2, WAVE);
      r += __shfl_xor(r, 4, WAVE);
      double res = m < 8 ? 0.0 : r;
#pragma unroll
      for (int j = 0; j < 7; ++j) {
        const double x = __shfl(tv[q], lbase + j, WAVE);
        if (j < tcnt) res += x;
      }
      if (leaf && sub == 0) nd_v[q][k] = res;
    }
  }
  __syncthreads();
  for (int lev = nlev - 2; lev >= 0; --lev) {
    for (int k = lev_start[lev] + threadIdx.x; k < lev_start[lev + 1]; k += blockDim.x) {
      const int c = nd_ch[k];
      if (c >= 0)
        for (int q = 0; q < K; ++q) nd_v[q][k] = nd_v[q][c] + nd_v[q][c + 1];
    }
    __syncthreads();
  }
  for (int q = 0; q < K; ++q) out[q] = nd_v[q][0];
  __syncthreads();
}

// out[g] = np.sum(vals[start[g] .. start[g+1]))
__global__ __launch_bounds__(256) void k_group_np_sum(int G, const int* __restrict__ start,
                                                      const double* __restrict__ vals, int64_t stride,
                                                      double* __restrict__ out) {
  for (int g = blockIdx.x; g < G; g += gridDim.x) {
    const int lo = start[g], hi = start[g + 1];
    const double* a[1] = {vals + (int64_t)lo * stride};
    const int64_t st[1] = {stride};
    double s[1];
    block_np_sums<1, 1024>(a, st, hi - lo, s);
    if (threadIdx.x == 0) out[g] = s[0];
  }
}

// ------------------------------------------------------------ trim
__global__ void k_trim_flag(Count cnt, const double* __restrict__ d2acc, double band2, int* __restrict__ flag) {
  const int n = cnt.get();
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    flag[i] = d2acc[i] > band2 ? 1 : 0;
}
__global__ void k_trim_compact(Count cnt, const int* __restrict__ flag, const int* __restrict__ scan,
                               const int* __restrict__ blist, const double4* __restrict__ dpos,
                               const double4* __restrict__ dnml, int* __restrict__ tlist,
                               double4* __restrict__ txyz, double4* __restrict__ tnml) {
  const int n = cnt.get();
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    if (flag[i]) {
      const int o = scan[i];
      const int d = blist[i];
      tlist[o] = d;
      txyz[o] = dpos[d];
      tnml[o] = dnml[d];
    }
}

// Per trimmed dot: distance to the nearest trimmed dot of the other side and
// r = clip((n . n_other) * exp(-dist^2 * weight), +-0.999)  (_core.py:1134-1141)
__global__ void k_sc_terms(Count cnt, const int* __restrict__ tgrp_of, const int* __restrict__ tgrp,
                           const double* __restrict__ d2, const int* __restrict__ idx,
                           const double4* __restrict__ tnml, double weight,
                           double* __restrict__ dist, double* __restrict__ rr) {
  const int n = cnt.get();
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    if (idx[i] < 0) {  // the other side has no trimmed dots: statistics are skipped
      dist[i] = 0.0;
      rr[i] = 0.0;
      continue;
    }
    const int g = tgrp_of[i];
    const int h = g ^ 1;
    const int j = tgrp[h] + idx[i];
    const double dm = sqrt(d2[i]);
    const V3 a = ld3(tnml[i]), b = ld3(tnml[j]);
    double r = dot_alt(a, b);
    r = r * exp(-(dm * dm) * weight);
    r = fmin(fmax(r, -0.999), 0.999);
    dist[i] = dm;
    rr[i] = r;
  }
}

__global__ void k_group_of_items(int G, const int* __restrict__ start, int* __restrict__ gof) {
  for (int g = blockIdx.x; g < G; g += gridDim.x)
    for (int i = start[g] + threadIdx.x; i < start[g + 1]; i += blockDim.x) gof[i] = g;
}

// ------------------------------------------------------------ binned median
// Grouped-data median of v[0..n) (_core.py:1080-1104): bins floor(v/bw),
// first bin whose cumulative percentage reaches 50, linear interpolation.
// K arrays of the same length n at once (their passes share the loads'
// latency and the barriers).  Histograms of up to HBINS bins in LDS, a
// binary search on the bin value with block-wide counts for a wider range.
// Expects blockDim.x == 256 and n > 0.
constexpr int HBINS = 1024;
constexpr int MED_UNROLL = 4;
template <int K>
__device__ void block_binned_medians(const double* const (&v)[K], const double (&bw)[K], int n, double (&med)[K]) {
  static_assert(HBINS % 256 == 0, "bins per thread");
  __shared__ long long smin[K], smax[K];
  __shared__ int hist[K][HBINS];
  __shared__ int wsum[K][4];
  __shared__ unsigned long long fkey[K];
  __shared__ int bcount[K];
  if (threadIdx.x < K) {
    smin[threadIdx.x] = LLONG_MAX;
    smax[threadIdx.x] = LLONG_MIN;
    fkey[threadIdx.x] = ~0ull;
  }
  long long lmin[K], lmax[K];
  for (int q = 0; q < K; ++q) { lmin[q] = LLONG_MAX; lmax[q] = LLONG_MIN; }
  for (int i0 = threadIdx.x; i0 < n; i0 += MED_UNROLL * blockDim.x) {
    double x[K][MED_UNROLL];
#pragma unroll
    for (int u = 0; u < MED_UNROLL; ++u) {
      const int i = i0 + u * blockDim.x;
#pragma unroll
      for (int q = 0; q < K; ++q) x[q][u] = i < n ? v[q][i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < MED_UNROLL; ++u)
      if (i0 + u * blockDim.x < n)
#pragma unroll
        for (int q = 0; q < K; ++q) {
          const long long b = (long long)floor(x[q][u] / bw[q]);
          lmin[q] = b < lmin[q] ? b : lmin[q];
          lmax[q] = b > lmax[q] ? b : lmax[q];
        }
  }
#pragma unroll
  for (int q = 0; q < K; ++q)
    for (int o = 1; o < WAVE; o <<= 1) {
      const long long a = __shfl_xor(lmin[q], o, WAVE), b = __shfl_xor(lmax[q], o, WAVE);
      lmin[q] = a < lmin[q] ? a : lmin[q];
      lmax[q] = b > lmax[q] ? b : lmax[q];
    }
  __syncthreads();
  if (lane() == 0)
    for (int q = 0; q < K; ++q) {
      atomicMin(&smin[q], lmin[q]);
      atomicMax(&smax[q], lmax[q]);
    }
  __syncthreads();
  long long bmin[K];
  bool fits[K];
  bool any = false;
  for (int q = 0; q < K; ++q) {
    bmin[q] = smin[q];
    fits[q] = smax[q] - bmin[q] + 1 <= HBINS;
    any = any || fits[q];
  }
  const double pct = 100.0 / (double)n;
  long long fbin[K], fcum[K], fcnt[K];
  if (any) {
    for (int q = 0; q < K; ++q)
      if (fits[q])
        for (int i = threadIdx.x; i < HBINS; i += blockDim.x) hist[q][i] = 0;
    __syncthreads();
    for (int i0 = threadIdx.x; i0 < n; i0 += MED_UNROLL * blockDim.x) {
      double x[K][MED_UNROLL];
#pragma unroll
      for (int u = 0; u < MED_UNROLL; ++u) {
        const int i = i0 + u * blockDim.x;
#pragma unroll
        for (int q = 0; q < K; ++q) x[q][u] = i < n && fits[q] ? v[q][i] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < MED_UNROLL; ++u)
        if (i0 + u * blockDim.x < n)
#pragma unroll
          for (int q = 0; q < K; ++q)
            if (fits[q]) atomicAdd(&hist[q][(long long)floor(x[q][u] / bw[q]) - bmin[q]], 1);
    }
    __syncthreads();
    // each thread scans its HBINS / 256 consecutive bins; the first bin whose
    // cumulative count reaches 50 % is kept with its count as one key
    constexpr int per = HBINS / 256;
    const int b0 = threadIdx.x * per;
    const int wid = threadIdx.x / WAVE;
    int loc[K], inc[K];
    for (int q = 0; q < K; ++q) {
      loc[q] = 0;
      if (fits[q])
        for (int k = 0; k < per; ++k) loc[q] += hist[q][b0 + k];
      inc[q] = wave_incl_scan(loc[q]);
      if (lane() == WAVE - 1) wsum[q][wid] = inc[q];
    }
    __syncthreads();
    for (int q = 0; q < K; ++q) {
      if (!fits[q]) continue;
      const int nbin = (int)(smax[q] - bmin[q] + 1);
      int cum = inc[q] - loc[q];
      for (int k = 0; k < wid; ++k) cum += wsum[q][k];
      for (int k = 0; k < per; ++k) {
        const int b = b0 + k;
        const int c = hist[q][b];
        if (c > 0 && b < nbin) {
          cum += c;
          if ((double)cum * pct >= 50.0) {
            atomicMin(&fkey[q], ((unsigned long long)b << 32) | (unsigned)cum);
            break;
          }
        }
      }
    }
    __syncthreads();
    for (int q = 0; q < K; ++q) {
      if (!fits[q]) continue;
      const unsigned long long key = fkey[q];
      long long fb, cc;
      if (key == ~0ull) {  // no bin reaches 50%: last bin, which holds the n-th value
        fb = smax[q] - bmin[q];
        cc = n;
      } else {
        fb = (long long)(key >> 32);
        cc = (long long)(key & 0xffffffffull);
      }
      fbin[q] = fb + bmin[q];
      fcum[q] = cc;
      fcnt[q] = hist[q][fb];
    }
  }
  for (int q = 0; q < K; ++q) {
    if (fits[q]) continue;
    // wide range: binary search on the bin value with block-wide counts
    long long lo = bmin[q], hi = smax[q];   // smallest b with cum(b)*pct >= 50
    while (lo < hi) {
      const long long mid = lo + (hi - lo) / 2;
      if (threadIdx.x == 0) bcount[q] = 0;
      __syncthreads();
      int c = 0;
      for (int i = threadIdx.x; i < n; i += blockDim.x) c += (long long)floor(v[q][i] / bw[q]) <= mid;
      atomicAdd(&bcount[q], c);
      __syncthreads();
      const int cum = bcount[q];
      __syncthreads();
      if ((double)cum * pct >= 50.0) hi = mid; else lo = mid + 1;
    }
    if (threadIdx.x == 0) { bcount[q] = 0; fkey[q] = 0; }
    __syncthreads();
    int c = 0, e = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      const long long b = (long long)floor(v[q][i] / bw[q]);
      c += b <= lo;
      e += b == lo;
    }
    atomicAdd(&bcount[q], c);
    atomicAdd(&fkey[q], (unsigned long long)e);
    __syncthreads();
    fbin[q] = lo;
    fcum[q] = bcount[q];
    fcnt[q] = (long long)fkey[q];
  }
  for (int q = 0; q < K; ++q) {
    const double cumperc = (double)fcum[q] * pct;
    const double binperc = (double)fcnt[q] * pct;
    const double before = cumperc - binperc;
    med[q] = (double)fbin[q] * bw[q] + (50.0 - before) * bw[q] / binperc;
  }
  __syncthreads();
}

// Per molecule group (side): trimmed area, distance / normal statistics.
// side[g*8 + k]: d_mean, d_median, s_mean, s_median, trimmed_area,
// n_trimmed, n_dots, valid.  The three sums share one pairwise tree (same
// length), the two medians their passes.
__global__ __launch_bounds__(256) void k_sc_sides(int G, const int* __restrict__ tgrp,
                                                  const double4* __restrict__ txyz,
                                                  const double* __restrict__ dist,
                                                  const double* __restrict__ rr,
                                                  const int* __restrict__ dgrp, Settings s,
                                                  double* __restrict__ side) {
  for (int g = blockIdx.x; g < G; g += gridDim.x) {
    const int lo = tgrp[g], n = tgrp[g + 1] - lo;
    const int h = g ^ 1;
    const int no = tgrp[h + 1] - tgrp[h];
    // dist and rr hold 0 where the other side has no trimmed dots
    const double* a[3] = {&txyz[lo].w, dist + lo, rr + lo};
    const int64_t st[3] = {4, 1, 1};
    double sums[3];
    block_np_sums<3, 512>(a, st, n, sums);
    double dmean = 0.0, dmed = 0.0, smean = 0.0, smed = 0.0;
    if (n > 0 && no > 0) {
      const double* mv[2] = {dist + lo, rr + lo};
      const double bw[2] = {s.binwidth_dist, s.binwidth_norm};
      double med[2];
      block_binned_medians<2>(mv, bw, n, med);
      dmed = med[0];
      dmean = sums[1] / (double)n;
      smean = -(sums[2] / (double)n);
      smed = -med[1];
    }
    if (threadIdx.x == 0) {
      double* o = side + 8 * (int64_t)g;
      o[0] = dmean; o[1] = dmed; o[2] = smean; o[3] = smed;
      o[4] = sums[0]; o[5] = (double)n; o[6] = (double)(dgrp[g + 1] - dgrp[g]); o[7] = 1.0;
    }
    __syncthreads();
  }
}

// (_core.py:849-861) averages of the two sides per complex.
// res[c*CMS_SC_NRES + k]: sc, area, distance, d_mean, s_mean, then the two
// sides' 8 values at 8 and 16.
__global__ void k_sc_combine(int B, const double* __restrict__ side, double* __restrict__ res) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < B; c += gridDim.x * blockDim.x) {
    const double* s0 = side + 16 * (int64_t)c;
    const double* s1 = s0 + 8;
    double* o = res + 24 * (int64_t)c;
    o[0] = (s0[3] + s1[3]) / 2;
    o[1] = s0[4] + s1[4];
    o[2] = (s0[1] + s1[1]) / 2;
    o[3] = (s0[0] + s1[0]) / 2;
    o[4] = (s0[2] + s1[2]) / 2;
    o[5] = 1.0;
    o[6] = s0[5] + s1[5];
    o[7] = s0[6] + s1[6];
    for (int k = 0; k < 8; ++k) {
      o[8 + k] = s0[k];
      o[16 + k] = s1[k];
    }
  }
}

template __global__ void k_nearest64<false>(const double4*, const int*, const double4*, const int*, int,
                                            const int*, int, int, int64_t, double*, int*);
template __global__ void k_nearest64<true>(const double4*, const int*, const double4*, const int*, int,
                                           const int*, int, int, int64_t, double*, int*);

}  // namespace cms

namespace cms {

// ------------------------------------------------------------ SC trim
// _core.py:1037-1077 keeps a buried dot when its squared distance to every
// accessible dot of the same molecule exceeds band^2 (the minimum of the
// cdist row against band^2).  Only that boolean is needed, so the search is
// a fixed-radius test on a per-group grid of the accessible dots with cell
// edge >= band: every accessible dot within band of a buried dot lies in the
// 27 cells around it, and d^2 = (dx^2 + dy^2) + dz^2 is the cdist value.

// Per group: bounding box of its accessible dots, cell edge (>= band, grown
// until the group has at most 2 n + 64 cells) and dims.
// Trim grid of each group: cells of edge >= band over the box of the
// group's buried dots widened by band + 1e-6 A (at most 2 n + 64 cells for
// n accessible dots, the edge grows otherwise).  An accessible dot outside
// that box is farther than the band from every buried dot along one axis,
// so only the dots inside are binned.  A group without buried dots gets an
// empty grid (x0 = +inf: no dot is inside).
__global__ __launch_bounds__(256) void k_dot_grid_setup(int G, const int* __restrict__ agrp,
                                                        const int* __restrict__ bgrp,
                                                        const double4* __restrict__ bxyz, double band,
                                                        GridInfo* __restrict__ grid, int* __restrict__ ncell) {
  __shared__ double red[6][256];
  for (int g = blockIdx.x; g < G; g += gridDim.x) {
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = bgrp[g] + threadIdx.x; i < bgrp[g + 1]; i += blockDim.x) {
      const double4 a = bxyz[i];
      mn[0] = fmin(mn[0], a.x); mn[1] = fmin(mn[1], a.y); mn[2] = fmin(mn[2], a.z);
      mx[0] = fmax(mx[0], a.x); mx[1] = fmax(mx[1], a.y); mx[2] = fmax(mx[2], a.z);
    }
    for (int k = 0; k < 3; ++k) { red[k][threadIdx.x] = mn[k]; red[3 + k][threadIdx.x] = mx[k]; }
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if (threadIdx.x < o) {
        for (int k = 0; k < 3; ++k) red[k][threadIdx.x] = fmin(red[k][threadIdx.x], red[k][threadIdx.x + o]);
        for (int k = 3; k < 6; ++k) red[k][threadIdx.x] = fmax(red[k][threadIdx.x], red[k][threadIdx.x + o]);
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      const int n = agrp[g + 1] - agrp[g];
      const bool none = bgrp[g + 1] <= bgrp[g];
      const double wide = band + 1e-6;
      double lo[3] = {red[0][0] - wide, red[1][0] - wide, red[2][0] - wide};
      double hi[3] = {red[3][0] + wide, red[4][0] + wide, red[5][0] + wide};
      if (none || !(hi[0] >= lo[0]) || !(hi[1] >= lo[1]) || !(hi[2] >= lo[2])) {
        lo[0] = lo[1] = lo[2] = 0.0;
        hi[0] = hi[1] = hi[2] = 0.0;
      }
      double h = band * (1.0 + 1e-9) + 1e-9;
      if (!(h > 0.0)) h = 1.0;
      int d[3];
      const int64_t cap = 2ll * n + 64;
      for (int it = 0; it < 64; ++it) {
        int64_t tot = 1;
        for (int k = 0; k < 3; ++k) {
          const double e = (hi[k] - lo[k]) / h;
          d[k] = e < 1e6 ? (int)floor(e) + 1 : 1000001;
          tot *= d[k];
        }
        if (tot <= cap) break;
        h *= 1.25;
      }
      GridInfo gi;
      gi.x0 = none ? INFINITY : lo[0]; gi.y0 = lo[1]; gi.z0 = lo[2]; gi.h = h;
      gi.nx = d[0]; gi.ny = d[1]; gi.nz = d[2];
      gi.cell0 = 0;
      gi.rmax = 0.0;
      grid[g] = gi;
      // one slot past the grid's cells: the end of the group's last cell
      ncell[g] = d[0] * d[1] * d[2] + 1;
    }
    __syncthreads();
  }
}

__device__ __forceinline__ int group_of_item(const int* __restrict__ grp, int G, int i) {
  int lo = 0, hi = G;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (grp[mid] <= i) lo = mid; else hi = mid;
  }
  return lo;
}
// The same for the items of a wave: the group of the wave's first item
// (first, wave-uniform) is tried before a search of the lane's own.
__device__ __forceinline__ int group_of_item_wave(const int* __restrict__ grp, int G, int i, int first) {
  const int g0 = group_of_item(grp, G, __builtin_amdgcn_readfirstlane(first));
  const int lo = __builtin_amdgcn_readfirstlane(grp[g0]), hi = __builtin_amdgcn_readfirstlane(grp[g0 + 1]);
  return (i >= lo && i < hi) ? g0 : group_of_item(grp, G, i);
}

__device__ __forceinline__ void dot_cell(const GridInfo& gi, double x, double y, double z, int* ix, int* iy,
                                         int* iz) {
  *ix = min(max((int)floor((x - gi.x0) / gi.h), 0), gi.nx - 1);
  *iy = min(max((int)floor((y - gi.y0) / gi.h), 0), gi.ny - 1);
  *iz = min(max((int)floor((z - gi.z0) / gi.h), 0), gi.nz - 1);
}
// The cell of a point inside the grid's box (false outside it, or for an
// empty grid).
__device__ __forceinline__ bool dot_cell_in(const GridInfo& gi, double x, double y, double z, int* ix, int* iy,
                                            int* iz) {
  const double fx = floor((x - gi.x0) / gi.h), fy = floor((y - gi.y0) / gi.h), fz = floor((z - gi.z0) / gi.h);
  if (!(fx >= 0.0 && fx < (double)gi.nx && fy >= 0.0 && fy < (double)gi.ny && fz >= 0.0 && fz < (double)gi.nz))
    return false;
  *ix = (int)fx;
  *iy = (int)fy;
  *iz = (int)fz;
  return true;
}

// The accessible dots of each group inside its trim grid, sorted by cell
// (sorted[tstart[cell] .. tstart[cell + 1])), one workgroup per group: a
// count per cell (LDS counters for grids of at most TG_LDS cells, otherwise
// device atomics on the zero-filled tcount), a scan over the group's own
// cells, the dots placed through the counters turned cursors.  A group's
// sorted dots sit at its own dot offset dgrp[g] (there are at most as many
// as it has dots), and the slot after its cells holds their end.  Within a
// cell the order is whatever the atomics give; the trim test only asks
// whether any dot lies within the band.
constexpr int TG_LDS = 8192;
constexpr int TG_NT = 256;
__global__ __launch_bounds__(TG_NT) void k_dot_grid_build(int G, const int* __restrict__ dgrp,
                                                          const double4* __restrict__ dpos,
                                                          const int8_t* __restrict__ dbur,
                                                          const GridInfo* __restrict__ grid,
                                                          int* __restrict__ tcount, int* __restrict__ tstart,
                                                          double4* __restrict__ sorted) {
  __shared__ int cnt[TG_LDS];
  __shared__ int wsum[TG_NT / WAVE];
  for (int g = blockIdx.x; g < G; g += gridDim.x) {
    const GridInfo gi = grid[g];
    const int nc = gi.nx * gi.ny * gi.nz;
    const int base = gi.cell0;
    const bool lds = nc <= TG_LDS;
    int* const c = lds ? cnt : tcount + base;
    if (lds)
      for (int k = threadIdx.x; k < nc; k += TG_NT) cnt[k] = 0;
    __syncthreads();
    const int d0 = dgrp[g], d1 = dgrp[g + 1];
    auto cell_of_dot = [&](int d, double4* a) {
      if (dbur[d]) return -1;
      *a = dpos[d];
      int ix, iy, iz;
      return dot_cell_in(gi, a->x, a->y, a->z, &ix, &iy, &iz) ? (ix * gi.ny + iy) * gi.nz + iz : -1;
    };
    for (int d = d0 + threadIdx.x; d < d1; d += TG_NT) {
      double4 a;
      const int k = cell_of_dot(d, &a);
      if (k >= 0) atomicAdd(&c[k], 1);
    }
    __syncthreads();
    int carry = d0;
    for (int k0 = 0; k0 < nc; k0 += TG_NT) {
      const int k = k0 + threadIdx.x;
      const int v = k < nc ? (lds ? cnt[k] : __hip_atomic_load(&c[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                           : 0;
      const int inc = wave_incl_scan(v);
      if (lane() == WAVE - 1) wsum[threadIdx.x / WAVE] = inc;
      __syncthreads();
      int off = 0, tot = 0;
      for (int w = 0; w < TG_NT / WAVE; ++w) {
        off += w < (int)(threadIdx.x / WAVE) ? wsum[w] : 0;
        tot += wsum[w];
      }
      if (k < nc) {
        const int start = carry + off + inc - v;
        tstart[base + k] = start;
        c[k] = start;
      }
      carry += tot;
      __syncthreads();
    }
    if (threadIdx.x == 0) tstart[base + nc] = carry;
    __syncthreads();
    for (int d = d0 + threadIdx.x; d < d1; d += TG_NT) {
      double4 a;
      const int k = cell_of_dot(d, &a);
      if (k >= 0) sorted[atomicAdd(&c[k], 1)] = a;
    }
    __syncthreads();
  }
}

// Trim flag of every buried dot: 1 = kept (no accessible dot of its group
// within band), 0 = trimmed.  A group without accessible dots keeps all.
// L consecutive lanes test one buried dot: the dot is trimmed when an
// accessible dot of its group lies within the band.  The cell edge is at
// least the band, so the 27 cells around the dot's cell hold every such dot.
template <int L>
__global__ __launch_bounds__(256) void k_trim_grid(Count nb, const double4* __restrict__ bxyz,
                                                   const int* __restrict__ bgrp, int G,
                                                   const GridInfo* __restrict__ grid, const int* __restrict__ cstart,
                                                   const double4* __restrict__ sorted, double band2,
                                                   int* __restrict__ flag) {
  static_assert(WAVE % L == 0, "lanes per dot must divide the wave");
  const int n = nb.get();
  const int sub = lane() % L;
  const uint64_t submask = L == WAVE ? ~0ull : (((1ull << L) - 1) << (lane() - sub));
  const int64_t nthreads = (int64_t)gridDim.x * blockDim.x;
  // the trip count is uniform across a wave so that the group lookup may use
  // wave-level operations; lanes past the end idle
  const int64_t nq_pad = (((int64_t)n * L + WAVE - 1) / WAVE) * WAVE;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t - lane() < nq_pad; t += nthreads) {
    const int i = (int)(t / L);
    const int first = (int)((t - lane()) / L);
    if (i >= n) continue;
    const GridInfo gi = grid[group_of_item_wave(bgrp, G, i, first)];
    const double4 q = bxyz[i];
    int ix, iy, iz;
    dot_cell(gi, q.x, q.y, q.z, &ix, &iy, &iz);
    bool hit = false;
    // only cells whose box comes within the band of the dot, one z-run per
    // (x, y) column
    const double band = sqrt(band2);
    const double l2 = (band + 1e-6) * (band + 1e-6);
    const double fx = q.x - (gi.x0 + ix * gi.h), fy = q.y - (gi.y0 + iy * gi.h), fz = q.z - (gi.z0 + iz * gi.h);
    for (int col = 0; col < 9 && !hit; ++col) {
      int zlo, zhi;
      if (!column_within(gi, fx, fy, fz, ix, iy, iz, col / 3, col % 3, l2, &zlo, &zhi)) continue;
      const int cx = ix + col / 3 - 1, cy = iy + col % 3 - 1;
      const int c0 = gi.cell0 + (cx * gi.ny + cy) * gi.nz;
      const int lo = cstart[c0 + zlo], hi = cstart[c0 + zhi + 1];
      if constexpr (L == 1) {
        for (int u = lo; u < hi; ++u) {
          const double4 a = sorted[u];
          const double ddx = q.x - a.x, ddy = q.y - a.y, ddz = q.z - a.z;
          if ((ddx * ddx + ddy * ddy) + ddz * ddz <= band2) { hit = true; break; }
        }
      } else {
        // uniform trip count over the L lanes of the dot
        for (int u0 = lo; u0 < hi; u0 += L) {
          const int u = u0 + sub;
          bool h = false;
          if (u < hi) {
            const double4 a = sorted[u];
            const double ddx = q.x - a.x, ddy = q.y - a.y, ddz = q.z - a.z;
            h = (ddx * ddx + ddy * ddy) + ddz * ddz <= band2;
          }
          if (ballot(h) & submask) { hit = true; break; }
        }
      }
    }
    if (sub == 0) flag[i] = hit ? 0 : 1;
  }
}
template __global__ void k_trim_grid<1>(Count, const double4*, const int*, int, const GridInfo*, const int*,
                                        const double4*, double, int*);
template __global__ void k_trim_grid<4>(Count, const double4*, const int*, int, const GridInfo*, const int*,
                                        const double4*, double, int*);
template __global__ void k_trim_grid<32>(Count, const double4*, const int*, int, const GridInfo*, const int*,
                                         const double4*, double, int*);

}  // namespace cms
