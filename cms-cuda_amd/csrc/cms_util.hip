// Device-wide utilities: exclusive scan, fills, group offsets.
#include "cms_device.h"
#include "cms_internal.h"
#include "cms_util.h"

namespace cms {

constexpr int SCAN_THREADS = 256;

__device__ __forceinline__ int block_excl_scan(int v, int* total) {
  // 256 threads = 4 waves
  __shared__ int wsum[SCAN_THREADS / WAVE];
  const int l = lane(), wid = threadIdx.x / WAVE;
  const int inc = wave_incl_scan(v);
  if (l == WAVE - 1) wsum[wid] = inc;
  __syncthreads();
  int off = 0, tot = 0;
  for (int i = 0; i < SCAN_THREADS / WAVE; ++i) {
    if (i < wid) off += wsum[i];
    tot += wsum[i];
  }
  __syncthreads();
  *total = tot;
  return off + inc - v;
}

// Multi-block scan: nb blocks, each owns a contiguous chunk of whole tiles
// (SCAN_PER items per thread, 4,096 per tile) covering n, read on the device.
constexpr int SCAN_PER = 16;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_PER;
__device__ __forceinline__ int64_t scan_chunk(int64_t n, int nb) {
  const int64_t c = (n + nb - 1) / nb;
  return (c + SCAN_TILE - 1) / SCAN_TILE * SCAN_TILE;
}

// The SCAN_PER items of thread t in the tile at base (zero past hi); vector
// loads when the run is whole (base and the buffers are 16-byte aligned).
__device__ __forceinline__ void scan_load(const int* __restrict__ in, int64_t base, int64_t hi, int* v) {
  const int64_t i0 = base + (int64_t)threadIdx.x * SCAN_PER;
  if (i0 + SCAN_PER <= hi) {
    const int4* p = reinterpret_cast<const int4*>(in + i0);
#pragma unroll
    for (int k = 0; k < SCAN_PER / 4; ++k) {
      const int4 q = p[k];
      v[4 * k] = q.x; v[4 * k + 1] = q.y; v[4 * k + 2] = q.z; v[4 * k + 3] = q.w;
    }
  } else {
#pragma unroll
    for (int k = 0; k < SCAN_PER; ++k) v[k] = i0 + k < hi ? in[i0 + k] : 0;
  }
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_partial(const int* __restrict__ in, Count cnt, int nb,
                                                              int* __restrict__ part) {
  const int64_t n = cnt.get();
  const int64_t chunk = scan_chunk(n, nb);
  const int64_t lo = (int64_t)blockIdx.x * chunk;
  const int64_t hi = lo + chunk < n ? lo + chunk : n;
  int acc = 0;
  for (int64_t base = lo; base < hi; base += SCAN_TILE) {
    int v[SCAN_PER];
    scan_load(in, base, hi, v);
#pragma unroll
    for (int k = 0; k < SCAN_PER; ++k) acc += v[k];
  }
  int tot;
  block_excl_scan(acc, &tot);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_final(const int* __restrict__ in, int* __restrict__ out,
                                                            Count cnt, int nb, const int* __restrict__ part) {
  const int64_t n = cnt.get();
  const int64_t chunk = scan_chunk(n, nb);
  const int64_t lo = (int64_t)blockIdx.x * chunk;
  const int64_t hi = lo + chunk < n ? lo + chunk : n;
  // the block's carry: the totals of the blocks before it (nb <= 1024, so
  // each block adds them up itself instead of a separate top-level pass)
  int before = 0, all = 0;
  for (int i = threadIdx.x; i < nb; i += SCAN_THREADS) {
    const int v = part[i];
    if (i < (int)blockIdx.x) before += v;
    all += v;
  }
  int carry, total;
  block_excl_scan(before, &carry);
  block_excl_scan(all, &total);
  for (int64_t base = lo; base < hi; base += SCAN_TILE) {
    int v[SCAN_PER];
    scan_load(in, base, hi, v);
    int run = 0;
#pragma unroll
    for (int k = 0; k < SCAN_PER; ++k) run += v[k];
    int tot;
    int e = carry + block_excl_scan(run, &tot);
    const int64_t i0 = base + (int64_t)threadIdx.x * SCAN_PER;
    if (i0 + SCAN_PER <= hi) {
      int4* p = reinterpret_cast<int4*>(out + i0);
#pragma unroll
      for (int k = 0; k < SCAN_PER / 4; ++k) {
        int4 q;
        q.x = e; e += v[4 * k];
        q.y = e; e += v[4 * k + 1];
        q.z = e; e += v[4 * k + 2];
        q.w = e; e += v[4 * k + 3];
        p[k] = q;
      }
    } else {
      for (int k = 0; k < SCAN_PER && i0 + k < hi; ++k) {
        out[i0 + k] = e;
        e += v[k];
      }
    }
    carry += tot;
  }
  if (blockIdx.x == nb - 1 && threadIdx.x == 0) out[n] = total;
}

// Small arrays: one workgroup of 1024 threads, each owning a contiguous run
// of at most 64 elements.
constexpr int SS_THREADS = 1024;
constexpr int64_t SS_MAX = 64 * SS_THREADS;
__global__ __launch_bounds__(SS_THREADS) void k_scan_single(const int* __restrict__ in, int* __restrict__ out,
                                                            Count cnt) {
  __shared__ int wsum[SS_THREADS / WAVE];
  const int64_t n = cnt.get();
  const int64_t per = (n + SS_THREADS - 1) / SS_THREADS;
  const int64_t lo = threadIdx.x * per;
  const int64_t hi = lo + per < n ? lo + per : n;
  int acc = 0;
  for (int64_t i = lo; i < hi; ++i) acc += in[i];
  const int l = lane(), wid = threadIdx.x / WAVE;
  const int inc = wave_incl_scan(acc);
  if (l == WAVE - 1) wsum[wid] = inc;
  __syncthreads();
  int off = 0, tot = 0;
  for (int k = 0; k < SS_THREADS / WAVE; ++k) {
    if (k < wid) off += wsum[k];
    tot += wsum[k];
  }
  int run = off + inc - acc;
  for (int64_t i = lo; i < hi; ++i) {
    const int v = in[i];
    out[i] = run;
    run += v;
  }
  if (threadIdx.x == 0) out[n] = tot;
}

void scan_excl(const int* in, int* out, Count n, int64_t cap, int* part, hipStream_t st) {
  // part must hold at least 1025 ints; cap bounds the count
  if (cap <= SS_MAX) {
    k_scan_single<<<1, SS_THREADS, 0, st>>>(in, out, n);
    return;
  }
  int nb = (int)((cap + 4095) / 4096);
  if (nb < 1) nb = 1;
  if (nb > 1024) nb = 1024;
  k_scan_partial<<<nb, SCAN_THREADS, 0, st>>>(in, n, nb, part);
  k_scan_final<<<nb, SCAN_THREADS, 0, st>>>(in, out, n, nb, part);
}

void scan_excl(const int* in, int* out, int64_t n, int* part, hipStream_t st) {
  scan_excl(in, out, host_count((int)n), n, part, st);
}

// Stage total: meta[slot] = total if it fits the capacity, else 0 with the
// overflow bit set and the stage's group offsets zeroed, so that every
// later kernel of the run sees an empty stage and stays inside its buffers.
__global__ __launch_bounds__(256) void k_stage_total(const int* __restrict__ scan, Count at, int64_t cap,
                                                     int* __restrict__ meta, int slot, int* __restrict__ grp,
                                                     int G) {
  __shared__ int ovf;
  if (threadIdx.x == 0) {
    const int v = scan[at.get()];
    ovf = (int64_t)v > cap ? 1 : 0;
    meta[M_RAW + slot] = v;
    meta[slot] = ovf ? 0 : v;
    if (ovf) meta[M_OVF] |= 1 << slot;
  }
  __syncthreads();
  if (ovf && grp)
    for (int g = threadIdx.x; g <= G; g += blockDim.x) grp[g] = 0;
}

// Buried / accessible dot totals of the run: nB = bscan[nD], nA = nD - nB.
__global__ void k_list_totals(const int* __restrict__ bscan, int* __restrict__ meta) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const int nD = meta[M_D];
    const int nB = bscan[nD];
    meta[M_B] = nB;
    meta[M_A] = nD - nB;
  }
}

__global__ void k_copy_word(const int* __restrict__ src, Count at, int* __restrict__ dst) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *dst = src[at.get()];
}

__global__ void k_fill_int(int* p, int64_t n, int v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void k_fill_u64(unsigned long long* p, int64_t n, unsigned long long v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void k_fill_int8(int8_t* p, int64_t n, int8_t v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void k_fill_double(double* p, int64_t n, double v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

void fill_int(int* p, int64_t n, int v, hipStream_t st) {
  if (n <= 0) return;
  int g = (int)((n + 255) / 256);
  if (g > 4096) g = 4096;
  k_fill_int<<<g, 256, 0, st>>>(p, n, v);
}
void fill_u64(unsigned long long* p, int64_t n, unsigned long long v, hipStream_t st) {
  if (n <= 0) return;
  int g = (int)((n + 255) / 256);
  if (g > 4096) g = 4096;
  k_fill_u64<<<g, 256, 0, st>>>(p, n, v);
}
__global__ void k_fill_int_count(int* p, Count cnt, int v) {
  const int64_t n = cnt.get();
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
void fill_int(int* p, Count n, int64_t cap, int v, hipStream_t st) {
  if (cap <= 0) return;
  int g = (int)((cap + 255) / 256);
  if (g > 4096) g = 4096;
  k_fill_int_count<<<g, 256, 0, st>>>(p, n, v);
}
void fill_int8(int8_t* p, int64_t n, int8_t v, hipStream_t st) {
  if (n <= 0) return;
  int g = (int)((n + 255) / 256);
  if (g > 4096) g = 4096;
  k_fill_int8<<<g, 256, 0, st>>>(p, n, v);
}
void fill_double(double* p, int64_t n, double v, hipStream_t st) {
  if (n <= 0) return;
  int g = (int)((n + 255) / 256);
  if (g > 4096) g = 4096;
  k_fill_double<<<g, 256, 0, st>>>(p, n, v);
}

struct FillArgs {
  FillSpec e[FILL_BATCH];
};
__global__ void k_fill_batch(FillArgs f) {
  const FillSpec& s = f.e[blockIdx.y];
  const int64_t c = s.n.get();
  const int64_t n = c < s.cap ? c : s.cap;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    if (s.bytes == 1)
      static_cast<uint8_t*>(s.p)[i] = (uint8_t)s.v;
    else if (s.bytes == 4)
      static_cast<uint32_t*>(s.p)[i] = (uint32_t)s.v;
    else
      static_cast<unsigned long long*>(s.p)[i] = s.v;
  }
}
void fill_batch(std::initializer_list<FillSpec> fills, hipStream_t st) {
  FillArgs a{};
  int k = 0;
  int64_t cap = 0;
  for (const FillSpec& f : fills) {
    if (f.cap <= 0) continue;
    if (k == FILL_BATCH) {
      // more entries than one launch holds: launch these, start again
      k_fill_batch<<<dim3((unsigned)std::min<int64_t>((cap + 255) / 256, 4096), k), 256, 0, st>>>(a);
      k = 0;
      cap = 0;
    }
    a.e[k++] = f;
    cap = std::max(cap, f.cap);
  }
  if (k) k_fill_batch<<<dim3((unsigned)std::min<int64_t>((cap + 255) / 256, 4096), k), 256, 0, st>>>(a);
}

// Pack (x, y, z) + radius into double4 and derive per-atom complex id and
// molecule from the complex offsets and binder counts.
__global__ void k_pack_atoms(int ntot, const double* __restrict__ xyz, const double* __restrict__ rad,
                             const double* __restrict__ dens_in, double dens_default,
                             const int* __restrict__ aoff, const int* __restrict__ nb0, int B, int full,
                             double4* __restrict__ atom, double* __restrict__ dens,
                             int* __restrict__ acplx, int8_t* __restrict__ amol) {
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < ntot; g += gridDim.x * blockDim.x) {
    if (!full) {
      atom[g] = make_double4(xyz[3 * (int64_t)g], xyz[3 * (int64_t)g + 1], xyz[3 * (int64_t)g + 2], rad[g]);
      continue;
    }
    int lo = 0, hi = B;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (aoff[mid] <= g) lo = mid; else hi = mid;
    }
    atom[g] = make_double4(xyz[3 * (int64_t)g], xyz[3 * (int64_t)g + 1], xyz[3 * (int64_t)g + 2], rad[g]);
    dens[g] = dens_in ? dens_in[g] : dens_default;
    acplx[g] = lo;
    amol[g] = (g - aoff[lo]) >= nb0[lo] ? 1 : 0;
  }
}

// First index of each group in a list sorted by group (group of item i is
// key(i)); writes start[0..G], start[G] = n.
__global__ void k_group_starts_from_atoms(Count cnt, const int* __restrict__ items,
                                          const int* __restrict__ acplx, const int8_t* __restrict__ amol,
                                          int G, int* __restrict__ start) {
  const int n = cnt.get();
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += gridDim.x * blockDim.x) {
    const int k = i < n ? 2 * acplx[items[i]] + amol[items[i]] : G;
    const int kp = i > 0 ? 2 * acplx[items[i - 1]] + amol[items[i - 1]] : -1;
    for (int g = kp + 1; g <= k; ++g) start[g] = i;
  }
}

// out[g] = scan[idx[g] * mul] for g in [0, G]
__global__ void k_gather_scan(int G, const int* __restrict__ idx, int mul, const int* __restrict__ scan,
                              int* __restrict__ out) {
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g <= G; g += gridDim.x * blockDim.x)
    out[g] = scan[(int64_t)idx[g] * mul];
}

// Dot group bases: per group the four stage totals, group offsets (scan over
// G) and per-stage bases so that a unit's output offset is
// dbase[4g + stage] + stage_scan[unit].  k_dot_stage_starts gathers each
// group's four stage starts and ends (a thread per group: the gathers of all
// groups are in flight together); k_dot_bases, one workgroup of 1,024
// threads each taking a contiguous run of groups, scans the run totals and
// writes the bases.
__global__ void k_dot_stage_starts(int G, const int* __restrict__ rgrp, const int* __restrict__ pgrp,
                                   const int* __restrict__ prgrp, const int* __restrict__ cscan,
                                   const int* __restrict__ t1scan, const int* __restrict__ t2scan,
                                   const int* __restrict__ kscan, int4* __restrict__ ss) {
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < G; g += gridDim.x * blockDim.x) {
    ss[2 * g] = make_int4(cscan[rgrp[g]], t1scan[pgrp[g]], t2scan[pgrp[g]], kscan[prgrp[g]]);
    ss[2 * g + 1] = make_int4(cscan[rgrp[g + 1]], t1scan[pgrp[g + 1]], t2scan[pgrp[g + 1]], kscan[prgrp[g + 1]]);
  }
}
__global__ __launch_bounds__(DOT_BASES_THREADS) void k_dot_bases(int G, const int4* __restrict__ ss,
                                                                 int64_t* __restrict__ dbase,
                                                                 int* __restrict__ dgrp) {
  __shared__ int wsum[DOT_BASES_THREADS / WAVE];
  const int per = (G + DOT_BASES_THREADS - 1) / DOT_BASES_THREADS;
  const int g0 = min(G, (int)threadIdx.x * per), g1 = min(G, g0 + per);
  int S = 0;
  for (int g = g0; g < g1; ++g) {
    const int4 a = ss[2 * g], e = ss[2 * g + 1];
    S += (((e.x - a.x) + (e.y - a.y)) + (e.z - a.z)) + (e.w - a.w);
  }
  const int inc = wave_incl_scan(S);
  const int wid = threadIdx.x / WAVE;
  if (lane() == WAVE - 1) wsum[wid] = inc;
  __syncthreads();
  int off = 0;
  for (int i = 0; i < wid; ++i) off += wsum[i];
  int start = off + inc - S;
  for (int g = g0; g < g1; ++g) {
    const int4 a = ss[2 * g], e = ss[2 * g + 1];
    const int T0 = e.x - a.x, T1 = e.y - a.y, T2 = e.z - a.z, T3 = e.w - a.w;
    dgrp[g] = start;
    dbase[4 * g + 0] = (int64_t)start - a.x;
    dbase[4 * g + 1] = (int64_t)start + T0 - a.y;
    dbase[4 * g + 2] = (int64_t)start + T0 + T1 - a.z;
    dbase[4 * g + 3] = (int64_t)start + T0 + T1 + T2 - a.w;
    start += ((T0 + T1) + T2) + T3;
  }
  if (threadIdx.x == DOT_BASES_THREADS - 1) dgrp[G] = off + inc;
}

// ------------------------------------------------------------ failed complexes
// A complex whose status word flags coincident atoms or a neighbour / buried
// row over capacity gets empty rows right after the neighbour pass: no
// pairs, probes or dots, so it cannot disturb the rest of the batch.
__global__ void k_drop_failed_rows(Count cnt, const int* __restrict__ ratom, const int* __restrict__ acplx,
                                   const int* __restrict__ status, int* __restrict__ nn, int* __restrict__ nb,
                                   int* __restrict__ pcnt) {
  const int n = cnt.get();
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
    if (status[acplx[ratom[r]]] & (ST_COINCIDENT | ST_NBR_CAP | ST_BUR_CAP)) {
      nn[r] = 0;
      nb[r] = 0;
      pcnt[r] = 0;
    }
  }
}

// Outputs of every complex with a non-zero status word set to 0: per-atom
// values (width doubles per atom), totals (tstride per complex), SC records
// (sc_width per complex).  One workgroup per complex.
__global__ void k_zero_failed(int B, const int* __restrict__ status, const int* __restrict__ aoff, double* per_atom,
                              int width, double* totals, int tstride, double* sc, int sc_width) {
  for (int c = blockIdx.x; c < B; c += gridDim.x) {
    if (status[c] == 0) continue;
    if (per_atom)
      for (int64_t i = (int64_t)width * aoff[c] + threadIdx.x; i < (int64_t)width * aoff[c + 1]; i += blockDim.x)
        per_atom[i] = 0.0;
    if (totals && (int)threadIdx.x < tstride) totals[(int64_t)c * tstride + threadIdx.x] = 0.0;
    if (sc)
      for (int k = threadIdx.x; k < sc_width; k += blockDim.x) sc[(int64_t)c * sc_width + k] = 0.0;
  }
}

}  // namespace cms
