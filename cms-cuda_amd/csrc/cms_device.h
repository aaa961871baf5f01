// Device-side helpers for the CMS / SC kernels (gfx950, wave64).
//
// Floating-point contract: every kernel is compiled with -ffp-contract=off,
// so a*b + c is never fused, and each helper below performs its additions
// in a fixed order:
//   dot_seq : (a0*b0 + a1*b1) + a2*b2    (np.sum(axis=-1), linalg.norm, cdist)
//   dot_alt : (a0*b0 + a2*b2) + a1*b1    (np.einsum('ij,ij->i'))
// With correctly rounded +, -, *, / and sqrt this reproduces the NumPy
// reference bit for bit; only the transcendental functions (sin, cos,
// acos, atan2, exp) may differ in the last ulp.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cms {

constexpr int WAVE = 64;
constexpr int MAX_SUBDIV = 100;
constexpr int KCAP = 256;      // max neighbours per atom row
constexpr int BCAP = 256;      // max buried (other-molecule) partners per atom
constexpr int CMASK_WORDS = 16; // convex points per row whose collision outcome is kept (x 64)
constexpr int NEARCAP = 128;   // max near low probes per low probe
constexpr int8_t ATTEN_BLOCKER = 1;
constexpr int8_t ATTEN_2 = 2;
constexpr int8_t ATTEN_BURIED = 5;
constexpr double TWO_PI = 6.283185307179586;

// Dot descriptor kinds: what re-evaluates the dot from atom coordinates
// with its subdivision counts frozen (backward pass, frozen-topology CMS).
constexpr int DK_CONVEX = 1, DK_ARC1 = 2, DK_ARC2 = 3, DK_CONCAVE = 4;

// status bits (per complex)
constexpr int ST_COINCIDENT = 1;
constexpr int ST_NBR_CAP = 2;
constexpr int ST_BUR_CAP = 4;
constexpr int ST_NEAR_CAP = 8;
constexpr int ST_POOL = 16;      // a device pool was too small: grow and rerun

struct V3 {
  double x, y, z;
};

__host__ __device__ __forceinline__ V3 v3(double x, double y, double z) { return V3{x, y, z}; }
__host__ __device__ __forceinline__ V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ __forceinline__ V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ __forceinline__ V3 operator*(V3 a, double s) { return V3{a.x * s, a.y * s, a.z * s}; }
__host__ __device__ __forceinline__ V3 operator/(V3 a, double s) { return V3{a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ V3 neg(V3 a) { return V3{-a.x, -a.y, -a.z}; }

__host__ __device__ __forceinline__ double dot_seq(V3 a, V3 b) {
  double p0 = a.x * b.x, p1 = a.y * b.y, p2 = a.z * b.z;
  return (p0 + p1) + p2;
}
__host__ __device__ __forceinline__ double dot_alt(V3 a, V3 b) {
  double p0 = a.x * b.x, p1 = a.y * b.y, p2 = a.z * b.z;
  return (p0 + p2) + p1;
}
__host__ __device__ __forceinline__ double norm_seq(V3 a) { return sqrt(dot_seq(a, a)); }
// norm_seq(a) < lim, decided without the square root unless q = |a|^2 lies
// within a relative 2^-40 of lim^2.  Outside that band the rounded square
// root provably sits on the same side of lim: q < fl(fl(lim^2)(1 - 2^-40))
// implies sqrt(q) < lim (1 - 2^-42), whose rounding stays below lim, and
// q > fl(fl(lim^2)(1 + 2^-40)) implies sqrt(q) > lim (1 + 2^-43).  NaN, 0
// and infinite cases fall into the exact branch or agree with it.
__host__ __device__ __forceinline__ bool norm_lt(V3 a, double lim) {
  const double q = dot_seq(a, a);
  const double l2 = lim * lim;
  if (q < l2 * (1.0 - 0x1p-40)) return true;
  if (q > l2 * (1.0 + 0x1p-40)) return false;
  return sqrt(q) < lim;
}
// a x b with NumPy's component formulas
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
  return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ V3 unit(V3 a) { return a / norm_seq(a); }

__device__ __forceinline__ V3 ld3(const double4& a) { return V3{a.x, a.y, a.z}; }

// ---------------------------------------------------------------- wave ops
__device__ __forceinline__ int lane() { return __lane_id(); }
__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }
__device__ __forceinline__ int lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}
__device__ __forceinline__ int wave_sum_int(int v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
  return v;
}
__device__ __forceinline__ bool wave_any(bool p) { return __ballot(p) != 0ull; }
// inclusive prefix sum over the wave, by DPP moves (no LDS traffic): sums
// within each row of 16 lanes (row_shr 1, 2, 4, 8; lanes shifted in from
// outside the row read 0), then row 1 and row 3 add the last lane of the
// row before (row_bcast:15), rows 2 and 3 the last lane of row 1
// (row_bcast:31); rows outside a row_mask keep old = 0 and add nothing.
__device__ __forceinline__ int wave_incl_scan(int v) {
  v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, true);    // row_shr:1
  v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, true);    // row_shr:2
  v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, true);    // row_shr:4
  v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, true);    // row_shr:8
  v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);   // row_bcast:15
  v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);   // row_bcast:31
  return v;
}
__device__ __forceinline__ double wave_min_double(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, WAVE));
  return v;
}

// ------------------------------------------------------------ pairs
// Torus of the pair (i, j): axis u, centre t, radius rij, whether the centre
// lies between the atoms; false when the probe cannot touch both spheres.
__device__ __forceinline__ bool pair_torus(V3 xi, double ri, V3 xj, double rj, double rp, V3* u, V3* t,
                                           double* rij, bool* between) {
  const double eri = ri + rp, erj = rj + rp;
  const V3 diff = xj - xi;
  const double dsq = dot_alt(diff, diff);
  const double dij = sqrt(dsq);
  const double asymm = (eri * eri - erj * erj) / dij;
  *between = fabs(asymm) < dij;
  *u = diff / dij;
  *t = (xi + xj) * 0.5 + *u * (asymm * 0.5);
  const double far_sq = (eri + erj) * (eri + erj) - dsq;
  const double cont_sq = dsq - (ri - rj) * (ri - rj);
  const bool ok = (far_sq > 0.0) && (cont_sq > 0.0);
  *rij = ok ? (0.5 * sqrt(far_sq)) * sqrt(cont_sq) / dij : 0.0;
  return ok;
}

// pair_torus's decision alone (same expressions, so the same outcome): the
// probe can touch both spheres.
__device__ __forceinline__ bool pair_geom_ok(V3 xi, double ri, V3 xj, double rj, double rp) {
  const double eri = ri + rp, erj = rj + rp;
  const V3 diff = xj - xi;
  const double dsq = dot_alt(diff, diff);
  const double far_sq = (eri + erj) * (eri + erj) - dsq;
  const double cont_sq = dsq - (ri - rj) * (ri - rj);
  return (far_sq > 0.0) && (cont_sq > 0.0);
}

// A pair matters only for the toroid queue (a non-blocker end) or for probes
// (some atom of a triple i, j, k > j is a non-blocker); the toroid queue takes
// pairs with a non-blocker first atom, or a non-blocker second atom when the
// probe has a radius.
__device__ __forceinline__ bool pair_relevant(int8_t atti, int8_t attj, int nbmax, int j) {
  return atti > ATTEN_BLOCKER || attj > ATTEN_BLOCKER || nbmax > j;
}
__device__ __forceinline__ bool pair_toroid(int8_t atti, int8_t attj, double rp) {
  return (atti > ATTEN_BLOCKER) || ((attj > ATTEN_BLOCKER) && (rp > 0.0));
}

// ------------------------------------------------------- arc subdivision
// Number of points a_i = delta*(i+1/2) <= angle with i < MAX_SUBDIV, where
// delta = 1/(sqrt(density)*rad).  NaN angle or infinite delta give 0.
__host__ __device__ __forceinline__ int subdiv_count(double delta, double angle) {
  double raw = angle / delta;
  if (!(raw > 0.0) || !(raw < 1e300)) return 0;
  const int lim = (int)fmin((double)MAX_SUBDIV, ceil(raw) + 1.0);
  // a_i is non-decreasing in i (a rounded product of delta > 0 and an exact
  // i + 1/2), so the count is the first i < lim with a_i > angle: start
  // from the closed-form estimate and step to it with the exact test
  int n = (int)fmin(fmax(floor(raw - 0.5) + 1.0, 0.0), (double)lim);
  while (n > 0 && delta * ((double)(n - 1) + 0.5) > angle) --n;
  while (n < lim && delta * ((double)n + 0.5) <= angle) ++n;
  return n;
}

// cos(a), sin(a) and atan2(y, x) out of line: the polynomial constants of
// the library routines are then materialised inside the call instead of
// being held in registers across the callers' loops.  cos and sin come from
// one sincos call: the device library builds all three from the same
// argument reduction and the same pair of reduced-range polynomials, so the
// results are those of cos() and sin() bit for bit (compared on 5.4e8
// angles up to 1e9 rad), with the reduction done once (the step's trig-heavy
// passes 2-10 % faster).
__device__ __attribute__((noinline)) inline double2 cos_sin(double a) {
  double sn, cs;
  sincos(a, &sn, &cs);
  return make_double2(cs, sn);
}
__device__ __attribute__((noinline)) inline double atan2_ool(double y, double x) {
  return atan2(y, x);
}

// The last index k < n with a[k] <= q, for a non-decreasing a[0 .. n] with
// a[0] <= q < a[n]: binary lifting whose step count depends on n alone (a
// wave-uniform loop where n is uniform) and whose probes are clamped to n
// instead of branching on the bound (a[n] > q never moves the result).
__device__ __forceinline__ int last_le(const int* a, int n, int q) {
  int lo = 0;
  for (int step = n > 0 ? 1 << (31 - __builtin_clz(n)) : 0; step > 0; step >>= 1) {
    const int c = min(lo + step, n);
    lo = a[c] <= q ? c : lo;
  }
  return lo;
}

// Circle point i: cen + x*c + y*s with c = rad*cos(a), s = rad*sin(a)
__device__ __forceinline__ V3 subdiv_point(V3 cen, V3 x, V3 y, double rad, double delta, int i) {
  double a = delta * ((double)i + 0.5);
  const double2 cs = cos_sin(a);
  double c = rad * cs.x;
  double s = rad * cs.y;
  return (cen + x * c) + y * s;
}

// The same with cos(a) and sin(a) given
__device__ __forceinline__ V3 subdiv_point_cs(V3 cen, V3 x, V3 y, double rad, double cosa, double sina) {
  double c = rad * cosa;
  double s = rad * sina;
  return (cen + x * c) + y * s;
}

// Angle from x to v about axis in [0, 2pi), NaN if v is (numerically)
// perpendicular to the plane.  Also returns y = axis x x.
__device__ __forceinline__ double arc_angle(V3 x, V3 v, V3 axis, V3* y_out) {
  V3 y = cross(axis, x);
  double dt1 = dot_seq(v, x);
  double dt2 = dot_seq(v, y);
  double ang = atan2_ool(dt2, dt1);
  if (ang < 0.0) ang = ang + TWO_PI;
  if (fabs(dt1) <= 1e-8 && fabs(dt2) <= 1e-8) ang = __builtin_nan("");
  *y_out = y;
  return ang;
}

// In-plane basis of the circle about axis.
__device__ __forceinline__ void circle_frame(V3 axis, V3* xo, V3* yo) {
  axis = unit(axis);
  V3 v1 = v3(axis.y * axis.y + axis.z * axis.z, axis.x * axis.x + axis.z * axis.z,
             axis.x * axis.x + axis.y * axis.y);
  v1 = unit(v1);
  double dt = dot_seq(v1, axis);
  if (fabs(dt) > 0.99) v1 = v3(1.0, 0.0, 0.0);
  V3 v2 = unit(cross(axis, v1));
  V3 x = unit(cross(axis, v2));
  *xo = x;
  *yo = cross(axis, x);
}

__device__ __forceinline__ double dist_to_line(V3 cen, V3 axis, V3 p) {
  V3 v = p - cen;
  double dt = dot_seq(v, axis);
  double d2 = dot_seq(v, v) - dt * dt;
  return d2 < 0.0 ? 0.0 : sqrt(d2);
}

// Unit loop of the one-wavefront-per-unit kernels (unit count read from
// device memory).  XCD-aware: workgroup w is dispatched to XCD w % 8, and
// XCD x walks the x-th contiguous eighth of the units, so neighbouring units
// (which share atoms and neighbour rows) meet in the same L2.  With fewer
// than 8 workgroups each one takes a contiguous share.  Every unit is
// visited exactly once.
#define CMS_UNIT_LOOP(u, n)                                                                      \
  for (int64_t u##_n = (int64_t)(n), u##_nx = gridDim.x < 8 ? gridDim.x : 8,                    \
               u##_x = blockIdx.x % u##_nx, u##_per = (u##_n + u##_nx - 1) / u##_nx,            \
               u##_hi = (u##_x + 1) * u##_per < u##_n ? (u##_x + 1) * u##_per : u##_n,          \
               u##_step = (gridDim.x - u##_x + u##_nx - 1) / u##_nx,                            \
               u = u##_x * u##_per + blockIdx.x / u##_nx;                                       \
       u < u##_hi; u += u##_step)

// Groups the active lanes of a wave by key: each lane gets the lowest lane
// with its key (leader), its rank among those lanes and their number.  One
// ballot round per distinct key.  Call with all lanes of the wave.
__device__ __forceinline__ void wave_group_by_key(int key, bool active, int* leader, int* rank, int* cnt) {
  *leader = lane();
  *rank = 0;
  *cnt = 0;
  bool todo = active;
  for (;;) {
    const uint64_t m = ballot(todo);
    if (!m) break;
    const int l = __builtin_ctzll(m);
    const bool mine = todo && key == __builtin_amdgcn_readlane(key, l);
    const uint64_t sm = ballot(mine);
    if (mine) {
      *leader = l;
      *rank = lanes_below(sm);
      *cnt = __popcll(sm);
      todo = false;
    }
  }
}

// Every active lane adds 1 to counters[key]; the lanes of one key are served
// by one atomic (issued together for all keys), in lane order.  Returns the
// counter's value before this lane's increment.  Call with all lanes.
__device__ __forceinline__ int wave_counter_add(int* counters, int key, bool active) {
  int leader, rank, cnt;
  wave_group_by_key(key, active, &leader, &rank, &cnt);
  int base = 0;
  if (active && lane() == leader) base = atomicAdd(&counters[key], cnt);
  base = __shfl(base, leader, WAVE);
  return base + rank;
}

// The same without the result.
__device__ __forceinline__ void wave_counter_inc(int* counters, int key, bool active) {
  int leader, rank, cnt;
  wave_group_by_key(key, active, &leader, &rank, &cnt);
  if (active && lane() == leader) atomicAdd(&counters[key], cnt);
}

// Grid-stride loop in which every lane of a wave takes part in every
// iteration (i may be >= n: callers test it), for wave-level primitives.
#define CMS_WAVE_LOOP(i, n)                                                                   \
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x,                           \
               i##_end = (((int64_t)(n) + WAVE - 1) / WAVE) * WAVE;                         \
       i - lane() < i##_end; i += (int64_t)gridDim.x * blockDim.x)

// Width of the unit blocks of a sparse loop: 64 when there are enough units
// to give every workgroup several blocks, down to 1 for small launches.
__device__ __forceinline__ int64_t sparse_width(int64_t n, int64_t grid) {
  const int64_t w = n / (4 * grid);
  return w < 1 ? 1 : (w > WAVE ? WAVE : w);
}

// Sparse unit loop: units in blocks of sparse_width(n) (blocks visited like
// CMS_UNIT_LOOP); the lanes evaluate pred on the block's units at once and
// the wave runs the body, uniformly, for each unit whose pred holds, in
// increasing order.  `continue` in the body moves to the next such unit.
#define CMS_SPARSE_LOOP(u, n, pred)                                                               \
  for (int64_t u##_nn = (int64_t)(n), u##_w = sparse_width(u##_nn, gridDim.x), u##_one = 1; u##_one; \
       u##_one = 0)                                                                               \
    CMS_UNIT_LOOP(u##_blk, (u##_nn + u##_w - 1) / u##_w)                                          \
      for (uint64_t u##_m = ballot(lane() < u##_w && u##_blk * u##_w + lane() < u##_nn &&         \
                                   (pred)(u##_blk * u##_w + lane()));                             \
           u##_m;)                                                                                \
        for (int64_t u = u##_blk * u##_w + __builtin_ctzll(u##_m), u##_o = (u##_m &= u##_m - 1, 1); \
             u##_o; u##_o = 0)

}  // namespace cms
