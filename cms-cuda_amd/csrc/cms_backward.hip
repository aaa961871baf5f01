// Backward pass of the contact molecular surface at frozen topology.
//
// CMS of side m = sum over buried dots q of molecule m of
//     A_q * exp(-w |P_q - P_r(q)|^2)
// with r(q) the nearest buried dot of the other molecule.  With topology
// frozen (which dots exist, their burial, partners, every subdivision count
// and branch), P and A are smooth functions of the coordinates of the one to
// three atoms that generate the dot: the convex dot of atom a (pole from its
// nearest neighbour n0), the saddle dot of pair (i, j), the concave dot of
// probe (i, j, k).  eval_dot recomputes P and A from a dot's descriptor and
// the atom coordinates, operation by operation as the forward kernels do
// (so with T = double at the forward coordinates it reproduces the forward
// dots bit for bit -- cms_check_dots verifies that).  With T = Dual<3> it
// carries the derivatives with respect to one generating atom, which gives
// the vector-Jacobian products of the backward pass; at perturbed
// coordinates it gives the frozen-topology CMS used for finite differences.
#include "cms_device.h"
#include "cms_internal.h"
#include "cms_backward.h"

namespace cms {

// ------------------------------------------------------------ dual numbers
template <int N>
struct Dual {
  double v;
  double d[N];
  __device__ Dual() {}
  __device__ Dual(double x) : v(x) {
    for (int i = 0; i < N; ++i) d[i] = 0.0;
  }
};

template <int N> __device__ __forceinline__ Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r; r.v = a.v + b.v;
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i];
  return r;
}
template <int N> __device__ __forceinline__ Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r; r.v = a.v - b.v;
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i];
  return r;
}
template <int N> __device__ __forceinline__ Dual<N> operator-(const Dual<N>& a) {
  Dual<N> r; r.v = -a.v;
  for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
  return r;
}
template <int N> __device__ __forceinline__ Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r; r.v = a.v * b.v;
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
template <int N> __device__ __forceinline__ Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r; r.v = a.v / b.v;
  for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) / b.v;
  return r;
}
template <int N> __device__ __forceinline__ Dual<N> operator+(const Dual<N>& a, double b) { return a + Dual<N>(b); }
template <int N> __device__ __forceinline__ Dual<N> operator+(double a, const Dual<N>& b) { return Dual<N>(a) + b; }
template <int N> __device__ __forceinline__ Dual<N> operator-(const Dual<N>& a, double b) { return a - Dual<N>(b); }
template <int N> __device__ __forceinline__ Dual<N> operator-(double a, const Dual<N>& b) { return Dual<N>(a) - b; }
template <int N> __device__ __forceinline__ Dual<N> operator*(const Dual<N>& a, double b) {
  Dual<N> r; r.v = a.v * b;
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b;
  return r;
}
template <int N> __device__ __forceinline__ Dual<N> operator*(double a, const Dual<N>& b) {
  Dual<N> r; r.v = a * b.v;
  for (int i = 0; i < N; ++i) r.d[i] = a * b.d[i];
  return r;
}
template <int N> __device__ __forceinline__ Dual<N> operator/(const Dual<N>& a, double b) {
  Dual<N> r; r.v = a.v / b;
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] / b;
  return r;
}
template <int N> __device__ __forceinline__ Dual<N> operator/(double a, const Dual<N>& b) { return Dual<N>(a) / b; }

__device__ __forceinline__ double val(double x) { return x; }
template <int N> __device__ __forceinline__ double val(const Dual<N>& x) { return x.v; }

__device__ __forceinline__ double t_sqrt(double x) { return sqrt(x); }
__device__ __forceinline__ double t_sin(double x) { return sin(x); }
__device__ __forceinline__ double t_cos(double x) { return cos(x); }
__device__ __forceinline__ double t_atan2(double y, double x) { return atan2(y, x); }
__device__ __forceinline__ double t_acos(double x) { return acos(x); }
template <int N> __device__ __forceinline__ Dual<N> t_sqrt(const Dual<N>& a) {
  Dual<N> r; r.v = sqrt(a.v);
  const double f = 0.5 / r.v;
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * f;
  return r;
}
template <int N> __device__ __forceinline__ Dual<N> t_sin(const Dual<N>& a) {
  Dual<N> r;
  double f;
  sincos(a.v, &r.v, &f);   // one argument reduction, the bits of sin() and cos()
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * f;
  return r;
}
template <int N> __device__ __forceinline__ Dual<N> t_cos(const Dual<N>& a) {
  Dual<N> r;
  double sn;
  sincos(a.v, &sn, &r.v);
  const double f = -sn;
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * f;
  return r;
}
template <int N> __device__ __forceinline__ Dual<N> t_atan2(const Dual<N>& y, const Dual<N>& x) {
  Dual<N> r; r.v = atan2(y.v, x.v);
  const double q = x.v * x.v + y.v * y.v;
  for (int i = 0; i < N; ++i) r.d[i] = (x.v * y.d[i] - y.v * x.d[i]) / q;
  return r;
}
template <int N> __device__ __forceinline__ Dual<N> t_acos(const Dual<N>& a) {
  Dual<N> r; r.v = acos(a.v);
  const double f = -1.0 / sqrt(1.0 - a.v * a.v);
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * f;
  return r;
}

// ------------------------------------------------------------ vectors over T
template <class T> struct V3T { T x, y, z; };
template <class T> __device__ __forceinline__ V3T<T> mk(T x, T y, T z) { V3T<T> r; r.x = x; r.y = y; r.z = z; return r; }
template <class T> __device__ __forceinline__ V3T<T> operator+(const V3T<T>& a, const V3T<T>& b) { return mk<T>(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class T> __device__ __forceinline__ V3T<T> operator-(const V3T<T>& a, const V3T<T>& b) { return mk<T>(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class T, class S> __device__ __forceinline__ V3T<T> operator*(const V3T<T>& a, const S& s) { return mk<T>(a.x * s, a.y * s, a.z * s); }
template <class T, class S> __device__ __forceinline__ V3T<T> operator/(const V3T<T>& a, const S& s) { return mk<T>(a.x / s, a.y / s, a.z / s); }
template <class T> __device__ __forceinline__ V3T<T> neg3(const V3T<T>& a) { return mk<T>(-a.x, -a.y, -a.z); }
template <class T> __device__ __forceinline__ T dseq(const V3T<T>& a, const V3T<T>& b) {
  T p0 = a.x * b.x, p1 = a.y * b.y, p2 = a.z * b.z;
  return (p0 + p1) + p2;
}
template <class T> __device__ __forceinline__ T dalt(const V3T<T>& a, const V3T<T>& b) {
  T p0 = a.x * b.x, p1 = a.y * b.y, p2 = a.z * b.z;
  return (p0 + p2) + p1;
}
template <class T> __device__ __forceinline__ V3T<T> crs(const V3T<T>& a, const V3T<T>& b) {
  return mk<T>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <class T> __device__ __forceinline__ T nseq(const V3T<T>& a) { return t_sqrt(dseq(a, a)); }
template <class T> __device__ __forceinline__ V3T<T> unt(const V3T<T>& a) { return a / nseq(a); }
template <class T> __device__ __forceinline__ V3T<T> cst(double x, double y, double z) { return mk<T>(T(x), T(y), T(z)); }

template <class T>
__device__ __forceinline__ T arc_angle_t(const V3T<T>& x, const V3T<T>& v, const V3T<T>& axis, V3T<T>* y_out) {
  V3T<T> y = crs(axis, x);
  T dt1 = dseq(v, x);
  T dt2 = dseq(v, y);
  T ang = t_atan2(dt2, dt1);
  if (val(ang) < 0.0) ang = ang + TWO_PI;
  if (fabs(val(dt1)) <= 1e-8 && fabs(val(dt2)) <= 1e-8) ang = T(__builtin_nan(""));
  *y_out = y;
  return ang;
}

template <class T>
__device__ __forceinline__ void circle_frame_t(V3T<T> axis, V3T<T>* xo, V3T<T>* yo) {
  axis = unt(axis);
  V3T<T> v1 = mk<T>(axis.y * axis.y + axis.z * axis.z, axis.x * axis.x + axis.z * axis.z,
                    axis.x * axis.x + axis.y * axis.y);
  v1 = unt(v1);
  T dt = dseq(v1, axis);
  if (fabs(val(dt)) > 0.99) v1 = cst<T>(1.0, 0.0, 0.0);
  V3T<T> v2 = unt(crs(axis, v1));
  V3T<T> x = unt(crs(axis, v2));
  *xo = x;
  *yo = crs(axis, x);
}

template <class T, class R, class D>
__device__ __forceinline__ V3T<T> subdiv_point_t(const V3T<T>& cen, const V3T<T>& x, const V3T<T>& y, const R& rad,
                                                 const D& delta, int i) {
  const D a = delta * ((double)i + 0.5);
  const T c = T(rad) * t_cos(T(a));
  const T s = T(rad) * t_sin(T(a));
  return (cen + x * c) + y * s;
}

template <class T>
__device__ __forceinline__ T dist_to_line_t(const V3T<T>& cen, const V3T<T>& axis, const V3T<T>& p) {
  V3T<T> v = p - cen;
  T dt = dseq(v, axis);
  T d2 = dseq(v, v) - dt * dt;
  return val(d2) < 0.0 ? T(0.0) : t_sqrt(d2);
}

// atom coordinates as T; the derivative seeds (identity) when seed
template <class T> struct Seed;
template <> struct Seed<double> {
  static __device__ __forceinline__ V3T<double> at(const double4& a, bool) { return mk<double>(a.x, a.y, a.z); }
};
template <int N> struct Seed<Dual<N>> {
  static __device__ __forceinline__ V3T<Dual<N>> at(const double4& a, bool seed) {
    V3T<Dual<N>> r;
    r.x = Dual<N>(a.x); r.y = Dual<N>(a.y); r.z = Dual<N>(a.z);
    if (seed) { r.x.d[0] = 1.0; r.y.d[1] = 1.0; r.z.d[2] = 1.0; }
    return r;
  }
};

// Torus geometry of pair (i, j), as pair_torus computes it.
template <class T>
__device__ __forceinline__ void pair_geom(const V3T<T>& xi, const V3T<T>& xj, double ri, double rj, double rp,
                                          V3T<T>* u, V3T<T>* t, T* rij, bool* between) {
  const double eri = ri + rp, erj = rj + rp;
  const V3T<T> diff = xj - xi;
  const T dsq = dalt(diff, diff);
  const T dij = t_sqrt(dsq);
  const T asymm = (eri * eri - erj * erj) / dij;
  *between = fabs(val(asymm)) < val(dij);
  *u = diff / dij;
  *t = (xi + xj) * 0.5 + (*u) * (asymm * 0.5);
  const T far_sq = (eri + erj) * (eri + erj) - dsq;
  const T cont_sq = dsq - (ri - rj) * (ri - rj);
  *rij = (0.5 * t_sqrt(far_sq)) * t_sqrt(cont_sq) / dij;
}

// Position and area of one dot from its descriptor.  coords: atom
// coordinates to evaluate at; base: the forward atoms (radii); slot: which
// generating atom (0..2, descriptor order a, b, c) carries derivative
// seeds, -1 none.
template <class T>
__device__ void eval_dot(const BatchView& b, const Settings& s, const double4* __restrict__ coords, int4 d1,
                         int4 d2, int slot, V3T<T>* P, T* A) {
  const int kind = d1.x & 15;
  const double rp = s.rp;
  if (kind == DK_CONVEX) {
    const int g = d1.y, n0 = d1.z;
    const int t = d2.x, k = d2.y, L = d2.z, Pc = d2.w;
    const double ri = b.atom[g].w, eri = ri + rp, dens = b.dens[g], r0 = b.atom[n0].w;
    const V3T<T> xi = Seed<T>::at(coords[g], slot == 0);
    const V3T<T> x0 = Seed<T>::at(coords[n0], slot == 1);
    const V3T<T> north = unt(xi - x0);
    V3T<T> vt = mk<T>(north.y * north.y + north.z * north.z, north.x * north.x + north.z * north.z,
                      north.x * north.x + north.y * north.y);
    vt = unt(vt);
    if (fabs(val(dseq(vt, north))) > 0.99) vt = cst<T>(1.0, 0.0, 0.0);
    const V3T<T> eq = unt(crs(north, vt));
    const V3T<T> vql = crs(eq, north);
    const double erj = r0 + rp;
    const T dij = nseq(x0 - xi);
    const V3T<T> uij = (x0 - xi) / dij;
    const T asymm = (eri * eri - erj * erj) / dij;
    const V3T<T> tij = ((xi + x0) * 0.5) + uij * (asymm * 0.5);
    const T far_sq = (eri + erj) * (eri + erj) - dij * dij;
    const T cont_sq = dij * dij - (ri - r0) * (ri - r0);
    V3T<T> south = cst<T>(0.0, 0.0, -1.0);
    if (val(far_sq) > 0.0 && val(cont_sq) > 0.0) {
      const T rij = 0.5 * t_sqrt(far_sq) * t_sqrt(cont_sq) / dij;
      const V3T<T> pij = tij + vql * rij;
      south = (pij - xi) / eri;
    }
    V3T<T> y;
    const T ang = arc_angle_t(north, south, eq, &y);
    const double delta = 1.0 / (sqrt(dens) * ri);
    const T cs = (ri * ang) / (double)L;
    V3T<T> cx, cy;
    circle_frame_t(north, &cx, &cy);
    const V3T<T> lat = subdiv_point_t(cst<T>(0.0, 0.0, 0.0), north, y, ri, delta, t);
    const T dt = dseq(lat, north);
    const V3T<T> cen = xi + north * dt;
    const T rsq = ri * ri - dt * dt;
    const T rad = t_sqrt(rsq);
    const T cdelta = 1.0 / (sqrt(dens) * rad);
    const T ps = (rad * TWO_PI) / (double)Pc;
    *P = subdiv_point_t(cen, cx, cy, rad, cdelta, k);
    *A = ps * cs;
    return;
  }
  if (kind == DK_ARC1 || kind == DK_ARC2) {
    const int i = d1.y, j = d1.z;
    const int m = d2.x, k = d2.y, M = d2.z, nArc = d2.w;
    const double ri = b.atom[i].w, rj = b.atom[j].w;
    const double eri = ri + rp, erj = rj + rp;
    const V3T<T> xi = Seed<T>::at(coords[i], slot == 0);
    const V3T<T> xj = Seed<T>::at(coords[j], slot == 1);
    V3T<T> uij, tij;
    T rij;
    bool between;
    pair_geom(xi, xj, ri, rj, rp, &uij, &tij, &rij, &between);
    const double dens = (b.dens[i] + b.dens[j]) * 0.5;
    const T rci = rij * ri / eri;
    const T rcj = rij * rj / erj;
    const T rb = val(rij - rp) > 0.0 ? rij - rp : T(0.0);
    const T rs = (rci + 2 * rb + rcj) * 0.25;
    const T e = rs / rij;
    const T edens = e * e * dens;
    V3T<T> cx, cy;
    circle_frame_t(uij, &cx, &cy);
    const T tdelta = 1.0 / (t_sqrt(edens) * rij);
    const T ts = (rij * TWO_PI) / (double)M;
    const V3T<T> pm = subdiv_point_t(tij, cx, cy, rij, tdelta, m);
    const V3T<T> pi = (xi - pm) / eri;
    const V3T<T> pj = (xj - pm) / erj;
    const V3T<T> axis = unt(crs(pi, pj));
    const T dtq = rp * rp - rij * rij;
    V3T<T> pqi, pqj;
    if (val(dtq) > 0.0 && between) {
      const T dq = t_sqrt(dtq);
      const V3T<T> qij = tij - uij * dq;
      pqi = (qij - pm) / rp;
      pqj = cst<T>(0.0, 0.0, 0.0);
    } else {
      const V3T<T> tmp = unt(pi + pj);
      pqi = tmp;
      pqj = tmp;
    }
    const double adelta = 1.0 / (sqrt(dens) * rp);
    V3T<T> y, pt;
    T ang;
    if (kind == DK_ARC1) {
      ang = arc_angle_t(pi, pqi, axis, &y);
      pt = subdiv_point_t(pm, pi, y, rp, adelta, k);
    } else {
      ang = arc_angle_t(pqj, pj, axis, &y);
      pt = subdiv_point_t(pm, pqj, y, rp, adelta, k);
    }
    const T ps = (rp * ang) / (double)nArc;
    const T dist = dist_to_line_t(tij, uij, pt);
    *P = pt;
    *A = ps * ts * dist / rij;
    return;
  }
  // concave
  {
    const int mm = (d1.x >> 4) & 3;
    const int a0 = d1.y, a1 = d1.z, a2 = d1.w;
    const int t = d2.x, k = d2.y, L = d2.z, Pc = d2.w;
    const bool plus = a0 < a1;
    const V3T<T> x0 = Seed<T>::at(coords[a0], slot == 0);
    const V3T<T> x1 = Seed<T>::at(coords[a1], slot == 1);
    const V3T<T> x2 = Seed<T>::at(coords[a2], slot == 2);
    const V3T<T> xi = plus ? x0 : x1;
    const V3T<T> xj = plus ? x1 : x0;
    const V3T<T>& xk = x2;
    const double ri = b.atom[plus ? a0 : a1].w, rj = b.atom[plus ? a1 : a0].w, rk = b.atom[a2].w;
    const double eri = ri + rp, erk = rk + rp;
    V3T<T> uij, tij;
    T rij;
    bool between;
    pair_geom(xi, xj, ri, rj, rp, &uij, &tij, &rij, &between);
    const T dik = nseq(xk - xi);
    const V3T<T> uik = (xk - xi) / dik;
    const T dt = dalt(uij, uik);
    const T dtc = val(dt) < -1.0 ? T(-1.0) : (val(dt) > 1.0 ? T(1.0) : dt);
    const T wk = t_acos(dtc);
    const T sw = t_sin(wk);
    const V3T<T> uijk = crs(uij, uik) / sw;
    const V3T<T> utb = crs(uijk, uij);
    const T asymm = (eri * eri - erk * erk) / dik;
    const V3T<T> tik = (xi + xk) * 0.5 + uik * (asymm * 0.5);
    const T dtb = dalt(uik, tik - tij);
    const V3T<T> bijk = tij + utb * (dtb / sw);
    const V3T<T> bi_ = bijk - xi;
    const T hsq = eri * eri - dalt(bi_, bi_);
    const T h = t_sqrt(hsq);
    const V3T<T> p = plus ? bijk + uijk * h : bijk + uijk * (-h);
    const V3T<T> alt = plus ? uijk : neg3(uijk);
    const double dens = ((b.dens[a0] + b.dens[a1]) + b.dens[a2]) / 3.0;
    const V3T<T> xa = mm == 0 ? x0 : (mm == 1 ? x1 : x2);
    const V3T<T> top = unt(xa - p);
    const V3T<T> south = neg3(alt);
    const V3T<T> axis = unt(crs(top, south));
    V3T<T> y;
    const T ang = arc_angle_t(top, south, axis, &y);
    const double delta = 1.0 / (sqrt(dens) * rp);
    const T cs = (rp * ang) / (double)L;
    V3T<T> cx, cy;
    circle_frame_t(south, &cx, &cy);
    const V3T<T> lat = subdiv_point_t(cst<T>(0.0, 0.0, 0.0), top, y, rp, delta, t);
    const T dtl = dseq(lat, south);
    const V3T<T> cen = south * dtl;
    const T rsq = rp * rp - dtl * dtl;
    const T rad = val(rsq) > 0.0 ? t_sqrt(rsq) : T(0.0);
    const T cdelta = 1.0 / (sqrt(dens) * rad);
    const T ps = (rad * TWO_PI) / (double)Pc;
    const V3T<T> rel = subdiv_point_t(cen, cx, cy, rad, cdelta, k);
    *P = rel + p;
    *A = ps * cs;
  }
}

// ------------------------------------------------------------ kernels

// Forward dots re-evaluated from their descriptors must equal the stored
// ones bit for bit; counts mismatches and the largest deviation.
__global__ __launch_bounds__(64) void k_check_dots(BatchView b, Settings s, Count cnD, const double4* __restrict__ dpos,
                                                   const int4* __restrict__ dd1, const int4* __restrict__ dd2,
                                                   unsigned long long* __restrict__ nbad, double* __restrict__ maxdev) {
  const int nD = cnD.get();
  double dev = 0.0;
  unsigned long long bad = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nD; i += gridDim.x * blockDim.x) {
    V3T<double> P;
    double A;
    eval_dot<double>(b, s, b.atom, dd1[i], dd2[i], -1, &P, &A);
    const double4 f = dpos[i];
    const bool same = P.x == f.x && P.y == f.y && P.z == f.z && A == f.w;
    if (!same) {
      ++bad;
      dev = fmax(dev, fmax(fmax(fabs(P.x - f.x), fabs(P.y - f.y)), fmax(fabs(P.z - f.z), fabs(A - f.w))));
    }
  }
  if (bad) {
    atomicAdd(nbad, bad);
    atomicMax((unsigned long long*)maxdev, (unsigned long long)__double_as_longlong(dev));
  }
}

// Adjoint of every buried query dot and its partner:
//   W = g_total[side] + g_atom[owner];  e = exp(-w d^2)
//   dL/dA_q = W e,  dL/dP_q = -2 w W A_q e (P_q - P_r),  dL/dP_r -= the same.
__global__ void k_cms_adjoint(Count cnB, const int* __restrict__ blist, const int* __restrict__ bgrp_of,
                              const int* __restrict__ bgrp, const int* __restrict__ bidx,
                              const double* __restrict__ bmin, const double4* __restrict__ dpos,
                              const int* __restrict__ datom, const double* __restrict__ g_total,
                              const double* __restrict__ g_atom, double weight, double* __restrict__ adj) {
  const int nB = cnB.get();
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nB; k += gridDim.x * blockDim.x) {
    const int ri = bidx[k];
    if (ri < 0) continue;  // no buried dot on the other side: contribution 0
    const int g = bgrp_of[k];
    const int dq = blist[k];
    const int dr = blist[bgrp[g ^ 1] + ri];
    double W = g_total ? g_total[g] : 0.0;
    if (g_atom) W += g_atom[datom[dq]];
    if (W == 0.0) continue;
    const double4 q = dpos[dq], r = dpos[dr];
    const double e = exp(-bmin[k] * weight);
    const double f = -2.0 * weight * W * q.w * e;
    const double gx = f * (q.x - r.x), gy = f * (q.y - r.y), gz = f * (q.z - r.z);
    atomicAdd(&adj[4 * (int64_t)dq + 0], gx);
    atomicAdd(&adj[4 * (int64_t)dq + 1], gy);
    atomicAdd(&adj[4 * (int64_t)dq + 2], gz);
    atomicAdd(&adj[4 * (int64_t)dq + 3], W * e);
    atomicAdd(&adj[4 * (int64_t)dr + 0], -gx);
    atomicAdd(&adj[4 * (int64_t)dr + 1], -gy);
    atomicAdd(&adj[4 * (int64_t)dr + 2], -gz);
  }
}

// Vector-Jacobian product per dot: one forward-mode pass per generating atom.
__global__ __launch_bounds__(64) void k_dot_grad(BatchView b, Settings s, Count cnD, const int4* __restrict__ dd1,
                                                 const int4* __restrict__ dd2, const double* __restrict__ adj,
                                                 double* __restrict__ grad) {
  const int nD = cnD.get();
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nD; i += gridDim.x * blockDim.x) {
    const double ax = adj[4 * (int64_t)i], ay = adj[4 * (int64_t)i + 1], az = adj[4 * (int64_t)i + 2],
                 aa = adj[4 * (int64_t)i + 3];
    if (ax == 0.0 && ay == 0.0 && az == 0.0 && aa == 0.0) continue;
    const int4 d1 = dd1[i], d2 = dd2[i];
    const int kind = d1.x & 15;
    const int nslot = kind == DK_CONCAVE ? 3 : 2;
    const int atoms[3] = {d1.y, d1.z, d1.w};
    for (int sl = 0; sl < nslot; ++sl) {
      V3T<Dual<3>> P;
      Dual<3> A;
      eval_dot<Dual<3>>(b, s, b.atom, d1, d2, sl, &P, &A);
      const int a = atoms[sl];
      for (int c = 0; c < 3; ++c) {
        const double gc = ax * P.x.d[c] + ay * P.y.d[c] + az * P.z.d[c] + aa * A.d[c];
        if (gc != 0.0) atomicAdd(&grad[3 * (int64_t)a + c], gc);
      }
    }
  }
}

// Frozen-topology CMS at other coordinates: positions and areas of every
// buried dot re-evaluated, partners kept.
__global__ __launch_bounds__(64) void k_frozen_eval(BatchView b, Settings s, const double4* __restrict__ coords, Count cnB,
                                                    const int* __restrict__ blist, const int4* __restrict__ dd1,
                                                    const int4* __restrict__ dd2, double4* __restrict__ out) {
  const int nB = cnB.get();
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nB; k += gridDim.x * blockDim.x) {
    const int d = blist[k];
    V3T<double> P;
    double A;
    eval_dot<double>(b, s, coords, dd1[d], dd2[d], -1, &P, &A);
    out[k] = make_double4(P.x, P.y, P.z, A);
  }
}

__global__ void k_frozen_cms(Count cnB, const int* __restrict__ bgrp_of, const int* __restrict__ bgrp,
                             const int* __restrict__ bidx, const double4* __restrict__ ev, double weight,
                             double* __restrict__ contrib) {
  const int nB = cnB.get();
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nB; k += gridDim.x * blockDim.x) {
    const int ri = bidx[k];
    double v = 0.0;
    if (ri >= 0) {
      const int g = bgrp_of[k];
      const double4 q = ev[k], r = ev[bgrp[g ^ 1] + ri];
      const double dx = r.x - q.x, dy = r.y - q.y, dz = r.z - q.z;
      const double d2 = (dx * dx + dy * dy) + dz * dz;
      v = q.w * exp(-d2 * weight);
    }
    contrib[k] = v;
  }
}

}  // namespace cms
