// Surface generation kernels: attention, interface region, neighbour rows,
// torus pairs, probes, and convex / toroidal / concave dots with burial.
//
// Every kernel works on a batch of complexes laid out atom-contiguously
// (complex-major, binder atoms before target atoms) and loops grid-stride
// over its units with the unit count read from device memory, so the whole
// pipeline launches without host synchronisation.  Units whose output size
// is data dependent run twice: a COUNT pass and, after a device-wide scan,
// an EMIT pass that writes at the scanned offset.  Units are visited in the
// reference's emission order, so dot arrays come out in the same order as
// the NumPy calculator (cms-cuda/_core.py) produces them.
//
// One wavefront (64 lanes, one 64-thread workgroup) handles one unit; lanes
// split the unit's inner work (neighbour slots, circle points, latitudes).
#include "cms_device.h"
#include "cms_internal.h"

namespace cms {

__global__ void k_compact_index(int n, const int* __restrict__ flag, const int* __restrict__ scan,
                                int* __restrict__ out) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    if (flag[i]) out[scan[i]] = i;
}

// ------------------------------------------------------------ pairs
// _core.py:1351-1442.  Pair records of small batches: the forward pairs
// (j > i) of each row that the neighbour pass counted (valid torus,
// relevant), at the scanned offsets, with torus centre/radius and flags.
// Rows that take the single-neighbour shortcut have a count of 0 and are
// not visited.
__global__ __launch_bounds__(64) void k_pairs(BatchView b, Settings s, Work w) {
  const int nR = *w.nR;
  auto live = [&](int64_t q) { return w.pcnt[q] != 0; };
  CMS_SPARSE_LOOP(r, nR, live) {
    const int g = w.ratom[r];
    const int nn = w.nn[r];
    const double4 ai = b.atom[g];
    const V3 xi = ld3(ai);
    const double ri = ai.w;
    const int* row = w.nbr + (int64_t)r * KCAP;
    const int nbmax = w.nbmax[r];
    const int base = w.pscan[r];
    int total = 0;
    for (int s0 = 0; s0 < nn; s0 += WAVE) {
      const int sl = s0 + lane();
      bool relevant = false;
      int j = -1;
      V3 u{0, 0, 0}, t{0, 0, 0};
      double rij = 0.0;
      bool between = false;
      if (sl < nn) {
        j = row[sl];
        if (j > g) {
          const double4 aj = b.atom[j];
          relevant = pair_torus(xi, ri, ld3(aj), aj.w, s.rp, &u, &t, &rij, &between) &&
                     pair_relevant(w.atten[g], w.atten[j], nbmax, j);
        }
      }
      const uint64_t mk = ballot(relevant);
      if (relevant && base + total + lanes_below(mk) < w.pair_cap) {
        const int o = base + total + lanes_below(mk);
        w.pi[o] = g;
        w.pj[o] = j;
        w.pu[o] = make_double4(u.x, u.y, u.z, rij);
        w.pt[o] = make_double4(t.x, t.y, t.z, 0.0);
        const bool tor = pair_toroid(w.atten[g], w.atten[j], s.rp);
        w.pflag[o] = (int8_t)((between ? 1 : 0) | (tor ? 2 : 0));
      }
      total += __popcll(mk);
    }
  }
}

// ------------------------------------------------------------ probes
// _core.py:1445-1643.  For pair (i, j) and neighbour k of i with k > j:
// triangle tests, pair kill by degenerate triples, probe base/height, the
// two probe positions, and their collision with i's other neighbours.
// Probe order per molecule group: all "+" probes, then all "-" probes, each
// in (pair, k) order.
__device__ __forceinline__ int group_of_atom(const BatchView& b, int g) {
  return 2 * b.acplx[g] + b.amol[g];
}

// Uniform copy of lane l's value (v_readlane on both halves).
__device__ __forceinline__ double readlane_d(double v, int l) {
  const unsigned long long bits = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(bits & 0xffffffffull), l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(bits >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Triangle tests of slot k of atom i's row for the pair (i, j)
// (_core.py:1478-1500): k > j, d(j, k) < erj + erk, d(i, k) < eri + erk, not
// all three atoms blockers.
__device__ __forceinline__ bool probe_triple_valid(int k, int j, V3 xk, double erk, V3 xi, V3 xj, double eri,
                                                   double erj, bool bij, const int8_t* atten, double* dik) {
  if (k <= j) return false;
  if (!norm_lt(xk - xj, erj + erk)) return false;
  if (!norm_lt(xk - xi, eri + erk)) return false;
  *dik = norm_seq(xk - xi);
  return !(bij && atten[k] <= ATTEN_BLOCKER);
}

// Torus-circle containment.  The circle of the pair (i, j) -- centre t,
// axis u, radius rho -- holds every probe of the pair (a probe touches i
// and j) and every torus point.  If one of i's row neighbours s other than
// j holds the whole circle inside its sphere shrunk by COVER_MARGIN -- the
// farthest circle point from x_s, at distance sqrt((d + rho)^2 + a^2) with
// a and d the axial and radial offsets of x_s, lies within
// r_s + rp - COVER_MARGIN -- then every torus point collides with s, and so
// does every probe: s is neither j nor the probe's own third atom k (a probe
// of k sits on k's sphere, which therefore cannot hold the circle).  The
// pair then has neither probes nor torus points.  Computed torus points lie
// within 1e-12 A of the circle; a computed probe lies within 1e-4 A of it
// when sin w >= 1.4e-5 (arccos amplifies the rounding of the cosine by
// 1 / sin w, and the base point moves by at most twice the probe radius
// times that relative error), so the caller excludes pairs with a candidate
// closer to degenerate; the containment test's own rounding is below
// 1e-7 A.  All far inside the margin.
constexpr double COVER_MARGIN = 0.01;

__device__ __forceinline__ bool circle_in_one_sphere(const int* lj, const double4* lx, int nn, int j, V3 t, V3 u,
                                                     double rho, double rp) {
  bool hit = false;
  for (int sl = lane(); sl < nn; sl += WAVE) {
    if (lj[sl] == j) continue;
    const V3 v = ld3(lx[sl]) - t;
    const double a = dot_seq(v, u);
    const double d = sqrt(fmax(dot_seq(v, v) - a * a, 0.0));
    const double R = (lx[sl].w + rp) - COVER_MARGIN;
    const double far2 = (d + rho) * (d + rho) + a * a;
    hit = hit || (R > 0.0 && far2 < R * R);
  }
  return wave_any(hit);
}

// w = arccos(clip(dt)) and sin(w) (_core.py:1505-1512, 1545-1550), out of
// line: the polynomial constants of both are then materialised inside the
// call instead of being held in registers across the callers' loops.
__device__ __attribute__((noinline)) double2 acos_sin(double dt) {
  const double wk = acos(fmin(fmax(dt, -1.0), 1.0));
  return make_double2(wk, sin(wk));
}

// Probe positions of the non-degenerate triple (i, j, k) (_core.py:1545-1575):
// base point on the axis of the pair's torus circle, height h above the
// plane of the three centres, probes bijk +- h uijk.  False when h^2 <= 0.
// Both probe passes evaluate the triple through this function, so their
// collision outcomes agree bit for bit.
__device__ __forceinline__ bool probe_points(V3 xi, double eri, V3 xk, double erk, V3 uij, V3 tij, V3* pp, V3* pm,
                                             V3* uijk_out, double* h_out) {
  const double dik = norm_seq(xk - xi);
  const V3 uik = (xk - xi) / dik;
  const double dt = dot_alt(uij, uik);
  const double sw = acos_sin(dt).y;
  const V3 uijk = cross(uij, uik) / sw;
  const V3 utb = cross(uijk, uij);
  const double asymm = (eri * eri - erk * erk) / dik;
  const V3 tik = (xi + xk) * 0.5 + uik * (asymm * 0.5);
  const double dtb = dot_alt(uik, tik - tij);
  const V3 bijk = tij + utb * (dtb / sw);
  const V3 bi_ = bijk - xi;
  const double hsq = eri * eri - dot_alt(bi_, bi_);
  *uijk_out = uijk;
  if (!(hsq > 0.0)) return false;
  const double h = sqrt(hsq);
  *pp = bijk + uijk * h;
  *pm = bijk + uijk * (-h);
  *h_out = h;
  return true;
}

// dt = uij . (xk - xi) / |xk - xi| is decided against 1 - 1e-6 without the
// square root and division: with c = uij . v and vv = |v|^2 (v = xk - xi),
// c^2 < (1 - 1.000001e-6)^2 vv puts |c| / |v| below 1 - 1e-6 - 1e-12, and the
// computed dt differs from c / |v| by a few ulps, so it stays below 1 - 1e-6
// (not degenerate, not near it); only the few other triples compute dt.
constexpr double DT_CLEAR_SQ = (1.0 - 1.000001e-6) * (1.0 - 1.000001e-6);

// Degenerate test of a valid triple (_core.py:1505-1543): dt >= 1, dt <= -1,
// w <= 0 or sin w <= 0 with w = arccos(clip(dt)).  For |dt| < 1 - 1e-6, w
// lies in (1.4e-3, pi - 1.4e-3), so none of these holds and arccos / sin are
// evaluated only near |dt| = 1.
__device__ __forceinline__ bool triple_degenerate(double dt) {
  if (fabs(dt) < 1.0 - 1e-6) return false;
  const double2 ws = acos_sin(dt);
  const double wk = ws.x, sw = ws.y;
  return (dt >= 1.0) || (dt <= -1.0) || (wk <= 0.0) || (sw <= 0.0);
}

template <bool EMIT>
__global__ __launch_bounds__(64, 4) void k_probes(BatchView b, Settings s, Work w) {
  __shared__ int lj[KCAP];
  __shared__ double4 lx[KCAP];
  constexpr int NCH = KCAP / WAVE;
  const int nP = *w.nP;
  // the emit pass visits only pairs with surviving probes (flag bit 8); after
  // the per-row count pass those pairs are listed (elist, a few)
  const bool listed = EMIT && w.elist != nullptr;
  auto live = [&](int64_t q) { return listed || !EMIT || (w.pflag[q] & 8) != 0; };
  int staged = -1;   // neighbour row held in lj / lx (consecutive pairs share it)
  CMS_SPARSE_LOOP(u, listed ? *w.nelist : nP, live) {
    const int64_t p = listed ? w.elist[u] : u;
    const int i = w.pi[p], j = w.pj[p];
    const int grp = group_of_atom(b, i);
    const int g0 = w.pgrp[grp], g1 = w.pgrp[grp + 1];
    const int64_t vp = 2ll * g0 + (p - g0);
    const int64_t vm = 2ll * g0 + (g1 - g0) + (p - g0);
    const int r = w.rpos[i];
    const int nn = w.nn[r];
    if (r != staged) {
      const int* row = w.nbr + (int64_t)r * KCAP;
      for (int sl = lane(); sl < nn; sl += WAVE) {
        const int k = row[sl];
        lj[sl] = k;
        lx[sl] = b.atom[k];
      }
      __syncthreads();
      staged = r;
    }
    const double4 ai = b.atom[i], aj = b.atom[j];
    const V3 xi = ld3(ai), xj = ld3(aj);
    const double eri = ai.w + s.rp, erj = aj.w + s.rp;
    const double4 pu = w.pu[p], ptt = w.pt[p];
    const V3 uij = ld3(pu), tij = ld3(ptt);
    const double rij = pu.w;
    const bool bij = w.atten[i] <= ATTEN_BLOCKER && w.atten[j] <= ATTEN_BLOCKER;
    const int nch = (nn + WAVE - 1) / WAVE;

    // pass 1: valid, non-degenerate triples of every 64-slot chunk, and
    // whether a degenerate triple kills the pair (_core.py:1505-1543).
    uint64_t vmask[NCH];
    bool killed = false, near_deg = false;
    int ncand = 0;
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
      vmask[c] = 0;
      if (c < nch) {
        const int sl = c * WAVE + lane();
        bool cand = false, kill = false, nd = false;
        if (sl < nn) {
          const V3 xk = ld3(lx[sl]);
          const double erk = lx[sl].w + s.rp;
          double dik = 0.0;
          if (probe_triple_valid(lj[sl], j, xk, erk, xi, xj, eri, erj, bij, w.atten, &dik)) {
            const V3 uik = (xk - xi) / dik;
            const double dt = dot_alt(uij, uik);
            const bool deg = triple_degenerate(dt);
            if (deg) {
              const V3 dk = tij - xk;
              kill = dot_seq(dk, dk) < erk * erk - rij * rij;
            }
            cand = !deg;
            nd = cand && !(fabs(dt) < 1.0 - 1e-10);   // sin w below ~1.4e-5
          }
        }
        vmask[c] = ballot(cand);
        ncand += __popcll(vmask[c]);
        killed = killed || wave_any(kill);
        near_deg = near_deg || wave_any(nd);
      }
    }
    // pairs whose torus circle lies inside one neighbour's sphere have no
    // probes and no torus points (flag bit 4, read by the toroidal stage)
    if (!EMIT) {
      const bool tor = (w.pflag[p] & 2) != 0;
      if (!near_deg && (tor || (ncand > 0 && !killed)) && circle_in_one_sphere(lj, lx, nn, j, tij, uij, rij, s.rp)) {
        if (lane() == 0) {
          w.pflag[p] = (int8_t)(w.pflag[p] | 4);
          w.prcnt[vp] = 0;
          w.prcnt[vm] = 0;
        }
        __syncthreads();
        continue;
      }
    }

    // pass 2: probe positions of the candidates, then their collision with
    // i's other neighbours (_core.py:1545-1623): one candidate at a time,
    // the lanes spread over the neighbour row.
    int np_plus = 0, np_minus = 0;
    int ob_plus = 0, ob_minus = 0;
    if (EMIT) {
      ob_plus = w.prscan[vp];
      ob_minus = w.prscan[vm];
    }
    if (!killed) {
#pragma unroll
      for (int c = 0; c < NCH; ++c) {
        if (c >= nch || vmask[c] == 0) continue;
        const int sl = c * WAVE + lane();
        const bool mine = (vmask[c] >> lane()) & 1ull;
        bool okp = false, okm = false;
        int k = -1;
        V3 pp{0, 0, 0}, pm{0, 0, 0}, uijk{0, 0, 0};
        double h = 0.0;
        if (mine) {
          k = lj[sl];
          okp = okm = probe_points(xi, eri, ld3(lx[sl]), lx[sl].w + s.rp, uij, tij, &pp, &pm, &uijk, &h);
        }
        uint64_t cm = ballot(okp);
        while (cm) {
          const int cl = __builtin_ctzll(cm);
          cm &= cm - 1;
          const V3 P = v3(readlane_d(pp.x, cl), readlane_d(pp.y, cl), readlane_d(pp.z, cl));
          const V3 M = v3(readlane_d(pm.x, cl), readlane_d(pm.y, cl), readlane_d(pm.z, cl));
          const int kk = __builtin_amdgcn_readlane(k, cl);
          bool hp = false, hm = false;
          for (int t = lane(); t < nn; t += WAVE) {
            const int n = lj[t];
            if (n == j || n == kk) continue;
            const double rr = lx[t].w + s.rp;
            const double lim = rr * rr;
            const V3 xs = ld3(lx[t]);
            const V3 dp = P - xs, dm = M - xs;
            hp = hp || dot_alt(dp, dp) <= lim;
            hm = hm || dot_alt(dm, dm) <= lim;
          }
          const bool anyp = wave_any(hp), anym = wave_any(hm);
          if (lane() == cl) {
            okp = !anyp;
            okm = !anym;
          }
        }
        const uint64_t mp = ballot(okp), mm = ballot(okm);
        if (!EMIT) {
          if (okp || okm) { w.access[i] = 1; w.access[j] = 1; w.access[k] = 1; }
        } else {
          if (okp && ob_plus + np_plus + lanes_below(mp) < w.probe_cap) {
            const int o = ob_plus + np_plus + lanes_below(mp);
            w.patom[o] = make_int4(i, j, k, 0);
            w.ppos[o] = make_double4(pp.x, pp.y, pp.z, h);
            w.palt[o] = make_double4(uijk.x, uijk.y, uijk.z, 0.0);
          }
          if (okm && ob_minus + np_minus + lanes_below(mm) < w.probe_cap) {
            const int o = ob_minus + np_minus + lanes_below(mm);
            w.patom[o] = make_int4(j, i, k, 0);
            w.ppos[o] = make_double4(pm.x, pm.y, pm.z, h);
            w.palt[o] = make_double4(-uijk.x, -uijk.y, -uijk.z, 0.0);
          }
        }
        np_plus += __popcll(mp);
        np_minus += __popcll(mm);
      }
    }
    if (!EMIT && lane() == 0) {
      w.prcnt[vp] = np_plus;
      w.prcnt[vm] = np_minus;
      if (np_plus + np_minus > 0) w.pflag[p] = (int8_t)(w.pflag[p] | 8);
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------ torus geometry
// Shared by the toroidal passes and the per-row count pass.
struct TorusSub {
  V3 p, pi, pj, axis, pqi, pqj;
  double ang1, ang2;
  V3 y1, y2;
  int n1, n2;
};

// Torus points of the pair (i, j) (_core.py:1863-1885): the circle frame
// (cx, cy) about uij, the spacing tdelta of the M points, their arc length
// ts, and the angular spacing adelta of the saddle arcs.
struct TorusFrame {
  V3 cx, cy;
  double tdelta, ts, adelta;
  int M;
};
__device__ __forceinline__ TorusFrame torus_frame(double ri, double rj, double rp, double rij, V3 uij,
                                                  double dens) {
  const double eri = ri + rp, erj = rj + rp;
  const double rci = rij * ri / eri;
  const double rcj = rij * rj / erj;
  const double rb = fmax(rij - rp, 0.0);
  const double rs = (rci + 2 * rb + rcj) * 0.25;
  const double e = rs / rij;
  const double edens = e * e * dens;
  TorusFrame f;
  circle_frame(uij, &f.cx, &f.cy);
  f.tdelta = 1.0 / (sqrt(edens) * rij);
  f.M = subdiv_count(f.tdelta, TWO_PI);
  f.ts = f.M > 0 ? (rij * TWO_PI) / (double)f.M : 0.0;
  f.adelta = 1.0 / (sqrt(dens) * rp);
  return f;
}

// Saddle arcs of the free torus point pm (_core.py:1947-2044): the probe
// directions towards i and j, the arc axis, the arc ends (the cusp point
// below the circle when the probe cuts the axis between the atoms), and the
// arc dot counts of the arcs that are on.
__device__ __forceinline__ void torus_point_arcs(V3 pm, V3 xi, V3 xj, double eri, double erj, V3 tij, V3 uij,
                                                 double rij, double rp, bool between, bool arc1_on, bool arc2_on,
                                                 double adelta, TorusSub* t) {
  const double rp2 = rp * rp;
  t->p = pm;
  t->pi = (xi - pm) / eri;
  t->pj = (xj - pm) / erj;
  t->axis = unit(cross(t->pi, t->pj));
  const double dtq = rp2 - rij * rij;
  if (dtq > 0.0 && between) {
    const double dq = sqrt(dtq);
    const V3 qij = tij - uij * dq;
    t->pqi = (qij - pm) / rp;
    t->pqj = v3(0.0, 0.0, 0.0);
  } else {
    const V3 tmp = unit(t->pi + t->pj);
    t->pqi = tmp;
    t->pqj = tmp;
  }
  const double dt1 = dot_seq(t->pqi, t->pi);
  const double dt2 = dot_seq(t->pqj, t->pj);
  if (fabs(dt1) < 1.0 && fabs(dt2) < 1.0) {
    if (arc1_on) {
      t->ang1 = arc_angle(t->pi, t->pqi, t->axis, &t->y1);
      t->n1 = subdiv_count(adelta, t->ang1);
    }
    if (arc2_on) {
      t->ang2 = arc_angle(t->pqj, t->pj, t->axis, &t->y2);
      t->n2 = subdiv_count(adelta, t->ang2);
    }
  }
}

// Count pass of the probes stage, one wavefront per neighbour row: the
// row's forward pairs (i, j) are taken in batches, and each step spreads its
// work over the lanes across pair boundaries, so that lanes are not left
// idle by the few valid triples of a single pair:
//   1. triangle tests of the triples (i, j, k), k > j: the row's forward
//      atoms are ranked by atom index once, so each pair visits exactly the
//      atoms after j; valid triples are queued;
//   2. degenerate / kill tests of the queued triples, compacting the
//      non-degenerate ones (candidates) in place;
//   3. per (pair, row atom), for the pairs of the toroid queue: the
//      single-sphere containment test, and the row atoms that can reach the
//      pair's torus circle (the only ones a probe or torus point of the pair
//      can collide with); the other pairs' probes take every row atom (their
//      masks cost more than they save);
//   4. probe positions, one candidate per lane, and their collisions with the
//      reaching row atoms other than j and k;
//   5. the torus points of the toroid-queue pairs whose circle is not inside
//      a neighbour's sphere, one per lane, their collisions with the reaching
//      row atoms, and the saddle-arc dot counts of the free ones.
// Per pair it writes what the pair-per-wavefront probe and toroidal emit
// passes rely on: the probe counts of both signs, pflag bits 4 (circle
// inside one neighbour sphere) and 8 (probes), the access marks of the atoms
// of every surviving probe (_core.py:1445-1643), the arc dot counts of both
// atoms, whether any torus point is free and the free points among the
// first 64 (_core.py:1847-2044).
constexpr int PR_PAIRS = 32;            // pairs per batch (one per lane in the per-pair steps)
// Row capacity of the LDS staging (RC_SMALL, cms_internal.h): rows of at
// most RC_SMALL neighbours run in the RC_SMALL instantiation of the per-row
// kernels (half the LDS, so more workgroups per CU), longer rows (up to
// KCAP) in the KCAP one; each launch visits only its own rows.
template <int RC>
__device__ __forceinline__ bool row_in_class(int nn) {
  return RC < KCAP ? nn <= RC : nn > RC_SMALL;
}
// Reach margin: a computed probe of a non-near-degenerate triple lies within
// 1e-4 A of its pair's torus circle (see circle_in_one_sphere), and the
// squared-distance forms below round far below 1e-10 A.
constexpr double REACH_MARGIN = 0.01;

// pair-batch flag bits
// latt: attention value of the row atom (low bits) and whether it passes
// the triangle test's d(i, k) < eri + erk (forward atoms only)
constexpr int LATT_ATT = 0x3f, LATT_IK = 0x40;
constexpr int PB_KILLED = 1, PB_NEARDEG = 2, PB_CAND = 4, PB_TOR = 8, PB_BLOCKERS = 16, PB_COVERED = 32,
              PB_BETWEEN = 64, PB_ARC2 = 128, PB_EMIT = 256;

// Step 5 of k_probe_rows, in the LDS room of the triple queue: the circle
// frames, point prefix, free-point masks and arc counts of up to TL_PAIRS
// live toroid-queue pairs of the batch, their lanes in the batch, and the
// free points queued for their saddle arcs ((slot << 8) | point).
constexpr int TL_PAIRS = 8;
struct TorusLds {
  double cx[TL_PAIRS][3], cy[TL_PAIRS][3];
  double td[TL_PAIRS], ad[TL_PAIRS];
  unsigned long long mk0[TL_PAIRS], mk1[TL_PAIRS];
  int pre[TL_PAIRS], n1[TL_PAIRS], n2[TL_PAIRS];
  uint16_t fq[2 * WAVE];
  uint8_t pa[TL_PAIRS];
};

// Four wavefronts per SIMD (at most 128 VGPRs; LDS under 10 KiB per
// workgroup for the 128-slot rows): the kernel waits on LDS and global loads
// between dependent steps, so occupancy is what hides them.
template <int RC>
__global__ __launch_bounds__(64, 4) void k_probe_rows(BatchView b, Settings s, Work w) {
  constexpr int PR_RELW = RC / WAVE;      // reach mask words per pair
  constexpr int PR_QCAP = RC * 7 / 2;     // triangle-test items per batch, bounds the triple queue
  __shared__ int lj[RC];
  __shared__ double4 lx[RC];
  __shared__ int8_t latt[RC];
  __shared__ uint8_t fws[RC];   // forward slots (atom index > i) by ascending atom index
  __shared__ uint8_t frk[RC];   // forward rank of a forward slot
  __shared__ uint8_t pslots[RC];   // row slots of the row's pairs, in pair order
  __shared__ double4 bu[PR_PAIRS], bt[PR_PAIRS];
  __shared__ int bj[PR_PAIRS], bflag[PR_PAIRS], bnp[PR_PAIRS], bnm[PR_PAIRS], bdel[PR_PAIRS];
  __shared__ uint8_t bsl[PR_PAIRS];   // row slot of the pair's j
  __shared__ uint64_t brel[PR_PAIRS][PR_RELW];
  // the triple queue and the item start marks (four per word) of steps 1-4;
  // step 5 takes their room
  constexpr int QBYTES = 2 * PR_QCAP + 4 * ((PR_QCAP + 3) / 4);
  __shared__ double qraw[QBYTES / 8 + 1];
  uint16_t* const queue = reinterpret_cast<uint16_t*>(qraw);
  uint32_t* const smark_w = reinterpret_cast<uint32_t*>(queue + PR_QCAP);
  uint8_t* const smark = reinterpret_cast<uint8_t*>(smark_w);
  TorusLds& tl = *reinterpret_cast<TorusLds*>(qraw);
  static_assert(sizeof(TorusLds) <= QBYTES && PR_QCAP % 2 == 0, "step 5 fits the room of the triple queue");
  static_assert(RC <= 256 && PR_PAIRS <= 32 && RC <= PR_QCAP, "queue entries pack (pair << 8) | slot in 16 bits");
  const int nR = *w.nR;
  const int nP = *w.nP;
  auto live = [&](int64_t q) { return w.pcnt[q] != 0 && row_in_class<RC>(w.nn[q]); };
  CMS_SPARSE_LOOP(r, nR, live) {
    const int i = w.ratom[r];
    const int nn = w.nn[r];
    const int p_lo = w.pscan[r];
    const int p_hi = min(p_lo + w.pcnt[r], nP);
    if (p_lo >= p_hi || nn <= 0) continue;
    {
      const int* row = w.nbr + (int64_t)r * KCAP;
      for (int sl = lane(); sl < nn; sl += WAVE) {
        const int k = row[sl];
        lj[sl] = k;
        lx[sl] = b.atom[k];
        latt[sl] = w.atten[k];
      }
    }
    __syncthreads();
    // forward atoms by ascending index (ordered by the neighbour pass); the
    // triangle test's d(i, k) < eri + erk depends on k alone, so it is
    // taken once per forward atom (bit LATT_IK of latt, above the
    // attention values)
    const double4 ai = b.atom[i];
    const V3 xi = ld3(ai);
    const double eri = ai.w + s.rp;
    int F = 0;
    for (int s0 = 0; s0 < nn; s0 += WAVE) {
      const int sl = s0 + lane();
      F += __popcll(ballot(sl < nn && lj[sl] > i));
    }
    {
      const uint8_t* fw = w.fwd + (int64_t)r * KCAP;
      for (int f = lane(); f < F; f += WAVE) {
        const uint8_t sl = fw[f];
        fws[f] = sl;
        frk[sl] = (uint8_t)f;
        const double4 ak = lx[sl];
        if (norm_lt(ld3(ak) - xi, eri + (ak.w + s.rp))) latt[sl] = (int8_t)(latt[sl] | LATT_IK);
      }
    }
    const int grp = group_of_atom(b, i);
    const int g0 = w.pgrp[grp], g1 = w.pgrp[grp + 1];
    const int8_t iatt = w.atten[i];
    const bool iblock = iatt <= ATTEN_BLOCKER;
    const int nw = (nn + WAVE - 1) / WAVE;
    // the row's pairs as the pair pass numbers them: relevant forward
    // neighbours in row order (the pass's count for the row is p_hi - p_lo)
    {
      const int nbmax = w.nbmax[r];
      int np = 0;
      for (int s0 = 0; s0 < nn; s0 += WAVE) {
        const int sl = s0 + lane();
        bool rel = false;
        if (sl < nn && lj[sl] > i) {
          const double4 aj = lx[sl];
          rel = pair_geom_ok(xi, ai.w, ld3(aj), aj.w, s.rp) &&
                pair_relevant(iatt, (int8_t)(latt[sl] & LATT_ATT), nbmax, lj[sl]);
        }
        const uint64_t m = ballot(rel);
        if (rel) pslots[np + lanes_below(m)] = (uint8_t)sl;
        np += __popcll(m);
      }
    }
    __syncthreads();
    int prow = 0;   // surviving candidates offered to the row's probe slots
    for (int pb = p_lo; pb < p_hi;) {
      // the batch: consecutive pairs whose triangle items fit the queue
      const int avail = min(PR_PAIRS, p_hi - pb);
      int j = -1, first = F, cnt = 0;
      int slot = 0;
      if (lane() < avail) {
        // forward atoms after j: positions first .. F-1 of fws
        slot = pslots[pb - p_lo + lane()];
        j = lj[slot];
        first = frk[slot] + 1;
        cnt = F - first;
      }
      const int inc = wave_incl_scan(cnt);
      const int nb = __popcll(ballot(lane() < avail && inc <= PR_QCAP));   // >= 1: cnt < RC <= PR_QCAP
      const int items = __shfl(inc, nb - 1, WAVE);
      if (lane() < nb) {
        const double4 aj = lx[slot];
        V3 u, t;
        double rij;
        bool between;
        pair_torus(xi, ai.w, ld3(aj), aj.w, s.rp, &u, &t, &rij, &between);
        bj[lane()] = j;
        bu[lane()] = make_double4(u.x, u.y, u.z, rij);
        bt[lane()] = make_double4(t.x, t.y, t.z, 0.0);
        bsl[lane()] = (uint8_t)slot;
        const int8_t jatt = (int8_t)(latt[slot] & LATT_ATT);
        const bool bij = iblock && jatt <= ATTEN_BLOCKER;
        bflag[lane()] = (pair_toroid(iatt, jatt, s.rp) ? PB_TOR : 0) | (bij ? PB_BLOCKERS : 0) |
                        (between ? PB_BETWEEN : 0) | (jatt >= ATTEN_2 ? PB_ARC2 : 0);
        bnp[lane()] = 0;
        bnm[lane()] = 0;
        // every row atom but j (what the toroid-queue pairs narrow in step 3)
        for (int wd = 0; wd < PR_RELW; ++wd) {
          const int rem = nn - wd * WAVE;
          const uint64_t rowm = rem >= WAVE ? ~0ull : (rem > 0 ? (1ull << rem) - 1ull : 0ull);
          brel[lane()][wd] = rowm & ~(wd == slot / WAVE ? 1ull << (slot % WAVE) : 0ull);
        }
        // the pair's triangle items (pair, forward slot after j) are items
        // inc - cnt .. inc - 1; item it takes forward position it + bdel
        bdel[lane()] = first - (inc - cnt);
      }
      // start marks: pair + 1 at the first item of each pair with items
      for (int q = lane(); q < (items + 3) / 4; q += WAVE) smark_w[q] = 0u;
      __syncthreads();
      if (lane() < nb && cnt > 0) smark[inc - cnt] = (uint8_t)(lane() + 1);
      __syncthreads();

      // 1. triangle tests (_core.py:1478-1500), the lanes over the batch's
      // items in pair order (an item's pair is the last start mark at or
      // before it); valid triples are queued in the same order
      int nq = 0;
      int acur = 0;   // pair + 1 of the items just before the chunk
      for (int it0 = 0; it0 < items; it0 += WAVE) {
        const int it = it0 + lane();
        const int mk = it < items ? (int)smark[it] : 0;
        const uint64_t S = ballot(mk != 0);
        const uint64_t upto = S & ((2ull << lane()) - 1ull);   // marks at lanes <= this one
        const int src = upto ? 63 - __builtin_clzll(upto) : lane();
        const int mlast = __shfl(mk, src, WAVE);
        const int apl = upto ? mlast : acur;
        if (S) acur = __builtin_amdgcn_readlane(mk, 63 - __builtin_clzll(S));
        bool valid = false;
        int e = 0;
        if (it < items) {
          const int a = apl - 1;
          const int sl = fws[it + bdel[a]];
          e = (a << 8) | sl;
          const int lk = latt[sl];
          const double4 ak = lx[sl], aj = lx[bsl[a]];
          const double erk = ak.w + s.rp, erj = aj.w + s.rp;
          valid = (lk & LATT_IK) && !((bflag[a] & PB_BLOCKERS) && (lk & LATT_ATT) <= ATTEN_BLOCKER) &&
                  norm_lt(ld3(ak) - ld3(aj), erj + erk);
        }
        const uint64_t m = ballot(valid);
        if (valid) queue[nq + lanes_below(m)] = (uint16_t)e;
        nq += __popcll(m);
      }
      __syncthreads();

      // 2. degenerate triples kill the pair when the torus centre lies inside
      // k's sphere shrunk to the circle (_core.py:1505-1543); candidates are
      // compacted in place (each lane writes at or below the entry it read)
      int nc = 0;
      for (int q0 = 0; q0 < nq; q0 += WAVE) {
        const int q = q0 + lane();
        bool cand = false;
        int e = 0;
        if (q < nq) {
          e = queue[q];
          const int a = e >> 8, sl = e & 255;
          const double4 ak = lx[sl], pu = bu[a];
          const V3 xk = ld3(ak);
          const V3 v = xk - xi;
          const double vv = dot_seq(v, v);
          const double c = dot_alt(ld3(pu), v);
          if (c * c < DT_CLEAR_SQ * vv) {
            cand = true;   // |dt| < 1 - 1e-6: neither degenerate nor near it
          } else {
            const double erk = ak.w + s.rp;
            const double dik = sqrt(vv);   // norm_seq(xk - xi)
            const V3 uik = v / dik;
            const double dt = dot_alt(ld3(pu), uik);
            if (triple_degenerate(dt)) {
              const V3 dk = ld3(bt[a]) - xk;
              if (dot_seq(dk, dk) < erk * erk - pu.w * pu.w) atomicOr(&bflag[a], PB_KILLED);
            } else {
              cand = true;
              if (!(fabs(dt) < 1.0 - 1e-10)) atomicOr(&bflag[a], PB_NEARDEG);
            }
          }
        }
        const uint64_t m = ballot(cand);
        if (cand) queue[nc + lanes_below(m)] = (uint16_t)e;
        nc += __popcll(m);
      }
      __syncthreads();

      // 3. containment (circle_in_one_sphere, in squared form) and the row
      // atoms within reach of the circle, for the toroid-queue pairs: one
      // pair at a time, the lanes over the row slots, the reach mask words
      // by ballot; near-degenerate pairs take every row atom.  (Probes of a
      // pair inside one sphere all collide with it, so leaving the other
      // pairs unflagged changes no result.)
      {
        bool need = false;
        if (lane() < nb) need = (bflag[lane()] & PB_TOR) != 0;
        uint64_t m = ballot(need);
        while (m) {
          const int a = __builtin_ctzll(m);
          m &= m - 1;
          const int f = bflag[a];
          const int sj = __builtin_amdgcn_readlane(slot, a);   // j's row slot
          const double4 pu = bu[a], pt = bt[a];
          const double rho2 = pu.w * pu.w;
          bool cov = false;
#pragma unroll
          for (int wd = 0; wd < PR_RELW; ++wd) {
            const int sl = wd * WAVE + lane();
            bool hit = false;
            if (sl < nn && sl != sj) {
              const double4 ax = lx[sl];
              const V3 v = ld3(ax) - ld3(pt);
              const double aa = dot_seq(v, ld3(pu));
              const double d2 = fmax(dot_seq(v, v) - aa * aa, 0.0);
              const double rr = ax.w + s.rp;
              if (!(f & PB_NEARDEG)) {
                // farthest circle point within R: (d + rho)^2 + a^2 < R^2
                const double R = rr - COVER_MARGIN;
                const double D = ((R * R - aa * aa) - d2) - rho2;
                cov = cov || (R > 0.0 && D > 0.0 && 4.0 * d2 * rho2 < D * D);
              }
              // nearest circle point within reach: (d - rho)^2 + a^2 < reach^2
              const double reach = rr + REACH_MARGIN;
              const double E = ((d2 + rho2) + aa * aa) - reach * reach;
              hit = (f & PB_NEARDEG) || E < 0.0 || 4.0 * d2 * rho2 > E * E;
            }
            const uint64_t word = ballot(hit);
            if (lane() == 0) brel[a][wd] = word;
          }
          if (wave_any(cov) && lane() == 0) bflag[a] = f | PB_COVERED;
        }
      }
      __syncthreads();

      // the candidates of killed pairs and of pairs inside one sphere have
      // no probes: the others are compacted in place, in order, so that
      // the placement below runs on full wavefronts
      {
        int nl = 0;
        for (int q0 = 0; q0 < nc; q0 += WAVE) {
          const int q = q0 + lane();
          int e = 0;
          bool lv = false;
          if (q < nc) {
            e = queue[q];
            lv = !(bflag[e >> 8] & (PB_KILLED | PB_COVERED));
          }
          const uint64_t m = ballot(lv);
          if (lv) queue[nl + lanes_below(m)] = (uint16_t)e;
          nl += __popcll(m);
        }
        nc = nl;
      }
      __syncthreads();

      // 4. probes of the candidates of live pairs and their collisions
      // (_core.py:1545-1623)
      for (int q0 = 0; q0 < nc; q0 += WAVE) {
        const int q = q0 + lane();
        bool okp = false, okm = false;
        int a = 0, j = -1, k = -1, sl = 0;
        V3 pp{0, 0, 0}, pm{0, 0, 0}, uijk{0, 0, 0};
        double h = 0.0;
        if (q < nc) {
          const int e = queue[q];
          a = e >> 8;
          sl = e & 255;
          j = bj[a];
          k = lj[sl];
          const double4 ak = lx[sl], pu = bu[a], pt = bt[a];
          okp = okm = probe_points(xi, eri, ld3(ak), ak.w + s.rp, ld3(pu), ld3(pt), &pp, &pm, &uijk, &h);
        }
        for (int wd = 0; wd < nw && (okp || okm); ++wd) {
          uint64_t rm = brel[a][wd];
          // two row atoms per step, their loads issued together; k itself is
          // skipped (row slots hold distinct atoms)
          while (rm && (okp || okm)) {
            const int t0 = wd * WAVE + __builtin_ctzll(rm);
            rm &= rm - 1;
            const int t1 = rm ? wd * WAVE + __builtin_ctzll(rm) : t0;
            rm &= rm - 1;
            const double4 ax0 = lx[t0], ax1 = lx[t1];
            bool hp = false, hm = false;
            if (t0 != sl) {
              const double rr = ax0.w + s.rp;
              const double lim = rr * rr;
              const V3 xs = ld3(ax0);
              const V3 dp = pp - xs, dm = pm - xs;
              hp = dot_alt(dp, dp) <= lim;
              hm = dot_alt(dm, dm) <= lim;
            }
            if (t1 != sl) {
              const double rr = ax1.w + s.rp;
              const double lim = rr * rr;
              const V3 xs = ld3(ax1);
              const V3 dp = pp - xs, dm = pm - xs;
              hp = hp || dot_alt(dp, dp) <= lim;
              hm = hm || dot_alt(dm, dm) <= lim;
            }
            okp = okp && !hp;
            okm = okm && !hm;
          }
        }
        if (okp || okm) {
          w.access[i] = 1;
          w.access[j] = 1;
          w.access[k] = 1;
        }
        if (okp) atomicAdd(&bnp[a], 1);
        if (okm) atomicAdd(&bnm[a], 1);
        // survivors go to the row's probe slots for k_probe_place; a pair
        // with one past them is written by the emit pass instead
        const uint64_t sm = ballot(okp || okm);
        if (sm) {
          const int idx = prow + lanes_below(sm);
          prow += __popcll(sm);
          if (okp || okm) {
            if (idx < w.pslots) {
              ProbeRec pr;
              pr.pp = make_double4(pp.x, pp.y, pp.z, h);
              pr.pm = make_double4(pm.x, pm.y, pm.z, 0.0);
              pr.u = make_double4(uijk.x, uijk.y, uijk.z, 0.0);
              pr.p = pb + a;
              pr.k = k;
              pr.sl = sl;
              pr.f = (okp ? 1 : 0) | (okm ? 2 : 0);
              w.pstage[r * PSLOTS + idx] = pr;
            } else {
              atomicOr(&bflag[a], PB_EMIT);
            }
          }
        }
      }
      __syncthreads();

      // 5. torus points of the batch's live toroid-queue pairs (circle not
      // inside one neighbour sphere; _core.py:1863-2044), TL_PAIRS pairs at
      // a time: the circle frames, then the points spread over the lanes
      // across pair boundaries, each tested against the row atoms within
      // reach of its pair's circle (strict <, _core.py:1905-1917), and the
      // saddle-arc dot counts of the free ones, queued and taken a wave at
      // a time.  Writes the arc dot counts of both atoms, whether any point
      // is free and the free points among the first 128; the host zero-fills
      // them for the other pairs.
      {
        const bool tlv = lane() < nb && (bflag[lane()] & (PB_TOR | PB_COVERED)) == PB_TOR;
        const uint64_t tlm = ballot(tlv);
        const int ntl = __popcll(tlm);
        const int tix = lanes_below(tlm);
        const bool iarc = iatt >= ATTEN_2;
        for (int t0 = 0; t0 < ntl; t0 += TL_PAIRS) {
          const int nt = min(TL_PAIRS, ntl - t0);
          if (tlv && tix >= t0 && tix < t0 + nt) tl.pa[tix - t0] = (uint8_t)lane();
          __syncthreads();
          int M = 0;
          if (lane() < nt) {
            const int a = tl.pa[lane()];
            const double4 pu = bu[a], aj = lx[bsl[a]];
            const TorusFrame tf = torus_frame(ai.w, aj.w, s.rp, pu.w, ld3(pu), (b.dens[i] + b.dens[bj[a]]) * 0.5);
            tl.cx[lane()][0] = tf.cx.x;
            tl.cx[lane()][1] = tf.cx.y;
            tl.cx[lane()][2] = tf.cx.z;
            tl.cy[lane()][0] = tf.cy.x;
            tl.cy[lane()][1] = tf.cy.y;
            tl.cy[lane()][2] = tf.cy.z;
            tl.td[lane()] = tf.tdelta;
            tl.ad[lane()] = tf.adelta;
            tl.mk0[lane()] = 0ull;
            tl.mk1[lane()] = 0ull;
            tl.n1[lane()] = 0;
            tl.n2[lane()] = 0;
            M = tf.M;
          }
          const int incM = wave_incl_scan(M);
          if (lane() < nt) tl.pre[lane()] = incM - M;
          const int T = __shfl(incM, WAVE - 1, WAVE);
          __syncthreads();
          auto tpoint = [&](int q, int m) {
            const double4 pt = bt[tl.pa[q]];
            return subdiv_point(ld3(pt), v3(tl.cx[q][0], tl.cx[q][1], tl.cx[q][2]),
                                v3(tl.cy[q][0], tl.cy[q][1], tl.cy[q][2]), bu[tl.pa[q]].w, tl.td[q], m);
          };
          auto arcs = [&](int cnt) {
            if (lane() < cnt) {
              const int e = tl.fq[lane()];
              const int q = e >> 8;
              const int a = tl.pa[q];
              const double4 pu = bu[a], aj = lx[bsl[a]], pt = bt[a];
              const int f = bflag[a];
              const V3 pm = tpoint(q, e & 255);
              TorusSub ts_{};
              ts_.n1 = ts_.n2 = 0;
              torus_point_arcs(pm, xi, ld3(aj), eri, aj.w + s.rp, ld3(pt), ld3(pu), pu.w, s.rp, (f & PB_BETWEEN) != 0,
                               iarc, (f & PB_ARC2) != 0, tl.ad[q], &ts_);
              if (ts_.n1) atomicAdd(&tl.n1[q], ts_.n1);
              if (ts_.n2) atomicAdd(&tl.n2[q], ts_.n2);
            }
          };
          int nf = 0;   // queued free points
          for (int it0 = 0; it0 < T; it0 += WAVE) {
            const int it = it0 + lane();
            int fe = -1;
            if (it < T) {
              int q = 0;
#pragma unroll
              for (int step = TL_PAIRS / 2; step > 0; step >>= 1)
                if (q + step < nt && tl.pre[q + step] <= it) q += step;
              const int m = it - tl.pre[q];
              const int a = tl.pa[q];
              const V3 pm = tpoint(q, m);
              // two reaching atoms per step, their loads issued together
              // (the point is free when none of them holds it, in any order)
              bool free_pt = true;
              for (int wd = 0; wd < nw && free_pt; ++wd) {
                uint64_t rm = brel[a][wd];
                while (rm && free_pt) {
                  const int u0 = wd * WAVE + __builtin_ctzll(rm);
                  rm &= rm - 1;
                  const int u1 = rm ? wd * WAVE + __builtin_ctzll(rm) : u0;
                  rm &= rm - 1;
                  const double4 ax0 = lx[u0], ax1 = lx[u1];
                  const V3 d0 = pm - ld3(ax0), d1 = pm - ld3(ax1);
                  const double lim0 = ax0.w + s.rp, lim1 = ax1.w + s.rp;
                  if (dot_alt(d0, d0) < lim0 * lim0 || dot_alt(d1, d1) < lim1 * lim1) free_pt = false;
                }
              }
              if (free_pt) {
                if (m < WAVE) atomicOr(&tl.mk0[q], 1ull << m);
                else atomicOr(&tl.mk1[q], 1ull << (m - WAVE));
                fe = (q << 8) | m;
              }
            }
            // free points join the queue; a full wave of them gets its arcs
            const uint64_t fm = ballot(fe >= 0);
            if (fe >= 0) tl.fq[nf + lanes_below(fm)] = (uint16_t)fe;
            nf += __popcll(fm);
            if (nf >= WAVE) {
              __syncthreads();
              arcs(WAVE);
              // move the rest (fewer than WAVE) to the front
              const int rest = nf - WAVE;
              uint16_t e = 0;
              if (lane() < rest) e = tl.fq[WAVE + lane()];
              __syncthreads();
              if (lane() < rest) tl.fq[lane()] = e;
              nf = rest;
            }
          }
          __syncthreads();
          if (nf > 0) arcs(nf);
          __syncthreads();
          if (lane() < nt) {
            const int p = pb + tl.pa[lane()];
            const unsigned long long m0 = tl.mk0[lane()], m1 = tl.mk1[lane()];
            w.t1cnt[p] = tl.n1[lane()];
            w.t2cnt[p] = tl.n2[lane()];
            w.tflag[p] = (m0 | m1) ? 1 : 0;
            w.tmask[2 * (int64_t)p] = m0;
            w.tmask[2 * (int64_t)p + 1] = m1;
          }
          __syncthreads();
        }
      }

      // 6. per-pair results
      if (lane() < nb) {
        const int p = pb + lane();
        const int64_t vp = 2ll * g0 + (p - g0);
        const int64_t vm = 2ll * g0 + (g1 - g0) + (p - g0);
        const int f = bflag[lane()];
        int pf = ((f & PB_BETWEEN) ? 1 : 0) | ((f & PB_TOR) ? 2 : 0);
        int np = 0, nm = 0;
        if (f & PB_COVERED) {
          pf |= 4;
        } else if (!(f & PB_KILLED)) {
          np = bnp[lane()];
          nm = bnm[lane()];
          if (np + nm > 0 && (f & PB_EMIT)) {
            pf |= 8;
            w.elist[atomicAdd(w.nelist, 1)] = p;   // a pair with a probe past the slots: rare
          }
        }
        // the pair record (what the pair pass's emit would write); the
        // emit passes read the geometry only of pairs with probes and of
        // the live toroid-queue pairs
        const bool tor_live = (f & PB_TOR) && !(f & PB_COVERED);
        if (f & PB_ARC2) pf |= 16;
        w.pi[p] = i;
        w.pj[p] = bj[lane()];
        w.pflag[p] = (int8_t)pf;
        if (np + nm > 0 || tor_live) {
          w.pu[p] = bu[lane()];
          w.pt[p] = bt[lane()];
        }
        w.prcnt[vp] = np;
        w.prcnt[vm] = nm;
      }
      pb += nb;
      __syncthreads();
    }
    if (lane() == 0) w.nps[r] = min(prow, w.pslots);
    __syncthreads();
  }
}

// Probes staged by k_probe_rows, one thread per row slot, written where the
// emit pass would write them: the pair's "+" / "-" probes in row-slot order
// of k after the pair's scanned offsets, with the atom order, height and
// normal sign of k_probes<true> (_core.py:1590-1643).  Pairs flagged for the
// emit pass (a survivor past the slots) are left to it.
__global__ __launch_bounds__(256) void k_probe_place(BatchView b, Work w) {
  const int64_t n = (int64_t)*w.nR * PSLOTS;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = c / PSLOTS;
    const int e = (int)(c % PSLOTS);
    const int cnt = w.nps[r];
    if (e >= cnt) continue;
    const ProbeRec* rec = w.pstage + r * PSLOTS;
    const ProbeRec pr = rec[e];
    const int p = pr.p;
    if (w.pflag[p] & 8) continue;
    int rkp = 0, rkm = 0;   // earlier survivors of the pair (by row slot)
    for (int t = 0; t < cnt; ++t) {
      const ProbeRec& o = rec[t];
      if (o.p == p && o.sl < pr.sl) {
        rkp += o.f & 1;
        rkm += (o.f >> 1) & 1;
      }
    }
    const int i = w.pi[p], j = w.pj[p];
    const int grp = group_of_atom(b, i);
    const int g0 = w.pgrp[grp], g1 = w.pgrp[grp + 1];
    if (pr.f & 1) {
      const int o = w.prscan[2ll * g0 + (p - g0)] + rkp;
      if (o < w.probe_cap) {
        w.patom[o] = make_int4(i, j, pr.k, 0);
        w.ppos[o] = pr.pp;
        w.palt[o] = make_double4(pr.u.x, pr.u.y, pr.u.z, 0.0);
      }
    }
    if (pr.f & 2) {
      const int o = w.prscan[2ll * g0 + (g1 - g0) + (p - g0)] + rkm;
      if (o < w.probe_cap) {
        w.patom[o] = make_int4(j, i, pr.k, 0);
        w.ppos[o] = make_double4(pr.pm.x, pr.pm.y, pr.pm.z, pr.pp.w);
        w.palt[o] = make_double4(-pr.u.x, -pr.u.y, -pr.u.z, 0.0);
      }
    }
  }
}

// Low probes (h < rp) within 2rp of another low probe of the same complex
// (_core.py:2105-2134; self excluded).
__global__ __launch_bounds__(64) void k_near(BatchView b, Settings s, Work w) {
  const int nQ = *w.nQ;
  const double rp2 = s.rp * s.rp;
  const double lim = 4.0 * rp2;
  CMS_UNIT_LOOP(q, nQ) {
    const double4 pq = w.ppos[q];
    int cnt = 0;
    if (pq.w < s.rp) {
      const int c = b.acplx[w.patom[q].x];
      const int lo = w.prgrp[2 * c], hi = w.prgrp[2 * c + 2];
      const V3 xq = ld3(pq);
      int* out = w.near + (int64_t)q * NEARCAP;
      for (int l0 = lo; l0 < hi; l0 += WAVE) {
        const int l = l0 + lane();
        bool ok = false;
        if (l < hi && l != q) {
          const double4 pl = w.ppos[l];
          if (pl.w < s.rp) {
            const V3 d = xq - ld3(pl);
            ok = dot_seq(d, d) <= lim;
          }
        }
        const uint64_t mk = ballot(ok);
        if (ok) {
          const int o = cnt + lanes_below(mk);
          if (o < NEARCAP) out[o] = l;
        }
        cnt += __popcll(mk);
      }
      if (cnt > NEARCAP) {
        if (lane() == 0) atomicOr(&w.status[c], ST_NEAR_CAP);
        cnt = NEARCAP;
      }
    }
    if (lane() == 0) w.nnear[q] = cnt;
  }
}

// ------------------------------------------------------------ dot output
// Outward normal and burial test against the owning atom's buried row
// (_core.py:2337-2377), then the store.
__device__ __forceinline__ DotRec make_dot(const BatchView& b, const Settings& s, const Work& w, V3 pt, V3 pc,
                                           double area, int owner, int rowner, int8_t type, int4 desc,
                                           int4 desc2) {
  V3 nml = (s.rp <= 0.0) ? pt - ld3(b.atom[owner]) : (pc - pt) / s.rp;
  // every dot owner has a neighbour row; a missing one would be a bug, so
  // it is flagged (pool status) instead of read out of bounds
  int nbur = 0;
  if (rowner >= 0 && rowner < *w.nR && w.ratom[rowner] == owner) {
    nbur = w.nb[rowner];
  } else {
    atomicOr(&w.status[b.acplx[owner]], ST_POOL);
  }
  const int* brow = w.bur + (int64_t)rowner * BCAP;
  int8_t buried = 0;
  for (int t = 0; t < nbur; ++t) {
    const double4 ab = b.atom[brow[t]];
    const V3 d = ld3(ab) - pc;
    const double lim = ab.w + s.rp;
    if (dot_seq(d, d) <= lim * lim) { buried = 1; break; }
  }
  DotRec d;
  d.pos = make_double4(pt.x, pt.y, pt.z, area);
  d.nml = make_double4(nml.x, nml.y, nml.z, 0.0);
  d.desc = desc;
  d.desc2 = desc2;
  d.atom = owner;
  d.bur = buried;
  d.type = type;
  return d;
}
__device__ __forceinline__ void write_dot(const Work& w, int64_t o, const DotRec& d) {
  if (o >= w.dot_cap) return;
  w.dpos[o] = d.pos;
  w.dnml[o] = d.nml;
  w.dbur[o] = d.bur;
  w.dtype[o] = d.type;
  w.datom[o] = d.atom;
  if (w.ddesc) {
    w.ddesc[o] = d.desc;
    w.ddesc2[o] = d.desc2;
  }
}
__device__ __forceinline__ void store_dot(const BatchView& b, const Settings& s, const Work& w,
                                          int64_t o, V3 pt, V3 pc, double area, int owner,
                                          int rowner, int8_t type, int4 desc, int4 desc2) {
  write_dot(w, o, make_dot(b, s, w, pt, pc, area, owner, rowner, type, desc, desc2));
}

// ------------------------------------------------------------ convex
// _core.py:1646-1840.  Pole from the nearest neighbour, latitude arcs from
// the pole to the torus-derived south point, circles at each latitude,
// collision with every neighbour but the nearest one.

// The atom's frame (_core.py:1660-1712): pole direction north (away from
// the nearest neighbour x0), the latitude arc's second axis y, its spacing
// delta and count L, the arc length cs of a latitude step, and the
// latitude circles' frame (cx, cy).
struct ConvexFrame {
  V3 north, y, cx, cy;
  double delta, cs;
  int L;
};
__device__ __forceinline__ ConvexFrame convex_frame(V3 xi, double ri, double dens, double4 a0, double rp) {
  ConvexFrame f;
  const double eri = ri + rp;
  const V3 x0 = ld3(a0);
  const double r0 = a0.w;
  const V3 north = unit(xi - x0);
  V3 vt = v3(north.y * north.y + north.z * north.z, north.x * north.x + north.z * north.z,
             north.x * north.x + north.y * north.y);
  vt = unit(vt);
  if (fabs(dot_seq(vt, north)) > 0.99) vt = v3(1.0, 0.0, 0.0);
  const V3 eq = unit(cross(north, vt));
  const V3 vql = cross(eq, north);
  const double erj = r0 + rp;
  const double dij = norm_seq(x0 - xi);
  const V3 uij = (x0 - xi) / dij;
  const double asymm = (eri * eri - erj * erj) / dij;
  const V3 tij = ((xi + x0) * 0.5) + uij * (asymm * 0.5);
  const double far_sq = (eri + erj) * (eri + erj) - dij * dij;
  const double cont_sq = dij * dij - (ri - r0) * (ri - r0);
  V3 south = v3(0.0, 0.0, -1.0);
  if (far_sq > 0.0 && cont_sq > 0.0) {
    const double rij = 0.5 * sqrt(far_sq) * sqrt(cont_sq) / dij;
    const V3 pij = tij + vql * rij;
    south = (pij - xi) / eri;
  }
  const double ang = arc_angle(north, south, eq, &f.y);
  f.north = north;
  f.delta = 1.0 / (sqrt(dens) * ri);
  f.L = subdiv_count(f.delta, ang);
  f.cs = f.L > 0 ? (ri * ang) / (double)f.L : 0.0;
  circle_frame(north, &f.cx, &f.cy);
  return f;
}

// Frames of the rows the convex count pass visits (accessible non-blocker
// atoms with neighbours), one thread per row: the passes, one wavefront per
// row, would otherwise compute the same frame on every lane.  Four double4
// per row: (north, delta), (y, cs), (cx, L), (cy, 0).
__global__ __launch_bounds__(256) void k_convex_frames(BatchView b, Settings s, Work w) {
  const int nR = *w.nR;
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < nR; r += gridDim.x * blockDim.x) {
    const int g = w.ratom[r];
    if (!(w.nn[r] > 0 && w.access[g] != 0 && w.atten[g] > ATTEN_BLOCKER)) continue;
    const double4 ai = b.atom[g];
    const ConvexFrame f = convex_frame(ld3(ai), ai.w, b.dens[g], b.atom[w.nbr[(int64_t)r * KCAP]], s.rp);
    double4* o = w.cvf + 4 * (int64_t)r;
    o[0] = make_double4(f.north.x, f.north.y, f.north.z, f.delta);
    o[1] = make_double4(f.y.x, f.y.y, f.y.z, f.cs);
    o[2] = make_double4(f.cx.x, f.cx.y, f.cx.z, (double)f.L);
    o[3] = make_double4(f.cy.x, f.cy.y, f.cy.z, 0.0);
  }
}
template <bool EMIT, int RC>
__global__ __launch_bounds__(64, EMIT ? 4 : 5) void k_convex(BatchView b, Settings s, Work w) {
  // Occupancy: the count pass is held to 102 VGPRs (five wavefronts per
  // SIMD), the emit pass to 128 (four); rows of at most RC_SMALL neighbours
  // run in the RC_SMALL instantiation, whose neighbour staging takes half
  // the LDS (under 8 KiB per workgroup, twenty workgroups per CU), longer
  // rows in the KCAP one.  The point spacings exist in the emit pass only,
  // and the emit pass's list of kept points shares the neighbours' storage
  // (listed and placed before they load)
  __shared__ double4 lx[RC];
  uint16_t* const klist = reinterpret_cast<uint16_t*>(lx);   // emit: kept points under the count pass's mask
  static_assert(CMASK_WORDS * WAVE * sizeof(uint16_t) <= RC * sizeof(double4), "kept-point list fits");
  // latitude circle t: centre xi + north * lat_dt[t] (recomputed where used,
  // the same operations, so the same bits)
  __shared__ double lat_dt[MAX_SUBDIV];
  __shared__ double lat_rad[MAX_SUBDIV], lat_delta[MAX_SUBDIV], lat_ps[EMIT ? MAX_SUBDIV : 1];
  __shared__ int lat_start[MAX_SUBDIV + 1];
  const int nR = *w.nR;
  // count pass: accessible non-blocker atoms with neighbours (the counts of
  // the other rows are zero-filled before); emit pass: rows with dots
  auto live = [&](int64_t q) {
    if (EMIT) return w.ccnt[q] != 0 && row_in_class<RC>(w.nn[q]);
    const int g = w.ratom[q];
    return w.nn[q] > 0 && row_in_class<RC>(w.nn[q]) && w.access[g] != 0 && w.atten[g] > ATTEN_BLOCKER;
  };
  CMS_SPARSE_LOOP(r, nR, live) {
    const int g = w.ratom[r];
    const int nn = w.nn[r];
    int total = 0;
    {
      const int* row = w.nbr + (int64_t)r * KCAP;
      const double4 ai = b.atom[g];
      const V3 xi = ld3(ai);
      const double ri = ai.w, eri = ri + s.rp;
      const double dens = b.dens[g];
      // the row's frame from k_convex_frames
      const double4 f0 = w.cvf[4 * (int64_t)r], f1 = w.cvf[4 * (int64_t)r + 1];
      const double4 f2 = w.cvf[4 * (int64_t)r + 2], f3 = w.cvf[4 * (int64_t)r + 3];
      const V3 north = ld3(f0), y = ld3(f1), cx = ld3(f2), cy = ld3(f3);
      const double delta = f0.w, cs = f1.w;
      const int L = (int)f2.w;
      // per-latitude circles
      int run = 0;
      for (int t0 = 0; t0 < L; t0 += WAVE) {
        const int t = t0 + lane();
        int P = 0;
        if (t < L) {
          const V3 lat = subdiv_point(v3(0.0, 0.0, 0.0), north, y, ri, delta, t);
          const double dt = dot_seq(lat, north);
          const double rsq = ri * ri - dt * dt;
          double rad = 0.0, cdelta = 0.0, ps = 0.0;
          if (rsq > 0.0) {
            rad = sqrt(rsq);
            cdelta = 1.0 / (sqrt(dens) * rad);
            P = subdiv_count(cdelta, TWO_PI);
            ps = P > 0 ? (rad * TWO_PI) / (double)P : 0.0;
          }
          lat_dt[t] = dt;
          lat_rad[t] = rad; lat_delta[t] = cdelta;
          if (EMIT) lat_ps[t] = ps;
        }
        const int inc = wave_incl_scan(P);
        if (t < L) lat_start[t] = run + inc - P;
        run += __shfl(inc, WAVE - 1, WAVE);
      }
      if (lane() == 0) lat_start[L] = run;
      __syncthreads();
      const int C = run;
      int64_t obase = 0;
      if (EMIT) {
        const int grp = group_of_atom(b, g);
        obase = w.dbase[grp * 4 + 0] + w.cscan[r];
      }
      const double scale = eri / ri;
      // the emit pass takes the collision outcome of the first
      // CMASK_WORDS * 64 points from the count pass: their kept points are
      // listed (in point order) and placed one per lane; the points past
      // them are tested again below
      int q_from = 0;
      if (EMIT) {
        const int cm = min(C, CMASK_WORDS * WAVE);
        int nk = 0;
        for (int wd = 0; wd * WAVE < cm; ++wd) {
          const uint64_t m = w.cmask[(int64_t)r * CMASK_WORDS + wd];
          if ((m >> lane()) & 1ull) klist[nk + lanes_below(m)] = (uint16_t)(wd * WAVE + lane());
          nk += __popcll(m);
        }
        __syncthreads();
        for (int k0 = 0; k0 < nk; k0 += WAVE) {
          const int kq = k0 + lane();
          if (kq < nk) {
            const int q = klist[kq];
            const int t = last_le(lat_start, L, q);   // lat_start[t] <= q < lat_start[t + 1]
            const V3 pt = subdiv_point(xi + north * lat_dt[t], cx, cy, lat_rad[t], lat_delta[t],
                                       q - lat_start[t]);
            const V3 pc = xi + (pt - xi) * scale;
            store_dot(b, s, w, obase + kq, pt, pc, lat_ps[t] * cs, g, r, 1, make_int4(DK_CONVEX, g, row[0], -1),
                      make_int4(t, q - lat_start[t], L, lat_start[t + 1] - lat_start[t]));
          }
        }
        total = nk;
        q_from = cm;
      }
      // the neighbours, for the collision tests the emit pass does not take
      // from the count pass (in the emit pass over the kept-point list,
      // which every lane has finished reading)
      if (!EMIT || C > CMASK_WORDS * WAVE) {
        if (EMIT) __syncthreads();
        for (int sl = lane(); sl < nn; sl += WAVE) lx[sl] = b.atom[row[sl]];
        __syncthreads();
      }
      for (int q0 = q_from; q0 < C; q0 += WAVE) {
        const int q = q0 + lane();
        bool keep = false;
        V3 pt{0, 0, 0}, pc{0, 0, 0};
        int t = 0;
        if (q < C) {
          t = last_le(lat_start, L, q);   // lat_start[t] <= q < lat_start[t + 1]
          const int kk = q - lat_start[t];
          pt = subdiv_point(xi + north * lat_dt[t], cx, cy, lat_rad[t], lat_delta[t], kk);
          pc = xi + (pt - xi) * scale;
          {
            keep = true;
            for (int sl = 1; sl < nn; ++sl) {
              const V3 d = ld3(lx[sl]) - pc;
              const double lim = lx[sl].w + s.rp;
              if (dot_seq(d, d) <= lim * lim) { keep = false; break; }
            }
          }
        }
        const uint64_t mk = ballot(keep);
        if (!EMIT && q0 < CMASK_WORDS * WAVE && lane() == 0) w.cmask[(int64_t)r * CMASK_WORDS + q0 / WAVE] = mk;
        if (EMIT && keep) {
          const int64_t o = obase + total + lanes_below(mk);
          store_dot(b, s, w, o, pt, pc, lat_ps[t] * cs, g, r, 1, make_int4(DK_CONVEX, g, row[0], -1),
                    make_int4(t, q - lat_start[t], L, lat_start[t + 1] - lat_start[t]));
        }
        total += __popcll(mk);
      }
    }
    if (!EMIT && lane() == 0) w.ccnt[r] = total;
    __syncthreads();
  }
}

// ------------------------------------------------------------ toroidal
// _core.py:1847-2067.  Torus circle of the pair, collision with atom 1's
// neighbours other than atom 2 (strict <), then the saddle arcs towards atom
// 1 (owner i) and atom 2 (owner j).  Cusp pairs keep pqj = 0, which leaves
// the atom-2 arc without an angle and hence without dots.
// Count pass for small batches, one wavefront per toroid-queue pair whose
// circle is not inside a neighbour's sphere (the counts of the other pairs
// are zero-filled before): the torus points on the lanes, their collisions
// with the neighbours of i that can reach the circle, and the saddle-arc
// dot counts of the free ones; tmask keeps the free points (two words per
// pair: MAX_SUBDIV <= 128).
__global__ __launch_bounds__(64) void k_toroid_count(BatchView b, Settings s, Work w) {
  __shared__ int lj[KCAP];
  __shared__ double4 lx[KCAP];
  __shared__ int rl[KCAP];
  static_assert(MAX_SUBDIV <= 2 * WAVE, "two mask words per pair");
  const int nP = *w.nP;
  auto live = [&](int64_t q) { return (w.pflag[q] & 6) == 2; };
  int staged = -1;   // neighbour row held in lj / lx
  CMS_SPARSE_LOOP(p, nP, live) {
    int tot1 = 0, tot2 = 0;
    uint64_t lmw[2] = {0ull, 0ull};
    const int i = w.pi[p], j = w.pj[p];
    const int ri_ = w.rpos[i];
    const int nn = w.nn[ri_];
    const double4 ai = b.atom[i], aj = b.atom[j];
    const V3 xi = ld3(ai), xj = ld3(aj);
    const double ri = ai.w, rj = aj.w;
    const double eri = ri + s.rp, erj = rj + s.rp;
    const double dens = (b.dens[i] + b.dens[j]) * 0.5;
    const double4 pu = w.pu[p], ptt = w.pt[p];
    const V3 uij = ld3(pu), tij = ld3(ptt);
    const double rij = pu.w;
    const bool between = (w.pflag[p] & 1) != 0;
    const TorusFrame tf = torus_frame(ri, rj, s.rp, rij, uij, dens);
    const bool arc1_on = w.atten[i] >= ATTEN_2, arc2_on = w.atten[j] >= ATTEN_2;
    if (ri_ != staged) {
      const int* row = w.nbr + (int64_t)ri_ * KCAP;
      for (int sl = lane(); sl < nn; sl += WAVE) {
        const int k = row[sl];
        lj[sl] = k;
        lx[sl] = b.atom[k];
      }
      __syncthreads();
      staged = ri_;
    }
    // Neighbours (other than j) that come within r + rp of the torus circle
    // (centre tij, axis uij, radius rij): only they can collide with a torus
    // point.  The 1e-4 A margin covers the rounding of the circle distance
    // (cancellation in the radial term near the axis) and of the torus
    // points, so the pruned set gives the same collision results.
    int nrel = 0;
    for (int s0 = 0; s0 < nn; s0 += WAVE) {
      const int sl = s0 + lane();
      bool rel = false;
      if (sl < nn && lj[sl] != j) {
        const V3 v = ld3(lx[sl]) - tij;
        const double ax = dot_seq(v, uij);
        const double rad = sqrt(fmax(dot_seq(v, v) - ax * ax, 0.0)) - rij;
        const double reach = (lx[sl].w + s.rp) + 1e-4;
        rel = rad * rad + ax * ax < reach * reach;
      }
      const uint64_t mk = ballot(rel);
      if (rel) rl[nrel + lanes_below(mk)] = sl;
      nrel += __popcll(mk);
    }
    __syncthreads();
    for (int m0 = 0; m0 < tf.M; m0 += WAVE) {
      const int m = m0 + lane();
      TorusSub ts_{};
      ts_.n1 = ts_.n2 = 0;
      bool free_pt = false;
      if (m < tf.M) {
        const V3 pm = subdiv_point(tij, tf.cx, tf.cy, rij, tf.tdelta, m);
        free_pt = true;
        for (int t = 0; t < nrel; ++t) {
          const int sl = rl[t];
          const V3 d = pm - ld3(lx[sl]);
          const double lim = lx[sl].w + s.rp;
          if (dot_alt(d, d) < lim * lim) { free_pt = false; break; }
        }
        if (free_pt)
          torus_point_arcs(pm, xi, xj, eri, erj, tij, uij, rij, s.rp, between, arc1_on, arc2_on, tf.adelta, &ts_);
      }
      const uint64_t lm = ballot(free_pt);
      if (m0 == 0) lmw[0] = lm; else lmw[1] = lm;
      tot1 += wave_sum_int(ts_.n1);
      tot2 += wave_sum_int(ts_.n2);
    }
    if (lane() == 0) {
      w.t1cnt[p] = tot1;
      w.t2cnt[p] = tot2;
      w.tflag[p] = (lmw[0] | lmw[1]) ? 1 : 0;
      w.tmask[2 * p] = lmw[0];
      w.tmask[2 * p + 1] = lmw[1];
    }
    __syncthreads();
  }
}

// Pairs with toroidal dots: flag, then a scan and a compaction give plist.
__global__ void k_toroid_dot_pairs(Count nP, const int* __restrict__ t1cnt, const int* __restrict__ t2cnt,
                                   int* __restrict__ flag) {
  const int n = nP.get();
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x)
    flag[p] = (t1cnt[p] | t2cnt[p]) != 0 ? 1 : 0;
}
__global__ void k_compact_count(Count n, const int* __restrict__ flag, const int* __restrict__ scan,
                                int* __restrict__ out) {
  const int c = n.get();
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < c; i += gridDim.x * blockDim.x)
    if (flag[i]) out[scan[i]] = i;
}

// Emit pass of the toroidal stage (_core.py:1940-2061) over the pairs with
// toroidal dots (plist, count at *nlist), TE_PAIRS consecutive pairs per
// wavefront: the free torus points of the batch (the count pass's masks) are
// spread over the lanes, and the dots of their saddle arcs likewise.  Dot k
// of a point's arc goes to the pair's offset plus the arc dots of the pair's
// earlier points plus k -- the order of the reference (pair, torus point,
// arc point).
constexpr int TE_PAIRS = 16;
struct ToroidBatch {
  int i[TE_PAIRS], j[TE_PAIRS], ri[TE_PAIRS], rj[TE_PAIRS], flag[TE_PAIRS], M[TE_PAIRS];
  int tot1[TE_PAIRS], tot2[TE_PAIRS];
  double4 xi[TE_PAIRS], xj[TE_PAIRS], u[TE_PAIRS], t[TE_PAIRS], cx[TE_PAIRS], cy[TE_PAIRS];
  double tdelta[TE_PAIRS], ts[TE_PAIRS], adelta[TE_PAIRS];
  int64_t ob1[TE_PAIRS], ob2[TE_PAIRS];
};

// Dots of one arc per lane (n dots, inclusive scan inc over the lanes; the
// point lies at offset off of its pair a's arc list), spread over the lanes:
// dot d belongs to the first lane src with inc[src] > d.
__device__ __forceinline__ void emit_arc_dots(const BatchView& b, const Settings& s, const Work& w,
                                              const ToroidBatch& tb, int arc, int n, int inc, V3 p, V3 x, V3 y,
                                              double ps, int off, int a, int m) {
  const int T = __shfl(inc, WAVE - 1, WAVE);
  for (int d0 = 0; d0 < T; d0 += WAVE) {
    const int d = d0 + lane();
    int src = 0;
#pragma unroll
    for (int step = WAVE / 2; step > 0; step >>= 1)
      if (__shfl(inc, src + step - 1, WAVE) <= d) src += step;
    const int sn = __shfl(n, src, WAVE);
    const int k = d - (__shfl(inc, src, WAVE) - sn);
    const V3 sp = v3(__shfl(p.x, src, WAVE), __shfl(p.y, src, WAVE), __shfl(p.z, src, WAVE));
    const V3 sx = v3(__shfl(x.x, src, WAVE), __shfl(x.y, src, WAVE), __shfl(x.z, src, WAVE));
    const V3 sy = v3(__shfl(y.x, src, WAVE), __shfl(y.y, src, WAVE), __shfl(y.z, src, WAVE));
    const double sps = __shfl(ps, src, WAVE);
    const int soff = __shfl(off, src, WAVE), sa = __shfl(a, src, WAVE), sm = __shfl(m, src, WAVE);
    if (d < T) {
      const double4 pu = tb.u[sa], pt = tb.t[sa];
      const V3 pt_ = subdiv_point(sp, sx, sy, s.rp, tb.adelta[sa], k);
      const double dist = dist_to_line(ld3(pt), ld3(pu), pt_);
      const double area = sps * tb.ts[sa] * dist / pu.w;
      const int owner = arc == 1 ? tb.i[sa] : tb.j[sa];
      const int rowner = arc == 1 ? tb.ri[sa] : tb.rj[sa];
      const int64_t ob = arc == 1 ? tb.ob1[sa] : tb.ob2[sa];
      store_dot(b, s, w, ob + soff + k, pt_, sp, area, owner, rowner, 2,
                make_int4(arc == 1 ? DK_ARC1 : DK_ARC2, tb.i[sa], tb.j[sa], -1), make_int4(sm, k, tb.M[sa], sn));
    }
  }
}

__global__ __launch_bounds__(64) void k_toroid_emit(BatchView b, Settings s, Work w, const int* __restrict__ plist,
                                                    const int* __restrict__ nlist) {
  __shared__ ToroidBatch tb;
  __shared__ uint16_t items[TE_PAIRS * 2 * WAVE];
  const int nE = *nlist;
  const int nbatch = (nE + TE_PAIRS - 1) / TE_PAIRS;
  CMS_UNIT_LOOP(bt, nbatch) {
    const int e0 = (int)bt * TE_PAIRS;
    const int np = min(TE_PAIRS, nE - e0);
    int lc = 0;
    uint64_t mw0 = 0ull, mw1 = 0ull;
    if (lane() < np) {
      const int a = lane();
      const int p = plist[e0 + a];
      const int i = w.pi[p], j = w.pj[p];
      const double4 ai = b.atom[i], aj = b.atom[j], pu = w.pu[p];
      const TorusFrame tf = torus_frame(ai.w, aj.w, s.rp, pu.w, ld3(pu), (b.dens[i] + b.dens[j]) * 0.5);
      tb.i[a] = i;
      tb.j[a] = j;
      tb.ri[a] = w.rpos[i];
      tb.rj[a] = w.rpos[j];
      tb.flag[a] = ((w.pflag[p] & 1) ? 1 : 0) | (w.atten[i] >= ATTEN_2 ? 2 : 0) | (w.atten[j] >= ATTEN_2 ? 4 : 0);
      tb.M[a] = tf.M;
      tb.tot1[a] = 0;
      tb.tot2[a] = 0;
      tb.xi[a] = ai;
      tb.xj[a] = aj;
      tb.u[a] = pu;
      tb.t[a] = w.pt[p];
      tb.cx[a] = make_double4(tf.cx.x, tf.cx.y, tf.cx.z, 0.0);
      tb.cy[a] = make_double4(tf.cy.x, tf.cy.y, tf.cy.z, 0.0);
      tb.tdelta[a] = tf.tdelta;
      tb.ts[a] = tf.ts;
      tb.adelta[a] = tf.adelta;
      const int grp = group_of_atom(b, i);
      tb.ob1[a] = w.dbase[grp * 4 + 1] + w.t1scan[p];
      tb.ob2[a] = w.dbase[grp * 4 + 2] + w.t2scan[p];
      mw0 = w.tmask[2 * (int64_t)p];
      mw1 = w.tmask[2 * (int64_t)p + 1];
      lc = __popcll(mw0) + __popcll(mw1);
    }
    const int inc = wave_incl_scan(lc);
    const int L = __shfl(inc, WAVE - 1, WAVE);
    if (lane() < np) {
      int o = inc - lc;
      for (; mw0; mw0 &= mw0 - 1) items[o++] = (uint16_t)((lane() << 8) | __builtin_ctzll(mw0));
      for (; mw1; mw1 &= mw1 - 1) items[o++] = (uint16_t)((lane() << 8) | (WAVE + __builtin_ctzll(mw1)));
    }
    __syncthreads();
    for (int q0 = 0; q0 < L; q0 += WAVE) {
      const int q = q0 + lane();
      int a = -1, m = 0;
      TorusSub t{};
      t.n1 = t.n2 = 0;
      if (q < L) {
        const int e = items[q];
        a = e >> 8;
        m = e & 255;
        const double4 pu = tb.u[a], pt = tb.t[a], ai = tb.xi[a], aj = tb.xj[a];
        const int f = tb.flag[a];
        const V3 pm = subdiv_point(ld3(pt), ld3(tb.cx[a]), ld3(tb.cy[a]), pu.w, tb.tdelta[a], m);
        torus_point_arcs(pm, ld3(ai), ld3(aj), ai.w + s.rp, aj.w + s.rp, ld3(pt), ld3(pu), pu.w, s.rp,
                         (f & 1) != 0, (f & 2) != 0, (f & 4) != 0, tb.adelta[a], &t);
      }
      // lanes of one pair are consecutive: per-lane offsets within the pair
      const int aprev = __shfl_up(a, 1, WAVE), anext = __shfl_down(a, 1, WAVE);
      const bool run_start = q < L && (lane() == 0 || aprev != a);
      const bool run_end = q < L && (lane() == WAVE - 1 || q + 1 >= L || anext != a);
      const uint64_t starts = ballot(run_start);
      const uint64_t below = lane() == WAVE - 1 ? ~0ull : ((2ull << lane()) - 1);
      const int first = (starts & below) ? 63 - __builtin_clzll(starts & below) : 0;
      const int inc1 = wave_incl_scan(t.n1), inc2 = wave_incl_scan(t.n2);
      const int b1 = __shfl(inc1 - t.n1, first, WAVE), b2 = __shfl(inc2 - t.n2, first, WAVE);
      int off1 = 0, off2 = 0;
      if (q < L) {
        off1 = tb.tot1[a] + (inc1 - t.n1) - b1;
        off2 = tb.tot2[a] + (inc2 - t.n2) - b2;
      }
      const double ps1 = t.n1 > 0 ? (s.rp * t.ang1) / (double)t.n1 : 0.0;
      const double ps2 = t.n2 > 0 ? (s.rp * t.ang2) / (double)t.n2 : 0.0;
      emit_arc_dots(b, s, w, tb, 1, t.n1, inc1, t.p, t.pi, t.y1, ps1, off1, a, m);
      emit_arc_dots(b, s, w, tb, 2, t.n2, inc2, t.p, t.pqj, t.y2, ps2, off2, a, m);
      __syncthreads();
      if (run_end) {
        tb.tot1[a] += inc1 - b1;
        tb.tot2[a] += inc2 - b2;
      }
      __syncthreads();
    }
  }
}

// ------------------------------------------------------------ concave
// _core.py:2069-2333.  Spherical triangle of each probe: latitude arcs from
// the highest vertex towards -alt, circles, the three plane tests, low-probe
// rejection, and the closest of the three atoms as owner.
constexpr int CC_BLOCK = 512;   // concave candidate points per prefilter block

// The probe's frame (_core.py:2140-2200): the three plane normals vec0-2 of
// the spherical triangle, the highest vertex top (index mm) with the
// latitude arc's second axis y, spacing delta, count L and step length cs,
// the latitude circles' frame (cx, cy) about south = -alt, and the mean
// density of the three atoms.
struct ConcaveFrame {
  V3 vec0, vec1, vec2, top, y, cx, cy;
  double delta, cs, dens;
  int L, mm;
};
__device__ __forceinline__ ConcaveFrame concave_frame(V3 p, V3 u, double4 A0, double4 A1, double4 A2, double dens,
                                                      double rp) {
  ConcaveFrame f;
  const V3 vp0 = unit(ld3(A0) - p), vp1 = unit(ld3(A1) - p), vp2 = unit(ld3(A2) - p);
  f.vec0 = unit(cross(vp0, vp1));
  f.vec1 = unit(cross(vp1, vp2));
  f.vec2 = unit(cross(vp2, vp0));
  int mm = 0;
  {
    // the first of the largest (strict >), as in a loop over the atoms
    double best = dot_seq(u, vp0);
    const double d1 = dot_seq(u, vp1);
    if (d1 > best) { best = d1; mm = 1; }
    if (dot_seq(u, vp2) > best) mm = 2;
  }
  f.mm = mm;
  f.top = mm == 0 ? vp0 : (mm == 1 ? vp1 : vp2);
  const V3 south = neg(u);
  const V3 axis = unit(cross(f.top, south));
  const double ang = arc_angle(f.top, south, axis, &f.y);
  f.delta = 1.0 / (sqrt(dens) * rp);
  f.L = subdiv_count(f.delta, ang);
  f.cs = f.L > 0 ? (rp * ang) / (double)f.L : 0.0;
  f.dens = dens;
  circle_frame(south, &f.cx, &f.cy);
  return f;
}

// Frames of every probe, one thread per probe (the concave passes, one
// wavefront per probe, would otherwise compute the same frame on every
// lane).  Seven double4 per probe: (vec0, delta), (vec1, cs), (vec2, dens),
// (top, L), (y, mm), (cx, 0), (cy, 0).
__global__ __launch_bounds__(256) void k_concave_frames(BatchView b, Settings s, Work w) {
  const int nQ = *w.nQ;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nQ; q += gridDim.x * blockDim.x) {
    const int4 at = w.patom[q];
    const double dens = ((b.dens[at.x] + b.dens[at.y]) + b.dens[at.z]) / 3.0;
    const ConcaveFrame f = concave_frame(ld3(w.ppos[q]), ld3(w.palt[q]), b.atom[at.x], b.atom[at.y], b.atom[at.z],
                                         dens, s.rp);
    double4* o = w.ccf + 7 * (int64_t)q;
    o[0] = make_double4(f.vec0.x, f.vec0.y, f.vec0.z, f.delta);
    o[1] = make_double4(f.vec1.x, f.vec1.y, f.vec1.z, f.cs);
    o[2] = make_double4(f.vec2.x, f.vec2.y, f.vec2.z, f.dens);
    o[3] = make_double4(f.top.x, f.top.y, f.top.z, (double)f.L);
    o[4] = make_double4(f.y.x, f.y.y, f.y.z, (double)f.mm);
    o[5] = make_double4(f.cx.x, f.cx.y, f.cx.z, 0.0);
    o[6] = make_double4(f.cy.x, f.cy.y, f.cy.z, 0.0);
  }
}

// The count pass is held to 168 VGPRs for three wavefronts per SIMD: the
// few spilled values cost less than the lost occupancy (not so in the emit
// pass).
template <bool EMIT>
__global__ __launch_bounds__(64, EMIT ? 2 : 3) void k_concave(BatchView b, Settings s, Work w) {
  __shared__ uint16_t clist[CC_BLOCK];
  __shared__ double lat_cx[MAX_SUBDIV], lat_cy[MAX_SUBDIV], lat_cz[MAX_SUBDIV];
  __shared__ double lat_rad[MAX_SUBDIV], lat_delta[MAX_SUBDIV], lat_ps[MAX_SUBDIV];
  __shared__ int lat_start[MAX_SUBDIV + 1];
  __shared__ double4 lnear[NEARCAP];
  const int nQ = *w.nQ;
  const double rp = s.rp, rp2 = s.rp * s.rp;
  // count pass: every probe, its first KSLOTS dots computed into the
  // probe's staging slots (placed by k_concave_place after the scan); emit
  // pass: only probes with more dots than that
  auto live = [&](int64_t u) { return !EMIT || w.kcnt[u] > KSLOTS; };
  CMS_SPARSE_LOOP(q, nQ, live) {
    const int4 at = w.patom[q];
    const double4 pq = w.ppos[q];
    const V3 p = ld3(pq);
    const double h = pq.w;
    const V3 u = ld3(w.palt[q]);
    // the three atoms in named registers: indexing arrays of them by the
    // runtime owner choice would put them in scratch memory
    const double4 A0 = b.atom[at.x], A1 = b.atom[at.y], A2 = b.atom[at.z];
    // the probe's frame from k_concave_frames
    const double4* fr = w.ccf + 7 * (int64_t)q;
    const double4 f0 = fr[0], f1 = fr[1], f2 = fr[2], f3 = fr[3], f4 = fr[4], f5 = fr[5], f6 = fr[6];
    const V3 vec0 = ld3(f0), vec1 = ld3(f1), vec2 = ld3(f2), top = ld3(f3), y = ld3(f4), cx = ld3(f5),
             cy = ld3(f6);
    const double delta = f0.w, cs = f1.w, dens = f2.w;
    const int L = (int)f3.w, mm = (int)f4.w;
    const V3 south = neg(u);
    const int nnear = (h < rp) ? w.nnear[q] : 0;
    for (int t = lane(); t < nnear; t += WAVE) lnear[t] = w.ppos[w.near[(int64_t)q * NEARCAP + t]];
    int run = 0;
    for (int t0 = 0; t0 < L; t0 += WAVE) {
      const int t = t0 + lane();
      int P = 0;
      if (t < L) {
        const V3 lat = subdiv_point(v3(0.0, 0.0, 0.0), top, y, rp, delta, t);
        const double dtl = dot_seq(lat, south);
        const V3 cen = south * dtl;
        const double rsq = rp2 - dtl * dtl;
        const double rad = sqrt(fmax(rsq, 0.0));
        double cdelta = 0.0, ps = 0.0;
        if (rad > 0.0) {
          cdelta = 1.0 / (sqrt(dens) * rad);
          P = subdiv_count(cdelta, TWO_PI);
          ps = P > 0 ? (rad * TWO_PI) / (double)P : 0.0;
        }
        lat_cx[t] = cen.x; lat_cy[t] = cen.y; lat_cz[t] = cen.z;
        lat_rad[t] = rad; lat_delta[t] = cdelta; lat_ps[t] = ps;
      }
      const int inc = wave_incl_scan(P);
      if (t < L) lat_start[t] = run + inc - P;
      run += __shfl(inc, WAVE - 1, WAVE);
    }
    if (lane() == 0) lat_start[L] = run;
    __syncthreads();
    const int C = run;
    int64_t obase = 0;
    if (EMIT) {
      const int grp = group_of_atom(b, at.x);
      obase = w.dbase[grp * 4 + 3] + w.kscan[q];
    }
    int total = 0;
    // candidates in blocks of CC_BLOCK: a float pass drops the points that
    // lie clearly outside one of the three planes, the rest are listed in
    // order in LDS and go through the exact tests.  The float point differs
    // from the computed one by at most rad * (|dcos| + |dsin|) <= 1.3e-6 rp
    // (float rounding of the angle and float cos / sin), far inside the
    // margin 1e-4 (1 + rp), so every dropped point fails a plane test.
    const double pf = 1e-4 * (1.0 + rp);
    for (int c0 = 0; c0 < C; c0 += CC_BLOCK) {
      const int c1 = min(C, c0 + CC_BLOCK);
      int ns = 0;
      for (int q0 = c0; q0 < c1; q0 += WAVE) {
        const int qq = q0 + lane();
        bool cand = false;
        if (qq < c1) {
          const int t = last_le(lat_start, L, qq);   // lat_start[t] <= qq < lat_start[t + 1]
          const double ang = lat_delta[t] * ((double)(qq - lat_start[t]) + 0.5);
          float sa, ca;
          sincosf((float)ang, &sa, &ca);
          const V3 rel = subdiv_point_cs(v3(lat_cx[t], lat_cy[t], lat_cz[t]), cx, cy, lat_rad[t], (double)ca,
                                         (double)sa);
          cand = !(dot_seq(rel, vec0) > pf || dot_seq(rel, vec1) > pf || dot_seq(rel, vec2) > pf);
        }
        const uint64_t cm = ballot(cand);
        if (cand) clist[ns + lanes_below(cm)] = (uint16_t)(qq - c0);
        ns += __popcll(cm);
      }
      __syncthreads();
      for (int s0 = 0; s0 < ns; s0 += WAVE) {
        const int si = s0 + lane();
        bool keep = false;
        V3 pts{0, 0, 0};
        int t = 0, qq = 0;
        if (si < ns) {
          qq = c0 + clist[si];
          t = last_le(lat_start, L, qq);   // lat_start[t] <= qq < lat_start[t + 1]
          const int kk = qq - lat_start[t];
          const V3 rel = subdiv_point(v3(lat_cx[t], lat_cy[t], lat_cz[t]), cx, cy, lat_rad[t],
                                      lat_delta[t], kk);
          const bool bail = dot_seq(rel, vec0) >= 0.0 || dot_seq(rel, vec1) >= 0.0 ||
                            dot_seq(rel, vec2) >= 0.0;
          if (!bail) {
            pts = rel + p;
            keep = true;
            for (int l = 0; l < nnear; ++l) {
              const V3 d = pts - ld3(lnear[l]);
              if (dot_alt(d, d) < rp2) { keep = false; break; }
            }
          }
        }
        const uint64_t mk = ballot(keep);
        const int k = total + lanes_below(mk);
        if (keep && (EMIT || k < KSLOTS)) {
          int mc = 0;
          {
            // the first of the closest (strict <)
            double dbest = norm_seq(pts - ld3(A0)) - A0.w;
            const double d1 = norm_seq(pts - ld3(A1)) - A1.w;
            if (d1 < dbest) { dbest = d1; mc = 1; }
            if (norm_seq(pts - ld3(A2)) - A2.w < dbest) mc = 2;
          }
          const int owner = mc == 0 ? at.x : (mc == 1 ? at.y : at.z);
          const DotRec d = make_dot(b, s, w, pts, p, lat_ps[t] * cs, owner, w.rpos[owner], 3,
                                    make_int4(DK_CONCAVE | (mm << 4), at.x, at.y, at.z),
                                    make_int4(t, qq - lat_start[t], L, lat_start[t + 1] - lat_start[t]));
          if (EMIT) {
            write_dot(w, obase + k, d);
          } else {
            w.kpos[q * KSLOTS + k] = d.pos;
            w.kinfo[q * KSLOTS + k] =
                (uint64_t)mc | ((uint64_t)(d.bur != 0) << 2) | ((uint64_t)mm << 3) | ((uint64_t)t << 8) |
                ((uint64_t)(qq - lat_start[t]) << 16) | ((uint64_t)L << 24) |
                ((uint64_t)(lat_start[t + 1] - lat_start[t]) << 32);
          }
        }
        total += __popcll(mk);
      }
      __syncthreads();
    }
    if (!EMIT && lane() == 0) w.kcnt[q] = total;
    __syncthreads();
  }
}

// explicit instantiations used by the host pipeline
template __global__ void k_probes<false>(BatchView, Settings, Work);
template __global__ void k_probes<true>(BatchView, Settings, Work);
template __global__ void k_convex<false, RC_SMALL>(BatchView, Settings, Work);
template __global__ void k_convex<true, RC_SMALL>(BatchView, Settings, Work);
template __global__ void k_convex<false, KCAP>(BatchView, Settings, Work);
template __global__ void k_convex<true, KCAP>(BatchView, Settings, Work);
template __global__ void k_concave<false>(BatchView, Settings, Work);
template __global__ void k_concave<true>(BatchView, Settings, Work);
template __global__ void k_probe_rows<RC_SMALL>(BatchView, Settings, Work);
template __global__ void k_probe_rows<KCAP>(BatchView, Settings, Work);

// Staged concave dots to their place in the dot arrays: one thread per
// staging slot of probes with at most KSLOTS dots.
__global__ void k_concave_place(BatchView b, Settings s, Work w) {
  const int nQ = *w.nQ;
  const int64_t n = (int64_t)nQ * KSLOTS;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int q = (int)(i / KSLOTS), k = (int)(i % KSLOTS);
    const int kc = w.kcnt[q];
    if (kc > KSLOTS || k >= kc) continue;
    const int4 at = w.patom[q];
    const int grp = group_of_atom(b, at.x);
    const uint64_t f = w.kinfo[i];
    const int mc = (int)(f & 3), mm = (int)((f >> 3) & 3);
    // the fields of make_dot: normal (pc - pt) / rp towards the probe
    // centre (concave dots exist only for rp > 0), descriptors, owner
    DotRec d;
    d.pos = w.kpos[i];
    const V3 pt = ld3(d.pos), pc = ld3(w.ppos[q]);
    const V3 nml = (pc - pt) / s.rp;
    d.nml = make_double4(nml.x, nml.y, nml.z, 0.0);
    d.desc = make_int4(DK_CONCAVE | (mm << 4), at.x, at.y, at.z);
    d.desc2 = make_int4((int)((f >> 8) & 255), (int)((f >> 16) & 255), (int)((f >> 24) & 255),
                        (int)((f >> 32) & 255));
    d.atom = mc == 0 ? at.x : mc == 1 ? at.y : at.z;
    d.bur = (int8_t)((f >> 2) & 1);
    d.type = 3;
    write_dot(w, w.dbase[grp * 4 + 3] + w.kscan[q] + k, d);
  }
}

}  // namespace cms
