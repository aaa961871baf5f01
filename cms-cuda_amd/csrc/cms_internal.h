// Internal data layout shared by the kernels and the host pipeline.
//
// HBM layout (one batch of B complexes, G = 2B molecule groups, group
// g = 2*complex + molecule):
//   atoms   : double4 (x, y, z, radius) per atom, complex-major, binder
//             (molecule 0) atoms before target (molecule 1) atoms
//   region  : ratom[nR] = atoms that need neighbour rows (ascending, hence
//             grouped by g), rpos[atom] = its row index
//   rows    : nbr[nR][KCAP], bur[nR][BCAP] global atom indices
//   pairs   : pi/pj int32, pu double4 (u, rij), pt double4 (t, -), pflag
//   probes  : patom int4 (a0, a1, a2, -), ppos double4 (p, height), palt
//             double4 (alt, -), grouped by g, "+" probes before "-" probes
//   dots    : dpos double4 (x, y, z, area), dnml double4 (normal, -),
//             dbur/dtype int8, datom int32, grouped by g, each group in
//             emission order convex | toroidal atom-1 arcs | toroidal
//             atom-2 arcs | concave
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cms {

// Item count of a launch: a host value, or a device word (times mul) that
// an earlier kernel of the same stream wrote.  Kernels size their loops with
// get(); the host sizes the grid from a capacity that bounds the count.
struct Count {
  int n;
  const int* p;
  int mul;
  __host__ __device__ __forceinline__ int get() const { return p ? *p * mul : n; }
};
inline Count host_count(int n) { return Count{n, nullptr, 1}; }
inline Count dev_count(const int* p, int mul = 1) { return Count{0, p, mul}; }

// Device words of the per-run count block ("meta"): clamped stage counts
// that kernels loop over, the raw totals, and the overflow bits (one per
// stage whose total exceeded the buffer capacity it was enqueued with).
// M_TE: the number of pairs with toroidal dots (a subset of the pairs, so it
// needs no capacity of its own).
constexpr int M_R = 0, M_P = 1, M_Q = 2, M_D = 3, M_B = 4, M_A = 5, M_T = 6, M_TE = 7;
constexpr int M_RAW = 8, M_OVF = 15, M_WORDS = 16;

struct Settings {
  double rp, band, sep, weight, binwidth_dist, binwidth_norm;
};

struct BatchView {
  const double4* atom;   // ntot
  const double* dens;    // ntot
  const int32_t* acplx;  // ntot complex id
  const int8_t* amol;    // ntot molecule (0 binder, 1 target)
  const int32_t* aoff;   // B+1 atom offsets
  int ntot;
  int B;
};

// One dot's output record (the fields of the dot arrays).
struct DotRec {
  double4 pos;    // x, y, z, area
  double4 nml;    // outward normal, 0
  int4 desc;
  int4 desc2;
  int atom;
  int8_t bur;
  int8_t type;
};
// Concave dots computed by the count pass per probe before the scan places
// them (probes with more are emitted again after the scan).  A staged dot
// is its position and area (kpos) and one word (kinfo) with what the
// placement needs besides the probe: owner choice (bits 0-1), burial (2),
// top vertex mm (3-4), latitude t (8-15), point on the latitude (16-23),
// latitude count L (24-31) and the latitude's point count (32-39).
constexpr int KSLOTS = 64;

// Surviving probe candidate staged by the per-row probe pass (large
// batches) for k_probe_place: both probe positions and the height, the
// unit normal of the triple, pair, third atom, its row slot and which signs
// survive (bit 0 "+", bit 1 "-").  PSLOTS per neighbour row; a pair with a
// survivor past them is left to the probe emit pass (pflag bit 8).
struct ProbeRec {
  double4 pp;   // + probe, height
  double4 pm;   // - probe, 0
  double4 u;    // uijk, 0
  int p, k, sl, f;
};
constexpr int PSLOTS = 8;

struct Work {
  // per atom
  int8_t* atten;
  int8_t* access;
  int* status;            // B
  // region
  const int* nR;          // device pointer to the region size
  int* ratom;
  int* rpos;              // exclusive scan of the region flag (ntot + 1)
  int* nn;
  int* nbr;
  int* nb;
  int* bur;
  int* nbmax;             // largest non-blocker neighbour index per row (-1: none)
  uint8_t* fwd;            // per row: slots of the neighbours with index > the row atom, by index
  // pairs
  int* pcnt;
  int* pscan;
  const int* nP;
  int* pi;
  int* pj;
  double4* pu;
  double4* pt;
  int8_t* pflag;          // bit0 between, bit1 toroid queue
  int* pgrp;              // G+1 pair offsets per group
  // probes
  int* prcnt;             // 2*nP, virtual (group, sign, pair) order
  int* prscan;
  const int* nQ;
  int4* patom;
  double4* ppos;
  double4* palt;
  int* prgrp;             // G+1 probe offsets per group
  int* near;
  int* nnear;
  ProbeRec* pstage;       // PSLOTS staged probe candidates per row (large batches)
  int* nps;               // per row: slots used
  int* elist;             // per-row pass: pairs left to the emit pass (null: by flag)
  int* nelist;            // their number
  int pslots;             // slots per row in use (<= PSLOTS)
  // dot counts
  int* ccnt;
  double4* cvf;           // convex frames, 4 per row (k_convex_frames)
  int* cscan;
  uint64_t* cmask;        // count pass: kept convex points, CMASK_WORDS words per row
  int* t1cnt;
  int* t1scan;
  int* t2cnt;
  int* t2scan;
  int8_t* tflag;
  uint64_t* tmask;        // count pass: free torus points of each pair (two words)
  int* kcnt;
  double4* ccf;           // concave frames, 7 per probe (k_concave_frames)
  double4* kpos;          // KSLOTS staged concave dots per probe
  uint64_t* kinfo;
  int* kscan;
  int64_t* dbase;         // G*4 stage bases
  // dots
  double4* dpos;
  double4* dnml;
  int8_t* dbur;
  int8_t* dtype;
  int* datom;
  int4* ddesc;            // (kind | mm << 4, atom a, atom b, atom c)
  int4* ddesc2;           // (latitude or torus point, point, count 1, count 2)
  // buffer capacities: writes at or beyond them are dropped (the stage
  // total then exceeds the capacity and the run is repeated, larger)
  int64_t pair_cap;
  int64_t probe_cap;
  int64_t dot_cap;
};

__global__ void k_pairs(BatchView, Settings, Work);
template <bool EMIT> __global__ void k_probes(BatchView, Settings, Work);
// LDS row capacity of the common instantiation of the per-row kernels
constexpr int RC_SMALL = 128;
template <int RC> __global__ void k_probe_rows(BatchView, Settings, Work);
__global__ void k_probe_place(BatchView, Work);
template <bool EMIT, int RC> __global__ void k_convex(BatchView, Settings, Work);
__global__ void k_convex_frames(BatchView, Settings, Work);
__global__ void k_toroid_count(BatchView, Settings, Work);
__global__ void k_toroid_dot_pairs(Count, const int*, const int*, int*);
__global__ void k_compact_count(Count, const int*, const int*, int*);
__global__ void k_toroid_emit(BatchView, Settings, Work, const int*, const int*);
template <bool EMIT> __global__ void k_concave(BatchView, Settings, Work);
__global__ void k_concave_frames(BatchView, Settings, Work);
__global__ void k_concave_place(BatchView, Settings, Work);
__global__ void k_compact_index(int, const int*, const int*, int*);
__global__ void k_near(BatchView, Settings, Work);

}  // namespace cms
