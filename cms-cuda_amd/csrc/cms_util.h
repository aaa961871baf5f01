#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <initializer_list>
#include <string>

#include "cms_internal.h"

namespace cms {

// Exclusive scan of in[0..n) into out[0..n], out[n] = total.  part needs
// 1025 ints of scratch.
void scan_excl(const int* in, int* out, int64_t n, int* part, hipStream_t st);
// Same with the count read on the device; cap bounds it and picks the grid.
void scan_excl(const int* in, int* out, Count n, int64_t cap, int* part, hipStream_t st);
__global__ void k_stage_total(const int* scan, Count at, int64_t cap, int* meta, int slot, int* grp, int G);
__global__ void k_list_totals(const int* bscan, int* meta);
__global__ void k_copy_word(const int* src, Count at, int* dst);
void fill_int(int* p, int64_t n, int v, hipStream_t st);
void fill_u64(unsigned long long* p, int64_t n, unsigned long long v, hipStream_t st);
void fill_int8(int8_t* p, int64_t n, int8_t v, hipStream_t st);
// fill of a device-counted range; cap bounds the count and sizes the grid
void fill_int(int* p, Count n, int64_t cap, int v, hipStream_t st);
void fill_double(double* p, int64_t n, double v, hipStream_t st);
// Several fills in one launch: each entry sets p[0 .. n) to v, an element
// of `bytes` bytes (1, 4 or 8); n is read on the device when it is a device
// count, cap bounds it (and sizes the grid).
struct FillSpec {
  void* p;
  Count n;
  int64_t cap;
  unsigned long long v;
  int bytes;
};
constexpr int FILL_BATCH = 6;
void fill_batch(std::initializer_list<FillSpec> fills, hipStream_t st);
inline FillSpec fill_of(int* p, int64_t n, int v) { return FillSpec{p, host_count((int)n), n, (unsigned)v, 4}; }
inline FillSpec fill_of(int* p, Count n, int64_t cap, int v) { return FillSpec{p, n, cap, (unsigned)v, 4}; }
inline FillSpec fill_of(int8_t* p, int64_t n, int8_t v) { return FillSpec{p, host_count((int)n), n, (uint8_t)v, 1}; }
inline FillSpec fill_of(unsigned long long* p, int64_t n, unsigned long long v) {
  return FillSpec{p, host_count((int)n), n, v, 8};
}

__global__ void k_pack_atoms(int ntot, const double* xyz, const double* rad, const double* dens_in,
                             double dens_default, const int* aoff, const int* nb0, int B, int full,
                             double4* atom, double* dens, int* acplx, int8_t* amol);
__global__ void k_group_starts_from_atoms(Count n, const int* items, const int* acplx,
                                          const int8_t* amol, int G, int* start);
__global__ void k_gather_scan(int G, const int* idx, int mul, const int* scan, int* out);
__global__ void k_pack3(int n, const double* xyz, double4* out);
__global__ void k_points_frame(int nq, const double4* q, int nr, const double4* r, double4* frame);
constexpr int DOT_BASES_THREADS = 1024;   // k_dot_bases runs as one workgroup of this size
__global__ void k_dot_stage_starts(int G, const int* rgrp, const int* pgrp, const int* prgrp, const int* cscan,
                                   const int* t1scan, const int* t2scan, const int* kscan, int4* ss);
__global__ void k_dot_bases(int G, const int4* ss, int64_t* dbase, int* dgrp);

__global__ void k_drop_failed_rows(Count cnt, const int* ratom, const int* acplx, const int* status, int* nn,
                                   int* nb, int* pcnt);
__global__ void k_zero_failed(int B, const int* status, const int* aoff, double* per_atom, int width,
                              double* totals, int tstride, double* sc, int sc_width);

// allocation-free nearest search (cms_noalloc.hip)
constexpr int64_t NOALLOC_MAX_POINTS = 131072;
int noalloc_setup(int device, std::string* err);
int nearest_noalloc(int device, hipStream_t st, int precision, const double* q, int64_t nq, const double* r,
                    int64_t nr, double* out_d2, int32_t* out_idx, std::string* err);

}  // namespace cms
