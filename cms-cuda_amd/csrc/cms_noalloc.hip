// cms_nearest_noalloc: the nearest-point search of cms_nearest without any
// device allocation, for callers that cannot allocate (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc).
//
// cms_nearest keeps its working buffers (packed double4 copies of the
// points, their float copies for the FP32 search, the split partial results,
// the group / tile offsets, the float frame and the rescoring counter) in the
// context's grow-only workspace.  Here they live in one file-scope device
// array, `scratch`, of SCRATCH_BYTES: every GPU the library runs on holds its
// own copy until the process ends.  Host code never passes the array's name
// to a kernel (on the host it is only a stand-in symbol); cms_create looks up
// the device address once per device (noalloc_setup, with that device
// current) and the entry point uses that address.
//
// Limits: at most NOALLOC_MAX_POINTS query and NOALLOC_MAX_POINTS reference
// points per call.  Every call on one GPU shares that GPU's scratch and
// leaves its contents behind; each call initialises what it reads (offsets,
// counter) with a kernel, not a memset.  Two calls that overlap in time on
// one GPU (different streams or host threads) race on the scratch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "cms_hip.h"
#include "cms_internal.h"
#include "cms_readout.h"
#include "cms_util.h"

namespace cms {

constexpr int64_t NOALLOC_MAX_PARTIAL = 786432;     // split partials: nsplit * nq
constexpr int64_t MIB = 1 << 20;
// byte offsets inside the scratch array
constexpr int64_t OFF_Q = 0;                                   // double4 queries
constexpr int64_t OFF_R = OFF_Q + 32 * NOALLOC_MAX_POINTS;     // double4 references
constexpr int64_t OFF_Q32 = OFF_R + 32 * NOALLOC_MAX_POINTS;   // float4 queries
constexpr int64_t OFF_R32 = OFF_Q32 + 16 * NOALLOC_MAX_POINTS; // float4 references
constexpr int64_t OFF_PD2 = OFF_R32 + 16 * NOALLOC_MAX_POINTS; // double partial d2
constexpr int64_t OFF_PIDX = OFF_PD2 + 8 * NOALLOC_MAX_PARTIAL; // int partial index
constexpr int64_t OFF_MISC = OFF_PIDX + 4 * NOALLOC_MAX_PARTIAL;
constexpr int64_t OFF_GRP = OFF_MISC;          // int[4]: query / reference group bounds
constexpr int64_t OFF_TOFF = OFF_MISC + 256;   // int[2]: tile offsets
constexpr int64_t OFF_FRAME = OFF_MISC + 512;  // double4[2]: float frame
constexpr int64_t OFF_NRES = OFF_MISC + 768;   // u64[2]: rescoring counter
constexpr int64_t SCRATCH_BYTES = OFF_MISC + 4096;
static_assert(SCRATCH_BYTES == 21 * MIB + 4096, "scratch layout: 21 MiB + 4 KiB");
static_assert(OFF_PIDX % 256 == 0 && OFF_MISC % 256 == 0, "256-byte aligned regions");

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
unsigned char* g_scratch[MAX_DEVICES] = {};

__global__ void k_noalloc_setup(int nq, int nr, int tq, int* grp, int* toff, unsigned long long* nres) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    grp[0] = 0;
    grp[1] = nq;
    grp[2] = 0;
    grp[3] = nr;
    toff[0] = 0;
    toff[1] = (nq + tq - 1) / tq;
    nres[0] = 0ull;
    nres[1] = 0ull;
  }
}

inline int blocks(int64_t n, int per_block, int cap) {
  int64_t g = (n + per_block - 1) / per_block;
  return (int)std::min<int64_t>(std::max<int64_t>(g, 1), cap);
}
}  // namespace

// Device address of this device's scratch; called by cms_create with the
// device current, once per device (later calls return the kept address).
int noalloc_setup(int device, std::string* err) {
  if (device < 0 || device >= MAX_DEVICES) {
    *err = "cms_create: device index beyond the scratch table";
    return CMS_ERR_INVALID;
  }
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_scratch[device]) return CMS_OK;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    *err = std::string("cms_create: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    return CMS_ERR_HIP;
  }
  g_scratch[device] = static_cast<unsigned char*>(p);
  return CMS_OK;
}

static unsigned char* scratch_of(int device) {
  if (device < 0 || device >= MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lock(g_mu);
  return g_scratch[device];
}

// The launches of cms_nearest with every buffer in the scratch: the same
// kernels, tile sizes and split rule (the split count only bounded so the
// partials fit), hence the same results bit for bit.
int nearest_noalloc(int device, hipStream_t st, int precision, const double* q, int64_t nq, const double* r,
                    int64_t nr, double* out_d2, int32_t* out_idx, std::string* err) {
  unsigned char* base = scratch_of(device);
  if (!base) {
    *err = "cms_nearest_noalloc: no scratch address for this device (context not created on it)";
    return CMS_ERR_STATE;
  }
  double4* qp = reinterpret_cast<double4*>(base + OFF_Q);
  double4* rp = reinterpret_cast<double4*>(base + OFF_R);
  float4* q32 = reinterpret_cast<float4*>(base + OFF_Q32);
  float4* r32 = reinterpret_cast<float4*>(base + OFF_R32);
  double* pd2 = reinterpret_cast<double*>(base + OFF_PD2);
  int* pidx = reinterpret_cast<int*>(base + OFF_PIDX);
  int* grp = reinterpret_cast<int*>(base + OFF_GRP);
  int* toff = reinterpret_cast<int*>(base + OFF_TOFF);
  double4* frame = reinterpret_cast<double4*>(base + OFF_FRAME);
  unsigned long long* nres = reinterpret_cast<unsigned long long*>(base + OFF_NRES);

  const bool fp32 = precision == CMS_PREC_FP32_RESCORE;
  const int tq = fp32 ? NT32 : NT64;
  k_noalloc_setup<<<1, 64, 0, st>>>((int)nq, (int)nr, tq, grp, toff, nres);
  k_pack3<<<blocks(nq, 256, 4096), 256, 0, st>>>((int)nq, q, qp);
  if (nr > 0) k_pack3<<<blocks(nr, 256, 4096), 256, 0, st>>>((int)nr, r, rp);
  if (fp32) {
    k_points_frame<<<1, 256, 0, st>>>((int)nq, qp, (int)nr, rp, frame);
    k_to_float<<<blocks(nq, 256, 4096), 256, 0, st>>>(host_count((int)nq), qp, nullptr, 1, 0, frame, q32);
    if (nr > 0)
      k_to_float<<<blocks(nr, 256, 4096), 256, 0, st>>>(host_count((int)nr), rp, nullptr, 1, 0, frame, r32);
  }
  // split rule of cms_nearest for one group, bounded by the partial space
  const int64_t tiles_est = nq / tq + 1;
  int nsplit = 1;
  if (tiles_est < 2048) {
    nsplit = (int)std::min<int64_t>((2048 + tiles_est - 1) / tiles_est, 64);
    nsplit = (int)std::max<int64_t>(1, std::min<int64_t>(nsplit, std::max<int64_t>(1, nr) / 128));
  }
  nsplit = (int)std::max<int64_t>(1, std::min<int64_t>(nsplit, NOALLOC_MAX_PARTIAL / std::max<int64_t>(nq, 1)));
  const int grid = blocks(tiles_est * nsplit, 1, 16384);
  const int64_t ps = nsplit == 1 ? 0 : nq;
  double* d2o = nsplit > 1 ? pd2 : out_d2;
  int* ido = out_idx ? (nsplit > 1 ? pidx : out_idx) : nullptr;
  if (fp32) {
    if (ido)
      k_nearest32<true><<<grid, TPB32, 0, st>>>(q32, qp, grp, r32, rp, grp + 2, 0, toff, 1, nsplit, ps, frame, d2o,
                                                ido, nres);
    else
      k_nearest32<false><<<grid, TPB32, 0, st>>>(q32, qp, grp, r32, rp, grp + 2, 0, toff, 1, nsplit, ps, frame,
                                                 d2o, nullptr, nres);
  } else {
    if (ido)
      k_nearest64<true><<<grid, NT64, 0, st>>>(qp, grp, rp, grp + 2, 0, toff, 1, nsplit, ps, d2o, ido);
    else
      k_nearest64<false><<<grid, NT64, 0, st>>>(qp, grp, rp, grp + 2, 0, toff, 1, nsplit, ps, d2o, nullptr);
  }
  if (nsplit > 1)
    k_nearest_combine<<<blocks(nq, 256, 4096), 256, 0, st>>>(host_count((int)nq), nsplit, ps, d2o, ido, out_d2,
                                                             out_idx);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    *err = std::string("cms_nearest_noalloc: kernel launch: ") + hipGetErrorString(e);
    return CMS_ERR_HIP;
  }
  return CMS_OK;
}

}  // namespace cms
