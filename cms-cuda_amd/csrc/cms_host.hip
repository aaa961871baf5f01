// Host pipeline and C ABI of libcms_hip.so.
//
// A context owns a stream, the loaded batch and a grow-only workspace.
// Surface generation issues ~40 kernels; the host waits on the stream only
// where a total decides a buffer size (region atoms, pairs, probes, dots)
// and once at the end to read the per-complex status words.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cms_hip.h"
#include "cms_backward.h"
#include "cms_cells.h"
#include "cms_device.h"
#include "cms_internal.h"
#include "cms_readout.h"
#include "cms_util.h"

using namespace cms;

namespace {

struct DBuf {
  void* p = nullptr;
  size_t cap = 0;
  hipError_t ensure(size_t bytes) {
    if (bytes <= cap && p) return hipSuccess;
    if (p) {
      hipError_t e = hipFree(p);
      if (e != hipSuccess) return e;
      p = nullptr;
      cap = 0;
    }
    size_t want = std::max<size_t>(bytes + bytes / 4, 256);
    hipError_t e = hipMalloc(&p, want);
    if (e == hipSuccess) cap = want;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
};

inline int grid_for(int64_t n, int per_block, int cap = 16384) {
  int64_t g = (n + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}

}  // namespace

struct cms_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  // batch
  int B = 0, ntot = 0;
  std::vector<int> h_aoff, h_nb0, h_agrp;
  Settings s{};
  double density = 15.0;
  DBuf atom, dens, acplx, amol, aoff, nb0, agrp;
  // surface state
  bool generated = false;
  int gen_all_atoms = -1;
  int nR = 0, nP = 0, nQ = 0, nD = 0;
  std::vector<int> h_dgrp, h_prgrp;
  DBuf atten, access, status, coinc, flag, nbscan, nbl, need, rpos, ratom, part;
  DBuf grid, gframe, ncellc, ncscan, acell, ccount, cstart, ccur, catom;
  DBuf nn, nbr, nb, bur, nbmax;
  DBuf pcnt, pscan, pi, pj, pu, pt, pflag, rgrp, pgrp;
  DBuf prcnt, prscan, patom, ppos, palt, prgrp, nearl, nnear;
  DBuf ccnt, cscan, t1cnt, t1scan, t2cnt, t2scan, tflag, kcnt, kscan, dbase, dgrp;
  DBuf dpos, dnml, dbur, dtype, datom, ddesc, ddesc2;
  // readout state
  bool lists_ready = false;
  int nB = 0, nA = 0;
  DBuf bflag, bscan, blist, bxyz, bgrp, aflag, ascan, alist, axyz, agrpd;
  DBuf tcnt, toff, bmin, wv, ocnt, ostart, ocur, operm;
  DBuf bacc, trflag, trscan, tlist, txyz, tnml, tgrp, tmin, tidx, tgof, tdist, trr, side;
  DBuf qpack, rpack, grp1, gtot, npd2, npidx;
  DBuf frame, bxyz32, axyz32, txyz32, qpack32, rpack32, nresc;
  DBuf bidx, bgof, adj, eatom, evpos, chk;
  bool cms_ready = false;
  bool float_lists = false;
  // kernel timing with HIP events on the context's stream
  std::string prof_filter;  // "" = off, "*" = every instrumented kernel
  struct Ev { std::string name; hipEvent_t a, b; };
  std::vector<Ev> prof_events;
};

// Records a HIP event pair around one kernel launch when profiling is on
// for that kernel name.
struct ProfScope {
  cms_ctx* ctx;
  bool on = false;
  hipEvent_t a{}, b{};
  const char* name;
  ProfScope(cms_ctx* c, const char* n) : ctx(c), name(n) {
    if (c->prof_filter.empty()) return;
    if (c->prof_filter != "*" && c->prof_filter != n) return;
    if (hipEventCreate(&a) != hipSuccess) return;
    if (hipEventCreate(&b) != hipSuccess) { (void)hipEventDestroy(a); return; }
    on = hipEventRecord(a, c->stream) == hipSuccess;
  }
  ~ProfScope() {
    if (!on) return;
    (void)hipEventRecord(b, ctx->stream);
    ctx->prof_events.push_back({name, a, b});
  }
};
#define TIMED(name, ...)          \
  do {                            \
    ProfScope ps_(ctx, name);     \
    __VA_ARGS__;                  \
  } while (0)

static thread_local std::string g_create_err;

#define CK(expr)                                                                      \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess) {                                                           \
      ctx->err = std::string(#expr) + ": " + hipGetErrorString(e_);                   \
      return CMS_ERR_HIP;                                                             \
    }                                                                                 \
  } while (0)

#define CKL()                                                                         \
  do {                                                                                \
    hipError_t e_ = hipGetLastError();                                                \
    if (e_ != hipSuccess) {                                                           \
      ctx->err = std::string("kernel launch: ") + hipGetErrorString(e_);              \
      return CMS_ERR_HIP;                                                             \
    }                                                                                 \
  } while (0)

static int fail(cms_ctx* ctx, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  ctx->err = buf;
  return code;
}

static int read_int(cms_ctx* ctx, const int* dptr, int* out) {
  CK(hipMemcpyAsync(out, dptr, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  CK(hipStreamSynchronize(ctx->stream));
  return CMS_OK;
}

static BatchView view(cms_ctx* ctx) {
  BatchView b;
  b.atom = ctx->atom.as<double4>();
  b.dens = ctx->dens.as<double>();
  b.acplx = ctx->acplx.as<int32_t>();
  b.amol = ctx->amol.as<int8_t>();
  b.aoff = ctx->aoff.as<int32_t>();
  b.ntot = ctx->ntot;
  b.B = ctx->B;
  return b;
}

static Work work(cms_ctx* ctx) {
  Work w{};
  w.atten = ctx->atten.as<int8_t>();
  w.access = ctx->access.as<int8_t>();
  w.status = ctx->status.as<int>();
  w.nR = ctx->rpos.as<int>() + ctx->ntot;
  w.ratom = ctx->ratom.as<int>();
  w.rpos = ctx->rpos.as<int>();
  w.nn = ctx->nn.as<int>();
  w.nbr = ctx->nbr.as<int>();
  w.nb = ctx->nb.as<int>();
  w.bur = ctx->bur.as<int>();
  w.nbmax = ctx->nbmax.as<int>();
  w.pcnt = ctx->pcnt.as<int>();
  w.pscan = ctx->pscan.as<int>();
  w.nP = ctx->pscan.as<int>() + ctx->nR;
  w.pi = ctx->pi.as<int>();
  w.pj = ctx->pj.as<int>();
  w.pu = ctx->pu.as<double4>();
  w.pt = ctx->pt.as<double4>();
  w.pflag = ctx->pflag.as<int8_t>();
  w.pgrp = ctx->pgrp.as<int>();
  w.prcnt = ctx->prcnt.as<int>();
  w.prscan = ctx->prscan.as<int>();
  w.nQ = ctx->prscan.as<int>() + 2 * (int64_t)ctx->nP;
  w.patom = ctx->patom.as<int4>();
  w.ppos = ctx->ppos.as<double4>();
  w.palt = ctx->palt.as<double4>();
  w.prgrp = ctx->prgrp.as<int>();
  w.near = ctx->nearl.as<int>();
  w.nnear = ctx->nnear.as<int>();
  w.ccnt = ctx->ccnt.as<int>();
  w.cscan = ctx->cscan.as<int>();
  w.t1cnt = ctx->t1cnt.as<int>();
  w.t1scan = ctx->t1scan.as<int>();
  w.t2cnt = ctx->t2cnt.as<int>();
  w.t2scan = ctx->t2scan.as<int>();
  w.tflag = ctx->tflag.as<int8_t>();
  w.kcnt = ctx->kcnt.as<int>();
  w.kscan = ctx->kscan.as<int>();
  w.dbase = ctx->dbase.as<int64_t>();
  w.dpos = ctx->dpos.as<double4>();
  w.dnml = ctx->dnml.as<double4>();
  w.dbur = ctx->dbur.as<int8_t>();
  w.dtype = ctx->dtype.as<int8_t>();
  w.datom = ctx->datom.as<int>();
  w.ddesc = ctx->ddesc.as<int4>();
  w.ddesc2 = ctx->ddesc2.as<int4>();
  w.dot_cap = ctx->nD;
  return w;
}

extern "C" {

int cms_version(void) { return 100; }

int cms_create(int device, void* stream, cms_ctx** out) {
  if (!out) {
    g_create_err = "cms_create: out is NULL";
    return CMS_ERR_INVALID;
  }
  *out = nullptr;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0) {
    g_create_err = std::string("cms_create: no HIP device: ") + hipGetErrorString(e);
    return CMS_ERR_HIP;
  }
  if (device < 0 || device >= ndev) {
    g_create_err = "cms_create: device index out of range";
    return CMS_ERR_INVALID;
  }
  e = hipSetDevice(device);
  if (e != hipSuccess) {
    g_create_err = std::string("cms_create: hipSetDevice: ") + hipGetErrorString(e);
    return CMS_ERR_HIP;
  }
  cms_ctx* ctx = new cms_ctx();
  ctx->device = device;
  ctx->stream = static_cast<hipStream_t>(stream);
  if (ctx->nresc.ensure(sizeof(unsigned long long) * 2) != hipSuccess) {
    g_create_err = "cms_create: out of device memory";
    delete ctx;
    return CMS_ERR_HIP;
  }
  fill_u64(ctx->nresc.as<unsigned long long>(), 2, 0ull, ctx->stream);
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
    g_create_err = "cms_create: stream synchronisation failed";
    ctx->nresc.release();
    delete ctx;
    return CMS_ERR_HIP;
  }
  *out = ctx;
  return CMS_OK;
}

int cms_destroy(cms_ctx* ctx) {
  if (!ctx) return CMS_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  DBuf* bufs[] = {&ctx->atom, &ctx->dens, &ctx->acplx, &ctx->amol, &ctx->aoff, &ctx->nb0,
                  &ctx->agrp, &ctx->atten, &ctx->access, &ctx->status, &ctx->coinc,
                  &ctx->flag, &ctx->nbscan, &ctx->nbl, &ctx->grid, &ctx->gframe, &ctx->ncellc,
                  &ctx->ncscan, &ctx->acell, &ctx->ccount, &ctx->cstart, &ctx->ccur, &ctx->catom, &ctx->need, &ctx->rpos, &ctx->ratom,
                  &ctx->part, &ctx->nn, &ctx->nbmax, &ctx->nbr, &ctx->nb, &ctx->bur, &ctx->pcnt,
                  &ctx->pscan, &ctx->pi, &ctx->pj, &ctx->pu, &ctx->pt, &ctx->pflag,
                  &ctx->rgrp, &ctx->pgrp, &ctx->prcnt, &ctx->prscan, &ctx->patom,
                  &ctx->ppos, &ctx->palt, &ctx->prgrp, &ctx->nearl, &ctx->nnear, &ctx->ccnt,
                  &ctx->cscan, &ctx->t1cnt, &ctx->t1scan, &ctx->t2cnt, &ctx->t2scan,
                  &ctx->tflag, &ctx->kcnt, &ctx->kscan, &ctx->dbase, &ctx->dgrp, &ctx->dpos,
                  &ctx->dnml, &ctx->dbur, &ctx->dtype, &ctx->datom, &ctx->ddesc, &ctx->ddesc2, &ctx->bflag,
                  &ctx->bscan, &ctx->blist, &ctx->bxyz, &ctx->bgrp, &ctx->aflag,
                  &ctx->ascan, &ctx->alist, &ctx->axyz, &ctx->agrpd, &ctx->tcnt, &ctx->toff,
                  &ctx->bmin, &ctx->wv, &ctx->ocnt, &ctx->ostart, &ctx->ocur, &ctx->operm,
                  &ctx->bacc, &ctx->trflag, &ctx->trscan, &ctx->tlist, &ctx->txyz,
                  &ctx->tnml, &ctx->tgrp, &ctx->tmin, &ctx->tidx, &ctx->tgof, &ctx->tdist,
                  &ctx->trr, &ctx->side, &ctx->qpack, &ctx->rpack, &ctx->grp1, &ctx->gtot, &ctx->npd2, &ctx->npidx, &ctx->frame,
                  &ctx->bxyz32, &ctx->axyz32, &ctx->txyz32, &ctx->qpack32, &ctx->rpack32, &ctx->nresc, &ctx->bidx,
                  &ctx->bgof, &ctx->adj, &ctx->eatom, &ctx->evpos, &ctx->chk};
  for (DBuf* b : bufs) b->release();
  delete ctx;
  return CMS_OK;
}

const char* cms_last_error(const cms_ctx* ctx) {
  if (!ctx) return g_create_err.c_str();
  return ctx->err.c_str();
}

int cms_synchronize(cms_ctx* ctx) {
  if (!ctx) return CMS_ERR_INVALID;
  CK(hipStreamSynchronize(ctx->stream));
  return CMS_OK;
}

int cms_load(cms_ctx* ctx, int n_complex, const int32_t* n_binder, const int32_t* n_target,
             const double* xyz, const double* radius, const double* density,
             const cms_settings* settings) {
  if (!ctx) return CMS_ERR_INVALID;
  if (n_complex < 1 || !n_binder || !n_target || !settings)
    return fail(ctx, CMS_ERR_INVALID, "cms_load: invalid arguments");
  CK(hipSetDevice(ctx->device));
  ctx->B = n_complex;
  ctx->h_aoff.assign(n_complex + 1, 0);
  ctx->h_nb0.assign(n_complex, 0);
  ctx->h_agrp.assign(2 * n_complex + 1, 0);
  int64_t tot = 0;
  for (int c = 0; c < n_complex; ++c) {
    if (n_binder[c] < 0 || n_target[c] < 0)
      return fail(ctx, CMS_ERR_INVALID, "cms_load: negative atom count");
    ctx->h_aoff[c] = (int)tot;
    ctx->h_nb0[c] = n_binder[c];
    ctx->h_agrp[2 * c] = (int)tot;
    ctx->h_agrp[2 * c + 1] = (int)(tot + n_binder[c]);
    tot += (int64_t)n_binder[c] + n_target[c];
    if (tot > (1ll << 30)) return fail(ctx, CMS_ERR_INVALID, "cms_load: too many atoms");
  }
  ctx->h_aoff[n_complex] = (int)tot;
  ctx->h_agrp[2 * n_complex] = (int)tot;
  ctx->ntot = (int)tot;
  if (tot > 0 && (!xyz || !radius)) return fail(ctx, CMS_ERR_INVALID, "cms_load: NULL coordinates");
  ctx->s.rp = settings->rp;
  ctx->s.band = settings->band;
  ctx->s.sep = settings->sep;
  ctx->s.weight = settings->weight;
  ctx->s.binwidth_dist = settings->binwidth_dist;
  ctx->s.binwidth_norm = settings->binwidth_norm;
  ctx->density = settings->density;
  const int n = ctx->ntot;
  CK(ctx->atom.ensure(sizeof(double4) * (n + 1)));
  CK(ctx->dens.ensure(sizeof(double) * (n + 1)));
  CK(ctx->acplx.ensure(sizeof(int) * (n + 1)));
  CK(ctx->amol.ensure(n + 1));
  CK(ctx->aoff.ensure(sizeof(int) * (n_complex + 1)));
  CK(ctx->nb0.ensure(sizeof(int) * n_complex));
  CK(ctx->agrp.ensure(sizeof(int) * (2 * n_complex + 1)));
  CK(hipMemcpyAsync(ctx->aoff.p, ctx->h_aoff.data(), sizeof(int) * (n_complex + 1),
                    hipMemcpyHostToDevice, ctx->stream));
  CK(hipMemcpyAsync(ctx->nb0.p, ctx->h_nb0.data(), sizeof(int) * n_complex, hipMemcpyHostToDevice,
                    ctx->stream));
  CK(hipMemcpyAsync(ctx->agrp.p, ctx->h_agrp.data(), sizeof(int) * (2 * n_complex + 1),
                    hipMemcpyHostToDevice, ctx->stream));
  if (n > 0) {
    k_pack_atoms<<<grid_for(n, 256, 4096), 256, 0, ctx->stream>>>(
        n, xyz, radius, density, ctx->density, ctx->aoff.as<int>(), ctx->nb0.as<int>(), n_complex,
        ctx->atom.as<double4>(), ctx->dens.as<double>(), ctx->acplx.as<int>(), ctx->amol.as<int8_t>());
    CKL();
  }
  CK(ctx->nresc.ensure(sizeof(unsigned long long) * 2));
  fill_u64(ctx->nresc.as<unsigned long long>(), 2, 0ull, ctx->stream);
  // host arrays above are copied asynchronously: wait before returning
  CK(hipStreamSynchronize(ctx->stream));
  ctx->generated = false;
  ctx->gen_all_atoms = -1;
  ctx->lists_ready = false;
  ctx->float_lists = false;
  ctx->cms_ready = false;
  return CMS_OK;
}

int cms_generate(cms_ctx* ctx, int all_atoms, int full_state) {
  if (!ctx) return CMS_ERR_INVALID;
  if (ctx->B < 1) return fail(ctx, CMS_ERR_INVALID, "cms_generate: nothing loaded");
  if (ctx->generated) {
    if (ctx->gen_all_atoms != (all_atoms ? 1 : 0))
      return fail(ctx, CMS_ERR_STATE,
                  "This MolecularSurfaceCalculator instance already generated its molecular surface "
                  "with all_atoms=%s, but was asked to reuse it with all_atoms=%s.",
                  ctx->gen_all_atoms ? "True" : "False", all_atoms ? "True" : "False");
    return CMS_OK;
  }
  CK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const int B = ctx->B, G = 2 * B, n = ctx->ntot;
  BatchView b = view(ctx);
  const Settings s = ctx->s;
  ctx->lists_ready = false;
  ctx->float_lists = false;
  ctx->cms_ready = false;

  CK(ctx->atten.ensure(n + 1));
  CK(ctx->access.ensure(n + 1));
  CK(ctx->status.ensure(sizeof(int) * B));
  CK(ctx->coinc.ensure(sizeof(unsigned long long) * B));
  CK(ctx->flag.ensure(sizeof(int) * (n + 1)));
  CK(ctx->nbscan.ensure(sizeof(int) * (n + 1)));
  CK(ctx->nbl.ensure(sizeof(int) * (n + 1)));
  CK(ctx->need.ensure(sizeof(int) * (n + 1)));
  CK(ctx->rpos.ensure(sizeof(int) * (n + 1)));
  CK(ctx->ratom.ensure(sizeof(int) * (n + 1)));
  CK(ctx->part.ensure(sizeof(int) * 1040));
  CK(ctx->rgrp.ensure(sizeof(int) * (G + 1)));
  CK(ctx->pgrp.ensure(sizeof(int) * (G + 1)));
  CK(ctx->prgrp.ensure(sizeof(int) * (G + 1)));
  CK(ctx->dgrp.ensure(sizeof(int) * (G + 1)));
  CK(ctx->dbase.ensure(sizeof(int64_t) * 4 * G));
  int* part = ctx->part.as<int>();

  fill_int(ctx->status.as<int>(), B, 0, st);
  fill_u64(ctx->coinc.as<unsigned long long>(), B, ~0ull, st);
  if (n > 0) {
    // cell lists: per-complex grids, atoms sorted by cell
    const int64_t cell_cap = 8ll * n + 64ll * B;
    CK(ctx->grid.ensure(sizeof(GridInfo) * (B + 1)));
    CK(ctx->gframe.ensure(sizeof(double4) * (B + 1)));
    CK(ctx->ncellc.ensure(sizeof(int) * (B + 1)));
    CK(ctx->ncscan.ensure(sizeof(int) * (B + 1)));
    CK(ctx->acell.ensure(sizeof(int) * (n + 1)));
    CK(ctx->catom.ensure(sizeof(int) * (n + 1)));
    CK(ctx->ccount.ensure(sizeof(int) * (cell_cap + 1)));
    CK(ctx->cstart.ensure(sizeof(int) * (cell_cap + 1)));
    CK(ctx->ccur.ensure(sizeof(int) * (cell_cap + 1)));
    GridInfo* grid = ctx->grid.as<GridInfo>();
    TIMED("grid", k_grid_setup<<<grid_for(B, 1, 4096), 256, 0, st>>>(B, ctx->aoff.as<int>(), ctx->atom.as<double4>(), s,
                                                                      grid, ctx->gframe.as<double4>(),
                                                                      ctx->ncellc.as<int>()));
    scan_excl(ctx->ncellc.as<int>(), ctx->ncscan.as<int>(), B, part, st);
    k_grid_offsets<<<grid_for(B, 256, 4096), 256, 0, st>>>(B, ctx->ncscan.as<int>(), grid);
    fill_int(ctx->ccount.as<int>(), cell_cap, 0, st);
    fill_int(ctx->ccur.as<int>(), cell_cap, 0, st);
    k_cell_count<<<grid_for(n, 256, 4096), 256, 0, st>>>(b, grid, ctx->acell.as<int>(), ctx->ccount.as<int>());
    scan_excl(ctx->ccount.as<int>(), ctx->cstart.as<int>(), cell_cap, part, st);
    k_cell_scatter<<<grid_for(n, 256, 4096), 256, 0, st>>>(n, ctx->acell.as<int>(), ctx->cstart.as<int>(),
                                                            ctx->ccur.as<int>(), ctx->catom.as<int>());
    // small batches: one wavefront per atom (parallelism); large batches:
    // one thread per cell-sorted atom (cell data shared within a wavefront)
    const bool per_thread = n >= 200000;
    const int* cst = ctx->cstart.as<int>();
    const int* cat = ctx->catom.as<int>();
    if (per_thread) {
      TIMED("attention", k_attention_cells_t<<<grid_for(n, 256, 65536), 256, 0, st>>>(
          b, s, all_atoms, grid, cst, cat, ctx->atten.as<int8_t>(), ctx->access.as<int8_t>(),
          ctx->coinc.as<unsigned long long>(), ctx->status.as<int>()));
      TIMED("region", k_region_cells_t<<<grid_for(n, 256, 65536), 256, 0, st>>>(
          b, s, full_state || all_atoms, grid, cst, cat, ctx->atten.as<int8_t>(), ctx->need.as<int>()));
    } else {
      TIMED("attention", k_attention_cells<<<grid_for(n, 1, 65536), 64, 0, st>>>(
          b, s, all_atoms, grid, cst, cat, ctx->atten.as<int8_t>(), ctx->access.as<int8_t>(),
          ctx->coinc.as<unsigned long long>(), ctx->status.as<int>()));
      TIMED("region", k_region_cells<<<grid_for(n, 1, 65536), 64, 0, st>>>(
          b, s, full_state || all_atoms, grid, cst, cat, ctx->atten.as<int8_t>(), ctx->need.as<int>()));
    }
    scan_excl(ctx->need.as<int>(), ctx->rpos.as<int>(), n, part, st);
    k_compact_index<<<grid_for(n, 256, 4096), 256, 0, st>>>(n, ctx->need.as<int>(),
                                                             ctx->rpos.as<int>(), ctx->ratom.as<int>());
    CKL();
  } else {
    fill_int(ctx->rpos.as<int>(), 1, 0, st);
  }
  int r = read_int(ctx, ctx->rpos.as<int>() + n, &ctx->nR);
  if (r) return r;
  const int nR = ctx->nR;
  CK(ctx->nn.ensure(sizeof(int) * (nR + 1)));
  CK(ctx->nb.ensure(sizeof(int) * (nR + 1)));
  CK(ctx->nbmax.ensure(sizeof(int) * (nR + 1)));
  CK(ctx->nbr.ensure(sizeof(int) * ((int64_t)nR * KCAP + 1)));
  CK(ctx->bur.ensure(sizeof(int) * ((int64_t)nR * BCAP + 1)));
  CK(ctx->pcnt.ensure(sizeof(int) * (nR + 1)));
  CK(ctx->pscan.ensure(sizeof(int) * (nR + 1)));
  CK(ctx->ccnt.ensure(sizeof(int) * (nR + 1)));
  CK(ctx->cscan.ensure(sizeof(int) * (nR + 1)));
  Work w = work(ctx);
  if (nR > 0) {
    TIMED("neighbors", k_neighbors_cells<<<grid_for(nR, 1), 64, 0, st>>>(b, s, w, ctx->grid.as<GridInfo>(),
                                                                          ctx->cstart.as<int>(), ctx->catom.as<int>()));
    TIMED("pairs_count", k_pairs<false><<<grid_for(nR, 1), 64, 0, st>>>(b, s, w));
    CKL();
  }
  scan_excl(ctx->pcnt.as<int>(), ctx->pscan.as<int>(), nR, part, st);
  k_group_starts_from_atoms<<<grid_for(nR + 1, 256, 4096), 256, 0, st>>>(
      nR, ctx->ratom.as<int>(), ctx->acplx.as<int>(), ctx->amol.as<int8_t>(), G, ctx->rgrp.as<int>());
  k_gather_scan<<<grid_for(G + 1, 256, 4096), 256, 0, st>>>(G, ctx->rgrp.as<int>(), 1,
                                                             ctx->pscan.as<int>(), ctx->pgrp.as<int>());
  CKL();
  r = read_int(ctx, ctx->pscan.as<int>() + nR, &ctx->nP);
  if (r) return r;
  const int nP = ctx->nP;
  CK(ctx->pi.ensure(sizeof(int) * (nP + 1)));
  CK(ctx->pj.ensure(sizeof(int) * (nP + 1)));
  CK(ctx->pu.ensure(sizeof(double4) * (nP + 1)));
  CK(ctx->pt.ensure(sizeof(double4) * (nP + 1)));
  CK(ctx->pflag.ensure(nP + 1));
  CK(ctx->prcnt.ensure(sizeof(int) * (2 * (int64_t)nP + 1)));
  CK(ctx->prscan.ensure(sizeof(int) * (2 * (int64_t)nP + 1)));
  CK(ctx->t1cnt.ensure(sizeof(int) * (nP + 1)));
  CK(ctx->t1scan.ensure(sizeof(int) * (nP + 1)));
  CK(ctx->t2cnt.ensure(sizeof(int) * (nP + 1)));
  CK(ctx->t2scan.ensure(sizeof(int) * (nP + 1)));
  CK(ctx->tflag.ensure(nP + 1));
  w = work(ctx);
  if (nP > 0) {
    TIMED("pairs_emit", k_pairs<true><<<grid_for(nR, 1), 64, 0, st>>>(b, s, w));
    TIMED("probes_count", k_probes<false><<<grid_for(nP, 1), 64, 0, st>>>(b, s, w));
    CKL();
  }
  scan_excl(ctx->prcnt.as<int>(), ctx->prscan.as<int>(), 2 * (int64_t)nP, part, st);
  k_gather_scan<<<grid_for(G + 1, 256, 4096), 256, 0, st>>>(G, ctx->pgrp.as<int>(), 2,
                                                             ctx->prscan.as<int>(), ctx->prgrp.as<int>());
  CKL();
  r = read_int(ctx, ctx->prscan.as<int>() + 2 * (int64_t)nP, &ctx->nQ);
  if (r) return r;
  const int nQ = ctx->nQ;
  CK(ctx->patom.ensure(sizeof(int4) * (nQ + 1)));
  CK(ctx->ppos.ensure(sizeof(double4) * (nQ + 1)));
  CK(ctx->palt.ensure(sizeof(double4) * (nQ + 1)));
  CK(ctx->nearl.ensure(sizeof(int) * ((int64_t)nQ * NEARCAP + 1)));
  CK(ctx->nnear.ensure(sizeof(int) * (nQ + 1)));
  CK(ctx->kcnt.ensure(sizeof(int) * (nQ + 1)));
  CK(ctx->kscan.ensure(sizeof(int) * (nQ + 1)));
  w = work(ctx);
  if (nQ > 0) {
    TIMED("probes_emit", k_probes<true><<<grid_for(nP, 1), 64, 0, st>>>(b, s, w));
    TIMED("near", k_near<<<grid_for(nQ, 1), 64, 0, st>>>(b, s, w));
    CKL();
  }
  if (nR > 0) TIMED("convex_count", k_convex<false><<<grid_for(nR, 1), 64, 0, st>>>(b, s, w));
  if (nP > 0) TIMED("toroid_count", k_toroid<false><<<grid_for(nP, 1), 64, 0, st>>>(b, s, w));
  if (nQ > 0 && s.rp > 0.0) TIMED("concave_count", k_concave<false><<<grid_for(nQ, 1), 64, 0, st>>>(b, s, w));
  if (nQ > 0 && !(s.rp > 0.0)) fill_int(ctx->kcnt.as<int>(), nQ, 0, st);
  CKL();
  scan_excl(ctx->ccnt.as<int>(), ctx->cscan.as<int>(), nR, part, st);
  scan_excl(ctx->t1cnt.as<int>(), ctx->t1scan.as<int>(), nP, part, st);
  scan_excl(ctx->t2cnt.as<int>(), ctx->t2scan.as<int>(), nP, part, st);
  scan_excl(ctx->kcnt.as<int>(), ctx->kscan.as<int>(), nQ, part, st);
  k_dot_bases<<<1, 256, 0, st>>>(G, ctx->rgrp.as<int>(), ctx->pgrp.as<int>(), ctx->prgrp.as<int>(),
                                 ctx->cscan.as<int>(), ctx->t1scan.as<int>(), ctx->t2scan.as<int>(),
                                 ctx->kscan.as<int>(), ctx->dbase.as<int64_t>(), ctx->dgrp.as<int>());
  CKL();
  r = read_int(ctx, ctx->dgrp.as<int>() + G, &ctx->nD);
  if (r) return r;
  const int nD = ctx->nD;
  CK(ctx->dpos.ensure(sizeof(double4) * (nD + 1)));
  CK(ctx->dnml.ensure(sizeof(double4) * (nD + 1)));
  CK(ctx->dbur.ensure(nD + 1));
  CK(ctx->dtype.ensure(nD + 1));
  CK(ctx->datom.ensure(sizeof(int) * (nD + 1)));
  CK(ctx->ddesc.ensure(sizeof(int4) * (nD + 1)));
  CK(ctx->ddesc2.ensure(sizeof(int4) * (nD + 1)));
  w = work(ctx);
  if (nD > 0) {
    if (nR > 0) TIMED("convex_emit", k_convex<true><<<grid_for(nR, 1), 64, 0, st>>>(b, s, w));
    if (nP > 0) TIMED("toroid_emit", k_toroid<true><<<grid_for(nP, 1), 64, 0, st>>>(b, s, w));
    if (nQ > 0 && s.rp > 0.0) TIMED("concave_emit", k_concave<true><<<grid_for(nQ, 1), 64, 0, st>>>(b, s, w));
    CKL();
  }
  ctx->h_dgrp.assign(G + 1, 0);
  ctx->h_prgrp.assign(G + 1, 0);
  CK(hipMemcpyAsync(ctx->h_dgrp.data(), ctx->dgrp.p, sizeof(int) * (G + 1), hipMemcpyDeviceToHost, st));
  CK(hipMemcpyAsync(ctx->h_prgrp.data(), ctx->prgrp.p, sizeof(int) * (G + 1), hipMemcpyDeviceToHost, st));
  std::vector<int> stat(B);
  CK(hipMemcpyAsync(stat.data(), ctx->status.p, sizeof(int) * B, hipMemcpyDeviceToHost, st));
  CK(hipStreamSynchronize(st));
  int worst = 0;
  for (int c = 0; c < B; ++c) {
    if (stat[c] & ST_COINCIDENT) return fail(ctx, CMS_ERR_COINCIDENT, "coincident atoms in complex %d", c);
    if (stat[c]) worst = stat[c];
  }
  if (worst)
    return fail(ctx, CMS_ERR_CAPACITY, "row capacity exceeded (status %d): more than %d neighbours, "
                "%d buried partners or %d near low probes", worst, KCAP, BCAP, NEARCAP);
  ctx->generated = true;
  ctx->gen_all_atoms = all_atoms ? 1 : 0;
  return CMS_OK;
}

int cms_status(cms_ctx* ctx, int32_t* status_host) {
  if (!ctx || !status_host) return CMS_ERR_INVALID;
  CK(hipMemcpyAsync(status_host, ctx->status.p, sizeof(int) * ctx->B, hipMemcpyDeviceToHost, ctx->stream));
  CK(hipStreamSynchronize(ctx->stream));
  return CMS_OK;
}

int cms_coincident(cms_ctx* ctx, int c, int32_t* i, int32_t* j) {
  if (!ctx || c < 0 || c >= ctx->B || !i || !j) return CMS_ERR_INVALID;
  unsigned long long key = ~0ull;
  CK(hipMemcpyAsync(&key, ctx->coinc.as<unsigned long long>() + c, sizeof key, hipMemcpyDeviceToHost,
                    ctx->stream));
  CK(hipStreamSynchronize(ctx->stream));
  if (key == ~0ull) {
    *i = *j = -1;
  } else {
    *i = (int32_t)(key >> 32);
    *j = (int32_t)(key & 0xffffffffu);
  }
  return CMS_OK;
}

// ------------------------------------------------------------ readouts

static int prepare_lists(cms_ctx* ctx) {
  if (ctx->lists_ready) return CMS_OK;
  hipStream_t st = ctx->stream;
  const int G = 2 * ctx->B, nD = ctx->nD;
  int* part = ctx->part.as<int>();
  CK(ctx->bflag.ensure(sizeof(int) * (nD + 1)));
  CK(ctx->bscan.ensure(sizeof(int) * (nD + 1)));
  CK(ctx->aflag.ensure(sizeof(int) * (nD + 1)));
  CK(ctx->ascan.ensure(sizeof(int) * (nD + 1)));
  CK(ctx->bgrp.ensure(sizeof(int) * (G + 1)));
  CK(ctx->agrpd.ensure(sizeof(int) * (G + 1)));
  if (nD > 0) {
    k_dot_flag<<<grid_for(nD, 256, 4096), 256, 0, st>>>(nD, ctx->dbur.as<int8_t>(), 1, ctx->bflag.as<int>());
    k_dot_flag<<<grid_for(nD, 256, 4096), 256, 0, st>>>(nD, ctx->dbur.as<int8_t>(), 0, ctx->aflag.as<int>());
  }
  scan_excl(ctx->bflag.as<int>(), ctx->bscan.as<int>(), nD, part, st);
  k_gather_scan<<<grid_for(G + 1, 256, 4096), 256, 0, st>>>(G, ctx->dgrp.as<int>(), 1, ctx->bscan.as<int>(),
                                                             ctx->bgrp.as<int>());
  scan_excl(ctx->aflag.as<int>(), ctx->ascan.as<int>(), nD, part, st);
  k_gather_scan<<<grid_for(G + 1, 256, 4096), 256, 0, st>>>(G, ctx->dgrp.as<int>(), 1, ctx->ascan.as<int>(),
                                                             ctx->agrpd.as<int>());
  CKL();
  int r = read_int(ctx, ctx->bscan.as<int>() + nD, &ctx->nB);
  if (r) return r;
  ctx->nA = nD - ctx->nB;
  CK(ctx->blist.ensure(sizeof(int) * (ctx->nB + 1)));
  CK(ctx->bxyz.ensure(sizeof(double4) * (ctx->nB + 1)));
  CK(ctx->alist.ensure(sizeof(int) * (ctx->nA + 1)));
  CK(ctx->axyz.ensure(sizeof(double4) * (ctx->nA + 1)));
  if (nD > 0) {
    k_dot_compact<<<grid_for(nD, 256, 4096), 256, 0, st>>>(nD, ctx->bflag.as<int>(), ctx->bscan.as<int>(),
                                                            ctx->dpos.as<double4>(), ctx->blist.as<int>(),
                                                            ctx->bxyz.as<double4>());
    k_dot_compact<<<grid_for(nD, 256, 4096), 256, 0, st>>>(nD, ctx->aflag.as<int>(), ctx->ascan.as<int>(),
                                                            ctx->dpos.as<double4>(), ctx->alist.as<int>(),
                                                            ctx->axyz.as<double4>());
    CKL();
  }
  ctx->lists_ready = true;
  return CMS_OK;
}

// nearest search of query groups qgrp over reference groups rgrp (g ^ xmask).
// Small query sets split every group's references over several workgroups
// so that a single complex still spreads over the GPU.
static int launch_nearest(cms_ctx* ctx, int precision, const double4* Q, const float4* Q32, const int* qgrp,
                          int nq, int nr_total, const double4* R, const float4* R32, const int* rgrp, int xmask,
                          int G, double* out_d2, int* out_idx) {
  hipStream_t st = ctx->stream;
  CK(ctx->tcnt.ensure(sizeof(int) * (G + 1)));
  CK(ctx->toff.ensure(sizeof(int) * (G + 1)));
  k_tile_counts<<<grid_for(G, 256, 4096), 256, 0, st>>>(G, qgrp, 256, ctx->tcnt.as<int>());
  scan_excl(ctx->tcnt.as<int>(), ctx->toff.as<int>(), G, ctx->part.as<int>(), st);
  const int64_t tiles_est = (int64_t)nq / 256 + G;
  const int64_t target = 2048;
  int nsplit = 1;
  if (tiles_est < target) {
    nsplit = (int)std::min<int64_t>((target + tiles_est - 1) / tiles_est, 64);
    // at least 128 references per split on average
    const int64_t per_group_refs = std::max<int64_t>(1, (int64_t)nr_total / std::max(G, 1));
    nsplit = (int)std::max<int64_t>(1, std::min<int64_t>(nsplit, per_group_refs / 128));
  }
  const int grid = grid_for(tiles_est * nsplit, 1, 16384);
  const int64_t ps = nsplit == 1 ? 0 : nq;
  double* d2o = out_d2;
  int* ido = out_idx;
  if (nsplit > 1) {
    CK(ctx->npd2.ensure(sizeof(double) * ((int64_t)nsplit * ps + 1)));
    CK(ctx->npidx.ensure(sizeof(int) * ((int64_t)nsplit * ps + 1)));
    d2o = ctx->npd2.as<double>();
    ido = out_idx ? ctx->npidx.as<int>() : nullptr;
  }
  const int* toff = ctx->toff.as<int>();
  if (precision == CMS_PREC_FP32_RESCORE) {
    unsigned long long* nr = ctx->nresc.as<unsigned long long>();
    if (ido)
      TIMED("nearest32", k_nearest32<true><<<grid, 256, 0, st>>>(Q32, Q, qgrp, R32, R, rgrp, xmask, toff, G, nsplit, ps,
                                                                 ctx->frame.as<double4>(), d2o, ido, nr));
    else
      TIMED("nearest32", k_nearest32<false><<<grid, 256, 0, st>>>(Q32, Q, qgrp, R32, R, rgrp, xmask, toff, G, nsplit, ps,
                                                                  ctx->frame.as<double4>(), d2o, nullptr, nr));
  } else {
    if (ido)
      TIMED("nearest", k_nearest64<true><<<grid, 256, 0, st>>>(Q, qgrp, R, rgrp, xmask, toff, G, nsplit, ps, d2o, ido));
    else
      TIMED("nearest", k_nearest64<false><<<grid, 256, 0, st>>>(Q, qgrp, R, rgrp, xmask, toff, G, nsplit, ps, d2o, nullptr));
  }
  if (nsplit > 1 && nq > 0)
    k_nearest_combine<<<grid_for(nq, 256, 4096), 256, 0, st>>>(nq, nsplit, ps, d2o, ido, out_d2, out_idx);
  CKL();
  return CMS_OK;
}

// float, recentred copies of the buried / accessible dot lists
static int prepare_float_lists(cms_ctx* ctx) {
  if (ctx->float_lists) return CMS_OK;
  hipStream_t st = ctx->stream;
  const int B = ctx->B, G = 2 * B;
  CK(ctx->frame.ensure(sizeof(double4) * (B + 1)));
  CK(ctx->nresc.ensure(sizeof(unsigned long long) * 2));
  CK(ctx->bxyz32.ensure(sizeof(float4) * (ctx->nB + 1)));
  CK(ctx->axyz32.ensure(sizeof(float4) * (ctx->nA + 1)));
  k_complex_frame<<<grid_for(B, 1, 4096), 256, 0, st>>>(B, ctx->aoff.as<int>(), ctx->atom.as<double4>(), ctx->s.rp,
                                                         ctx->frame.as<double4>());
  if (ctx->nB > 0)
    k_to_float<<<grid_for(ctx->nB, 256, 4096), 256, 0, st>>>(ctx->nB, ctx->bxyz.as<double4>(), ctx->bgrp.as<int>(), G, 1,
                                                              ctx->frame.as<double4>(), ctx->bxyz32.as<float4>());
  if (ctx->nA > 0)
    k_to_float<<<grid_for(ctx->nA, 256, 4096), 256, 0, st>>>(ctx->nA, ctx->axyz.as<double4>(), ctx->agrpd.as<int>(), G, 1,
                                                              ctx->frame.as<double4>(), ctx->axyz32.as<float4>());
  CKL();
  ctx->float_lists = true;
  return CMS_OK;
}

__global__ void k_pick_stride(int n, const double* in, int stride, double* out) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    out[i] = in[(int64_t)i * stride];
}

// per-atom sums in item order + per-group np.sum totals
static int owner_sums(cms_ctx* ctx, int nitems, const int* list, const double* val, int64_t vstride,
                      double* per_atom, double* totals, int G_tot, int gstride) {
  hipStream_t st = ctx->stream;
  const int n = ctx->ntot;
  CK(ctx->ocnt.ensure(sizeof(int) * (n + 1)));
  CK(ctx->ostart.ensure(sizeof(int) * (n + 1)));
  CK(ctx->ocur.ensure(sizeof(int) * (n + 1)));
  CK(ctx->operm.ensure(sizeof(int) * (nitems + 1)));
  fill_int(ctx->ocnt.as<int>(), n, 0, st);
  fill_int(ctx->ocur.as<int>(), n, 0, st);
  if (nitems > 0)
    k_owner_count<<<grid_for(nitems, 256, 4096), 256, 0, st>>>(nitems, list, ctx->datom.as<int>(),
                                                                ctx->ocnt.as<int>());
  scan_excl(ctx->ocnt.as<int>(), ctx->ostart.as<int>(), n, ctx->part.as<int>(), st);
  if (nitems > 0)
    k_owner_scatter<<<grid_for(nitems, 256, 4096), 256, 0, st>>>(nitems, list, ctx->datom.as<int>(),
                                                                  ctx->ostart.as<int>(), ctx->ocur.as<int>(),
                                                                  ctx->operm.as<int>());
  if (n > 0)
    TIMED("owner_sums", k_owner_sums<<<grid_for(n, 1, 16384), 64, 0, st>>>(n, ctx->ostart.as<int>(), ctx->operm.as<int>(), val,
                                                          vstride, per_atom));
  CKL();
  // np.sum of the per-atom values of every molecule group; gstride 1 keeps
  // all 2B group totals, gstride 2 only the molecule-0 ones (max-CMS)
  CK(ctx->gtot.ensure(sizeof(double) * (G_tot + 1)));
  k_group_np_sum<<<grid_for(G_tot, 1, 4096), 256, 0, st>>>(G_tot, ctx->agrp.as<int>(), per_atom, 1,
                                                            ctx->gtot.as<double>());
  k_pick_stride<<<grid_for(G_tot / gstride, 256, 4096), 256, 0, st>>>(G_tot / gstride, ctx->gtot.as<double>(),
                                                                       gstride, totals);
  CKL();
  return CMS_OK;
}

int cms_contact_ms(cms_ctx* ctx, int precision, double* per_atom, double* totals) {
  if (!ctx || !per_atom || !totals) return CMS_ERR_INVALID;
  if (!ctx->generated || ctx->gen_all_atoms != 0)
    return fail(ctx, CMS_ERR_STATE, "cms_contact_ms: surface not generated in CMS mode");
  if (precision != CMS_PREC_FP64 && precision != CMS_PREC_FP32_RESCORE)
    return fail(ctx, CMS_ERR_INVALID, "cms_contact_ms: unsupported precision");
  CK(hipSetDevice(ctx->device));
  int r = prepare_lists(ctx);
  if (r) return r;
  if (precision == CMS_PREC_FP32_RESCORE && (r = prepare_float_lists(ctx))) return r;
  const int G = 2 * ctx->B, nB = ctx->nB;
  hipStream_t st = ctx->stream;
  CK(ctx->bmin.ensure(sizeof(double) * (nB + 1)));
  CK(ctx->bidx.ensure(sizeof(int) * (nB + 1)));
  CK(ctx->wv.ensure(sizeof(double) * (nB + 1)));
  r = launch_nearest(ctx, precision, ctx->bxyz.as<double4>(), ctx->bxyz32.as<float4>(), ctx->bgrp.as<int>(), nB, nB,
                     ctx->bxyz.as<double4>(), ctx->bxyz32.as<float4>(), ctx->bgrp.as<int>(), 1, G,
                     ctx->bmin.as<double>(), ctx->bidx.as<int>());
  if (r) return r;
  if (nB > 0)
    k_cms_weights<<<grid_for(nB, 256, 4096), 256, 0, st>>>(nB, ctx->bxyz.as<double4>(), ctx->bmin.as<double>(),
                                                            ctx->s.weight, ctx->wv.as<double>());
  CKL();
  r = owner_sums(ctx, nB, ctx->blist.as<int>(), ctx->wv.as<double>(), 1, per_atom, totals, G, 1);
  if (r) return r;
  ctx->cms_ready = true;
  return CMS_OK;
}

int cms_max_contact_ms(cms_ctx* ctx, double* per_atom, double* totals) {
  if (!ctx || !per_atom || !totals) return CMS_ERR_INVALID;
  if (!ctx->generated || ctx->gen_all_atoms != 1)
    return fail(ctx, CMS_ERR_STATE, "cms_max_contact_ms: surface not generated in all_atoms mode");
  CK(hipSetDevice(ctx->device));
  // items: all dots; value: dpos[i].w (area)
  const double* val = reinterpret_cast<const double*>(ctx->dpos.as<double4>()) + 3;
  return owner_sums(ctx, ctx->nD, nullptr, val, 4, per_atom, totals, 2 * ctx->B, 2);
}

int cms_shape_complementarity(cms_ctx* ctx, int precision, double* results) {
  if (!ctx || !results) return CMS_ERR_INVALID;
  if (!ctx->generated || ctx->gen_all_atoms != 0)
    return fail(ctx, CMS_ERR_STATE, "cms_shape_complementarity: surface not generated in CMS mode");
  if (precision != CMS_PREC_FP64 && precision != CMS_PREC_FP32_RESCORE)
    return fail(ctx, CMS_ERR_INVALID, "cms_shape_complementarity: unsupported precision");
  CK(hipSetDevice(ctx->device));
  int r = prepare_lists(ctx);
  if (r) return r;
  if (precision == CMS_PREC_FP32_RESCORE && (r = prepare_float_lists(ctx))) return r;
  hipStream_t st = ctx->stream;
  const int B = ctx->B, G = 2 * B, nB = ctx->nB;
  int* part = ctx->part.as<int>();
  CK(ctx->bacc.ensure(sizeof(double) * (nB + 1)));
  CK(ctx->trflag.ensure(sizeof(int) * (nB + 1)));
  CK(ctx->trscan.ensure(sizeof(int) * (nB + 1)));
  CK(ctx->tgrp.ensure(sizeof(int) * (G + 1)));
  CK(ctx->side.ensure(sizeof(double) * 8 * G));
  // trim: buried dots of g against accessible dots of g
  r = launch_nearest(ctx, precision, ctx->bxyz.as<double4>(), ctx->bxyz32.as<float4>(), ctx->bgrp.as<int>(), nB,
                     ctx->nA, ctx->axyz.as<double4>(), ctx->axyz32.as<float4>(), ctx->agrpd.as<int>(), 0, G,
                     ctx->bacc.as<double>(), nullptr);
  if (r) return r;
  if (nB > 0)
    k_trim_flag<<<grid_for(nB, 256, 4096), 256, 0, st>>>(nB, ctx->bacc.as<double>(), ctx->s.band * ctx->s.band,
                                                          ctx->trflag.as<int>());
  scan_excl(ctx->trflag.as<int>(), ctx->trscan.as<int>(), nB, part, st);
  k_gather_scan<<<grid_for(G + 1, 256, 4096), 256, 0, st>>>(G, ctx->bgrp.as<int>(), 1, ctx->trscan.as<int>(),
                                                             ctx->tgrp.as<int>());
  CKL();
  int nT = 0;
  r = read_int(ctx, ctx->trscan.as<int>() + nB, &nT);
  if (r) return r;
  CK(ctx->tlist.ensure(sizeof(int) * (nT + 1)));
  CK(ctx->txyz.ensure(sizeof(double4) * (nT + 1)));
  CK(ctx->tnml.ensure(sizeof(double4) * (nT + 1)));
  CK(ctx->tmin.ensure(sizeof(double) * (nT + 1)));
  CK(ctx->tidx.ensure(sizeof(int) * (nT + 1)));
  CK(ctx->tgof.ensure(sizeof(int) * (nT + 1)));
  CK(ctx->tdist.ensure(sizeof(double) * (nT + 1)));
  CK(ctx->trr.ensure(sizeof(double) * (nT + 1)));
  if (nB > 0)
    k_trim_compact<<<grid_for(nB, 256, 4096), 256, 0, st>>>(nB, ctx->trflag.as<int>(), ctx->trscan.as<int>(),
                                                             ctx->blist.as<int>(), ctx->dpos.as<double4>(),
                                                             ctx->dnml.as<double4>(), ctx->tlist.as<int>(),
                                                             ctx->txyz.as<double4>(), ctx->tnml.as<double4>());
  CKL();
  if (precision == CMS_PREC_FP32_RESCORE) {
    CK(ctx->txyz32.ensure(sizeof(float4) * (nT + 1)));
    if (nT > 0)
      k_to_float<<<grid_for(nT, 256, 4096), 256, 0, st>>>(nT, ctx->txyz.as<double4>(), ctx->tgrp.as<int>(), G, 1,
                                                           ctx->frame.as<double4>(), ctx->txyz32.as<float4>());
  }
  r = launch_nearest(ctx, precision, ctx->txyz.as<double4>(), ctx->txyz32.as<float4>(), ctx->tgrp.as<int>(), nT, nT,
                     ctx->txyz.as<double4>(), ctx->txyz32.as<float4>(), ctx->tgrp.as<int>(), 1, G,
                     ctx->tmin.as<double>(), ctx->tidx.as<int>());
  if (r) return r;
  k_group_of_items<<<grid_for(G, 1, 4096), 256, 0, st>>>(G, ctx->tgrp.as<int>(), ctx->tgof.as<int>());
  if (nT > 0)
    k_sc_terms<<<grid_for(nT, 256, 4096), 256, 0, st>>>(nT, ctx->tgof.as<int>(), ctx->tgrp.as<int>(),
                                                         ctx->tmin.as<double>(), ctx->tidx.as<int>(),
                                                         ctx->tnml.as<double4>(), ctx->s.weight,
                                                         ctx->tdist.as<double>(), ctx->trr.as<double>());
  TIMED("sc_sides", k_sc_sides<<<grid_for(G, 1, 4096), 256, 0, st>>>(G, ctx->tgrp.as<int>(), ctx->txyz.as<double4>(),
                                                    ctx->tdist.as<double>(), ctx->trr.as<double>(),
                                                    ctx->dgrp.as<int>(), ctx->s, ctx->side.as<double>()));
  k_sc_combine<<<grid_for(B, 256, 4096), 256, 0, st>>>(B, ctx->side.as<double>(), results);
  CKL();
  return CMS_OK;
}

// ------------------------------------------------------------ introspection

__global__ void k_unpack_dots(int n, int lo, int aoff, const double4* dpos, const double4* dnml,
                              const int8_t* dbur, const int8_t* dtype, const int* datom, double* coor,
                              double* nml, double* area, int8_t* buried, int8_t* type, int32_t* atom) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double4 p = dpos[lo + i];
    const double4 q = dnml[lo + i];
    if (coor) { coor[3 * i] = p.x; coor[3 * i + 1] = p.y; coor[3 * i + 2] = p.z; }
    if (nml) { nml[3 * i] = q.x; nml[3 * i + 1] = q.y; nml[3 * i + 2] = q.z; }
    if (area) area[i] = p.w;
    if (buried) buried[i] = dbur[lo + i];
    if (type) type[i] = dtype[lo + i];
    if (atom) atom[i] = datom[lo + i] - aoff;
  }
}

int cms_dot_count(cms_ctx* ctx, int c, int mol, int64_t* n) {
  if (!ctx || !n || c < 0 || c >= ctx->B || mol < 0 || mol > 1) return CMS_ERR_INVALID;
  if (!ctx->generated) return fail(ctx, CMS_ERR_STATE, "cms_dot_count: no surface");
  const int g = 2 * c + mol;
  *n = ctx->h_dgrp[g + 1] - ctx->h_dgrp[g];
  return CMS_OK;
}

int cms_copy_dots(cms_ctx* ctx, int c, int mol, double* coor, double* nml, double* area, int8_t* buried,
                  int8_t* type, int32_t* atom) {
  int64_t n = 0;
  int r = cms_dot_count(ctx, c, mol, &n);
  if (r) return r;
  if (n == 0) return CMS_OK;
  const int g = 2 * c + mol;
  k_unpack_dots<<<grid_for(n, 256, 4096), 256, 0, ctx->stream>>>(
      (int)n, ctx->h_dgrp[g], ctx->h_aoff[c], ctx->dpos.as<double4>(), ctx->dnml.as<double4>(),
      ctx->dbur.as<int8_t>(), ctx->dtype.as<int8_t>(), ctx->datom.as<int>(), coor, nml, area, buried, type, atom);
  CKL();
  return CMS_OK;
}

int cms_copy_atom_state(cms_ctx* ctx, int8_t* atten, int8_t* access) {
  if (!ctx) return CMS_ERR_INVALID;
  if (!ctx->generated) return fail(ctx, CMS_ERR_STATE, "cms_copy_atom_state: no surface");
  if (atten) CK(hipMemcpyAsync(atten, ctx->atten.p, ctx->ntot, hipMemcpyDeviceToDevice, ctx->stream));
  if (access) CK(hipMemcpyAsync(access, ctx->access.p, ctx->ntot, hipMemcpyDeviceToDevice, ctx->stream));
  return CMS_OK;
}

__global__ void k_unpack_probes(int n, int lo, int aoff, const int4* patom, const double4* ppos,
                                const double4* palt, int32_t* atoms, double* height, double* point,
                                double* alt) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int4 a = patom[lo + i];
    const double4 p = ppos[lo + i];
    const double4 u = palt[lo + i];
    if (atoms) { atoms[3 * i] = a.x - aoff; atoms[3 * i + 1] = a.y - aoff; atoms[3 * i + 2] = a.z - aoff; }
    if (height) height[i] = p.w;
    if (point) { point[3 * i] = p.x; point[3 * i + 1] = p.y; point[3 * i + 2] = p.z; }
    if (alt) { alt[3 * i] = u.x; alt[3 * i + 1] = u.y; alt[3 * i + 2] = u.z; }
  }
}

int cms_probe_count(cms_ctx* ctx, int c, int64_t* n) {
  if (!ctx || !n || c < 0 || c >= ctx->B) return CMS_ERR_INVALID;
  if (!ctx->generated) return fail(ctx, CMS_ERR_STATE, "cms_probe_count: no surface");
  *n = ctx->h_prgrp[2 * c + 2] - ctx->h_prgrp[2 * c];
  return CMS_OK;
}

int cms_copy_probes(cms_ctx* ctx, int c, int32_t* atoms, double* height, double* point, double* alt) {
  int64_t n = 0;
  int r = cms_probe_count(ctx, c, &n);
  if (r) return r;
  if (n == 0) return CMS_OK;
  k_unpack_probes<<<grid_for(n, 256, 4096), 256, 0, ctx->stream>>>(
      (int)n, ctx->h_prgrp[2 * c], ctx->h_aoff[c], ctx->patom.as<int4>(), ctx->ppos.as<double4>(),
      ctx->palt.as<double4>(), atoms, height, point, alt);
  CKL();
  return CMS_OK;
}

int cms_counts(cms_ctx* ctx, int64_t* out) {
  if (!ctx || !out) return CMS_ERR_INVALID;
  if (!ctx->generated) return fail(ctx, CMS_ERR_STATE, "cms_counts: no surface");
  const int B = ctx->B, G = 2 * B;
  hipStream_t st = ctx->stream;
  int r = prepare_lists(ctx);
  if (r) return r;
  std::vector<int> bg(G + 1), rg(G + 1), pg(G + 1), prg(G + 1);
  std::vector<int> cs(ctx->nR + 1), t1(ctx->nP + 1), t2(ctx->nP + 1), ks(ctx->nQ + 1);
  std::vector<int8_t> att(ctx->ntot + 1);
  CK(hipMemcpyAsync(bg.data(), ctx->bgrp.p, sizeof(int) * (G + 1), hipMemcpyDeviceToHost, st));
  CK(hipMemcpyAsync(rg.data(), ctx->rgrp.p, sizeof(int) * (G + 1), hipMemcpyDeviceToHost, st));
  CK(hipMemcpyAsync(pg.data(), ctx->pgrp.p, sizeof(int) * (G + 1), hipMemcpyDeviceToHost, st));
  CK(hipMemcpyAsync(prg.data(), ctx->prgrp.p, sizeof(int) * (G + 1), hipMemcpyDeviceToHost, st));
  CK(hipMemcpyAsync(cs.data(), ctx->cscan.p, sizeof(int) * (ctx->nR + 1), hipMemcpyDeviceToHost, st));
  CK(hipMemcpyAsync(t1.data(), ctx->t1scan.p, sizeof(int) * (ctx->nP + 1), hipMemcpyDeviceToHost, st));
  CK(hipMemcpyAsync(t2.data(), ctx->t2scan.p, sizeof(int) * (ctx->nP + 1), hipMemcpyDeviceToHost, st));
  CK(hipMemcpyAsync(ks.data(), ctx->kscan.p, sizeof(int) * (ctx->nQ + 1), hipMemcpyDeviceToHost, st));
  if (ctx->ntot > 0) CK(hipMemcpyAsync(att.data(), ctx->atten.p, ctx->ntot, hipMemcpyDeviceToHost, st));
  CK(hipStreamSynchronize(st));
  for (int c = 0; c < B; ++c) {
    int64_t* o = out + (int64_t)c * CMS_NCOUNTS;
    const int g0 = 2 * c, g1 = 2 * c + 1, g2 = 2 * c + 2;
    o[0] = ctx->h_dgrp[g1] - ctx->h_dgrp[g0];
    o[1] = ctx->h_dgrp[g2] - ctx->h_dgrp[g1];
    o[2] = bg[g1] - bg[g0];
    o[3] = bg[g2] - bg[g1];
    o[4] = cs[rg[g2]] - cs[rg[g0]];
    o[5] = (t1[pg[g2]] - t1[pg[g0]]) + (t2[pg[g2]] - t2[pg[g0]]);
    o[6] = ks[prg[g2]] - ks[prg[g0]];
    o[7] = prg[g2] - prg[g0];
    o[8] = rg[g2] - rg[g0];
    o[9] = pg[g2] - pg[g0];
    int64_t bl0 = 0, bl1 = 0;
    for (int a = ctx->h_agrp[g0]; a < ctx->h_agrp[g1]; ++a) bl0 += att[a] == ATTEN_BLOCKER;
    for (int a = ctx->h_agrp[g1]; a < ctx->h_agrp[g2]; ++a) bl1 += att[a] == ATTEN_BLOCKER;
    o[10] = bl0;
    o[11] = bl1;
  }
  return CMS_OK;
}

__global__ void k_pack3(int n, const double* xyz, double4* out) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    out[i] = make_double4(xyz[3 * (int64_t)i], xyz[3 * (int64_t)i + 1], xyz[3 * (int64_t)i + 2], 0.0);
}

__global__ void k_points_frame(int nq, const double4* q, int nr, const double4* r, double4* frame) {
  __shared__ double s[6][256];
  double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int i = threadIdx.x; i < nq + nr; i += blockDim.x) {
    const double4 a = i < nq ? q[i] : r[i - nq];
    mn[0] = fmin(mn[0], a.x); mn[1] = fmin(mn[1], a.y); mn[2] = fmin(mn[2], a.z);
    mx[0] = fmax(mx[0], a.x); mx[1] = fmax(mx[1], a.y); mx[2] = fmax(mx[2], a.z);
  }
  for (int k = 0; k < 3; ++k) { s[k][threadIdx.x] = mn[k]; s[3 + k][threadIdx.x] = mx[k]; }
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      for (int k = 0; k < 3; ++k) s[k][threadIdx.x] = fmin(s[k][threadIdx.x], s[k][threadIdx.x + o]);
      for (int k = 3; k < 6; ++k) s[k][threadIdx.x] = fmax(s[k][threadIdx.x], s[k][threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double cx = 0.5 * (s[0][0] + s[3][0]), cy = 0.5 * (s[1][0] + s[4][0]), cz = 0.5 * (s[2][0] + s[5][0]);
    double half = fmax(fmax(s[3][0] - cx, s[4][0] - cy), s[5][0] - cz);
    if (!(half >= 0.0)) { cx = cy = cz = 0.0; half = 0.0; }
    frame[0] = make_double4(cx, cy, cz, half * 1.0000001 + 1e-6);
  }
}

int cms_nearest(cms_ctx* ctx, int precision, const double* q, int64_t nq, const double* r, int64_t nr,
                double* out_d2, int32_t* out_idx) {
  if (!ctx || nq < 0 || nr < 0 || (nq > 0 && (!q || !out_d2)) || (nr > 0 && !r)) return CMS_ERR_INVALID;
  if (nq > (1ll << 30) || nr > (1ll << 30)) return fail(ctx, CMS_ERR_INVALID, "cms_nearest: too many points");
  if (precision != CMS_PREC_FP64 && precision != CMS_PREC_FP32_RESCORE)
    return fail(ctx, CMS_ERR_INVALID, "cms_nearest: unsupported precision");
  CK(hipSetDevice(ctx->device));
  if (nq == 0) return CMS_OK;
  hipStream_t st = ctx->stream;
  CK(ctx->qpack.ensure(sizeof(double4) * (nq + 1)));
  CK(ctx->rpack.ensure(sizeof(double4) * (nr + 1)));
  CK(ctx->grp1.ensure(sizeof(int) * 4));
  CK(ctx->part.ensure(sizeof(int) * 1040));
  CK(ctx->frame.ensure(sizeof(double4) * 2));
  CK(ctx->nresc.ensure(sizeof(unsigned long long) * 2));
  k_pack3<<<grid_for(nq, 256, 4096), 256, 0, st>>>((int)nq, q, ctx->qpack.as<double4>());
  if (nr > 0) k_pack3<<<grid_for(nr, 256, 4096), 256, 0, st>>>((int)nr, r, ctx->rpack.as<double4>());
  int g[4] = {0, (int)nq, 0, (int)nr};
  CK(hipMemcpyAsync(ctx->grp1.p, g, sizeof g, hipMemcpyHostToDevice, st));
  if (precision == CMS_PREC_FP32_RESCORE) {
    CK(ctx->qpack32.ensure(sizeof(float4) * (nq + 1)));
    CK(ctx->rpack32.ensure(sizeof(float4) * (nr + 1)));
    k_points_frame<<<1, 256, 0, st>>>((int)nq, ctx->qpack.as<double4>(), (int)nr, ctx->rpack.as<double4>(),
                                      ctx->frame.as<double4>());
    k_to_float<<<grid_for(nq, 256, 4096), 256, 0, st>>>((int)nq, ctx->qpack.as<double4>(), nullptr, 1, 0,
                                                         ctx->frame.as<double4>(), ctx->qpack32.as<float4>());
    if (nr > 0)
      k_to_float<<<grid_for(nr, 256, 4096), 256, 0, st>>>((int)nr, ctx->rpack.as<double4>(), nullptr, 1, 0,
                                                           ctx->frame.as<double4>(), ctx->rpack32.as<float4>());
  }
  int rc = launch_nearest(ctx, precision, ctx->qpack.as<double4>(), ctx->qpack32.as<float4>(), ctx->grp1.as<int>(),
                          (int)nq, (int)nr, ctx->rpack.as<double4>(), ctx->rpack32.as<float4>(),
                          ctx->grp1.as<int>() + 2, 0, 1, out_d2, out_idx);
  if (rc) return rc;
  CK(hipStreamSynchronize(st));
  return CMS_OK;
}

int cms_search_stats(cms_ctx* ctx, uint64_t* rescored) {
  if (!ctx || !rescored) return CMS_ERR_INVALID;
  *rescored = 0;
  if (!ctx->nresc.p) return CMS_OK;
  unsigned long long v = 0;
  CK(hipMemcpyAsync(&v, ctx->nresc.p, sizeof v, hipMemcpyDeviceToHost, ctx->stream));
  CK(hipStreamSynchronize(ctx->stream));
  *rescored = v;
  fill_u64(ctx->nresc.as<unsigned long long>(), 1, 0ull, ctx->stream);
  CK(hipStreamSynchronize(ctx->stream));
  return CMS_OK;
}

int cms_profile(cms_ctx* ctx, const char* filter) {
  if (!ctx) return CMS_ERR_INVALID;
  ctx->prof_filter = filter ? filter : "";
  return CMS_OK;
}

int cms_profile_read(cms_ctx* ctx, const char* name, double* total_ms, int64_t* count) {
  if (!ctx || !total_ms || !count) return CMS_ERR_INVALID;
  CK(hipStreamSynchronize(ctx->stream));
  const std::string want = name ? name : "*";
  double tot = 0.0;
  int64_t n = 0;
  std::vector<cms_ctx::Ev> keep;
  for (auto& e : ctx->prof_events) {
    if (want == "*" || e.name == want) {
      float ms = 0.f;
      CK(hipEventElapsedTime(&ms, e.a, e.b));
      tot += ms;
      ++n;
      (void)hipEventDestroy(e.a);
      (void)hipEventDestroy(e.b);
    } else {
      keep.push_back(e);
    }
  }
  ctx->prof_events.swap(keep);
  *total_ms = tot;
  *count = n;
  return CMS_OK;
}

// ------------------------------------------------------------ backward

static int buried_groups(cms_ctx* ctx) {
  const int G = 2 * ctx->B;
  CK(ctx->bgof.ensure(sizeof(int) * (ctx->nB + 1)));
  k_group_of_items<<<grid_for(G, 1, 4096), 256, 0, ctx->stream>>>(G, ctx->bgrp.as<int>(), ctx->bgof.as<int>());
  CKL();
  return CMS_OK;
}

int cms_backward_cms(cms_ctx* ctx, const double* g_totals, const double* g_per_atom, double* grad_xyz) {
  if (!ctx || !grad_xyz) return CMS_ERR_INVALID;
  if (!ctx->cms_ready)
    return fail(ctx, CMS_ERR_STATE, "cms_backward_cms: run cms_contact_ms on the current surface first");
  CK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const int nD = ctx->nD, nB = ctx->nB;
  int r = buried_groups(ctx);
  if (r) return r;
  CK(ctx->adj.ensure(sizeof(double) * 4 * ((int64_t)nD + 1)));
  fill_double(ctx->adj.as<double>(), 4 * (int64_t)nD, 0.0, st);
  fill_double(grad_xyz, 3 * (int64_t)ctx->ntot, 0.0, st);
  if (nB > 0)
    TIMED("cms_adjoint", k_cms_adjoint<<<grid_for(nB, 256, 4096), 256, 0, st>>>(
        nB, ctx->blist.as<int>(), ctx->bgof.as<int>(), ctx->bgrp.as<int>(), ctx->bidx.as<int>(),
        ctx->bmin.as<double>(), ctx->dpos.as<double4>(), ctx->datom.as<int>(), g_totals, g_per_atom,
        ctx->s.weight, ctx->adj.as<double>()));
  if (nD > 0)
    TIMED("dot_grad", k_dot_grad<<<grid_for(nD, 64, 16384), 64, 0, st>>>(view(ctx), ctx->s, nD,
                                                                          ctx->ddesc.as<int4>(), ctx->ddesc2.as<int4>(),
                                                                          ctx->adj.as<double>(), grad_xyz));
  CKL();
  return CMS_OK;
}

__global__ void k_pack_eval(int n, const double* xyz, const double4* atom, double4* out) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    out[i] = make_double4(xyz[3 * (int64_t)i], xyz[3 * (int64_t)i + 1], xyz[3 * (int64_t)i + 2], atom[i].w);
}

int cms_frozen_cms(cms_ctx* ctx, const double* xyz_eval, double* totals) {
  if (!ctx || !xyz_eval || !totals) return CMS_ERR_INVALID;
  if (!ctx->cms_ready)
    return fail(ctx, CMS_ERR_STATE, "cms_frozen_cms: run cms_contact_ms on the current surface first");
  CK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const int n = ctx->ntot, nB = ctx->nB, G = 2 * ctx->B;
  int r = buried_groups(ctx);
  if (r) return r;
  CK(ctx->eatom.ensure(sizeof(double4) * (n + 1)));
  CK(ctx->evpos.ensure(sizeof(double4) * (nB + 1)));
  if (n > 0)
    k_pack_eval<<<grid_for(n, 256, 4096), 256, 0, st>>>(n, xyz_eval, ctx->atom.as<double4>(), ctx->eatom.as<double4>());
  CK(ctx->wv.ensure(sizeof(double) * (nB + 1)));
  if (nB > 0) {
    k_frozen_eval<<<grid_for(nB, 64, 16384), 64, 0, st>>>(view(ctx), ctx->s, ctx->eatom.as<double4>(), nB,
                                                           ctx->blist.as<int>(), ctx->ddesc.as<int4>(),
                                                           ctx->ddesc2.as<int4>(), ctx->evpos.as<double4>());
    k_frozen_cms<<<grid_for(nB, 256, 4096), 256, 0, st>>>(nB, ctx->bgof.as<int>(), ctx->bgrp.as<int>(),
                                                           ctx->bidx.as<int>(), ctx->evpos.as<double4>(),
                                                           ctx->s.weight, ctx->wv.as<double>());
  }
  // per-group sums in buried-dot order: deterministic, so finite differences
  // see no summation-order noise
  k_group_np_sum<<<grid_for(G, 1, 4096), 256, 0, st>>>(G, ctx->bgrp.as<int>(), ctx->wv.as<double>(), 1, totals);
  CKL();
  return CMS_OK;
}

int cms_check_dots(cms_ctx* ctx, int64_t* mismatches, double* maxdev) {
  if (!ctx || !mismatches || !maxdev) return CMS_ERR_INVALID;
  if (!ctx->generated) return fail(ctx, CMS_ERR_STATE, "cms_check_dots: no surface");
  CK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  CK(ctx->chk.ensure(16));
  fill_u64(ctx->chk.as<unsigned long long>(), 2, 0ull, st);
  if (ctx->nD > 0)
    k_check_dots<<<grid_for(ctx->nD, 64, 16384), 64, 0, st>>>(view(ctx), ctx->s, ctx->nD, ctx->dpos.as<double4>(),
                                                               ctx->ddesc.as<int4>(), ctx->ddesc2.as<int4>(),
                                                               ctx->chk.as<unsigned long long>(),
                                                               reinterpret_cast<double*>(ctx->chk.as<unsigned long long>() + 1));
  CKL();
  unsigned long long h[2] = {0, 0};
  CK(hipMemcpyAsync(h, ctx->chk.p, 16, hipMemcpyDeviceToHost, st));
  CK(hipStreamSynchronize(st));
  *mismatches = (int64_t)h[0];
  double d;
  memcpy(&d, &h[1], sizeof d);
  *maxdev = d;
  return CMS_OK;
}

}  // extern "C"


