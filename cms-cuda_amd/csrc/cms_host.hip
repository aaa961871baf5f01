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
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "cms_hip.h"
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

// Workgroups for n units, per_block units each, at most cap.  The default
// cap is for the unit loops (one wavefront per row, pair or probe): their
// work per unit varies widely, and with many short workgroups the
// dispatcher balances it -- 65,536 of them take 6 % off the surface passes
// of a 2,112-complex batch against 16,384, and twice as many give no more.
inline int grid_for(int64_t n, int per_block, int cap = 65536) {
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
  DBuf atten, catt, access, status, coinc, rsel, rseln, need, rpos, ratom, part;
  DBuf grid, gframe, ncellc, ncscan, acell, arank, ccount, cstart, ccur, catom, cxyz;
  DBuf nn, nbr, nb, bur, nbmax, fwd;
  DBuf pcnt, pscan, pi, pj, pu, pt, pflag, rgrp, pgrp, pstage, nps;
  DBuf prcnt, prscan, patom, ppos, palt, prgrp, nearl, nnear;
  DBuf elist, nelist, ccnt, cvf, ccf, cscan, cmask, t1cnt, t1scan, t2cnt, t2scan, tflag, tmask, tef, tes, tel, kcnt, kpos, kinfo, kscan, dbase, dss, dgrp;
  DBuf dpos, dnml, dbur, dtype, datom, ddesc, ddesc2;
  // readout state
  bool lists_ready = false;
  int nB = 0, nA = 0;
  DBuf bflag, bscan, blist, bxyz, bgrp, axyz, agrpd;
  DBuf tcnt, toff, bmin, wv, ocnt, ostart, ocur, operm, omulti, onmulti;
  DBuf trflag, trscan, tlist, txyz, tnml, tgrp, tmin, tidx, tgof, tdist, trr, side;
  DBuf qpack, rpack, grp1, gtot, npd2, npidx;
  // scan and search scratch of the SC readout when it runs beside the CMS
  // readout on a side stream (swapped in for the length of that pass)
  DBuf part_b, tcnt_b, toff_b, npd2_b, npidx_b;
  DBuf frame, bxyz32, txyz32, qpack32, rpack32, nresc;   // frame: cms_nearest's point frame
  DBuf bidx, bgof, adj, eatom, evpos, chk, qlast, tqidx;
  DBuf tgrid, tncell, tncscan, tcount, tstart, tsorted;   // trim grid
  bool cms_ready = false;
  bool sc_ready = false;            // trimmed lists of the current surface exist
  bool host_tgroups = false;        // h_tgrp holds the trimmed-list offsets
  std::vector<int> h_tgrp, h_pgrp, h_tq;
  int h_tq_c = -1;                  // complex whose toroid-queue pairs h_tq lists
  bool float_lists = false;
  // per-run count block and capacities
  DBuf meta;
  int* hmeta = nullptr;             // pinned host copy: M_WORDS meta words, then B status words
  int64_t hmeta_words = 0;
  bool host_groups = false;         // h_dgrp / h_prgrp hold the current run's offsets
  int64_t cap[M_WORDS] = {};        // element capacity of the buffers of stages R, P, Q, D
  int64_t est[M_WORDS] = {};        // counts the current run's launch shapes are sized for
  double ratio[2][M_WORDS] = {};    // largest count per atom seen (CMS mode, all-atoms mode)
  bool learned[2] = {false, false};
  uint64_t layout_epoch = 0;        // bumped when the batch layout changes
  // one captured cms_run step
  bool use_graph = false;
  bool keep_desc = true;            // generation writes dot descriptors
  bool gen_desc = false;            // the current surface has them
  // the packed per-atom ids / densities came from default densities
  bool packed_default = false;
  double packed_density = 0.0;
  // probes count pass: one wavefront per pair suits small batches (more
  // units than rows), one per neighbour row large ones (full lanes); the
  // environment variable CMS_PROBE_COUNT = "pairs" / "rows" fixes the choice
  int probe_count_mode = 0;   // 0 by size, 1 pairs, 2 rows
  // CMS_PROBE_SLOTS = 0..PSLOTS: probe slots per row the per-row pass uses
  // (fewer sends more pairs through the emit pass; same results)
  int probe_slots = PSLOTS;
  bool capturing = false;
  hipStream_t gstream = nullptr;    // stream the graph is captured on and launched to
  hipEvent_t gev0 = nullptr, gev1 = nullptr;
  // side streams of the surface stage's concurrent branches (default
  // priority) and their fork / join events
  hipStream_t sstream[2] = {nullptr, nullptr};
  hipEvent_t fev[3] = {nullptr, nullptr, nullptr};
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<uint64_t> gkey, ekey;  // key of the captured graph / of the last eager run
  int64_t n_captures = 0, n_replays = 0, n_exact = 0, n_overflow = 0;
  int n_failed = 0;                 // complexes of the last run with a non-zero status word
  int trim_lanes = 0;               // lanes per buried dot in the SC trim (0: by batch size)
  int cell_sweep = 0;               // cell sweeps: 0 by batch size, 1 wavefront per atom, 2 thread per atom
  int search_split = 0;             // nearest search: work items the reference splits aim for (0: 2048)
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
    if (c->prof_filter.empty() || c->capturing) return;
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
static const char* const NO_DESC =
    "the surface was generated without dot descriptors (cms_set_keep_descriptors(ctx, 1) before generating)";

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
  w.nR = ctx->meta.as<int>() + M_R;
  w.ratom = ctx->ratom.as<int>();
  w.rpos = ctx->rpos.as<int>();
  w.nn = ctx->nn.as<int>();
  w.nbr = ctx->nbr.as<int>();
  w.nb = ctx->nb.as<int>();
  w.bur = ctx->bur.as<int>();
  w.nbmax = ctx->nbmax.as<int>();
  w.fwd = ctx->fwd.as<uint8_t>();
  w.pcnt = ctx->pcnt.as<int>();
  w.pscan = ctx->pscan.as<int>();
  w.nP = ctx->meta.as<int>() + M_P;
  w.pi = ctx->pi.as<int>();
  w.pj = ctx->pj.as<int>();
  w.pu = ctx->pu.as<double4>();
  w.pt = ctx->pt.as<double4>();
  w.pflag = ctx->pflag.as<int8_t>();
  w.pgrp = ctx->pgrp.as<int>();
  w.prcnt = ctx->prcnt.as<int>();
  w.prscan = ctx->prscan.as<int>();
  w.nQ = ctx->meta.as<int>() + M_Q;
  w.patom = ctx->patom.as<int4>();
  w.ppos = ctx->ppos.as<double4>();
  w.palt = ctx->palt.as<double4>();
  w.prgrp = ctx->prgrp.as<int>();
  w.near = ctx->nearl.as<int>();
  w.nnear = ctx->nnear.as<int>();
  w.pstage = ctx->pstage.as<ProbeRec>();
  w.nps = ctx->nps.as<int>();
  w.pslots = ctx->probe_slots;
  w.ccnt = ctx->ccnt.as<int>();
  w.cscan = ctx->cscan.as<int>();
  w.cmask = ctx->cmask.as<uint64_t>();
  w.elist = ctx->elist.as<int>();
  w.nelist = ctx->nelist.as<int>();
  w.cvf = ctx->cvf.as<double4>();
  w.t1cnt = ctx->t1cnt.as<int>();
  w.t1scan = ctx->t1scan.as<int>();
  w.t2cnt = ctx->t2cnt.as<int>();
  w.t2scan = ctx->t2scan.as<int>();
  w.tflag = ctx->tflag.as<int8_t>();
  w.tmask = ctx->tmask.as<uint64_t>();
  w.kcnt = ctx->kcnt.as<int>();
  w.ccf = ctx->ccf.as<double4>();
  w.kpos = ctx->kpos.as<double4>();
  w.kinfo = ctx->kinfo.as<uint64_t>();
  w.kscan = ctx->kscan.as<int>();
  w.dbase = ctx->dbase.as<int64_t>();
  w.dpos = ctx->dpos.as<double4>();
  w.dnml = ctx->dnml.as<double4>();
  w.dbur = ctx->dbur.as<int8_t>();
  w.dtype = ctx->dtype.as<int8_t>();
  w.datom = ctx->datom.as<int>();
  // descriptors only when asked for (the emit passes test the pointer)
  w.ddesc = ctx->keep_desc ? ctx->ddesc.as<int4>() : nullptr;
  w.ddesc2 = ctx->keep_desc ? ctx->ddesc2.as<int4>() : nullptr;
  w.pair_cap = ctx->cap[M_P];
  w.probe_cap = ctx->cap[M_Q];
  w.dot_cap = ctx->cap[M_D];
  return w;
}

static std::vector<DBuf*> all_bufs(cms_ctx* ctx) {
  return {&ctx->atom, &ctx->dens, &ctx->acplx, &ctx->amol, &ctx->aoff, &ctx->nb0,
                  &ctx->agrp, &ctx->atten, &ctx->catt, &ctx->access, &ctx->status, &ctx->coinc,
                  &ctx->rsel, &ctx->rseln, &ctx->grid, &ctx->gframe, &ctx->ncellc,
                  &ctx->ncscan, &ctx->acell, &ctx->arank, &ctx->ccount, &ctx->cstart, &ctx->ccur, &ctx->catom, &ctx->cxyz, &ctx->need, &ctx->rpos, &ctx->ratom,
                  &ctx->part, &ctx->nn, &ctx->nbmax, &ctx->fwd, &ctx->nbr, &ctx->nb, &ctx->bur, &ctx->pcnt,
                  &ctx->pscan, &ctx->pi, &ctx->pj, &ctx->pu, &ctx->pt, &ctx->pflag, &ctx->pstage, &ctx->nps,
                  &ctx->rgrp, &ctx->pgrp, &ctx->prcnt, &ctx->prscan, &ctx->patom,
                  &ctx->ppos, &ctx->palt, &ctx->prgrp, &ctx->nearl, &ctx->nnear, &ctx->elist, &ctx->nelist, &ctx->ccnt, &ctx->cvf,
                  &ctx->cscan, &ctx->cmask, &ctx->t1cnt, &ctx->t1scan, &ctx->t2cnt, &ctx->t2scan,
                  &ctx->tflag, &ctx->tmask, &ctx->tef, &ctx->tes, &ctx->tel, &ctx->kcnt, &ctx->ccf, &ctx->kpos, &ctx->kinfo, &ctx->kscan, &ctx->dbase, &ctx->dss, &ctx->dgrp, &ctx->dpos,
                  &ctx->dnml, &ctx->dbur, &ctx->dtype, &ctx->datom, &ctx->ddesc, &ctx->ddesc2, &ctx->bflag,
                  &ctx->bscan, &ctx->blist, &ctx->bxyz, &ctx->bgrp,
                  &ctx->axyz, &ctx->agrpd, &ctx->tcnt, &ctx->toff,
                  &ctx->bmin, &ctx->wv, &ctx->ocnt, &ctx->ostart, &ctx->ocur, &ctx->operm, &ctx->omulti, &ctx->onmulti,
                  &ctx->trflag, &ctx->trscan, &ctx->tlist, &ctx->txyz,
                  &ctx->tnml, &ctx->tgrp, &ctx->tmin, &ctx->tidx, &ctx->tgof, &ctx->tdist,
                  &ctx->trr, &ctx->side, &ctx->qpack, &ctx->rpack, &ctx->part_b, &ctx->tcnt_b, &ctx->toff_b, &ctx->npd2_b, &ctx->npidx_b, &ctx->grp1, &ctx->gtot, &ctx->npd2, &ctx->npidx, &ctx->frame,
                  &ctx->bxyz32, &ctx->txyz32, &ctx->qpack32, &ctx->rpack32, &ctx->nresc, &ctx->bidx,
                  &ctx->bgof, &ctx->adj, &ctx->eatom, &ctx->evpos, &ctx->chk, &ctx->qlast, &ctx->tqidx, &ctx->tgrid, &ctx->tncell, &ctx->tncscan,
          &ctx->tcount, &ctx->tstart, &ctx->tsorted, &ctx->meta};
}

// Identity of the workspace: a captured graph holds these addresses.
static uint64_t workspace_signature(cms_ctx* ctx) {
  uint64_t h = 1469598103934665603ull;
  for (DBuf* b : all_bufs(ctx)) {
    h = (h ^ reinterpret_cast<uintptr_t>(b->p)) * 1099511628211ull;
    h = (h ^ (uint64_t)b->cap) * 1099511628211ull;
  }
  return h;
}

static void drop_graph(cms_ctx* ctx) {
  if (ctx->gexec) (void)hipGraphExecDestroy(ctx->gexec);
  if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
  ctx->gexec = nullptr;
  ctx->graph = nullptr;
  ctx->gkey.clear();
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
  if (const char* pm = std::getenv("CMS_PROBE_COUNT")) {
    if (!std::strcmp(pm, "pairs")) ctx->probe_count_mode = 1;
    if (!std::strcmp(pm, "rows")) ctx->probe_count_mode = 2;
  }
  if (const char* tl = std::getenv("CMS_TRIM_LANES")) {
    const int v = std::atoi(tl);
    if (v == 1 || v == 4 || v == 32) ctx->trim_lanes = v;
  }
  if (const char* ps = std::getenv("CMS_PROBE_SLOTS")) {
    ctx->probe_slots = std::min(std::max(std::atoi(ps), 0), PSLOTS);
  }
  ctx->stream = static_cast<hipStream_t>(stream);
  if (ctx->nresc.ensure(sizeof(unsigned long long) * 2) != hipSuccess ||
      ctx->meta.ensure(sizeof(int) * M_WORDS) != hipSuccess) {
    g_create_err = "cms_create: out of device memory";
    ctx->nresc.release();
    ctx->meta.release();
    delete ctx;
    return CMS_ERR_HIP;
  }
  {
    // device address of this device's scratch for cms_nearest_noalloc
    std::string serr;
    if (noalloc_setup(device, &serr) != CMS_OK) {
      g_create_err = serr;
      ctx->nresc.release();
      ctx->meta.release();
      delete ctx;
      return CMS_ERR_HIP;
    }
  }
  fill_u64(ctx->nresc.as<unsigned long long>(), 2, 0ull, ctx->stream);
  fill_int(ctx->meta.as<int>(), M_WORDS, 0, ctx->stream);
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
    g_create_err = "cms_create: stream synchronisation failed";
    ctx->nresc.release();
    ctx->meta.release();
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
  drop_graph(ctx);
  if (ctx->gev0) (void)hipEventDestroy(ctx->gev0);
  if (ctx->gev1) (void)hipEventDestroy(ctx->gev1);
  if (ctx->gstream) (void)hipStreamDestroy(ctx->gstream);
  for (hipEvent_t e : ctx->fev)
    if (e) (void)hipEventDestroy(e);
  for (hipStream_t q : ctx->sstream)
    if (q) (void)hipStreamDestroy(q);
  if (ctx->hmeta) (void)hipHostFree(ctx->hmeta);
  for (DBuf* b : all_bufs(ctx)) b->release();
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
  if (ctx->capturing) return fail(ctx, CMS_ERR_STATE, "cms_load: not allowed during capture");
  CK(hipSetDevice(ctx->device));
  std::vector<int> aoff(n_complex + 1, 0), nb0(n_complex, 0), agrp(2 * n_complex + 1, 0);
  int64_t tot = 0;
  for (int c = 0; c < n_complex; ++c) {
    if (n_binder[c] < 0 || n_target[c] < 0)
      return fail(ctx, CMS_ERR_INVALID, "cms_load: negative atom count");
    aoff[c] = (int)tot;
    nb0[c] = n_binder[c];
    agrp[2 * c] = (int)tot;
    agrp[2 * c + 1] = (int)(tot + n_binder[c]);
    tot += (int64_t)n_binder[c] + n_target[c];
    if (tot > (1ll << 30)) return fail(ctx, CMS_ERR_INVALID, "cms_load: too many atoms");
  }
  aoff[n_complex] = (int)tot;
  agrp[2 * n_complex] = (int)tot;
  if (tot > 0 && (!xyz || !radius)) return fail(ctx, CMS_ERR_INVALID, "cms_load: NULL coordinates");
  // the same layout as the loaded batch (same complex sizes): the device
  // offset tables stay, only the atoms are packed again
  const bool same_layout = ctx->B == n_complex && aoff == ctx->h_aoff && nb0 == ctx->h_nb0 && ctx->aoff.p;
  ctx->B = n_complex;
  ctx->ntot = (int)tot;
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
  if (!same_layout) {
    ctx->h_aoff.swap(aoff);
    ctx->h_nb0.swap(nb0);
    ctx->h_agrp.swap(agrp);
    ++ctx->layout_epoch;
    CK(ctx->aoff.ensure(sizeof(int) * (n_complex + 1)));
    CK(ctx->nb0.ensure(sizeof(int) * n_complex));
    CK(ctx->agrp.ensure(sizeof(int) * (2 * n_complex + 1)));
    CK(hipMemcpyAsync(ctx->aoff.p, ctx->h_aoff.data(), sizeof(int) * (n_complex + 1),
                      hipMemcpyHostToDevice, ctx->stream));
    CK(hipMemcpyAsync(ctx->nb0.p, ctx->h_nb0.data(), sizeof(int) * n_complex, hipMemcpyHostToDevice,
                      ctx->stream));
    CK(hipMemcpyAsync(ctx->agrp.p, ctx->h_agrp.data(), sizeof(int) * (2 * n_complex + 1),
                      hipMemcpyHostToDevice, ctx->stream));
  }
  if (n > 0) {
    // complex ids, molecules and default densities of the same layout are
    // already in place: only coordinates and radii are packed again
    const bool coords_only = same_layout && !density && ctx->packed_default && ctx->packed_density == ctx->density;
    k_pack_atoms<<<grid_for(n, 256, 4096), 256, 0, ctx->stream>>>(
        n, xyz, radius, density, ctx->density, ctx->aoff.as<int>(), ctx->nb0.as<int>(), n_complex,
        coords_only ? 0 : 1, ctx->atom.as<double4>(), ctx->dens.as<double>(), ctx->acplx.as<int>(),
        ctx->amol.as<int8_t>());
    CKL();
    ctx->packed_default = !density;
    ctx->packed_density = ctx->density;
  }
  CK(ctx->nresc.ensure(sizeof(unsigned long long) * 2));
  fill_u64(ctx->nresc.as<unsigned long long>(), 2, 0ull, ctx->stream);
  // host tables copied asynchronously: wait before they can change again
  if (!same_layout) CK(hipStreamSynchronize(ctx->stream));
  ctx->generated = false;
  ctx->gen_all_atoms = -1;
  ctx->lists_ready = false;
  ctx->float_lists = false;
  ctx->cms_ready = false;
  ctx->sc_ready = false;
  ctx->host_tgroups = false;
  ctx->h_tq_c = -1;
  ctx->host_groups = false;
  return CMS_OK;
}

// ------------------------------------------------------------ capacities
//
// A run is enqueued in one of two ways.  Exact: after each stage the host
// reads the stage total and sizes the next buffers to it (four waits).
// Learned: every buffer is sized up front from the largest count per atom
// seen so far for this mode, the whole run is enqueued without a host wait,
// and k_stage_total checks each total against its buffer on the device; a
// run that overflows is repeated exactly.  Launch shapes come from the same
// estimates, so a learned run can be captured in a HIP graph.

static int64_t headroom(int64_t v) { return v + v / 4 + 64; }

// neighbour rows from which the probes count pass runs one wavefront per row
constexpr int64_t PROBE_ROWS_MIN = 4096;

static int size_rows(cms_ctx* ctx, int64_t c) {
  CK(ctx->nn.ensure(sizeof(int) * (c + 1)));
  CK(ctx->nb.ensure(sizeof(int) * (c + 1)));
  CK(ctx->nbmax.ensure(sizeof(int) * (c + 1)));
  CK(ctx->nbr.ensure(sizeof(int) * (c * KCAP + 1)));
  CK(ctx->fwd.ensure(c * KCAP + 1));
  CK(ctx->bur.ensure(sizeof(int) * (c * BCAP + 1)));
  CK(ctx->pcnt.ensure(sizeof(int) * (c + 1)));
  CK(ctx->pscan.ensure(sizeof(int) * (c + 1)));
  CK(ctx->ccnt.ensure(sizeof(int) * (c + 1)));
  CK(ctx->cscan.ensure(sizeof(int) * (c + 1)));
  CK(ctx->cmask.ensure(sizeof(uint64_t) * (c * CMASK_WORDS + 1)));
  CK(ctx->pstage.ensure(sizeof(ProbeRec) * PSLOTS * (c + 1)));
  CK(ctx->nps.ensure(sizeof(int) * (c + 1)));
  CK(ctx->cvf.ensure(sizeof(double4) * 4 * (c + 1)));
  ctx->cap[M_R] = std::max(ctx->cap[M_R], c);
  return CMS_OK;
}

static int size_pairs(cms_ctx* ctx, int64_t c) {
  CK(ctx->pi.ensure(sizeof(int) * (c + 1)));
  CK(ctx->pj.ensure(sizeof(int) * (c + 1)));
  CK(ctx->pu.ensure(sizeof(double4) * (c + 1)));
  CK(ctx->pt.ensure(sizeof(double4) * (c + 1)));
  CK(ctx->pflag.ensure(c + 1));
  CK(ctx->prcnt.ensure(sizeof(int) * (2 * c + 1)));
  CK(ctx->prscan.ensure(sizeof(int) * (2 * c + 1)));
  CK(ctx->t1cnt.ensure(sizeof(int) * (c + 1)));
  CK(ctx->t1scan.ensure(sizeof(int) * (c + 1)));
  CK(ctx->t2cnt.ensure(sizeof(int) * (c + 1)));
  CK(ctx->t2scan.ensure(sizeof(int) * (c + 1)));
  CK(ctx->tflag.ensure(c + 1));
  CK(ctx->tmask.ensure(sizeof(uint64_t) * 2 * (c + 1)));
  CK(ctx->elist.ensure(sizeof(int) * (c + 1)));
  CK(ctx->nelist.ensure(sizeof(int)));
  CK(ctx->tef.ensure(sizeof(int) * (c + 1)));
  CK(ctx->tes.ensure(sizeof(int) * (c + 1)));
  CK(ctx->tel.ensure(sizeof(int) * (c + 1)));
  ctx->cap[M_P] = std::max(ctx->cap[M_P], c);
  return CMS_OK;
}

static int size_probes(cms_ctx* ctx, int64_t c) {
  CK(ctx->patom.ensure(sizeof(int4) * (c + 1)));
  CK(ctx->ppos.ensure(sizeof(double4) * (c + 1)));
  CK(ctx->palt.ensure(sizeof(double4) * (c + 1)));
  CK(ctx->nearl.ensure(sizeof(int) * (c * NEARCAP + 1)));
  CK(ctx->nnear.ensure(sizeof(int) * (c + 1)));
  CK(ctx->kcnt.ensure(sizeof(int) * (c + 1)));
  CK(ctx->ccf.ensure(sizeof(double4) * 7 * (c + 1)));
  CK(ctx->kpos.ensure(sizeof(double4) * KSLOTS * (c + 1)));
  CK(ctx->kinfo.ensure(sizeof(uint64_t) * KSLOTS * (c + 1)));
  CK(ctx->kscan.ensure(sizeof(int) * (c + 1)));
  ctx->cap[M_Q] = std::max(ctx->cap[M_Q], c);
  return CMS_OK;
}

// dots and every per-dot list of the readouts
static int size_dots(cms_ctx* ctx, int64_t c) {
  DBuf* d4[] = {&ctx->dpos, &ctx->dnml, &ctx->bxyz, &ctx->txyz, &ctx->tnml};
  for (DBuf* b : d4) CK(b->ensure(sizeof(double4) * (c + 1)));
  DBuf* i4[] = {&ctx->ddesc, &ctx->ddesc2};
  for (DBuf* b : i4) CK(b->ensure(sizeof(int4) * (c + 1)));
  DBuf* f4[] = {&ctx->bxyz32, &ctx->txyz32};
  for (DBuf* b : f4) CK(b->ensure(sizeof(float4) * (c + 1)));
  DBuf* i1[] = {&ctx->datom, &ctx->bflag, &ctx->bscan, &ctx->blist,
                &ctx->bidx, &ctx->operm, &ctx->trflag, &ctx->trscan, &ctx->tlist, &ctx->tidx, &ctx->tgof,
                &ctx->bgof};
  for (DBuf* b : i1) CK(b->ensure(sizeof(int) * (c + 1)));
  DBuf* f8[] = {&ctx->bmin, &ctx->wv, &ctx->tmin, &ctx->tdist, &ctx->trr};
  CK(ctx->tsorted.ensure(sizeof(double4) * (c + 1)));
  // trim-grid cells: at most 2 per accessible dot + 64 per group
  const int64_t cells = 2 * c + 65ll * 2 * ctx->B + 1;
  CK(ctx->tcount.ensure(sizeof(int) * (cells + 1)));
  CK(ctx->tstart.ensure(sizeof(int) * (cells + 1)));
  for (DBuf* b : f8) CK(b->ensure(sizeof(double) * (c + 1)));
  CK(ctx->dbur.ensure(c + 1));
  CK(ctx->dtype.ensure(c + 1));
  ctx->cap[M_D] = std::max(ctx->cap[M_D], c);
  return CMS_OK;
}

// Side streams and events of the surface stage's concurrent branches,
// created once (outside any capture: the first learned run is eager).
static int ensure_side(cms_ctx* ctx) {
  if (ctx->capturing && !(ctx->sstream[0] && ctx->sstream[1] && ctx->fev[0] && ctx->fev[1] && ctx->fev[2]))
    return CMS_ERR_STATE;
  for (hipStream_t& q : ctx->sstream)
    if (!q) CK(hipStreamCreateWithFlags(&q, hipStreamNonBlocking));
  for (hipEvent_t& e : ctx->fev)
    if (!e) CK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  return CMS_OK;
}

// Read the stage total of an exact run and size the next buffers to it.
static int exact_total(cms_ctx* ctx, int slot, int* out) {
  CK(hipMemcpyAsync(out, ctx->meta.as<int>() + slot, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  CK(hipStreamSynchronize(ctx->stream));
  return CMS_OK;
}

// Learn from a run's counts (host copy of the meta words).
static void learn(cms_ctx* ctx, int mode, const int* hm) {
  const double n = std::max(ctx->ntot, 1);
  for (int k = M_R; k <= M_T; ++k) {
    const int v = std::max(hm[k], hm[M_RAW + k]);
    ctx->ratio[mode][k] = std::max(ctx->ratio[mode][k], v / n);
  }
}

static int ensure_hmeta(cms_ctx* ctx) {
  const int64_t words = M_WORDS + ctx->B;
  if (ctx->hmeta && ctx->hmeta_words >= words) return CMS_OK;
  if (ctx->hmeta) CK(hipHostFree(ctx->hmeta));
  ctx->hmeta = nullptr;
  ctx->hmeta_words = 0;
  void* p = nullptr;
  CK(hipHostMalloc(&p, sizeof(int) * (words + words / 4 + 64)));
  ctx->hmeta = static_cast<int*>(p);
  ctx->hmeta_words = words + words / 4 + 64;
  return CMS_OK;
}

static const int CMS_RETRY = -1;  // internal: a learned run overflowed a buffer

// Copy the count block and the status words to the host (one wait) and
// evaluate them.
static int finish_run(cms_ctx* ctx, hipStream_t st, int mode) {
  int r = ensure_hmeta(ctx);
  if (r) return r;
  CK(hipMemcpyAsync(ctx->hmeta, ctx->meta.p, sizeof(int) * M_WORDS, hipMemcpyDeviceToHost, st));
  CK(hipMemcpyAsync(ctx->hmeta + M_WORDS, ctx->status.p, sizeof(int) * ctx->B, hipMemcpyDeviceToHost, st));
  CK(hipStreamSynchronize(st));
  const int* hm = ctx->hmeta;
  learn(ctx, mode, hm);
  if (hm[M_OVF]) {
    ++ctx->n_overflow;
    return CMS_RETRY;
  }
  ctx->nR = hm[M_R];
  ctx->nP = hm[M_P];
  ctx->nQ = hm[M_Q];
  ctx->nD = hm[M_D];
  // a failed complex (coincident atoms, a row over capacity) has empty rows
  // and zeroed results, the rest of the batch is unaffected; the run fails
  // only when every complex did
  int worst = 0, first = -1, coincident = -1;
  ctx->n_failed = 0;
  for (int c = 0; c < ctx->B; ++c) {
    const int stc = hm[M_WORDS + c];
    if (!stc) continue;
    ++ctx->n_failed;
    if (first < 0) first = c;
    if ((stc & ST_COINCIDENT) && coincident < 0) coincident = c;
    worst |= stc;
  }
  if (ctx->B > 0 && ctx->n_failed == ctx->B) {
    if (coincident >= 0) return fail(ctx, CMS_ERR_COINCIDENT, "coincident atoms in complex %d", coincident);
    return fail(ctx, CMS_ERR_CAPACITY, "row capacity exceeded (status %d): more than %d neighbours, "
                "%d buried partners or %d near low probes", worst, KCAP, BCAP, NEARCAP);
  }
  return CMS_OK;
}

// Launch shapes and buffer sizes of a learned run.
static int plan_learned(cms_ctx* ctx, int mode) {
  const int n = ctx->ntot;
  int64_t* est = ctx->est;
  for (int k = M_R; k <= M_T; ++k) est[k] = (int64_t)std::ceil(ctx->ratio[mode][k] * n);
  for (int k = M_R; k <= M_D; ++k) est[k] = headroom(est[k]);
  est[M_R] = std::min<int64_t>(est[M_R], n);
  int r;
  if ((r = size_rows(ctx, est[M_R])) || (r = size_pairs(ctx, est[M_P])) || (r = size_probes(ctx, est[M_Q])) ||
      (r = size_dots(ctx, est[M_D])))
    return r;
  return CMS_OK;
}

static int generate_pass(cms_ctx* ctx, int all_atoms, int full_state, bool exact) {
  hipStream_t st = ctx->stream;
  const int B = ctx->B, G = 2 * B, n = ctx->ntot;
  const int mode = all_atoms ? 1 : 0;
  BatchView b = view(ctx);
  const Settings s = ctx->s;
  ctx->lists_ready = false;
  ctx->float_lists = false;
  ctx->cms_ready = false;
  ctx->host_groups = false;

  CK(ctx->atten.ensure(n + 1));
  CK(ctx->catt.ensure(n + 1));
  CK(ctx->access.ensure(n + 1));
  CK(ctx->status.ensure(sizeof(int) * B));
  CK(ctx->coinc.ensure(sizeof(unsigned long long) * B));
  CK(ctx->need.ensure(sizeof(int) * (n + 1)));
  CK(ctx->rsel.ensure(sizeof(int) * (n + 1)));
  CK(ctx->rseln.ensure(sizeof(int)));
  CK(ctx->rpos.ensure(sizeof(int) * (n + 1)));
  CK(ctx->ratom.ensure(sizeof(int) * (n + 1)));
  CK(ctx->part.ensure(sizeof(int) * 1040));
  CK(ctx->part_b.ensure(sizeof(int) * 1040));
  CK(ctx->rgrp.ensure(sizeof(int) * (G + 1)));
  CK(ctx->pgrp.ensure(sizeof(int) * (G + 1)));
  CK(ctx->prgrp.ensure(sizeof(int) * (G + 1)));
  CK(ctx->dgrp.ensure(sizeof(int) * (G + 1)));
  CK(ctx->dbase.ensure(sizeof(int64_t) * 4 * G));
  CK(ctx->dss.ensure(sizeof(int4) * 2 * G));
  // two slots (molecule 0, molecule 1) per cell
  const int64_t cell_cap = 2 * (8ll * n + 64ll * B);
  CK(ctx->grid.ensure(sizeof(GridInfo) * (B + 1)));
  CK(ctx->gframe.ensure(sizeof(double4) * (B + 1)));
  CK(ctx->ncellc.ensure(sizeof(int) * (B + 1)));
  CK(ctx->ncscan.ensure(sizeof(int) * (B + 1)));
  CK(ctx->acell.ensure(sizeof(int) * (n + 1)));
  CK(ctx->arank.ensure(sizeof(int) * (n + 1)));
  CK(ctx->catom.ensure(sizeof(int) * (n + 1)));
  CK(ctx->cxyz.ensure(sizeof(double4) * (n + 1)));
  CK(ctx->ccount.ensure(sizeof(int) * (cell_cap + 1)));
  CK(ctx->cstart.ensure(sizeof(int) * (cell_cap + 1)));
  CK(ctx->ccur.ensure(sizeof(int) * (cell_cap + 1)));
  int* part = ctx->part.as<int>();
  int* meta = ctx->meta.as<int>();
  int64_t* est = ctx->est;
  int r;
  // overflow checks of an exact run never fire: it sizes after each total
  const int64_t NOLIMIT = INT64_MAX;
  if (!exact && (r = plan_learned(ctx, mode))) return r;

  fill_batch({fill_of(ctx->status.as<int>(), B, 0), fill_of(ctx->coinc.as<unsigned long long>(), B, ~0ull),
              fill_of(meta, M_WORDS, 0)},
             st);
  if (n > 0) {
    // cell lists: per-complex grids, atoms sorted by cell
    GridInfo* grid = ctx->grid.as<GridInfo>();
    TIMED("grid", k_grid_setup<<<grid_for(B, 1, 4096), 256, 0, st>>>(B, ctx->aoff.as<int>(), ctx->atom.as<double4>(), s,
                                                                      grid, ctx->gframe.as<double4>(),
                                                                      ctx->ncellc.as<int>()));
    scan_excl(ctx->ncellc.as<int>(), ctx->ncscan.as<int>(), B, part, st);
    k_grid_offsets<<<grid_for(B, 256, 4096), 256, 0, st>>>(B, ctx->ncscan.as<int>(), grid);
    const Count cCells = dev_count(ctx->ncscan.as<int>() + B, 2);   // two slots per cell
    fill_batch({fill_of(ctx->ccount.as<int>(), cCells, cell_cap, 0), fill_of(ctx->ccur.as<int>(), cCells, cell_cap, 0)},
               st);
    // small batches: cell lists thread per atom and one wavefront per atom
    // in the sweeps (parallelism); large batches: cell lists workgroup per
    // complex (LDS counters), one thread per cell-sorted atom in the
    // attention sweep (cell data shared within a wavefront) and one
    // wavefront per listed non-blocker in the region sweep
    const bool per_thread = ctx->cell_sweep ? ctx->cell_sweep == 2 : n >= 200000;
    if (per_thread) {
      TIMED("cell_lists", k_cell_build<<<grid_for(B, 1, 65536), 512, 0, st>>>(
                              b, grid, ctx->ccount.as<int>(), ctx->acell.as<int>(), ctx->arank.as<int>(),
                              ctx->cstart.as<int>(), ctx->catom.as<int>(), ctx->cxyz.as<double4>()));
    } else {
      k_cell_count_atoms<<<grid_for(n, 256, 4096), 256, 0, st>>>(b, grid, ctx->acell.as<int>(),
                                                                 ctx->ccount.as<int>());
      scan_excl(ctx->ccount.as<int>(), ctx->cstart.as<int>(), cCells, cell_cap, part, st);
      k_cell_scatter_atoms<<<grid_for(n, 256, 4096), 256, 0, st>>>(n, ctx->acell.as<int>(), ctx->cstart.as<int>(),
                                                                   ctx->ccur.as<int>(), ctx->catom.as<int>(),
                                                                   ctx->atom.as<double4>(), ctx->cxyz.as<double4>());
    }
    const int* cst = ctx->cstart.as<int>();
    const int* cat = ctx->catom.as<int>();
    const double4* cxyz = ctx->cxyz.as<double4>();
    // every atom is in the region in all-atoms / full-state runs
    const int full = (full_state || all_atoms) ? 1 : 0;
    // the attention passes store only the non-blockers over these
    // (and the count of the region pass's list of non-blockers)
    fill_batch({fill_of(ctx->atten.as<int8_t>(), n, (int8_t)(all_atoms ? ATTEN_BURIED : ATTEN_BLOCKER)),
                fill_of(ctx->catt.as<int8_t>(), n, (int8_t)(all_atoms ? ATTEN_BURIED : ATTEN_BLOCKER)),
                fill_of(ctx->access.as<int8_t>(), n, (int8_t)0), fill_of(ctx->need.as<int>(), (int64_t)n, full),
                fill_of(ctx->rseln.as<int>(), 1, 0)},
               st);
    if (per_thread) {
      // the attention pass lists the non-blockers' cell slots (a few per
      // cent of the atoms) for the region pass
      int* list = full ? nullptr : ctx->rsel.as<int>();
      TIMED("attention", k_attention_cells_t<<<grid_for(n, 256, 65536), 256, 0, st>>>(
          b, s, all_atoms, full, grid, cst, cat, cxyz, ctx->atten.as<int8_t>(), ctx->catt.as<int8_t>(),
          ctx->access.as<int8_t>(), ctx->need.as<int>(), ctx->coinc.as<unsigned long long>(), ctx->status.as<int>(),
          list, ctx->rseln.as<int>()));
      if (!full)
        TIMED("region", k_region_cells<<<grid_for(n, 16, 65536), 64, 0, st>>>(
            b, s, grid, cst, cat, cxyz, ctx->atten.as<int8_t>(), list, ctx->rseln.as<int>(), ctx->need.as<int>()));
    } else {
      TIMED("attention", k_attention_cells<<<grid_for(n, 1, 65536), 64, 0, st>>>(
          b, s, all_atoms, full, grid, cst, cat, cxyz, ctx->atten.as<int8_t>(), ctx->catt.as<int8_t>(),
          ctx->access.as<int8_t>(), ctx->need.as<int>(), ctx->coinc.as<unsigned long long>(), ctx->status.as<int>()));
      if (!full)
        TIMED("region", k_region_cells<<<grid_for(n, 1, 65536), 64, 0, st>>>(
            b, s, grid, cst, cat, cxyz, ctx->atten.as<int8_t>(), nullptr, nullptr, ctx->need.as<int>()));
    }
    scan_excl(ctx->need.as<int>(), ctx->rpos.as<int>(), n, part, st);
    k_compact_index<<<grid_for(n, 256, 4096), 256, 0, st>>>(n, ctx->need.as<int>(),
                                                             ctx->rpos.as<int>(), ctx->ratom.as<int>());
  } else {
    fill_int(ctx->rpos.as<int>(), 1, 0, st);
  }
  CKL();
  // ---- region rows
  k_stage_total<<<1, 256, 0, st>>>(ctx->rpos.as<int>(), host_count(n), exact ? NOLIMIT : ctx->cap[M_R], meta, M_R,
                                    nullptr, 0);
  CKL();
  if (exact) {
    int v = 0;
    if ((r = exact_total(ctx, M_R, &v))) return r;
    est[M_R] = v;
    if ((r = size_rows(ctx, headroom(v)))) return r;
  }
  const Count cR = dev_count(meta + M_R), cP = dev_count(meta + M_P), cQ = dev_count(meta + M_Q);
  Work w = work(ctx);
  if (est[M_R] > 0) {
    TIMED("neighbors", k_neighbors_cells<<<grid_for(est[M_R], 1), 64, 0, st>>>(b, s, w, ctx->grid.as<GridInfo>(),
                                                                                ctx->cstart.as<int>(),
                                                                                ctx->catom.as<int>(),
                                                                                ctx->cxyz.as<double4>(),
                                                                                ctx->catt.as<int8_t>()));
    // (the neighbour pass counted each row's pairs)
    k_drop_failed_rows<<<grid_for(est[M_R], 256, 4096), 256, 0, st>>>(cR, ctx->ratom.as<int>(), ctx->acplx.as<int>(),
                                                                      ctx->status.as<int>(), ctx->nn.as<int>(),
                                                                      ctx->nb.as<int>(), ctx->pcnt.as<int>());
    CKL();
  }
  scan_excl(ctx->pcnt.as<int>(), ctx->pscan.as<int>(), cR, ctx->cap[M_R], part, st);
  k_group_starts_from_atoms<<<grid_for(est[M_R] + 1, 256, 4096), 256, 0, st>>>(
      cR, ctx->ratom.as<int>(), ctx->acplx.as<int>(), ctx->amol.as<int8_t>(), G, ctx->rgrp.as<int>());
  k_gather_scan<<<grid_for(G + 1, 256, 4096), 256, 0, st>>>(G, ctx->rgrp.as<int>(), 1,
                                                             ctx->pscan.as<int>(), ctx->pgrp.as<int>());
  // ---- pairs
  k_stage_total<<<1, 256, 0, st>>>(ctx->pscan.as<int>(), cR, exact ? NOLIMIT : ctx->cap[M_P], meta, M_P,
                                    ctx->pgrp.as<int>(), G);
  CKL();
  if (exact) {
    int v = 0;
    if ((r = exact_total(ctx, M_P, &v))) return r;
    est[M_P] = v;
    if ((r = size_pairs(ctx, headroom(v)))) return r;
    w = work(ctx);
  }
  bool fork = false;   // the surface stage's branches run on side streams
  // large batches count probes and torus points per neighbour row, and that
  // pass writes the pair records itself
  const bool by_row = ctx->probe_count_mode == 2 || (ctx->probe_count_mode == 0 && est[M_R] >= PROBE_ROWS_MIN);
  if (est[M_R] > 0 && !by_row) TIMED("pairs_emit", k_pairs<<<grid_for(est[M_R], 1), 64, 0, st>>>(b, s, w));
  if (est[M_P] > 0) {
    if (by_row) {
      // the row pass writes the toroidal counts of the live toroid-queue
      // pairs only
      fill_batch({fill_of(ctx->nps.as<int>(), cR, ctx->cap[M_R], 0), fill_of(ctx->nelist.as<int>(), 1, 0),
                  fill_of(ctx->t1cnt.as<int>(), cP, ctx->cap[M_P], 0),
                  fill_of(ctx->t2cnt.as<int>(), cP, ctx->cap[M_P], 0),
                  fill_of(ctx->tflag.as<int8_t>(), ctx->cap[M_P], 0)},
                 st);
      // rows of at most RC_SMALL neighbours, then the longer ones (LDS staging
      // sized for each); the pass also takes the torus points of the live
      // toroid-queue pairs
      TIMED("probes_count", (k_probe_rows<RC_SMALL><<<grid_for(est[M_R], 1), 64, 0, st>>>(b, s, w)));
      TIMED("probes_count_long", (k_probe_rows<KCAP><<<grid_for(est[M_R], 64, 2048), 64, 0, st>>>(b, s, w)));
      // From here two branches touch disjoint data: the convex counts (side
      // stream 1, below) and the probe placement, near lists and concave
      // counts (this stream); they join before the dot scans.  Run
      // concurrently, each fills the register and LDS room the other's
      // wavefronts leave on a CU.  Learned runs only (an exact run waits on
      // the host between stages), and not while kernels are timed one by one.
      fork = !exact && ctx->prof_filter.empty() && ensure_side(ctx) == CMS_OK;
      if (fork) {
        CK(hipEventRecord(ctx->fev[0], st));
        CK(hipStreamWaitEvent(ctx->sstream[1], ctx->fev[0], 0));
      }
    }
    else
      TIMED("probes_count", k_probes<false><<<grid_for(est[M_P], 1), 64, 0, st>>>(b, s, w));
  }
  CKL();
  scan_excl(ctx->prcnt.as<int>(), ctx->prscan.as<int>(), dev_count(meta + M_P, 2), 2 * ctx->cap[M_P], part, st);
  k_gather_scan<<<grid_for(G + 1, 256, 4096), 256, 0, st>>>(G, ctx->pgrp.as<int>(), 2,
                                                             ctx->prscan.as<int>(), ctx->prgrp.as<int>());
  // ---- probes
  k_stage_total<<<1, 256, 0, st>>>(ctx->prscan.as<int>(), dev_count(meta + M_P, 2),
                                    exact ? NOLIMIT : ctx->cap[M_Q], meta, M_Q, ctx->prgrp.as<int>(), G);
  CKL();
  if (exact) {
    int v = 0;
    if ((r = exact_total(ctx, M_Q, &v))) return r;
    est[M_Q] = v;
    if ((r = size_probes(ctx, headroom(v)))) return r;
    w = work(ctx);
  }
  // the row pass staged the probes of large batches (pairs it could not
  // stage keep flag bit 8 for the emit pass)
  if (est[M_P] > 0) {
    // the row pass listed those pairs; the per-pair count pass only flags them
    Work we = w;
    if (!by_row) we.elist = nullptr;
    TIMED("probes_emit", if (by_row) k_probe_place<<<grid_for(est[M_R] * PSLOTS, 256, 16384), 256, 0, st>>>(b, w);
                         k_probes<true><<<by_row ? grid_for(est[M_P] / 256 + 1, 1, 4096) : grid_for(est[M_P], 1), 64,
                                          0, st>>>(b, s, we));
  }
  if (est[M_Q] > 0) TIMED("near", k_near<<<grid_for(est[M_Q], 1), 64, 0, st>>>(b, s, w));
  CKL();
  // ---- dot counts
  if (est[M_R] > 0) {
    hipStream_t cst_ = fork ? ctx->sstream[1] : st;
    fill_int(ctx->ccnt.as<int>(), cR, ctx->cap[M_R], 0, cst_);
    TIMED("convex_frames", k_convex_frames<<<grid_for(est[M_R], 256, 8192), 256, 0, cst_>>>(b, s, w));
    // rows of at most RC_SMALL neighbours, then the longer ones
    TIMED("convex_count", (k_convex<false, RC_SMALL><<<grid_for(est[M_R], 1), 64, 0, cst_>>>(b, s, w)));
    TIMED("convex_count_long", (k_convex<false, KCAP><<<grid_for(est[M_R], 64, 2048), 64, 0, cst_>>>(b, s, w)));
  }
  if (fork) CK(hipEventRecord(ctx->fev[2], ctx->sstream[1]));
  if (est[M_P] > 0 && !by_row) {
    // the count pass visits only live toroid-queue pairs
    fill_int(ctx->t1cnt.as<int>(), ctx->cap[M_P], 0, st);
    fill_int(ctx->t2cnt.as<int>(), ctx->cap[M_P], 0, st);
    fill_int8(ctx->tflag.as<int8_t>(), ctx->cap[M_P], 0, st);
    TIMED("toroid_count", k_toroid_count<<<grid_for(est[M_P], 1), 64, 0, st>>>(b, s, w));
  }
  if (est[M_Q] > 0) {
    if (s.rp > 0.0) {
      TIMED("concave_frames", k_concave_frames<<<grid_for(est[M_Q], 256, 8192), 256, 0, st>>>(b, s, w));
      TIMED("concave_count", k_concave<false><<<grid_for(est[M_Q], 1), 64, 0, st>>>(b, s, w));
    } else {
      fill_int(ctx->kcnt.as<int>(), ctx->cap[M_Q], 0, st);
    }
  }
  if (fork) CK(hipStreamWaitEvent(st, ctx->fev[2], 0));
  CKL();
  scan_excl(ctx->ccnt.as<int>(), ctx->cscan.as<int>(), cR, ctx->cap[M_R], part, st);
  scan_excl(ctx->t1cnt.as<int>(), ctx->t1scan.as<int>(), cP, ctx->cap[M_P], part, st);
  scan_excl(ctx->t2cnt.as<int>(), ctx->t2scan.as<int>(), cP, ctx->cap[M_P], part, st);
  scan_excl(ctx->kcnt.as<int>(), ctx->kscan.as<int>(), cQ, ctx->cap[M_Q], part, st);
  k_dot_stage_starts<<<grid_for(G, 256, 4096), 256, 0, st>>>(G, ctx->rgrp.as<int>(), ctx->pgrp.as<int>(),
                                                              ctx->prgrp.as<int>(), ctx->cscan.as<int>(),
                                                              ctx->t1scan.as<int>(), ctx->t2scan.as<int>(),
                                                              ctx->kscan.as<int>(), ctx->dss.as<int4>());
  k_dot_bases<<<1, DOT_BASES_THREADS, 0, st>>>(G, ctx->dss.as<int4>(), ctx->dbase.as<int64_t>(), ctx->dgrp.as<int>());
  // ---- dots
  k_stage_total<<<1, 256, 0, st>>>(ctx->dgrp.as<int>(), host_count(G), exact ? NOLIMIT : ctx->cap[M_D], meta, M_D,
                                    ctx->dgrp.as<int>(), G);
  CKL();
  if (exact) {
    int v = 0;
    if ((r = exact_total(ctx, M_D, &v))) return r;
    est[M_D] = v;
    if ((r = size_dots(ctx, headroom(v)))) return r;
    w = work(ctx);
    // readout launch shapes: the previous run's counts per atom, or the dots
    for (int k = M_B; k <= M_T; ++k) {
      est[k] = (int64_t)std::ceil(ctx->ratio[mode][k] * n);
      if (est[k] == 0) est[k] = v / 2;
    }
  }
  // the three emit passes write disjoint stretches of the dot arrays (each
  // stage at its own group bases): on large batches the convex dots go to
  // side stream 0, the toroidal ones (with their pair list and its scan) to
  // side stream 1, the concave ones stay here, and this stream waits for
  // both before the readouts
  const bool fork_emit = by_row && ctx->prof_filter.empty() && ensure_side(ctx) == CMS_OK;
  if (fork_emit) {
    CK(hipEventRecord(ctx->fev[0], st));
    CK(hipStreamWaitEvent(ctx->sstream[0], ctx->fev[0], 0));
    CK(hipStreamWaitEvent(ctx->sstream[1], ctx->fev[0], 0));
  }
  const hipStream_t cvs = fork_emit ? ctx->sstream[0] : st, tos = fork_emit ? ctx->sstream[1] : st;
  if (est[M_R] > 0) {
    TIMED("convex_emit", (k_convex<true, RC_SMALL><<<grid_for(est[M_R], 1), 64, 0, cvs>>>(b, s, w)));
    TIMED("convex_emit_long", (k_convex<true, KCAP><<<grid_for(est[M_R], 64, 2048), 64, 0, cvs>>>(b, s, w)));
  }
  if (est[M_P] > 0) {
    // the pairs with toroidal dots, then their dots in batches of pairs
    const int gP = grid_for(est[M_P], 256, 4096);
    k_toroid_dot_pairs<<<gP, 256, 0, tos>>>(cP, ctx->t1cnt.as<int>(), ctx->t2cnt.as<int>(), ctx->tef.as<int>());
    scan_excl(ctx->tef.as<int>(), ctx->tes.as<int>(), cP, ctx->cap[M_P], part, tos);
    k_compact_count<<<gP, 256, 0, tos>>>(cP, ctx->tef.as<int>(), ctx->tes.as<int>(), ctx->tel.as<int>());
    k_copy_word<<<1, 64, 0, tos>>>(ctx->tes.as<int>(), cP, meta + M_TE);
    TIMED("toroid_emit", k_toroid_emit<<<grid_for(est[M_P] / 16 + 1, 1), 64, 0, tos>>>(b, s, w, ctx->tel.as<int>(),
                                                                                       meta + M_TE));
  }
  if (est[M_Q] > 0 && s.rp > 0.0) {
    TIMED("concave_emit", (k_concave_place<<<grid_for(est[M_Q] * KSLOTS, 256, 8192), 256, 0, st>>>(b, s, w),
                           k_concave<true><<<grid_for(est[M_Q] / 64 + 1, 1), 64, 0, st>>>(b, s, w)));
  }
  if (fork_emit) {
    CK(hipEventRecord(ctx->fev[1], ctx->sstream[0]));
    CK(hipEventRecord(ctx->fev[2], ctx->sstream[1]));
    CK(hipStreamWaitEvent(st, ctx->fev[1], 0));
    CK(hipStreamWaitEvent(st, ctx->fev[2], 0));
  }
  CKL();
  ctx->generated = true;
  ctx->gen_all_atoms = mode;
  ctx->gen_desc = ctx->keep_desc;
  if (exact) {
    ++ctx->n_exact;
    ctx->learned[mode] = true;
  }
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
  if (ctx->capturing) return fail(ctx, CMS_ERR_STATE, "cms_generate: not allowed during capture");
  CK(hipSetDevice(ctx->device));
  const int mode = all_atoms ? 1 : 0;
  int r = generate_pass(ctx, all_atoms, full_state, !ctx->learned[mode]);
  if (!r) r = finish_run(ctx, ctx->stream, mode);
  if (r == CMS_RETRY) {
    r = generate_pass(ctx, all_atoms, full_state, true);
    if (!r) r = finish_run(ctx, ctx->stream, mode);
    if (r == CMS_RETRY) r = fail(ctx, CMS_ERR_HIP, "cms_generate: exact run overflowed");
  }
  if (r) {
    ctx->generated = false;
    ctx->gen_all_atoms = -1;
  }
  return r;
}

int cms_status(cms_ctx* ctx, int32_t* status_host) {
  if (!ctx || !status_host) return CMS_ERR_INVALID;
  CK(hipMemcpyAsync(status_host, ctx->status.p, sizeof(int) * ctx->B, hipMemcpyDeviceToHost, ctx->stream));
  CK(hipStreamSynchronize(ctx->stream));
  return CMS_OK;
}

int cms_set_tuning(cms_ctx* ctx, const cms_tuning* t) {
  if (!ctx || !t) return CMS_ERR_INVALID;
  if (!(t->trim_lanes == 0 || t->trim_lanes == 1 || t->trim_lanes == 4 || t->trim_lanes == 32) ||
      t->probe_count < 0 || t->probe_count > 2 || t->probe_slots < 0 || t->probe_slots > PSLOTS ||
      t->cell_sweep < 0 || t->cell_sweep > 2 || t->search_split < 0 || t->search_split > 65536)
    return fail(ctx, CMS_ERR_INVALID, "cms_set_tuning: value out of range");
  ctx->trim_lanes = t->trim_lanes;
  ctx->probe_count_mode = t->probe_count;
  ctx->probe_slots = t->probe_slots ? t->probe_slots : PSLOTS;
  ctx->cell_sweep = t->cell_sweep;
  ctx->search_split = t->search_split;
  return CMS_OK;
}

int cms_get_tuning(cms_ctx* ctx, cms_tuning* t) {
  if (!ctx || !t) return CMS_ERR_INVALID;
  t->trim_lanes = ctx->trim_lanes;
  t->probe_count = ctx->probe_count_mode;
  t->probe_slots = ctx->probe_slots;
  t->cell_sweep = ctx->cell_sweep;
  t->search_split = ctx->search_split;
  return CMS_OK;
}

int cms_failed_count(cms_ctx* ctx, int32_t* n_failed) {
  if (!ctx || !n_failed) return CMS_ERR_INVALID;
  *n_failed = ctx->n_failed;
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
  const int G = 2 * ctx->B;
  const int64_t capD = ctx->cap[M_D], estD = ctx->est[M_D];
  int* part = ctx->part.as<int>();
  int* meta = ctx->meta.as<int>();
  const Count cD = dev_count(meta + M_D);
  CK(ctx->bgrp.ensure(sizeof(int) * (G + 1)));
  CK(ctx->agrpd.ensure(sizeof(int) * (G + 1)));
  if (estD > 0)
    k_dot_flag<<<grid_for(estD, 256, 4096), 256, 0, st>>>(cD, ctx->dbur.as<int8_t>(), 1, ctx->bflag.as<int>());
  scan_excl(ctx->bflag.as<int>(), ctx->bscan.as<int>(), cD, capD, part, st);
  k_gather_scan<<<grid_for(G + 1, 256, 4096), 256, 0, st>>>(G, ctx->dgrp.as<int>(), 1, ctx->bscan.as<int>(),
                                                             ctx->bgrp.as<int>());
  k_gather_compl<<<grid_for(G + 1, 256, 4096), 256, 0, st>>>(G, ctx->dgrp.as<int>(), ctx->bscan.as<int>(),
                                                              ctx->agrpd.as<int>());
  k_list_totals<<<1, 64, 0, st>>>(ctx->bscan.as<int>(), meta);
  if (estD > 0)
    k_dot_split<<<grid_for(estD, 256, 4096), 256, 0, st>>>(cD, ctx->dbur.as<int8_t>(), ctx->bscan.as<int>(),
                                                            ctx->dpos.as<double4>(), ctx->blist.as<int>(),
                                                            ctx->bxyz.as<double4>());
  CKL();
  ctx->lists_ready = true;
  return CMS_OK;
}

// nearest search of query groups qgrp over reference groups rgrp (g ^ xmask).
// nq is read on the device; qcap bounds it.  Small query sets split every
// group's references over several workgroups so that a single complex
// still spreads over the GPU (the split comes from the expected sizes
// nq_hint / nr_hint; any split gives the same result).
static int launch_nearest(cms_ctx* ctx, int precision, const double4* Q, const float4* Q32, const int* qgrp,
                          Count nq, int64_t qcap, int64_t nq_hint, int64_t nr_hint, const double4* R,
                          const float4* R32, const int* rgrp, int xmask, int G, const double4* frame,
                          double* out_d2, int* out_idx) {
  hipStream_t st = ctx->stream;
  CK(ctx->tcnt.ensure(sizeof(int) * (G + 1)));
  CK(ctx->toff.ensure(sizeof(int) * (G + 1)));
  // queries per tile: 128 for the FP32 search (two per thread), 256 for FP64
  const int tq = precision == CMS_PREC_FP32_RESCORE ? NT32 : NT64;
  k_tile_counts<<<grid_for(G, 256, 4096), 256, 0, st>>>(G, qgrp, tq, ctx->tcnt.as<int>());
  scan_excl(ctx->tcnt.as<int>(), ctx->toff.as<int>(), G, ctx->part.as<int>(), st);
  const int64_t tiles_est = nq_hint / tq + G;
  const int64_t target = ctx->search_split ? ctx->search_split : 2048;
  int nsplit = 1;
  if (tiles_est < target) {
    nsplit = (int)std::min<int64_t>((target + tiles_est - 1) / tiles_est, 64);
    // at least 128 references per split on average
    const int64_t per_group_refs = std::max<int64_t>(1, nr_hint / std::max(G, 1));
    nsplit = (int)std::max<int64_t>(1, std::min<int64_t>(nsplit, per_group_refs / 128));
  }
  const int grid = grid_for(tiles_est * nsplit, 1, 16384);
  const int64_t ps = nsplit == 1 ? 0 : qcap;
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
      // 64 threads per block, two queries per thread (128-query tiles)
      TIMED("nearest32", k_nearest32<true><<<grid, TPB32, 0, st>>>(Q32, Q, qgrp, R32, R, rgrp, xmask, toff, G, nsplit, ps,
                                                                 frame, d2o, ido, nr));
    else
      TIMED("nearest32", k_nearest32<false><<<grid, TPB32, 0, st>>>(Q32, Q, qgrp, R32, R, rgrp, xmask, toff, G, nsplit, ps,
                                                                  frame, d2o, nullptr, nr));
  } else {
    if (ido)
      TIMED("nearest", k_nearest64<true><<<grid, NT64, 0, st>>>(Q, qgrp, R, rgrp, xmask, toff, G, nsplit, ps, d2o, ido));
    else
      TIMED("nearest", k_nearest64<false><<<grid, NT64, 0, st>>>(Q, qgrp, R, rgrp, xmask, toff, G, nsplit, ps, d2o, nullptr));
  }
  if (nsplit > 1)
    k_nearest_combine<<<grid_for(std::max<int64_t>(nq_hint, 1), 256, 4096), 256, 0, st>>>(nq, nsplit, ps, d2o, ido,
                                                                                           out_d2, out_idx);
  CKL();
  return CMS_OK;
}

// float, recentred copies of the buried / accessible dot lists
static int prepare_float_lists(cms_ctx* ctx) {
  if (ctx->float_lists) return CMS_OK;
  hipStream_t st = ctx->stream;
  const int B = ctx->B, G = 2 * B;
  const int* meta = ctx->meta.as<int>();
  CK(ctx->nresc.ensure(sizeof(unsigned long long) * 2));
  // the per-complex frames come from the grid setup of the generation
  if (ctx->est[M_D] > 0) {
    k_to_float<<<grid_for(ctx->est[M_B] + 1, 256, 4096), 256, 0, st>>>(
        dev_count(meta + M_B), ctx->bxyz.as<double4>(), ctx->bgrp.as<int>(), G, 1, ctx->gframe.as<double4>(),
        ctx->bxyz32.as<float4>());
  }
  CKL();
  ctx->float_lists = true;
  return CMS_OK;
}

__global__ void k_pick_stride(int n, const double* in, int stride, double* out) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    out[i] = in[(int64_t)i * stride];
}

// per-atom sums in item order + per-group np.sum totals
static int owner_sums(cms_ctx* ctx, Count nitems, int64_t est_items, const int* list, const double* val,
                      int64_t vstride, double* per_atom, double* totals, int G_tot, int gstride) {
  hipStream_t st = ctx->stream;
  const int n = ctx->ntot;
  CK(ctx->ocnt.ensure(sizeof(int) * (n + 1)));
  CK(ctx->ostart.ensure(sizeof(int) * (n + 1)));
  CK(ctx->ocur.ensure(sizeof(int) * (n + 1)));
  CK(ctx->omulti.ensure(sizeof(int) * (n + 1)));
  CK(ctx->onmulti.ensure(sizeof(int)));
  fill_batch({fill_of(ctx->ocnt.as<int>(), n, 0), fill_of(ctx->ocur.as<int>(), n, 0),
              fill_of(ctx->onmulti.as<int>(), 1, 0)},
             st);
  const int ig = grid_for(std::max<int64_t>(est_items, 1), 256, 4096);
  k_owner_count<<<ig, 256, 0, st>>>(nitems, list, ctx->datom.as<int>(), ctx->ocnt.as<int>());
  scan_excl(ctx->ocnt.as<int>(), ctx->ostart.as<int>(), n, ctx->part.as<int>(), st);
  k_owner_scatter<<<ig, 256, 0, st>>>(nitems, list, ctx->datom.as<int>(), ctx->ostart.as<int>(),
                                      ctx->ocur.as<int>(), ctx->operm.as<int>());
  if (n > 0) {
    k_owner_small<<<grid_for(n, 256, 4096), 256, 0, st>>>(n, ctx->ostart.as<int>(), ctx->operm.as<int>(), val, vstride,
                                                           per_atom, ctx->omulti.as<int>(), ctx->onmulti.as<int>());
    TIMED("owner_sums", k_owner_sums<<<grid_for(std::max<int64_t>(est_items / 2, 1), 1, 16384), 64, 0, st>>>(
                            ctx->omulti.as<int>(), ctx->onmulti.as<int>(), ctx->ostart.as<int>(),
                            ctx->operm.as<int>(), val, vstride, per_atom));
  }
  CKL();
  // np.sum of the per-atom values of every molecule group; gstride 1 keeps
  // all 2B group totals, gstride 2 only the molecule-0 ones (max-CMS)
  CK(ctx->gtot.ensure(sizeof(double) * (G_tot + 1)));
  TIMED("group_sums", k_group_np_sum<<<grid_for(G_tot, 1, 4096), 256, 0, st>>>(G_tot, ctx->agrp.as<int>(),
                                                                                per_atom, 1, ctx->gtot.as<double>()));
  k_pick_stride<<<grid_for(G_tot / gstride, 256, 4096), 256, 0, st>>>(G_tot / gstride, ctx->gtot.as<double>(),
                                                                       gstride, totals);
  CKL();
  return CMS_OK;
}

// Outputs of the complexes with a non-zero status word set to zero (a
// kernel, so the step stays capturable).
static void zero_failed(cms_ctx* ctx, double* per_atom, int width, double* totals, int tstride, double* sc) {
  k_zero_failed<<<grid_for(ctx->B, 1, 4096), 256, 0, ctx->stream>>>(ctx->B, ctx->status.as<int>(),
                                                                     ctx->aoff.as<int>(), per_atom, width, totals,
                                                                     tstride, sc, CMS_SC_NRES);
}

static int contact_ms_pass(cms_ctx* ctx, int precision, double* per_atom, double* totals) {
  int r = prepare_lists(ctx);
  if (r) return r;
  if (precision == CMS_PREC_FP32_RESCORE && (r = prepare_float_lists(ctx))) return r;
  const int G = 2 * ctx->B;
  hipStream_t st = ctx->stream;
  const int* meta = ctx->meta.as<int>();
  const Count cB = dev_count(meta + M_B);
  const int64_t estB = ctx->est[M_B];
  r = launch_nearest(ctx, precision, ctx->bxyz.as<double4>(), ctx->bxyz32.as<float4>(), ctx->bgrp.as<int>(), cB,
                     ctx->cap[M_D], estB, estB, ctx->bxyz.as<double4>(), ctx->bxyz32.as<float4>(),
                     ctx->bgrp.as<int>(), 1, G, ctx->gframe.as<double4>(), ctx->bmin.as<double>(),
                     ctx->bidx.as<int>());
  if (r) return r;
  k_cms_weights<<<grid_for(std::max<int64_t>(estB, 1), 256, 4096), 256, 0, st>>>(
      cB, ctx->bxyz.as<double4>(), ctx->bmin.as<double>(), ctx->s.weight, ctx->wv.as<double>());
  CKL();
  r = owner_sums(ctx, cB, estB, ctx->blist.as<int>(), ctx->wv.as<double>(), 1, per_atom, totals, G, 1);
  if (r) return r;
  zero_failed(ctx, per_atom, 1, totals, 2, nullptr);
  ctx->cms_ready = true;
  return CMS_OK;
}

int cms_contact_ms(cms_ctx* ctx, int precision, double* per_atom, double* totals) {
  if (!ctx || !per_atom || !totals) return CMS_ERR_INVALID;
  // any generated surface, like the reference's readout of whatever dots
  // exist (cms-cuda/_core.py:946-1004)
  if (!ctx->generated) return fail(ctx, CMS_ERR_STATE, "cms_contact_ms: no surface");
  if (precision != CMS_PREC_FP64 && precision != CMS_PREC_FP32_RESCORE)
    return fail(ctx, CMS_ERR_INVALID, "cms_contact_ms: unsupported precision");
  CK(hipSetDevice(ctx->device));
  return contact_ms_pass(ctx, precision, per_atom, totals);
}

int cms_max_contact_ms(cms_ctx* ctx, double* per_atom, double* totals) {
  if (!ctx || !per_atom || !totals) return CMS_ERR_INVALID;
  // any generated surface (cms-cuda/_core.py:1007-1034)
  if (!ctx->generated) return fail(ctx, CMS_ERR_STATE, "cms_max_contact_ms: no surface");
  CK(hipSetDevice(ctx->device));
  // items: all dots; value: dpos[i].w (area)
  const double* val = reinterpret_cast<const double*>(ctx->dpos.as<double4>()) + 3;
  int r = owner_sums(ctx, dev_count(ctx->meta.as<int>() + M_D), ctx->est[M_D], nullptr, val, 4, per_atom, totals,
                     2 * ctx->B, 2);
  if (r) return r;
  zero_failed(ctx, per_atom, 1, totals, 1, nullptr);
  CKL();
  return CMS_OK;
}

static int sc_pass(cms_ctx* ctx, int precision, double* results) {
  int r = prepare_lists(ctx);
  if (r) return r;
  if (precision == CMS_PREC_FP32_RESCORE && (r = prepare_float_lists(ctx))) return r;
  hipStream_t st = ctx->stream;
  const int B = ctx->B, G = 2 * B;
  int* part = ctx->part.as<int>();
  int* meta = ctx->meta.as<int>();
  const Count cB = dev_count(meta + M_B), cT = dev_count(meta + M_T);
  const int64_t capD = ctx->cap[M_D], estB = ctx->est[M_B], estT = ctx->est[M_T];
  CK(ctx->tgrp.ensure(sizeof(int) * (G + 1)));
  CK(ctx->side.ensure(sizeof(double) * 8 * G));
  // trim: buried dots of g against a grid of the accessible dots of g
  CK(ctx->tgrid.ensure(sizeof(GridInfo) * (G + 1)));
  CK(ctx->tncell.ensure(sizeof(int) * (G + 1)));
  CK(ctx->tncscan.ensure(sizeof(int) * (G + 1)));
  const int64_t cells_cap = 2 * capD + 65ll * G;
  const int gB = grid_for(std::max<int64_t>(estB, 1), 256, 4096);
  GridInfo* tg = ctx->tgrid.as<GridInfo>();
  k_dot_grid_setup<<<grid_for(G, 1, 4096), 256, 0, st>>>(G, ctx->agrpd.as<int>(), ctx->bgrp.as<int>(),
                                                          ctx->bxyz.as<double4>(), ctx->s.band, tg,
                                                          ctx->tncell.as<int>());
  scan_excl(ctx->tncell.as<int>(), ctx->tncscan.as<int>(), G, part, st);
  k_grid_offsets<<<grid_for(G, 256, 4096), 256, 0, st>>>(G, ctx->tncscan.as<int>(), tg);
  const Count cCells = dev_count(ctx->tncscan.as<int>() + G);
  // (device counters only for grids too large for LDS)
  fill_int(ctx->tcount.as<int>(), cCells, cells_cap, 0, st);
  // the accessible dots read in place from the dot arrays, binned per group
  TIMED("trim_grid", k_dot_grid_build<<<grid_for(G, 1, 65536), 256, 0, st>>>(
                         G, ctx->dgrp.as<int>(), ctx->dpos.as<double4>(), ctx->dbur.as<int8_t>(), tg,
                         ctx->tcount.as<int>(), ctx->tstart.as<int>(), ctx->tsorted.as<double4>()));
  {
    // lanes per buried dot: enough threads to fill the GPU for small
    // batches; four read a dot's candidates together (measured on
    // 2,112 complexes: 374 us against 451, 406, 424 and 596 us with 1, 2, 8
    // and 16 lanes per dot)
    const int64_t nb = std::max<int64_t>(estB, 1);
    const int L = ctx->trim_lanes ? ctx->trim_lanes : nb >= (1 << 23) ? 1 : nb >= (1 << 14) ? 4 : 32;
    const int gL = grid_for(nb * L, 256, 8192);
    const double band2 = ctx->s.band * ctx->s.band;
    const int* cs = ctx->tstart.as<int>();
    const double4* srt = ctx->tsorted.as<double4>();
    int* fl = ctx->trflag.as<int>();
    if (L == 1)
      TIMED("trim", (k_trim_grid<1><<<gL, 256, 0, st>>>(cB, ctx->bxyz.as<double4>(), ctx->bgrp.as<int>(), G, tg, cs, srt,
                                                        band2, fl)));
    else if (L == 4)
      TIMED("trim", (k_trim_grid<4><<<gL, 256, 0, st>>>(cB, ctx->bxyz.as<double4>(), ctx->bgrp.as<int>(), G, tg, cs, srt,
                                                        band2, fl)));
    else
      TIMED("trim", (k_trim_grid<32><<<gL, 256, 0, st>>>(cB, ctx->bxyz.as<double4>(), ctx->bgrp.as<int>(), G, tg, cs,
                                                         srt, band2, fl)));
  }
  CKL();
  scan_excl(ctx->trflag.as<int>(), ctx->trscan.as<int>(), cB, capD, part, st);
  k_gather_scan<<<grid_for(G + 1, 256, 4096), 256, 0, st>>>(G, ctx->bgrp.as<int>(), 1, ctx->trscan.as<int>(),
                                                             ctx->tgrp.as<int>());
  k_copy_word<<<1, 64, 0, st>>>(ctx->trscan.as<int>(), cB, meta + M_T);
  k_trim_compact<<<gB, 256, 0, st>>>(cB, ctx->trflag.as<int>(), ctx->trscan.as<int>(), ctx->blist.as<int>(),
                                     ctx->dpos.as<double4>(), ctx->dnml.as<double4>(), ctx->tlist.as<int>(),
                                     ctx->txyz.as<double4>(), ctx->tnml.as<double4>());
  CKL();
  const int gT = grid_for(std::max<int64_t>(estT, 1), 256, 4096);
  if (precision == CMS_PREC_FP32_RESCORE)
    k_to_float<<<gT, 256, 0, st>>>(cT, ctx->txyz.as<double4>(), ctx->tgrp.as<int>(), G, 1, ctx->gframe.as<double4>(),
                                   ctx->txyz32.as<float4>());
  r = launch_nearest(ctx, precision, ctx->txyz.as<double4>(), ctx->txyz32.as<float4>(), ctx->tgrp.as<int>(), cT, capD,
                     estT, estT, ctx->txyz.as<double4>(), ctx->txyz32.as<float4>(), ctx->tgrp.as<int>(), 1, G, ctx->gframe.as<double4>(),
                     ctx->tmin.as<double>(), ctx->tidx.as<int>());
  if (r) return r;
  k_group_of_items<<<grid_for(G, 1, 4096), 256, 0, st>>>(G, ctx->tgrp.as<int>(), ctx->tgof.as<int>());
  k_sc_terms<<<gT, 256, 0, st>>>(cT, ctx->tgof.as<int>(), ctx->tgrp.as<int>(), ctx->tmin.as<double>(),
                                 ctx->tidx.as<int>(), ctx->tnml.as<double4>(), ctx->s.weight, ctx->tdist.as<double>(),
                                 ctx->trr.as<double>());
  TIMED("sc_sides", k_sc_sides<<<grid_for(G, 1, 4096), 256, 0, st>>>(G, ctx->tgrp.as<int>(), ctx->txyz.as<double4>(),
                                                    ctx->tdist.as<double>(), ctx->trr.as<double>(),
                                                    ctx->dgrp.as<int>(), ctx->s, ctx->side.as<double>()));
  k_sc_combine<<<grid_for(B, 256, 4096), 256, 0, st>>>(B, ctx->side.as<double>(), results);
  zero_failed(ctx, nullptr, 0, nullptr, 0, results);
  CKL();
  ctx->sc_ready = true;
  ctx->host_tgroups = false;
  return CMS_OK;
}

int cms_shape_complementarity(cms_ctx* ctx, int precision, double* results) {
  if (!ctx || !results) return CMS_ERR_INVALID;
  if (!ctx->generated || ctx->gen_all_atoms != 0)
    return fail(ctx, CMS_ERR_STATE, "cms_shape_complementarity: surface not generated in CMS mode");
  if (precision != CMS_PREC_FP64 && precision != CMS_PREC_FP32_RESCORE)
    return fail(ctx, CMS_ERR_INVALID, "cms_shape_complementarity: unsupported precision");
  CK(hipSetDevice(ctx->device));
  return sc_pass(ctx, precision, results);
}

// ------------------------------------------------------------ one step

// The SC readout's scratch swapped with the CMS readout's (see part_b).
static void swap_readout_scratch(cms_ctx* ctx) {
  std::swap(ctx->part, ctx->part_b);
  std::swap(ctx->tcnt, ctx->tcnt_b);
  std::swap(ctx->toff, ctx->toff_b);
  std::swap(ctx->npd2, ctx->npd2_b);
  std::swap(ctx->npidx, ctx->npidx_b);
}

// Batches with at least this many buried dots run the SC readout beside the
// CMS readout (below it the two are too short to overlap usefully).
constexpr int64_t READOUT_FORK_MIN = 65536;

static int readouts_pass(cms_ctx* ctx, int precision, double* per_atom, double* totals, double* sc) {
  const bool fork = sc && ctx->prof_filter.empty() && ctx->est[M_B] >= READOUT_FORK_MIN && ensure_side(ctx) == CMS_OK;
  if (!fork) {
    int r = contact_ms_pass(ctx, precision, per_atom, totals);
    if (!r && sc) r = sc_pass(ctx, precision, sc);
    return r;
  }
  // both readouts start from the buried / accessible lists; the SC readout
  // (trim grid, trim, trimmed search, side statistics) then runs on side
  // stream 0 with its own scan and search scratch while the CMS readout
  // (search, weights, per-atom sums) runs on this stream, and this stream
  // waits for it
  int r = prepare_lists(ctx);
  if (!r && precision == CMS_PREC_FP32_RESCORE) r = prepare_float_lists(ctx);
  if (r) return r;
  hipStream_t st = ctx->stream;
  CK(hipEventRecord(ctx->fev[0], st));
  CK(hipStreamWaitEvent(ctx->sstream[0], ctx->fev[0], 0));
  swap_readout_scratch(ctx);
  ctx->stream = ctx->sstream[0];
  r = sc_pass(ctx, precision, sc);
  ctx->stream = st;
  swap_readout_scratch(ctx);
  CK(hipEventRecord(ctx->fev[1], ctx->sstream[0]));
  if (!r) r = contact_ms_pass(ctx, precision, per_atom, totals);
  CK(hipStreamWaitEvent(st, ctx->fev[1], 0));
  return r;
}

// Everything the launches of a learned step depend on besides device data.
static std::vector<uint64_t> run_key(cms_ctx* ctx, int precision, double* per_atom, double* totals, double* sc) {
  std::vector<uint64_t> k = {(uint64_t)ctx->B,        (uint64_t)ctx->ntot,     (uint64_t)precision,
                             (uint64_t)(uintptr_t)per_atom, (uint64_t)(uintptr_t)totals,
                             (uint64_t)(uintptr_t)sc,  ctx->layout_epoch,       workspace_signature(ctx)};
  for (int i = M_R; i <= M_T; ++i) k.push_back((uint64_t)ctx->est[i]);
  for (int v : {ctx->trim_lanes, ctx->probe_count_mode, ctx->probe_slots, ctx->cell_sweep, ctx->search_split,
                (int)ctx->keep_desc})
    k.push_back((uint64_t)(int64_t)v);
  const double sv[6] = {ctx->s.rp, ctx->s.band, ctx->s.sep, ctx->s.weight, ctx->s.binwidth_dist,
                        ctx->s.binwidth_norm};
  for (double v : sv) {
    uint64_t u;
    memcpy(&u, &v, sizeof u);
    k.push_back(u);
  }
  return k;
}

static int ensure_gstream(cms_ctx* ctx) {
  if (!ctx->gstream) CK(hipStreamCreateWithFlags(&ctx->gstream, hipStreamNonBlocking));
  if (!ctx->gev0) CK(hipEventCreateWithFlags(&ctx->gev0, hipEventDisableTiming));
  if (!ctx->gev1) CK(hipEventCreateWithFlags(&ctx->gev1, hipEventDisableTiming));
  return ensure_side(ctx);
}

// Capture one learned step on the graph stream.  Any failure (an
// allocation or a wait that capture refuses) leaves no graph behind and the
// caller runs the step eagerly.
static bool capture_step(cms_ctx* ctx, int precision, double* per_atom, double* totals, double* sc,
                         const std::vector<uint64_t>& key) {
  drop_graph(ctx);
  if (ensure_gstream(ctx)) return false;
  hipStream_t saved = ctx->stream;
  ctx->stream = ctx->gstream;
  ctx->capturing = true;
  int r = hipStreamBeginCapture(ctx->gstream, hipStreamCaptureModeThreadLocal) == hipSuccess ? CMS_OK : CMS_ERR_HIP;
  const bool began = r == CMS_OK;
  if (!r) r = generate_pass(ctx, 0, 0, false);
  if (!r) r = readouts_pass(ctx, precision, per_atom, totals, sc);
  hipGraph_t g = nullptr;
  hipError_t e = began ? hipStreamEndCapture(ctx->gstream, &g) : hipErrorUnknown;
  ctx->capturing = false;
  ctx->stream = saved;
  (void)hipGetLastError();
  const bool same_ws = run_key(ctx, precision, per_atom, totals, sc) == key;
  if (r || e != hipSuccess || !g || !same_ws) {
    if (g) (void)hipGraphDestroy(g);
    return false;
  }
  hipGraphExec_t ex = nullptr;
  if (hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) != hipSuccess || !ex) {
    (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    return false;
  }
  ctx->graph = g;
  ctx->gexec = ex;
  ctx->gkey = key;
  ++ctx->n_captures;
  return true;
}

int cms_run(cms_ctx* ctx, int precision, double* per_atom, double* totals, double* sc_results) {
  if (!ctx || !per_atom || !totals) return CMS_ERR_INVALID;
  if (ctx->B < 1) return fail(ctx, CMS_ERR_INVALID, "cms_run: nothing loaded");
  if (precision != CMS_PREC_FP64 && precision != CMS_PREC_FP32_RESCORE)
    return fail(ctx, CMS_ERR_INVALID, "cms_run: unsupported precision");
  if (ctx->capturing) return fail(ctx, CMS_ERR_STATE, "cms_run: not allowed during capture");
  CK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  ctx->generated = false;
  int r;
  if (!ctx->learned[0]) {
    r = generate_pass(ctx, 0, 0, true);
    if (!r) r = readouts_pass(ctx, precision, per_atom, totals, sc_results);
    if (!r) r = finish_run(ctx, st, 0);
  } else {
    if ((r = plan_learned(ctx, 0))) return r;
    const std::vector<uint64_t> key = run_key(ctx, precision, per_atom, totals, sc_results);
    bool replay = false;
    if (ctx->use_graph) {
      if (ctx->gexec && key == ctx->gkey) replay = true;
      else if (key == ctx->ekey) replay = capture_step(ctx, precision, per_atom, totals, sc_results, key);
    }
    hipStream_t fin = st;
    if (replay) {
      CK(hipEventRecord(ctx->gev0, st));
      CK(hipStreamWaitEvent(ctx->gstream, ctx->gev0, 0));
      CK(hipGraphLaunch(ctx->gexec, ctx->gstream));
      ++ctx->n_replays;
      fin = ctx->gstream;
      ctx->generated = true;
      ctx->gen_all_atoms = 0;
      ctx->gen_desc = ctx->keep_desc;
      ctx->lists_ready = true;
      ctx->float_lists = precision == CMS_PREC_FP32_RESCORE;
      ctx->cms_ready = true;
      ctx->sc_ready = sc_results != nullptr;
      ctx->host_tgroups = false;
      ctx->h_tq_c = -1;
      ctx->host_groups = false;
    } else {
      r = generate_pass(ctx, 0, 0, false);
      if (!r) r = readouts_pass(ctx, precision, per_atom, totals, sc_results);
      if (r) {
        ctx->generated = false;
        return r;
      }
      ctx->ekey = key;
    }
    r = finish_run(ctx, fin, 0);
    if (r == CMS_RETRY) {
      // a buffer was too small for this batch: repeat exactly (learns the
      // new sizes; the next step captures again)
      drop_graph(ctx);
      ctx->ekey.clear();
      r = generate_pass(ctx, 0, 0, true);
      if (!r) r = readouts_pass(ctx, precision, per_atom, totals, sc_results);
      if (!r) r = finish_run(ctx, st, 0);
    }
  }
  if (r == CMS_RETRY) r = fail(ctx, CMS_ERR_HIP, "cms_run: exact run overflowed");
  if (r) {
    ctx->generated = false;
    ctx->gen_all_atoms = -1;
  }
  return r;
}

int cms_set_graph(cms_ctx* ctx, int enable) {
  if (!ctx) return CMS_ERR_INVALID;
  ctx->use_graph = enable != 0;
  if (!ctx->use_graph) drop_graph(ctx);
  return CMS_OK;
}

int cms_set_keep_descriptors(cms_ctx* ctx, int keep) {
  if (!ctx) return CMS_ERR_INVALID;
  ctx->keep_desc = keep != 0;
  return CMS_OK;
}

int cms_run_stats(cms_ctx* ctx, int64_t* out) {
  if (!ctx || !out) return CMS_ERR_INVALID;
  out[0] = ctx->n_captures;
  out[1] = ctx->n_replays;
  out[2] = ctx->n_exact;
  out[3] = ctx->n_overflow;
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

// Host copies of the dot and probe group offsets, fetched on first use.
static int ensure_host_groups(cms_ctx* ctx) {
  if (ctx->host_groups) return CMS_OK;
  const int G = 2 * ctx->B;
  ctx->h_dgrp.assign(G + 1, 0);
  ctx->h_prgrp.assign(G + 1, 0);
  CK(hipMemcpyAsync(ctx->h_dgrp.data(), ctx->dgrp.p, sizeof(int) * (G + 1), hipMemcpyDeviceToHost, ctx->stream));
  CK(hipMemcpyAsync(ctx->h_prgrp.data(), ctx->prgrp.p, sizeof(int) * (G + 1), hipMemcpyDeviceToHost, ctx->stream));
  CK(hipStreamSynchronize(ctx->stream));
  ctx->host_groups = true;
  return CMS_OK;
}

int cms_dot_count(cms_ctx* ctx, int c, int mol, int64_t* n) {
  if (!ctx || !n || c < 0 || c >= ctx->B || mol < 0 || mol > 1) return CMS_ERR_INVALID;
  if (!ctx->generated) return fail(ctx, CMS_ERR_STATE, "cms_dot_count: no surface");
  int r = ensure_host_groups(ctx);
  if (r) return r;
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

__global__ void k_unpack_desc(int n, int lo, const int4* d1, const int4* d2, int32_t* o1, int32_t* o2) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int4 a = d1[lo + i], c = d2[lo + i];
    o1[4 * i] = a.x; o1[4 * i + 1] = a.y; o1[4 * i + 2] = a.z; o1[4 * i + 3] = a.w;
    o2[4 * i] = c.x; o2[4 * i + 1] = c.y; o2[4 * i + 2] = c.z; o2[4 * i + 3] = c.w;
  }
}

int cms_copy_dot_desc(cms_ctx* ctx, int c, int mol, int32_t* d1, int32_t* d2) {
  int64_t n = 0;
  int r = cms_dot_count(ctx, c, mol, &n);
  if (r) return r;
  if (!d1 || !d2) return CMS_ERR_INVALID;
  if (!ctx->gen_desc) return fail(ctx, CMS_ERR_STATE, "cms_copy_dot_desc: %s", NO_DESC);
  if (n == 0) return CMS_OK;
  const int g = 2 * c + mol;
  k_unpack_desc<<<grid_for(n, 256, 4096), 256, 0, ctx->stream>>>((int)n, ctx->h_dgrp[g], ctx->ddesc.as<int4>(),
                                                                   ctx->ddesc2.as<int4>(), d1, d2);
  CKL();
  return CMS_OK;
}

// The reference finishes access with two fancy-index ORs over the toroid
// queue (cms-cuda/_core.py:1247-1248): access[atom1s] |= has1 and then
// access[atom2s] |= has2, where has = the pair has a torus point clear of
// its neighbours.  With repeated indices NumPy keeps the last write, so each
// atom takes the flag of the last queue pair naming it as atom 1, then of
// the last naming it as atom 2.  Queue pairs are the toroid-flagged pairs in
// pair order.
__global__ void k_queue_last(Count np_, const int* pi, const int* pj, const int8_t* pflag, int* last1,
                             int* last2) {
  const int n = np_.get();
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x)
    if (pflag[p] & 2) {
      atomicMax(&last1[pi[p]], p);
      atomicMax(&last2[pj[p]], p);
    }
}

__global__ void k_access_final(int n, const int8_t* access, const int* last1, const int* last2,
                               const int8_t* live, int8_t* out) {
  for (int a = blockIdx.x * blockDim.x + threadIdx.x; a < n; a += gridDim.x * blockDim.x) {
    int8_t v = access[a];
    if (last1[a] >= 0 && live[last1[a]]) v = 1;
    if (last2[a] >= 0 && live[last2[a]]) v = 1;
    out[a] = v;
  }
}

int cms_copy_atom_state(cms_ctx* ctx, int8_t* atten, int8_t* access) {
  if (!ctx) return CMS_ERR_INVALID;
  if (!ctx->generated) return fail(ctx, CMS_ERR_STATE, "cms_copy_atom_state: no surface");
  hipStream_t st = ctx->stream;
  const int n = ctx->ntot;
  if (atten) CK(hipMemcpyAsync(atten, ctx->atten.p, n, hipMemcpyDeviceToDevice, st));
  if (access && n > 0) {
    CK(ctx->qlast.ensure(sizeof(int) * 2 * ((int64_t)n + 1)));
    int* last1 = ctx->qlast.as<int>();
    int* last2 = last1 + n + 1;
    fill_int(last1, 2 * ((int64_t)n + 1), -1, st);
    k_queue_last<<<grid_for(std::max<int64_t>(ctx->est[M_P], 1), 256, 4096), 256, 0, st>>>(
        dev_count(ctx->meta.as<int>() + M_P), ctx->pi.as<int>(), ctx->pj.as<int>(), ctx->pflag.as<int8_t>(), last1,
        last2);
    k_access_final<<<grid_for(n, 256, 4096), 256, 0, st>>>(n, ctx->access.as<int8_t>(), last1, last2,
                                                            ctx->tflag.as<int8_t>(), access);
    CKL();
  }
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
  int r = ensure_host_groups(ctx);
  if (r) return r;
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

// ------------------------------------------------------------ run-state copies
// Neighbour (which = 0) or buried (which = 1) rows of complex c's atoms: row
// length and entries as local atom indices, -1 past the row's end (the
// reference's neighbor_array / buried_array natom tables).
__global__ void k_copy_rows(int n, int a0, const int* rpos, const int* cnt, const int* rows, int rowcap, int width,
                            int32_t* counts, int32_t* natom) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int a = a0 + i;
    const int r = rpos[a];
    const int k = (rpos[a + 1] > r) ? min(cnt[r], rowcap) : 0;
    if (counts) counts[i] = k;
    if (natom)
      for (int t = 0; t < width; ++t) natom[(int64_t)i * width + t] = t < k ? rows[(int64_t)r * rowcap + t] - a0 : -1;
  }
}

int cms_copy_rows(cms_ctx* ctx, int c, int which, int32_t* counts, int32_t* natom, int width) {
  if (!ctx || c < 0 || c >= ctx->B || which < 0 || which > 1 || (natom && width < 0)) return CMS_ERR_INVALID;
  if (!ctx->generated) return fail(ctx, CMS_ERR_STATE, "cms_copy_rows: no surface");
  const int a0 = ctx->h_aoff[c], n = ctx->h_aoff[c + 1] - a0;
  if (n == 0 || (!counts && !natom)) return CMS_OK;
  k_copy_rows<<<grid_for(n, 256, 4096), 256, 0, ctx->stream>>>(
      n, a0, ctx->rpos.as<int>(), which ? ctx->nb.as<int>() : ctx->nn.as<int>(),
      which ? ctx->bur.as<int>() : ctx->nbr.as<int>(), which ? BCAP : KCAP, width, counts, natom);
  CKL();
  return CMS_OK;
}

static int ensure_host_tgroups(cms_ctx* ctx) {
  if (ctx->host_tgroups) return CMS_OK;
  const int G = 2 * ctx->B;
  ctx->h_tgrp.assign(G + 1, 0);
  CK(hipMemcpyAsync(ctx->h_tgrp.data(), ctx->tgrp.p, sizeof(int) * (G + 1), hipMemcpyDeviceToHost, ctx->stream));
  CK(hipStreamSynchronize(ctx->stream));
  ctx->host_tgroups = true;
  return CMS_OK;
}

int cms_trimmed_count(cms_ctx* ctx, int c, int mol, int64_t* n) {
  if (!ctx || !n || c < 0 || c >= ctx->B || mol < 0 || mol > 1) return CMS_ERR_INVALID;
  if (!ctx->generated || !ctx->sc_ready)
    return fail(ctx, CMS_ERR_STATE, "cms_trimmed_count: run cms_shape_complementarity on the current surface first");
  int r = ensure_host_tgroups(ctx);
  if (r) return r;
  const int g = 2 * c + mol;
  *n = ctx->h_tgrp[g + 1] - ctx->h_tgrp[g];
  return CMS_OK;
}

__global__ void k_unpack_dot_list(int n, const int* list, int lo, int aoff, const double4* dpos, const double4* dnml,
                                  const int8_t* dbur, const int8_t* dtype, const int* datom, double* coor, double* nml,
                                  double* area, int8_t* buried, int8_t* type, int32_t* atom) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int d = list[lo + i];
    const double4 p = dpos[d], q = dnml[d];
    if (coor) { coor[3 * i] = p.x; coor[3 * i + 1] = p.y; coor[3 * i + 2] = p.z; }
    if (nml) { nml[3 * i] = q.x; nml[3 * i + 1] = q.y; nml[3 * i + 2] = q.z; }
    if (area) area[i] = p.w;
    if (buried) buried[i] = dbur[d];
    if (type) type[i] = dtype[d];
    if (atom) atom[i] = datom[d] - aoff;
  }
}

int cms_copy_trimmed_dots(cms_ctx* ctx, int c, int mol, double* coor, double* nml, double* area, int8_t* buried,
                          int8_t* type, int32_t* atom) {
  int64_t n = 0;
  int r = cms_trimmed_count(ctx, c, mol, &n);
  if (r) return r;
  if (n == 0) return CMS_OK;
  const int g = 2 * c + mol;
  k_unpack_dot_list<<<grid_for(n, 256, 4096), 256, 0, ctx->stream>>>(
      (int)n, ctx->tlist.as<int>(), ctx->h_tgrp[g], ctx->h_aoff[c], ctx->dpos.as<double4>(), ctx->dnml.as<double4>(),
      ctx->dbur.as<int8_t>(), ctx->dtype.as<int8_t>(), ctx->datom.as<int>(), coor, nml, area, buried, type, atom);
  CKL();
  return CMS_OK;
}

// Toroid-queue pairs of complex c (pair flag bit 2), in pair order: the
// reference's queue order (atoms ascending, each atom's forward neighbours
// in row order, cms-cuda/_core.py:1419-1431).
static int toroid_pairs(cms_ctx* ctx, int c) {
  if (ctx->h_tq_c == c) return CMS_OK;
  const int G = 2 * ctx->B;
  ctx->h_pgrp.assign(G + 1, 0);
  CK(hipMemcpyAsync(ctx->h_pgrp.data(), ctx->pgrp.p, sizeof(int) * (G + 1), hipMemcpyDeviceToHost, ctx->stream));
  CK(hipStreamSynchronize(ctx->stream));
  const int p0 = ctx->h_pgrp[2 * c], p1 = ctx->h_pgrp[2 * c + 2];
  std::vector<int8_t> fl(std::max(p1 - p0, 0));
  if (p1 > p0) {
    CK(hipMemcpyAsync(fl.data(), ctx->pflag.as<int8_t>() + p0, p1 - p0, hipMemcpyDeviceToHost, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
  }
  ctx->h_tq.clear();
  for (int p = p0; p < p1; ++p)
    if (fl[p - p0] & 2) ctx->h_tq.push_back(p);
  ctx->h_tq_c = c;
  return CMS_OK;
}

int cms_toroid_count(cms_ctx* ctx, int c, int64_t* n) {
  if (!ctx || !n || c < 0 || c >= ctx->B) return CMS_ERR_INVALID;
  if (!ctx->generated) return fail(ctx, CMS_ERR_STATE, "cms_toroid_count: no surface");
  int r = toroid_pairs(ctx, c);
  if (r) return r;
  *n = (int64_t)ctx->h_tq.size();
  return CMS_OK;
}

// Queue entries from the pairs' atoms with the pair geometry of the
// surface kernels (pair_torus), so they equal what those kernels used.
__global__ void k_toroid_entries(int n, const int* idx, const int* pi, const int* pj, const double4* atom, double rp,
                                 int aoff, int32_t* atoms, double* uij, double* tij, double* rij, int8_t* between) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int p = idx[i];
    const int a = pi[p], b = pj[p];
    const double4 xa = atom[a], xb = atom[b];
    V3 u, t;
    double rr = 0.0;
    bool bt = false;
    pair_torus(ld3(xa), xa.w, ld3(xb), xb.w, rp, &u, &t, &rr, &bt);
    if (atoms) { atoms[2 * i] = a - aoff; atoms[2 * i + 1] = b - aoff; }
    if (uij) { uij[3 * i] = u.x; uij[3 * i + 1] = u.y; uij[3 * i + 2] = u.z; }
    if (tij) { tij[3 * i] = t.x; tij[3 * i + 1] = t.y; tij[3 * i + 2] = t.z; }
    if (rij) rij[i] = rr;
    if (between) between[i] = bt ? 1 : 0;
  }
}

int cms_copy_toroid_queue(cms_ctx* ctx, int c, int32_t* atoms, double* uij, double* tij, double* rij,
                          int8_t* between) {
  int64_t n = 0;
  int r = cms_toroid_count(ctx, c, &n);
  if (r) return r;
  if (n == 0) return CMS_OK;
  CK(ctx->tqidx.ensure(sizeof(int) * n));
  CK(hipMemcpyAsync(ctx->tqidx.p, ctx->h_tq.data(), sizeof(int) * n, hipMemcpyHostToDevice, ctx->stream));
  k_toroid_entries<<<grid_for(n, 256, 4096), 256, 0, ctx->stream>>>(
      (int)n, ctx->tqidx.as<int>(), ctx->pi.as<int>(), ctx->pj.as<int>(), ctx->atom.as<double4>(), ctx->s.rp,
      ctx->h_aoff[c], atoms, uij, tij, rij, between);
  CKL();
  CK(hipStreamSynchronize(ctx->stream));
  return CMS_OK;
}

int cms_counts(cms_ctx* ctx, int64_t* out) {
  if (!ctx || !out) return CMS_ERR_INVALID;
  if (!ctx->generated) return fail(ctx, CMS_ERR_STATE, "cms_counts: no surface");
  const int B = ctx->B, G = 2 * B;
  hipStream_t st = ctx->stream;
  int r = prepare_lists(ctx);
  if (!r) r = ensure_host_groups(ctx);
  if (r) return r;
  std::vector<int> bg(G + 1), rg(G + 1), pg(G + 1), prg(G + 1);
  std::vector<int> cs(ctx->nR + 1), t1(ctx->nP + 1), t2(ctx->nP + 1), ks(ctx->nQ + 1);
  std::vector<int> hnn(ctx->nR + 1), hnb(ctx->nR + 1), hpc(ctx->nR + 1);
  std::vector<int8_t> att(ctx->ntot + 1);
  CK(hipMemcpyAsync(bg.data(), ctx->bgrp.p, sizeof(int) * (G + 1), hipMemcpyDeviceToHost, st));
  CK(hipMemcpyAsync(rg.data(), ctx->rgrp.p, sizeof(int) * (G + 1), hipMemcpyDeviceToHost, st));
  CK(hipMemcpyAsync(pg.data(), ctx->pgrp.p, sizeof(int) * (G + 1), hipMemcpyDeviceToHost, st));
  CK(hipMemcpyAsync(prg.data(), ctx->prgrp.p, sizeof(int) * (G + 1), hipMemcpyDeviceToHost, st));
  CK(hipMemcpyAsync(cs.data(), ctx->cscan.p, sizeof(int) * (ctx->nR + 1), hipMemcpyDeviceToHost, st));
  CK(hipMemcpyAsync(t1.data(), ctx->t1scan.p, sizeof(int) * (ctx->nP + 1), hipMemcpyDeviceToHost, st));
  CK(hipMemcpyAsync(t2.data(), ctx->t2scan.p, sizeof(int) * (ctx->nP + 1), hipMemcpyDeviceToHost, st));
  CK(hipMemcpyAsync(ks.data(), ctx->kscan.p, sizeof(int) * (ctx->nQ + 1), hipMemcpyDeviceToHost, st));
  if (ctx->nR > 0) {
    CK(hipMemcpyAsync(hnn.data(), ctx->nn.p, sizeof(int) * ctx->nR, hipMemcpyDeviceToHost, st));
    CK(hipMemcpyAsync(hnb.data(), ctx->nb.p, sizeof(int) * ctx->nR, hipMemcpyDeviceToHost, st));
    CK(hipMemcpyAsync(hpc.data(), ctx->pcnt.p, sizeof(int) * ctx->nR, hipMemcpyDeviceToHost, st));
  }
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
    int64_t en = 0, eb = 0, pn = 0;
    for (int rr = rg[g0]; rr < rg[g2]; ++rr) {
      en += hnn[rr];
      eb += hnb[rr];
      pn += (int64_t)hpc[rr] * hnn[rr];
    }
    o[12] = en;
    o[13] = eb;
    o[14] = pn;
  }
  return CMS_OK;
}

}  // extern "C"

namespace cms {

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

}  // namespace cms

extern "C" {

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
    k_to_float<<<grid_for(nq, 256, 4096), 256, 0, st>>>(host_count((int)nq), ctx->qpack.as<double4>(), nullptr, 1, 0,
                                                         ctx->frame.as<double4>(), ctx->qpack32.as<float4>());
    if (nr > 0)
      k_to_float<<<grid_for(nr, 256, 4096), 256, 0, st>>>(host_count((int)nr), ctx->rpack.as<double4>(), nullptr, 1, 0,
                                                           ctx->frame.as<double4>(), ctx->rpack32.as<float4>());
  }
  int rc = launch_nearest(ctx, precision, ctx->qpack.as<double4>(), ctx->qpack32.as<float4>(), ctx->grp1.as<int>(),
                          host_count((int)nq), nq, nq, nr, ctx->rpack.as<double4>(), ctx->rpack32.as<float4>(),
                          ctx->grp1.as<int>() + 2, 0, 1, ctx->frame.as<double4>(), out_d2, out_idx);
  if (rc) return rc;
  CK(hipStreamSynchronize(st));
  return CMS_OK;
}

int cms_nearest_noalloc(cms_ctx* ctx, int precision, const double* q, int64_t nq, const double* r, int64_t nr,
                        double* out_d2, int32_t* out_idx) {
  if (!ctx || nq < 0 || nr < 0 || (nq > 0 && (!q || !out_d2)) || (nr > 0 && !r)) return CMS_ERR_INVALID;
  if (nq > NOALLOC_MAX_POINTS || nr > NOALLOC_MAX_POINTS)
    return fail(ctx, CMS_ERR_INVALID, "cms_nearest_noalloc: %lld queries / %lld references exceed the scratch limit "
                "of %lld points per side", (long long)nq, (long long)nr, (long long)NOALLOC_MAX_POINTS);
  if (precision != CMS_PREC_FP64 && precision != CMS_PREC_FP32_RESCORE)
    return fail(ctx, CMS_ERR_INVALID, "cms_nearest_noalloc: unsupported precision");
  CK(hipSetDevice(ctx->device));
  if (nq == 0) return CMS_OK;
  std::string err;
  const int rc = nearest_noalloc(ctx->device, ctx->stream, precision, q, nq, r, nr, out_d2, out_idx, &err);
  if (rc != CMS_OK) {
    ctx->err = err;
    return rc;
  }
  CK(hipStreamSynchronize(ctx->stream));
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
  k_group_of_items<<<grid_for(G, 1, 4096), 256, 0, ctx->stream>>>(G, ctx->bgrp.as<int>(), ctx->bgof.as<int>());
  CKL();
  return CMS_OK;
}

int cms_backward_cms(cms_ctx* ctx, const double* g_totals, const double* g_per_atom, double* grad_xyz) {
  if (!ctx || !grad_xyz) return CMS_ERR_INVALID;
  if (!ctx->cms_ready)
    return fail(ctx, CMS_ERR_STATE, "cms_backward_cms: run cms_contact_ms on the current surface first");
  if (!ctx->gen_desc) return fail(ctx, CMS_ERR_STATE, "cms_backward_cms: %s", NO_DESC);
  CK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const int* meta = ctx->meta.as<int>();
  const int64_t capD = ctx->cap[M_D], estD = ctx->est[M_D], estB = ctx->est[M_B];
  int r = buried_groups(ctx);
  if (r) return r;
  CK(ctx->adj.ensure(sizeof(double) * 4 * (capD + 1)));
  fill_double(ctx->adj.as<double>(), 4 * capD, 0.0, st);
  fill_double(grad_xyz, 3 * (int64_t)ctx->ntot, 0.0, st);
  TIMED("cms_adjoint", k_cms_adjoint<<<grid_for(std::max<int64_t>(estB, 1), 256, 4096), 256, 0, st>>>(
      dev_count(meta + M_B), ctx->blist.as<int>(), ctx->bgof.as<int>(), ctx->bgrp.as<int>(), ctx->bidx.as<int>(),
      ctx->bmin.as<double>(), ctx->dpos.as<double4>(), ctx->datom.as<int>(), g_totals, g_per_atom,
      ctx->s.weight, ctx->adj.as<double>()));
  TIMED("dot_grad", k_dot_grad<<<grid_for(std::max<int64_t>(estD, 1), 64, 16384), 64, 0, st>>>(
      view(ctx), ctx->s, dev_count(meta + M_D), ctx->ddesc.as<int4>(), ctx->ddesc2.as<int4>(),
      ctx->adj.as<double>(), grad_xyz));
  zero_failed(ctx, grad_xyz, 3, nullptr, 0, nullptr);
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
  if (!ctx->gen_desc) return fail(ctx, CMS_ERR_STATE, "cms_frozen_cms: %s", NO_DESC);
  CK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const int n = ctx->ntot, G = 2 * ctx->B;
  const Count cB = dev_count(ctx->meta.as<int>() + M_B);
  const int64_t estB = std::max<int64_t>(ctx->est[M_B], 1);
  int r = buried_groups(ctx);
  if (r) return r;
  CK(ctx->eatom.ensure(sizeof(double4) * (n + 1)));
  CK(ctx->evpos.ensure(sizeof(double4) * (ctx->cap[M_D] + 1)));
  if (n > 0)
    k_pack_eval<<<grid_for(n, 256, 4096), 256, 0, st>>>(n, xyz_eval, ctx->atom.as<double4>(), ctx->eatom.as<double4>());
  k_frozen_eval<<<grid_for(estB, 64, 16384), 64, 0, st>>>(view(ctx), ctx->s, ctx->eatom.as<double4>(), cB,
                                                           ctx->blist.as<int>(), ctx->ddesc.as<int4>(),
                                                           ctx->ddesc2.as<int4>(), ctx->evpos.as<double4>());
  k_frozen_cms<<<grid_for(estB, 256, 4096), 256, 0, st>>>(cB, ctx->bgof.as<int>(), ctx->bgrp.as<int>(),
                                                           ctx->bidx.as<int>(), ctx->evpos.as<double4>(),
                                                           ctx->s.weight, ctx->wv.as<double>());
  // per-group sums in buried-dot order: deterministic, so finite differences
  // see no summation-order noise
  k_group_np_sum<<<grid_for(G, 1, 4096), 256, 0, st>>>(G, ctx->bgrp.as<int>(), ctx->wv.as<double>(), 1, totals);
  zero_failed(ctx, nullptr, 0, totals, 2, nullptr);
  CKL();
  return CMS_OK;
}

int cms_check_dots(cms_ctx* ctx, int64_t* mismatches, double* maxdev) {
  if (!ctx || !mismatches || !maxdev) return CMS_ERR_INVALID;
  if (!ctx->generated) return fail(ctx, CMS_ERR_STATE, "cms_check_dots: no surface");
  if (!ctx->gen_desc) return fail(ctx, CMS_ERR_STATE, "cms_check_dots: %s", NO_DESC);
  CK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  CK(ctx->chk.ensure(16));
  fill_u64(ctx->chk.as<unsigned long long>(), 2, 0ull, st);
  if (ctx->nD > 0)
    k_check_dots<<<grid_for(ctx->nD, 64, 16384), 64, 0, st>>>(view(ctx), ctx->s, host_count(ctx->nD), ctx->dpos.as<double4>(),
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


