/*
 * cms_hip.h -- C ABI of libcms_hip.so, the MI355X (gfx950) implementation of
 * the contact-molecular-surface (CMS) and Lawrence-Colman shape
 * complementarity (SC) pipeline of cms-cuda's MolecularSurfaceCalculator.
 *
 * Plain C: device pointers, host pointers and sizes only.  All device work
 * is issued on the stream given to cms_create(); functions that return
 * sizes or status synchronise that stream, the others are asynchronous
 * unless stated.  The library never frees caller memory; it owns a
 * workspace per context that grows on demand.  One context per host thread.
 *
 * Batches: a call works on B >= 1 complexes at once.  Atoms are passed
 * complex-major; within a complex the binder (molecule 0) atoms come first,
 * then the target (molecule 1) atoms -- the order add_binder_and_target()
 * builds (cms-cuda/_core.py:918-943).  Atom indices reported back (dot
 * owners, coincident pairs) are local to their complex, numbered like the
 * reference's natom: binder 0..N0-1, target N0..N0+N1-1.
 */
#ifndef CMS_HIP_H
#define CMS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* return codes */
#define CMS_OK 0
#define CMS_ERR_INVALID 1    /* bad argument or call order                      */
#define CMS_ERR_HIP 2        /* HIP runtime error (message in cms_last_error)    */
#define CMS_ERR_COINCIDENT 3 /* two atoms of one molecule within 0.01 A          */
#define CMS_ERR_CAPACITY 4   /* a neighbour / near-probe row exceeded its limit  */
#define CMS_ERR_STATE 5      /* surface cached in the other all_atoms mode       */

/* precision of the nearest-dot searches */
#define CMS_PREC_FP64 0        /* plain float64 search                                */
#define CMS_PREC_FP32_RESCORE 1 /* float32 search on recentred coordinates, float64
                                   rescoring of the winner and of every candidate
                                   within the proven error margin                    */

/* per-complex counters written by cms_counts() */
#define CMS_NCOUNTS 15
/* [0] dots mol 0, [1] dots mol 1, [2] buried dots mol 0, [3] buried dots mol 1,
   [4] convex, [5] toroidal, [6] concave, [7] probes, [8] region atoms,
   [9] pairs, [10] blocked atoms mol 0, [11] blocked atoms mol 1,
   [12] neighbour-row entries, [13] buried-row entries, [14] sum over pairs
   (i, j) of the length of atom i's neighbour row                          */

/* per-complex SC record written by cms_shape_complementarity() */
#define CMS_SC_NRES 24
/* [0] sc  [1] area  [2] distance  [3] d_mean  [4] s_mean  [5] valid
   [6] trimmed dots  [7] all dots
   [8..15]  molecule 0 side: d_mean, d_median, s_mean, s_median,
            trimmed_area, n_trimmed, n_dots, valid
   [16..23] molecule 1 side, same fields                                   */

/* Calculator settings; defaults rp 1.7, density 15, band 1.5, sep 8,
   weight 0.5, binwidth_dist 0.02, binwidth_norm 0.02
   (replaces MolecularSurfaceCalculator.settings, cms-cuda/_core.py:766-773). */
typedef struct cms_settings {
  double rp;
  double density;
  double band;
  double sep;
  double weight;
  double binwidth_dist;
  double binwidth_norm;
} cms_settings;

typedef struct cms_ctx cms_ctx;

/* Launch-shape choices of a context (0 in a field = the built-in rule by
   batch size).  Every choice gives bit-identical results; they only change
   speed.  trim_lanes: lanes per buried dot in the SC trim (1, 4 or 32);
   probe_count: probe / torus-point count pass per pair (1) or per
   neighbour row (2); probe_slots: surviving probes staged per row by the
   per-row pass (1..8); cell_sweep: attention and region sweeps with a
   wavefront per atom (1) or a thread per cell-sorted atom (2);
   search_split: work items the nearest-dot searches split their reference
   ranges to fill (1..65536).  No reference counterpart (the reference
   fixes its launch shapes, cms-cuda/_gpu_kernels.py:34, 504-505). */
typedef struct cms_tuning {
  int trim_lanes;
  int probe_count;
  int probe_slots;
  int cell_sweep;
  int search_split;
} cms_tuning;

/* Library version (major*10000 + minor*100 + patch). */
int cms_version(void);

/* Create a context bound to HIP device `device` and stream `stream`
   (NULL = the legacy default stream).  Replaces the per-thread stream and
   lazily compiled module of cms-cuda/_gpu.py:70-83, _gpu_kernels.py:407-417. */
int cms_create(int device, void* stream, cms_ctx** out);
int cms_destroy(cms_ctx* ctx);
/* Message of the last error on this context (or of cms_create when ctx is
   NULL).  Never NULL. */
const char* cms_last_error(const cms_ctx* ctx);
/* Wait for all work issued on the context's stream. */
int cms_synchronize(cms_ctx* ctx);

/* Load B complexes.  xyz: device, 3*ntot float64; radius: device, ntot;
   density: device, ntot per-atom dot densities or NULL for
   settings->density; n_binder / n_target: host arrays of B counts.
   Replaces reset + AddMolecule(0, ...) + AddMolecule(1, ...)
   (cms-cuda/_core.py:918-943; _gpu.py:469-497).  Invalidates any cached
   surface.  Synchronous with respect to the host arrays only. */
int cms_load(cms_ctx* ctx, int n_complex, const int32_t* n_binder, const int32_t* n_target,
             const double* xyz, const double* radius, const double* density,
             const cms_settings* settings);

/* Attention numbers, neighbour rows, probes and convex / toroidal / concave
   dots with burial flags for every loaded complex.  all_atoms = 0 is the
   CMS / SC mode, 1 the max-CMS mode (every atom attention 5).  full_state
   = 1 builds neighbour rows and pairs for every atom instead of only the
   interface region (identical dots; needed only for full run-state
   inspection).  Cached: a second call with the same all_atoms is a no-op,
   one with the other mode returns CMS_ERR_STATE.  Replaces
   _ensure_surfaces_generated (cms-cuda/_core.py:865-909, 1197-2377;
   _gpu.py:421-461, 691-1765).  A complex holding coincident atoms (see
   cms_coincident) or a row over capacity fails alone: its status word is
   set, it gets no dots and zero results in every readout, and the rest of
   the batch is unaffected (cms_failed_count).  The call returns
   CMS_ERR_COINCIDENT / CMS_ERR_CAPACITY only when every complex of the
   batch failed -- for one complex, the reference's RuntimeError
   (cms-cuda/_core.py:1285-1293). */
int cms_generate(cms_ctx* ctx, int all_atoms, int full_state);

/* Per-complex status word after cms_generate (bit 1 coincident atoms, 2 / 4
   neighbour / buried row over capacity, 8 near-probe list over capacity). */
int cms_status(cms_ctx* ctx, int32_t* status_host);

/* Set / read the context's launch-shape choices (a captured graph is
   re-captured after a change). */
int cms_set_tuning(cms_ctx* ctx, const cms_tuning* tuning);
int cms_get_tuning(cms_ctx* ctx, cms_tuning* tuning);

/* Number of complexes of the last cms_generate / cms_run whose status word
   is non-zero (their results are zero).  No host wait. */
int cms_failed_count(cms_ctx* ctx, int32_t* n_failed);

/* First coincident pair (row-major, as np.argwhere) of complex c, local
   atom indices; -1/-1 if none.  Mirrors the RuntimeError of
   cms-cuda/_core.py:1285-1293. */
int cms_coincident(cms_ctx* ctx, int complex_index, int32_t* i, int32_t* j);

/* Counters per complex, CMS_NCOUNTS int64 each (host array of
   B*CMS_NCOUNTS).  Counterpart of calc.run.results
   (cms-cuda/_core.py:272-304). */
int cms_counts(cms_ctx* ctx, int64_t* out_host);

/* Contact molecular surface of every complex, both sides at once.
   per_atom: device, ntot float64 -- the CMS of each atom as a query atom
   (binder atoms: binder side; target atoms: target side); totals: device,
   2*B float64, totals[2c] binder side, totals[2c+1] target side (the value
   CalcLoaded returns).  Replaces calc_contact_molecular_surface(target_side)
   (cms-cuda/_core.py:946-1004; _gpu.py:515-571). */
int cms_contact_ms(cms_ctx* ctx, int precision, double* per_atom, double* totals);

/* Maximum possible CMS (surface area of molecule 0) from the current
   surface's dots -- the max-CMS value on a surface generated with
   all_atoms = 1; like the reference, any generated surface is read.  per_atom: device, ntot; totals: device, B.
   Replaces calc_max_possible_contact_molecular_surface
   (cms-cuda/_core.py:1007-1034; _gpu.py:573-600). */
int cms_max_contact_ms(cms_ctx* ctx, double* per_atom, double* totals);

/* Shape complementarity of every complex into results (device,
   B*CMS_SC_NRES float64).  Replaces CalcLoadedSC
   (cms-cuda/_core.py:807-863, 1037-1147; _gpu.py:350-419, 604-686). */
int cms_shape_complementarity(cms_ctx* ctx, int precision, double* results);

/* One step of the hot path over the loaded batch: surface (CMS mode), CMS
   of both sides into per_atom / totals (as cms_contact_ms) and, when
   sc_results is not NULL, SC (as cms_shape_complementarity).  The same
   results as cms_generate(ctx, 0, 0) + cms_contact_ms +
   cms_shape_complementarity, issued without host waits once the context
   has seen a batch of this kind (buffer sizes learned per atom; a batch
   that outgrows them is repeated exactly), and replayed from a captured
   HIP graph when cms_set_graph(ctx, 1) is on and the batch layout, output
   pointers and sizes are those of the previous step.  Returns after one
   synchronisation, with the same error codes as cms_generate.  The
   batched, graph-captured entry point the reference's per-complex
   _ensure_surfaces_generated + calc_contact_molecular_surface +
   CalcLoadedSC sequence (cms-cuda/_gpu.py:421-461, 515-571, 350-419)
   becomes. */
int cms_run(cms_ctx* ctx, int precision, double* per_atom, double* totals, double* sc_results);
/* Enable (1) or disable (0) graph capture / replay in cms_run. */
int cms_set_graph(cms_ctx* ctx, int enable);
/* Whether surface generation writes each dot's descriptors (default 1).
   Only cms_copy_dot_desc, cms_backward_cms, cms_frozen_cms and
   cms_check_dots read them; they return CMS_ERR_STATE for a surface
   generated with 0.  A forward-only caller saves 32 B of writes per dot.
   No reference counterpart (the reference has no backward pass). */
int cms_set_keep_descriptors(cms_ctx* ctx, int keep);
/* out[4]: graphs captured, graph replays, exact (host-sized) surface runs,
   learned runs repeated after a buffer overflow. */
int cms_run_stats(cms_ctx* ctx, int64_t* out);

/* Number of dots of (complex, molecule) and a copy of them into device
   buffers (coor, nml: 3n float64; area: n; buried, type: n int8; atom: n
   int32 local atom index).  Any pointer may be NULL.  Counterpart of
   calc.run.dots[mol] (cms-cuda/_core.py:454-519). */
int cms_dot_count(cms_ctx* ctx, int complex_index, int mol, int64_t* n);
int cms_copy_dots(cms_ctx* ctx, int complex_index, int mol, double* coor, double* nml,
                  double* area, int8_t* buried, int8_t* type, int32_t* atom);

/* Frozen-topology descriptors of the dots of (complex, molecule), in the
   order of cms_copy_dots: d1, d2 device int32 4n each.  d1 = (kind | mm << 4,
   atom a, atom b, atom c) with kind 1 convex (a = owner, b = its nearest
   neighbour), 2 / 3 toroidal arc of atom i / j (a = i, b = j), 4 concave
   (a, b, c = the probe's atoms, mm = the vertex the latitudes start from);
   d2 = (latitude or torus point, point on it, latitude / torus point count,
   point count).  Atom indices are batch-global.  No reference counterpart
   (what the backward pass re-evaluates a dot from). */
int cms_copy_dot_desc(cms_ctx* ctx, int complex_index, int mol, int32_t* d1, int32_t* d2);

/* Per-atom attention numbers and access flags (device, ntot int8 each) as
   the reference leaves them after surface generation, including its
   toroidal-stage merge access[atom1s] |= has1; access[atom2s] |= has2 with
   NumPy's last-write-wins on repeated atoms.  Atoms outside the interface
   region are only visited by a full_state generation, so for every atom to
   match, generate with full_state = 1.  Counterpart of calc.run.atoms.atten
   / .access (cms-cuda/_core.py:1185-1188, 1247-1248, 1315, 1396-1400). */
int cms_copy_atom_state(cms_ctx* ctx, int8_t* atten, int8_t* access);

/* Probes of complex c (host count; device copies: atoms 3n int32 local,
   height n, point 3n, alt 3n).  Counterpart of calc.run.probes
   (cms-cuda/_core.py:522-590), ordered within each molecule as the
   reference orders them. */
int cms_probe_count(cms_ctx* ctx, int complex_index, int64_t* n);
int cms_copy_probes(cms_ctx* ctx, int complex_index, int32_t* atoms, double* height,
                    double* point, double* alt);

/* Neighbour rows (which = 0) or buried rows (which = 1) of complex c's
   atoms: counts (device, n_atoms(c) int32, or NULL) and natom (device,
   n_atoms(c) * width int32, or NULL) with each row's atoms as local
   indices, -1 past its end; rows longer than width are cut.  Neighbour
   rows are ordered by (d^2, index), buried rows by index.  Atoms outside
   the interface region have rows only in a full_state generation.
   Counterpart of calc.run.neighbor_array / .buried_array .nneighbors and
   .natom (cms-cuda/_core.py:1318-1348; _gpu.py:270-296). */
int cms_copy_rows(cms_ctx* ctx, int complex_index, int which, int32_t* counts, int32_t* natom, int width);

/* Trimmed dots of (complex, molecule) -- the buried dots farther than the
   band from every accessible dot of their molecule, in dot order -- after
   cms_shape_complementarity (or cms_run with SC) on the current surface;
   fields as cms_copy_dots.  Counterpart of calc.run.trimmed_dots
   (cms-cuda/_core.py:1037-1077, 828-838). */
int cms_trimmed_count(cms_ctx* ctx, int complex_index, int mol, int64_t* n);
int cms_copy_trimmed_dots(cms_ctx* ctx, int complex_index, int mol, double* coor, double* nml, double* area,
                          int8_t* buried, int8_t* type, int32_t* atom);

/* Toroid queue of complex c in the reference's order: atoms (device, 2n
   int32 local indices: atom 1, atom 2), uij (3n), tij (3n), rij (n) float64,
   between (n int8); any pointer may be NULL.  Synchronous.  Counterpart of
   calc.run.toroid_queue (cms-cuda/_core.py:1419-1431; _gpu.py:262-296). */
int cms_toroid_count(cms_ctx* ctx, int complex_index, int64_t* n);
int cms_copy_toroid_queue(cms_ctx* ctx, int complex_index, int32_t* atoms, double* uij, double* tij, double* rij,
                          int8_t* between);

/* Standalone nearest-point search: for each of nq query points (device,
   3*nq float64) the squared distance to the nearest of nr reference
   points (device, 3*nr) and, if out_idx != NULL, its first index.  The
   primitive behind every readout (replaces _gpu_kernels.min_sqdist,
   cms-cuda/_gpu_kernels.py:508-566). */
int cms_nearest(cms_ctx* ctx, int precision, const double* q, int64_t nq, const double* r,
                int64_t nr, double* out_d2, int32_t* out_idx);

/* cms_nearest without any device allocation.  No counterpart in the
   reference: it is for callers that cannot allocate device memory (code
   under HIP graph capture, tight serving loops, harnesses that forbid
   hipMalloc).  The working buffers live in one fixed device array of
   21 MiB per GPU (held until the process ends), so nq and nr are each at
   most 131,072 (CMS_ERR_INVALID beyond).  Results equal cms_nearest's bit
   for bit.  All calls on one GPU share that array: calls that overlap in
   time on one GPU (other streams or host threads) race.  Synchronises the
   context's stream before returning. */
int cms_nearest_noalloc(cms_ctx* ctx, int precision, const double* q, int64_t nq, const double* r, int64_t nr,
                        double* out_d2, int32_t* out_idx);

/* Backward pass of cms_contact_ms at frozen topology (dot existence,
   burial, nearest partners and every subdivision count kept from the
   forward pass; dot positions and areas differentiated through the one to
   three atoms that generate each dot, recomputed here rather than stored).
   g_totals: device, 2B upstream gradients of the totals (or NULL);
   g_per_atom: device, ntot upstream gradients of the per-atom CMS (or
   NULL); grad_xyz: device, 3*ntot, overwritten with dL/dxyz.  Requires the
   last call on the context to have been cms_contact_ms on the current
   surface.  No reference counterpart (the reference has no gradient). */
int cms_backward_cms(cms_ctx* ctx, const double* g_totals, const double* g_per_atom, double* grad_xyz);

/* CMS totals (device, 2B) at other atom coordinates xyz_eval (device,
   3*ntot) with the topology of the current surface frozen -- the function
   cms_backward_cms differentiates; used for finite-difference checks. */
int cms_frozen_cms(cms_ctx* ctx, const double* xyz_eval, double* totals);

/* Re-evaluates every dot from its descriptor at the forward coordinates
   and counts those that differ from the stored dot (expected 0). */
int cms_check_dots(cms_ctx* ctx, int64_t* mismatches, double* maxdev);

/* Number of references rescored in FP64 by CMS_PREC_FP32_RESCORE searches
   since the last call (then reset).  No reference counterpart (the
   reference searches in FP64 only). */
int cms_search_stats(cms_ctx* ctx, uint64_t* rescored);

/* Kernel timing with HIP events on the context's stream.  filter: NULL or
   "" = off, "*" = every instrumented kernel, otherwise one kernel name
   ("attention", "neighbors", "pairs_count", "probes_count", "convex_emit",
   "toroid_emit", "concave_emit", "nearest", "owner_sums", "sc_sides", ...).
   cms_profile_read synchronises, returns the summed milliseconds and the
   number of timed launches of `name` ("*" = all) recorded since the last
   read, and forgets them.  No reference counterpart (the reference times
   stages with NVTX ranges, cms-cuda/_gpu.py:30). */
int cms_profile(cms_ctx* ctx, const char* filter);
int cms_profile_read(cms_ctx* ctx, const char* name, double* total_ms, int64_t* count);

#ifdef __cplusplus
}
#endif

#endif /* CMS_HIP_H */
