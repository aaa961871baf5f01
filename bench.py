#!/usr/bin/env python
"""Benchmark: complexes/s of CMS (both sides) + SC on one MI355X.

Default (the headline): the c4 workload -- 2,048 mixed complexes (binders
log-uniform 20-1,800 atoms, targets 800-6,000, gaps 3.5-5.0 A) plus 64
complexes of 400 + 1,000 atoms -- in one batch, nearest-dot searches in FP32
with FP64 rescoring, every step replayed from a captured HIP graph.  Beside
it, in the same run:
  * c3: the 2,048-complex batch in plain FP64, graph-replayed (what batching
    alone gives, without the FP32 search);
  * the plain per-complex FP64 path on the 1,800 + 4,000-atom complex (c2),
    one complex per step -- the "vs plain FP64" of the metric;
  * the drop-in calls (calculate_contact_ms, calculate_shape_complementarity,
    calculate_maximum_possible_contact_ms) timed end to end on c1 and c2,
    host arrays in and out, warm;
  * the differentiation workload (c5: 256 complexes, 1,500-atom targets,
    binders 80-600 atoms): FP32-search forward + hand-written backward, with
    a finite-difference check of the gradient;
  * FP32-search vs FP64 results on the whole c4 batch;
  * the NumPy oracle on a size-stratified sample of the batch, one core and
    a process pool over all cores, BLAS threads pinned to one.

A step is one pass of the hot path over one batch whose coordinates and
radii are already resident in HBM: pack -> attention -> neighbours -> probes
-> convex / toroidal / concave dots with burial -> CMS both sides -> SC
(-> backward for c5).  Prints one JSON line.
"""

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "cms-cuda_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

# MI355X published peaks.  FP32 without packed instructions (disabled in
# this build) issues at the FP64 rate: 256 CUs x 4 SIMDs x 16 lanes x 2 (FMA)
# x 2.4 GHz = 78.6 TFLOP/s; one wave64 VALU instruction occupies a SIMD for
# four cycles.
PEAK_HBM_GBS = 8000.0
PEAK_FP64_TFLOPS = 78.6
PEAK_VALU_TFLOPS = 78.6
N_SIMD = 1024
CLOCK_HZ = 2.4e9
VALU_ISSUE_PER_S = N_SIMD * CLOCK_HZ / 4.0

METRIC = "complexes/s CMS+SC(+grad): batched-FP32 vs plain FP64, 1 MI355X @≤1e-6 rel"

WORKLOAD_TEXT = {
    "c1": "c1: one 400 + 1,000-atom complex (gap 4.0 A)",
    "c2": "c2: one 1,800 + 4,000-atom complex (scFv-like binder 1:1:1.6, gap 4.0 A)",
    "c3": "c3: 2,048 complexes, binders log-uniform 20-1,800 atoms, targets 800-6,000, gaps 3.5-5.0 A",
    "c4": "c4: the c3 batch + 64 complexes of 400 + 1,000 atoms (2,112 complexes)",
    "c5": "c5: 256 complexes, 1,500-atom targets, binders log-uniform 80-600 atoms, gaps 3.5-5.0 A, "
          "with backward dCMS(target side)/dxyz",
    "mixed": "mixed: %d complexes drawn like c3",
}

KERNELS = ["grid", "cell_lists", "attention", "region", "neighbors", "pairs_emit", "probes_count", "probes_count_long", "probes_emit", "near",
           "convex_frames", "convex_count", "convex_count_long", "convex_emit", "convex_emit_long", "concave_frames", "toroid_count", "toroid_emit", "concave_count", "concave_emit",
           "nearest", "nearest32", "trim_grid", "trim", "owner_sums", "group_sums", "sc_sides", "cms_adjoint", "dot_grad"]


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def workers():
    return max(1, min(16, len(os.sched_getaffinity(0))))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def make_workload(name, batch, seed):
    from cms_cuda import synth
    if name == "c1":
        return [synth.config1(seed=1 + b) for b in range(batch)]
    if name == "c2":
        return [synth.config2(seed=2 + b) for b in range(batch)]
    if name == "c3":
        return synth.c3_batch(seed, workers())
    if name == "c4":
        return synth.c4_batch(seed, workers())
    if name == "c5":
        return synth.c5_batch(seed + 2, workers())
    return synth.mixed_batch(batch, seed=seed, workers=workers())


class Leg:
    """One workload resident on the GPU behind one context."""

    def __init__(self, cs, precision, graph, grad=False, tuned=True):
        import torch
        from cms_cuda import _lib, tuning
        from cms_cuda._calc import Settings
        self.tuning = tuning
        self.torch = torch
        self._lib = _lib
        dev = torch.device("cuda", 0)
        self.cs = cs
        self.B = len(cs)
        self.nb = np.array([len(c["binder_radii"]) for c in cs], dtype=np.int32)
        self.nt = np.array([len(c["target_radii"]) for c in cs], dtype=np.int32)
        xyz = np.concatenate([np.concatenate([c["binder_xyz"], c["target_xyz"]]) for c in cs])
        rad = np.concatenate([np.concatenate([c["binder_radii"], c["target_radii"]]) for c in cs])
        self.ntot = len(rad)
        self.d_xyz = torch.from_numpy(np.ascontiguousarray(xyz)).to(dev)
        self.d_rad = torch.from_numpy(np.ascontiguousarray(rad)).to(dev)
        self.per_atom = torch.empty(self.ntot, dtype=torch.float64, device=dev)
        self.totals = torch.empty(2 * self.B, dtype=torch.float64, device=dev)
        self.scres = torch.empty(self.B * _lib.SC_NRES, dtype=torch.float64, device=dev)
        self.grad = grad
        if grad:
            g = np.zeros(2 * self.B)
            g[1::2] = 1.0       # d(target-side CMS)/dxyz, summed over the batch
            self.g_tot = torch.from_numpy(g).to(dev)
            self.g_xyz = torch.empty(3 * self.ntot, dtype=torch.float64, device=dev)
        self.stream = torch.cuda.current_stream(dev)
        self.ctx = _lib.Context(0, self.stream.cuda_stream)
        # dot descriptors only for the backward pass
        self.ctx("cms_set_keep_descriptors", 1 if grad else 0)
        self.st = Settings().to_c()
        self.nb_c = (ctypes.c_int32 * self.B)(*self.nb.tolist())
        self.nt_c = (ctypes.c_int32 * self.B)(*self.nt.tolist())
        self.prec = precision
        self.graph = 1 if graph else 0
        self.ctx("cms_set_graph", self.graph)
        self.bucket = tuning.bucket(self.ntot)
        self.choice = tuning.choice(self.ntot) if tuned else tuning.default()
        tuning.apply(self.ctx, self.choice)

    def timed_default(self, steps, warmup):
        """The same timing with the library's built-in launch-shape rules."""
        self.tuning.apply(self.ctx, self.tuning.default())
        try:
            return self.timed(steps, warmup)
        finally:
            self.tuning.apply(self.ctx, self.choice)

    def tuning_record(self, value, steps, warmup):
        dv, dms = self.timed_default(steps, warmup)
        return {"bucket": self.bucket, "choice": self.choice, "value_default_rules": dv,
                "ms_per_step_default_rules": dms, "tuned_over_default": value / dv}

    @staticmethod
    def P(t):
        return ctypes.c_void_p(t.data_ptr())

    def step(self, precision=None):
        P = self.P
        prec = self.prec if precision is None else precision
        self.ctx("cms_load", self.B, self.nb_c, self.nt_c, P(self.d_xyz), P(self.d_rad), None, ctypes.byref(self.st))
        self.ctx("cms_run", prec, P(self.per_atom), P(self.totals), P(self.scres))
        if self.grad:
            self.ctx("cms_backward_cms", P(self.g_tot), None, P(self.g_xyz))

    def timed(self, steps, warmup):
        torch = self.torch
        for _ in range(warmup):
            self.step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            self.step()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        return self.B * steps / (t1 - t0), (t1 - t0) * 1e3 / steps

    def breakdown(self):
        """Per-kernel microseconds of one eagerly enqueued step (a graph
        replay records no events)."""
        self.ctx("cms_set_graph", 0)
        self.ctx("cms_profile", b"*")
        self.step()
        self.torch.cuda.synchronize()
        out = {}
        for nm in KERNELS:
            ms, cnt = ctypes.c_double(), ctypes.c_int64()
            self.ctx("cms_profile_read", nm.encode(), ctypes.byref(ms), ctypes.byref(cnt))
            if cnt.value:
                out[nm] = round(ms.value * 1e3, 2)
        self.ctx("cms_profile", b"")
        self.ctx("cms_set_graph", self.graph)
        return out

    def kernel_time(self, name, steps):
        """Average launch duration (s) of one kernel over eagerly enqueued
        steps, from HIP events around each launch on the context's stream."""
        self.ctx("cms_set_graph", 0)
        self.ctx("cms_profile", name.encode())
        for _ in range(steps):
            self.step()
        ms, cnt = ctypes.c_double(), ctypes.c_int64()
        self.ctx("cms_profile_read", name.encode(), ctypes.byref(ms), ctypes.byref(cnt))
        self.ctx("cms_profile", b"")
        self.ctx("cms_set_graph", self.graph)
        return ms.value / max(cnt.value, 1) / 1e3, cnt.value

    def results(self):
        sc = self.scres.cpu().numpy().reshape(self.B, self._lib.SC_NRES).copy()
        return self.totals.cpu().numpy().reshape(self.B, 2).copy(), self.per_atom.cpu().numpy().copy(), sc

    def counts(self):
        out = np.zeros(self.B * self._lib.NCOUNTS, dtype=np.int64)
        self.ctx("cms_counts", ctypes.c_void_p(out.ctypes.data))
        return out.reshape(self.B, self._lib.NCOUNTS)

    def stats(self):
        s = (ctypes.c_int64 * 4)()
        self.ctx("cms_run_stats", s)
        return {"captures": s[0], "replays": s[1], "exact_runs": s[2], "overflow_reruns": s[3]}

    def close(self):
        self.ctx.close()


def stage_models(counts, sc, nb, nt, fp32):
    """Algorithmic work of the stages for the loaded batch (the per-unit
    figures of the scope table, with the run's own counters): name ->
    (bound, work per launch, unit, peak, definition, dense_equivalent)."""
    N0 = nb.astype(np.float64)
    N1 = nt.astype(np.float64)
    R = counts[:, 8].astype(np.float64)
    En = counts[:, 12].astype(np.float64)
    Eb = counts[:, 13].astype(np.float64)
    Np = counts[:, 9].astype(np.float64)
    B0, B1 = counts[:, 2].astype(np.float64), counts[:, 3].astype(np.float64)
    A0, A1 = counts[:, 0] - counts[:, 2], counts[:, 1] - counts[:, 3]
    T0, T1 = sc[:, 13], sc[:, 21]
    # the two nearest-dot searches of a step: CMS partners (buried vs the
    # other side's buried) and SC partners (trimmed vs the other side's trimmed)
    search = 8.0 * (2 * B0 * B1 + 2 * T0 * T1)
    trim = 8.0 * (B0 * A0 + B1 * A1)
    m = {
        "attention": ("valu", float(np.sum(8.0 * N0 * N1)), "TFLOP/s", PEAK_FP64_TFLOPS,
                      "8 flops per binder-target atom pair (the dense minimum distance the cell sweep replaces)", True),
        "neighbors": ("hbm", float(np.sum(34.0 * R + 12.0 * En + 4.0 * Eb)), "GB/s", PEAK_HBM_GBS,
                      "34 B per row atom + 12 B per neighbour entry + 4 B per buried entry", False),
        "probes_count": ("hbm", float(np.sum(36.0 * En + 108.0 * Np)), "GB/s", PEAK_HBM_GBS,
                         "36 B per neighbour entry (index + coordinates, rows read once) + 108 B per pair "
                         "(at most: pair record 17 B, circle 64 B, toroidal counts, flag and free-point masks "
                         "25 B out)", False),
        "trim": ("valu", float(np.sum(trim)), "TFLOP/s", PEAK_FP64_TFLOPS,
                 "8 flops per (buried, accessible) dot pair of a side (the dense search the cell-grid trim "
                 "replaces)", True),
    }
    # two searches per step share the name; brute force, so the modelled
    # pairs are the pairs evaluated
    name = "nearest32" if fp32 else "nearest"
    m[name] = ("valu", float(np.sum(search)) / 2.0, "TFLOP/s", PEAK_VALU_TFLOPS,
               "8 flops per (query, reference) dot pair over the two searches of a step", False)
    return m


def pipeline_roofline(counts, sc, nb, nt, exec_flops):
    """T_roof = sum over stages of max(bytes / BW, flops / peak) with the
    scope table's per-stage work (S1-S5); also with the attention and trim
    flops replaced by the FP64 flops their kernels executed (from counters),
    when known."""
    N0 = nb.astype(np.float64)
    N1 = nt.astype(np.float64)
    N = N0 + N1
    R = counts[:, 8].astype(np.float64)
    En = counts[:, 12].astype(np.float64)
    Eb = counts[:, 13].astype(np.float64)
    P = counts[:, 9].astype(np.float64)
    D0, D1 = counts[:, 0].astype(np.float64), counts[:, 1].astype(np.float64)
    B0, B1 = counts[:, 2].astype(np.float64), counts[:, 3].astype(np.float64)
    A0, A1 = D0 - B0, D1 - B1
    T0, T1 = sc[:, 13], sc[:, 21]
    bw = PEAK_HBM_GBS * 1e9
    fp = PEAK_FP64_TFLOPS * 1e12
    s1 = float(np.sum(8.0 * N0 * N1))
    s2 = float(np.sum(34.0 * N + 12.0 * En + 4.0 * Eb))
    s3 = float(np.sum(34.0 * N + 36.0 * En + 68.0 * P + 62.0 * (D0 + D1)))
    s4_search = float(np.sum(8.0 * (2 * B0 * B1 + 2 * T0 * T1)))
    s4_trim = float(np.sum(8.0 * (B0 * A0 + B1 * A1)))
    s5 = float(np.sum(20.0 * (B0 + B1) + 8.0 * N + 32.0 * (T0 + T1)))
    t = s1 / fp + s2 / bw + s3 / bw + (s4_search + s4_trim) / fp + s5 / bw
    out = {"t_roof_ms": t * 1e3,
           "stages": {"S1_attention_flops": s1, "S2_neighbours_bytes": s2, "S3_surface_bytes": s3,
                      "S4_search_flops": s4_search, "S4_trim_flops": s4_trim, "S5_readout_bytes": s5}}
    if exec_flops.get("attention") is not None and exec_flops.get("trim") is not None:
        te = exec_flops["attention"] / fp + s2 / bw + s3 / bw + (s4_search + exec_flops["trim"]) / fp + s5 / bw
        out["t_roof_executed_ms"] = te * 1e3
    return out


def load_profile(name, workload, prec_name):
    path = os.path.join(ROOT, "profiles", "%s_%s.json" % (name, workload))
    if not os.path.exists(path):
        return {}
    with open(path) as fh:
        d = json.load(fh)
    return d if d.get("precision") == prec_name else {}


def valu_figures(c, avg_s):
    """Executed-work figures of one kernel from its SQ counters (per
    launch) and its live average duration: share of the chip's VALU issue
    slots used, lanes active per VALU instruction, executed FP64 + FP32
    flops (counted per wave instruction, times 64 lanes times the lane
    activity), and LDS bank-conflict cycles per LDS-active cycle."""
    if not c:
        return None
    lanes = c["SQ_THREAD_CYCLES_VALU"] / (64.0 * c["SQ_ACTIVE_INST_VALU"]) if c.get("SQ_ACTIVE_INST_VALU") else None
    flops = (c.get("SQ_INSTS_VALU_FLOPS_FP64", 0.0) + c.get("SQ_INSTS_VALU_FLOPS_FP32", 0.0)) * 64.0 * (lanes or 0.0)
    out = {"valu_issue_frac": c["SQ_INSTS_VALU"] / (VALU_ISSUE_PER_S * avg_s),
           "lanes_active": lanes,
           "executed_tflops": flops / avg_s / 1e12,
           "executed_flops_per_launch": flops,
           "frac_of_valu_peak": flops / avg_s / 1e12 / PEAK_VALU_TFLOPS,
           "vgprs": c.get("vgpr"), "scratch_bytes": c.get("scratch")}
    if c.get("SQ_ACTIVE_INST_LDS"):
        out["lds_bank_conflict_per_active"] = c.get("SQ_LDS_BANK_CONFLICT", 0.0) / c["SQ_ACTIVE_INST_LDS"]
    return out


# ------------------------------------------------------------ CPU baseline

def _oracle_one(c):
    from oracle.cms_oracle import OracleCalculator
    o = OracleCalculator()
    o.add_binder_and_target(c["binder_xyz"], c["binder_radii"], c["target_xyz"], c["target_radii"])
    ot, _ = o.contact_ms(True)
    ob, _ = o.contact_ms(False)
    osc = o.shape_complementarity()
    return [ot, ob, osc[0], osc[1], osc[2]]


def _single_core_run(payload):
    """Oracle over the sequence for up to budget seconds in this process."""
    seq, budget = payload
    out = []
    t0 = time.perf_counter()
    for idx, c in seq:
        out.append((idx, _oracle_one(c)))
        if time.perf_counter() - t0 >= budget:
            break
    return out, time.perf_counter() - t0


def cpu_baseline(cs, budget, gpu_tot, gpu_sc, all_cores):
    """The NumPy oracle (the CPU restatement of the reference's calculator)
    on a size-stratified sample of the workload, with BLAS / OpenMP threads
    pinned to one: one core in a child process, then (all_cores) a process
    pool with one worker per available core."""
    import concurrent.futures
    import multiprocessing
    sizes = np.array([len(c["binder_radii"]) + len(c["target_radii"]) for c in cs])
    order = np.argsort(sizes, kind="stable")
    stride = max(1, len(cs) // 16)
    seq = [int(order[(k * stride + k // 16) % len(cs)]) for k in range(len(cs))]
    keys = ("binder_xyz", "binder_radii", "target_xyz", "target_radii")
    items = [(i, {k: cs[i][k] for k in keys}) for i in seq]
    saved = {v: os.environ.get(v) for v in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS")}
    for v in saved:
        os.environ[v] = "1"
    ctx = multiprocessing.get_context("spawn")
    errs = []

    def check(idx, ref):
        got = [gpu_tot[idx, 1], gpu_tot[idx, 0], gpu_sc[idx, 0], gpu_sc[idx, 1], gpu_sc[idx, 2]]
        errs.extend(abs(a - b) / max(1.0, abs(a), abs(b)) for a, b in zip(got, ref))

    try:
        # one core: a child process (BLAS threads pinned before numpy loads);
        # the sample is cut to what the budget can take so that few
        # complexes travel to the child
        with concurrent.futures.ProcessPoolExecutor(max_workers=1, mp_context=ctx) as ex:
            res, el = ex.submit(_single_core_run, (items[:400], budget)).result()
        for idx, ref in res:
            check(idx, ref)
        one = {"value": len(res) / el, "complexes": len(res), "seconds": el,
               "mean_atoms": float(np.mean([sizes[i] for i, _ in res]))}
        log("CPU oracle, 1 core: %d complexes in %.1f s" % (len(res), el))
        pool = None
        if all_cores:
            nw = len(os.sched_getaffinity(0))
            done = []
            t0 = time.perf_counter()
            with concurrent.futures.ProcessPoolExecutor(max_workers=nw, mp_context=ctx) as ex:
                it = iter(items)
                ids = {}

                def submit():
                    nxt = next(it, None)
                    if nxt is not None:
                        f = ex.submit(_oracle_one, nxt[1])
                        ids[f] = nxt[0]

                for _ in range(2 * nw):
                    submit()
                pending = set(ids)
                # keep 2 tasks per worker in flight until the budget is spent,
                # then let the ones in flight finish
                while pending:
                    fin, pending = concurrent.futures.wait(pending, return_when=concurrent.futures.FIRST_COMPLETED)
                    for f in fin:
                        done.append((ids[f], f.result()))
                        if time.perf_counter() - t0 < budget:
                            before = set(ids)
                            submit()
                            pending |= set(ids) - before
                el_all = time.perf_counter() - t0
            for idx, ref in done:
                check(idx, ref)
            pool = {"value": len(done) / el_all, "complexes": len(done), "seconds": el_all, "cores": nw,
                    "mean_atoms": float(np.mean([sizes[i] for i, _ in done])) if done else 0.0}
            log("CPU oracle, %d cores: %d complexes in %.1f s" % (nw, len(done), el_all))
    finally:
        for v, val in saved.items():
            if val is None:
                os.environ.pop(v, None)
            else:
                os.environ[v] = val
    return one, pool, (max(errs) if errs else None)


# ------------------------------------------------------------ other legs

def latency_leg(reps=5):
    """The reference's own measurement: warm per-call time of the drop-in
    calls on one complex, host arrays in and results out (minimum of reps
    calls after one warm-up call)."""
    import cms_cuda
    from cms_cuda import synth
    out = {}
    for name, c in (("c1", synth.config1(seed=1)), ("c2", synth.config2(seed=2))):
        args = (c["binder_xyz"], c["binder_radii"], c["target_xyz"], c["target_radii"])
        calls = {
            "calculate_contact_ms": lambda: cms_cuda.calculate_contact_ms(*args),
            "calculate_shape_complementarity": lambda: cms_cuda.calculate_shape_complementarity(*args),
            "calculate_maximum_possible_contact_ms":
                lambda: cms_cuda.calculate_maximum_possible_contact_ms(c["target_xyz"], c["target_radii"]),
        }
        rec = {"atoms": "%d + %d" % (len(c["binder_radii"]), len(c["target_radii"]))}
        for k, fn in calls.items():
            fn()
            best = float("inf")
            for _ in range(reps):
                t0 = time.perf_counter()
                fn()
                best = min(best, time.perf_counter() - t0)
            rec[k + "_ms"] = best * 1e3
        out[name] = rec
        log("drop-in calls %s: %s" % (name, json.dumps(rec)))
    return out


def fd_check(cs, atoms_per_complex=8, seed=7):
    """Gradient of the summed target- and binder-side CMS of a batch against
    central differences of the frozen-topology CMS (FP64, h = 1e-4 A), one
    perturbed atom in every complex per evaluation."""
    import torch
    from cms_cuda import _lib
    from cms_cuda.autograd import _ContactMS
    nb = np.array([len(c["binder_radii"]) for c in cs])
    nt = np.array([len(c["target_radii"]) for c in cs])
    xyz = np.concatenate([np.concatenate([c["binder_xyz"], c["target_xyz"]]) for c in cs])
    rad = np.concatenate([np.concatenate([c["binder_radii"], c["target_radii"]]) for c in cs])
    B = len(cs)
    x0 = torch.from_numpy(xyz).cuda()
    st = _ContactMS(x0, torch.from_numpy(rad).cuda(), nb, nt, None, _lib.PREC_FP32_RESCORE)
    grad = st.backward(torch.ones(2 * B, dtype=torch.float64, device="cuda"), None).cpu().numpy()
    off = np.concatenate([[0], np.cumsum(nb + nt)])
    rng = np.random.default_rng(seed)
    picks = np.zeros((B, atoms_per_complex), dtype=np.int64)
    for c in range(B):
        act = off[c] + np.nonzero(np.abs(grad[off[c]:off[c + 1]]).sum(1) > 0)[0]
        a = rng.choice(act, size=min(atoms_per_complex, len(act)), replace=False) if len(act) else np.zeros(0, int)
        rest = rng.choice(np.setdiff1d(np.arange(off[c], off[c + 1]), a), size=atoms_per_complex - len(a),
                          replace=False)
        picks[c] = np.concatenate([a, rest])
    h = 1e-4

    def total(x):
        t = st.frozen_totals(x).cpu().numpy().reshape(B, 2)
        return t[:, 0] + t[:, 1]

    fd = np.zeros((B, atoms_per_complex, 3))
    rich = np.zeros_like(fd)
    for k in range(atoms_per_complex):
        idx = torch.from_numpy(picks[:, k]).cuda()
        for d in range(3):
            v = []
            for hh in (h, h / 2):
                xp = x0.clone()
                xp[idx, d] += hh
                xm = x0.clone()
                xm[idx, d] -= hh
                v.append((total(xp) - total(xm)) / (2 * hh))
            fd[:, k, d] = v[0]
            rich[:, k, d] = (4 * v[1] - v[0]) / 3
    st.close()
    g = grad[picks]
    scale = np.maximum(np.maximum(np.abs(g), np.abs(rich)), 1e-3)
    gn = np.linalg.norm(g.reshape(B, -1), axis=1)
    live = gn > 0
    e_norm = np.linalg.norm((g - fd).reshape(B, -1)[live], axis=1) / gn[live]
    return {"h": h, "atoms_per_complex": atoms_per_complex, "components": int(g.size),
            "max_rel_err_plain_per_component": float((np.abs(g - fd) / scale).max()),
            "max_rel_err_richardson_per_component": float((np.abs(g - rich) / scale).max()),
            "max_rel_err_plain_2norm_per_complex": float(e_norm.max()) if live.any() else 0.0,
            "relative_to": "max(|g|, |FD|, 1e-3) per component; the gradient's 2-norm per complex"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c4", choices=["c1", "c2", "c3", "c4", "c5", "mixed"])
    ap.add_argument("--batch", type=int, default=1, help="complexes per step for c1 / c2 / mixed")
    ap.add_argument("--precision", default="auto", choices=["auto", "fp64", "fp32"])
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--no-graph", action="store_true", help="enqueue every step instead of replaying a HIP graph")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain per-complex FP64 c2 leg")
    ap.add_argument("--no-c3", action="store_true", help="skip the batched FP64 c3 leg")
    ap.add_argument("--no-latency", action="store_true", help="skip the drop-in per-call latency leg")
    ap.add_argument("--no-grad", action="store_true", help="skip the c5 forward + backward leg")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-cpu-pool", action="store_true", help="CPU baseline on one core only")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of each CPU baseline")
    ap.add_argument("--profile-run", action="store_true",
                    help="only the main workload: warm-up and --steps eagerly enqueued steps (for profilers)")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="also run a 512 MiB read + 512 MiB write elementwise kernel three times (the reference "
                         "for scaling FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU; the CMS/SC path shards by complex across GPUs "
              "but no multi-GPU figure is produced here", file=sys.stderr)
        sys.exit(2)

    import torch
    from cms_cuda import _lib

    assert torch.cuda.is_available(), "no GPU"
    torch.cuda.set_device(0)
    batched = args.workload in ("c3", "c4", "c5", "mixed")
    prec_name = args.precision
    if prec_name == "auto":
        prec_name = "fp32" if args.workload in ("c4", "c5") else "fp64"
    prec = _lib.PREC_FP32_RESCORE if prec_name == "fp32" else _lib.PREC_FP64
    graph = not args.no_graph and not args.profile_run

    log("generating the %s workload" % args.workload)
    t0 = time.perf_counter()
    cs = make_workload(args.workload, args.batch, args.seed)
    log("%d complexes in %.1f s" % (len(cs), time.perf_counter() - t0))
    main_leg = Leg(cs, prec, graph, grad=(args.workload == "c5"))

    if args.profile_run:
        value, ms_per_step = main_leg.timed(args.steps, args.warmup)
        if args.traffic_probe:
            xs = torch.ones(1 << 26, dtype=torch.float64, device="cuda")
            for _ in range(3):
                ys = xs * 2.0
            torch.cuda.synchronize()
            del xs, ys
        print(json.dumps({"metric": METRIC, "value": value, "unit": "complexes/s", "n_gpus": 1,
                          "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
                          "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                          "dtype": "f64" if prec_name == "fp64" else "f32-search+f64", "data": "synthetic",
                          "config": {"workload": WORKLOAD_TEXT.get(args.workload, args.workload),
                                     "complexes": len(cs), "precision": prec_name, "graph": False,
                                     "profile_run": True}}), flush=True)
        main_leg.close()
        return

    for _ in range(args.warmup):
        main_leg.step()
    torch.cuda.synchronize()
    breakdown = main_leg.breakdown()
    log("per-kernel us in one step: %s" % json.dumps(breakdown))
    log("timing %d steps (graph %s)" % (args.steps, "on" if graph else "off"))
    value, ms_per_step = main_leg.timed(args.steps, 2)
    log("%.1f complexes/s, %.3f ms/step" % (value, ms_per_step))
    main_tuning = main_leg.tuning_record(value, args.steps, 2)
    log("tuning %s: %s" % (main_leg.bucket, json.dumps(main_tuning)))
    gstats = main_leg.stats()
    tot, per, sc = main_leg.results()
    counts = main_leg.counts()
    models = stage_models(counts, sc, main_leg.nb, main_leg.nt, prec_name == "fp32")
    # HBM bytes per launch (rocprofv3 FETCH_SIZE / WRITE_SIZE, scaled by a
    # copy kernel of known size) and SQ counters per launch, as recorded in
    # profiles/ for this workload and precision
    traffic = load_profile("traffic", args.workload, prec_name).get("bytes_per_launch", {})
    sq = load_profile("sq", args.workload, prec_name).get("per_launch", {})
    roofline = None
    by_kernel = {}
    exec_flops = {}
    for k in models:
        if k not in breakdown:
            continue
        avg_s, launches = main_leg.kernel_time(k, args.steps)
        bound, work, unit, peak, definition, dense = models[k]
        scale = 1e12 if unit == "TFLOP/s" else 1e9
        vf = valu_figures(sq.get(k), avg_s)
        rec = {"bound": bound, "unit": unit, "peak": peak, "traffic": traffic.get(k),
               "avg_launch_us": avg_s * 1e6, "launches": launches, "work": definition}
        if vf:
            rec["valu"] = vf
            exec_flops[k] = vf["executed_flops_per_launch"]
        if dense:
            # the cell sweeps do not execute the dense work; the executed
            # FP64 flops from the counters are the achieved figure
            if work is not None:
                rec["dense_equivalent_per_launch"] = work
                rec["dense_equivalent_tflops"] = work / avg_s / 1e12
            if vf:
                rec["achieved"] = vf["executed_tflops"]
                rec["frac"] = vf["executed_tflops"] / peak
                rec["work_per_launch"] = vf["executed_flops_per_launch"]
            else:
                rec["achieved"] = rec["frac"] = rec["work_per_launch"] = None
        else:
            rec["achieved"] = work / avg_s / scale
            rec["frac"] = rec["achieved"] / peak
            rec["work_per_launch"] = work
        by_kernel[k] = rec
    kname = max((k for k in by_kernel), key=lambda k: by_kernel[k]["avg_launch_us"]) if by_kernel else None
    if kname:
        r = by_kernel[kname]
        roofline = {"bound": r["bound"], "achieved": r["achieved"], "peak": r["peak"], "unit": r["unit"],
                    "frac": r["frac"], "traffic": r["traffic"], "kernel": kname,
                    "avg_launch_us": r["avg_launch_us"], "work_per_launch": r["work_per_launch"], "work": r["work"]}
        if "valu" in r:
            roofline["valu"] = r["valu"]
    pipe = pipeline_roofline(counts, sc, main_leg.nb, main_leg.nt, exec_flops)
    pipe["t_measured_ms"] = ms_per_step
    pipe["frac"] = pipe["t_roof_ms"] / ms_per_step
    if "t_roof_executed_ms" in pipe:
        pipe["frac_executed"] = pipe["t_roof_executed_ms"] / ms_per_step

    if args.traffic_probe:
        # x * 2 over 2^26 float64: 512 MiB read, 512 MiB written per launch
        xs = torch.ones(1 << 26, dtype=torch.float64, device="cuda")
        for _ in range(3):
            ys = xs * 2.0
        torch.cuda.synchronize()
        del xs, ys
        log("traffic probe: 3 launches of a 512 MiB -> 512 MiB elementwise kernel")

    # FP32 search against FP64 on the same batch (surface identical by
    # construction: only the searches change precision)
    fp32_vs_fp64 = None
    if prec_name == "fp32":
        main_leg.ctx("cms_set_graph", 0)
        main_leg.step(_lib.PREC_FP64)
        t64, p64, s64 = main_leg.results()
        c64 = main_leg.counts()
        main_leg.step()
        t32, p32, s32 = main_leg.results()

        def rel(a, b):
            a, b = np.asarray(a, float), np.asarray(b, float)
            return float(np.max(np.abs(a - b) / np.maximum(1.0, np.maximum(np.abs(a), np.abs(b))))) if a.size else 0.0
        fp32_vs_fp64 = {"dot_counts_identical": bool(np.array_equal(c64, main_leg.counts())),
                        "max_rel_err_cms": max(rel(t32, t64), rel(p32, p64)),
                        "max_rel_err_sc": rel(s32[:, :3], s64[:, :3]),
                        "complexes": main_leg.B}
        log("fp32 vs fp64: %s" % json.dumps(fp32_vs_fp64))
        main_leg.ctx("cms_set_graph", 1 if graph else 0)
    main_leg.close()
    torch.cuda.empty_cache()

    c3 = None
    if args.workload == "c4" and not args.no_c3:
        log("c3 leg: the 2,048-complex batch, plain FP64, graph replay")
        leg = Leg(cs[:2048], _lib.PREC_FP64, graph)
        v3, ms3 = leg.timed(args.steps, args.warmup)
        c3 = {"workload": WORKLOAD_TEXT["c3"], "value": v3, "unit": "complexes/s", "ms_per_step": ms3,
              "precision": "fp64", "graph": graph, "tuning": leg.tuning_record(v3, args.steps, 2)}
        leg.close()
        torch.cuda.empty_cache()
        log("c3: %.1f complexes/s" % v3)

    plain = None
    if batched and not args.no_plain:
        from cms_cuda import synth
        log("plain FP64 leg: c2, one complex per step, no graph")
        pl = Leg([synth.config2(seed=2)], _lib.PREC_FP64, False)
        pv, pms = pl.timed(args.steps, args.warmup)
        plain = {"workload": WORKLOAD_TEXT["c2"], "value": pv, "unit": "complexes/s", "ms_per_step": pms,
                 "precision": "fp64", "graph": False}
        pl.close()
        log("plain: %.1f complexes/s" % pv)

    latency = None
    if not args.no_latency:
        log("drop-in per-call latency, c1 and c2")
        latency = latency_leg()

    with_grad = None
    if args.workload != "c5" and batched and not args.no_grad:
        log("c5 leg: FP32-search forward + backward")
        gcs = make_workload("c5", 0, args.seed)
        gl = Leg(gcs, _lib.PREC_FP32_RESCORE, graph, grad=True)
        gv, gms = gl.timed(args.steps, args.warmup)
        g_tuning = gl.tuning_record(gv, args.steps, 2)
        bd = gl.breakdown()
        gl.close()
        log("with grad: %.1f complexes/s; finite-difference check" % gv)
        fd = fd_check(gcs)
        with_grad = {"workload": WORKLOAD_TEXT["c5"], "value": gv, "unit": "complexes/s", "ms_per_step": gms,
                     "precision": "fp32-search+f64",
                     "backward_us": round(bd.get("cms_adjoint", 0) + bd.get("dot_grad", 0), 1),
                     "fd_check": fd, "tuning": g_tuning}
        log("fd check: %s" % json.dumps(fd))

    cpu = None
    cpu_pool = None
    max_err = None
    if not args.no_cpu:
        log("CPU baseline (oracle) for up to %.0f s per leg" % args.cpu_budget)
        one, pool, max_err = cpu_baseline(cs, args.cpu_budget, tot, sc, not args.no_cpu_pool)
        model = cpu_model()
        cpu = {"value": one["value"], "unit": "complexes/s", "cores": 1, "kind": "port",
               "cpu": model, "threads_pinned": "OMP/OPENBLAS/MKL_NUM_THREADS=1",
               "sample": "%d of the %d %s complexes (size-stratified, mean %.0f atoms), CMS both sides + SC "
                         "through the NumPy oracle in %.1f s on one core%s"
                         % (one["complexes"], len(cs), args.workload, one["mean_atoms"], one["seconds"],
                            "; no gradient on the CPU" if args.workload == "c5" else "")}
        if pool:
            cpu_pool = {"value": pool["value"], "unit": "complexes/s", "cores": pool["cores"], "kind": "port",
                        "cpu": model,
                        "sample": "%d complexes of the same stratified order (mean %.0f atoms) through a process "
                                  "pool of %d workers in %.1f s" % (pool["complexes"], pool["mean_atoms"],
                                                                    pool["cores"], pool["seconds"])}

    wl = WORKLOAD_TEXT[args.workload]
    if args.workload == "mixed":
        wl = wl % len(cs)
    elif args.workload in ("c1", "c2") and len(cs) > 1:
        wl = "%s x %d per step" % (wl, len(cs))
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "complexes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec_name == "fp64" else "f32-search+f64",
        "data": "synthetic",
        "config": {"workload": wl, "complexes": len(cs), "precision": prec_name, "graph": graph,
                   "atoms": int(main_leg.ntot)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "cpu_baseline_all_cores": cpu_pool,
        "plain_fp64": plain,
        "speedup_vs_plain_fp64": (value / plain["value"]) if plain else None,
        "batched_fp64_c3": c3,
        "speedup_fp32_search_vs_batched_fp64": (value / c3["value"]) if c3 else None,
        "drop_in_latency": latency,
        "with_grad": with_grad,
        "fp32_vs_fp64": fp32_vs_fp64,
        "gpu_vs_oracle_max_rel_err": max_err,
        "pipeline_roofline": pipe,
        "kernel_us_per_step": breakdown,
        "roofline_by_kernel": by_kernel,
        "graph": gstats,
        "tuning": main_tuning,
        "counts_total": counts.sum(axis=0).tolist(),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
