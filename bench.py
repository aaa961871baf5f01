#!/usr/bin/env python
"""Benchmark: complexes/s of CMS (both sides) + SC on one MI355X.

Default (the headline): the c4 workload -- 2,048 mixed complexes (binders
log-uniform 20-1,800 atoms, targets 800-6,000, gaps 3.5-5.0 A) plus 64
complexes of 400 + 1,000 atoms -- in one batch, nearest-dot searches in FP32
with FP64 rescoring, every step replayed from a captured HIP graph.  Beside
it, in the same run:
  * the plain per-complex FP64 path on the 1,800 + 4,000-atom complex (c2),
    one complex per step -- the "vs plain FP64" of the metric;
  * the differentiation workload (c5: 256 complexes, 1,500-atom targets,
    binders 80-600 atoms): FP32-search forward + hand-written backward;
  * FP32-search vs FP64 results on the whole c4 batch;
  * the NumPy oracle on a size-stratified sample of the batch (1 core).

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

# MI355X published peaks
PEAK_HBM_GBS = 8000.0
PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3

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

KERNELS = ["attention", "region", "neighbors", "pairs_emit", "probes_count", "probes_emit", "near",
           "convex_count", "convex_emit", "toroid_count", "toroid_emit", "concave_count", "concave_emit",
           "nearest", "nearest32", "trim", "owner_sums", "sc_sides", "cms_adjoint", "dot_grad"]


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def workers():
    return max(1, min(16, len(os.sched_getaffinity(0))))


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

    def __init__(self, cs, precision, graph, grad=False):
        import torch
        from cms_cuda import _lib
        from cms_cuda._calc import Settings
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
        self.st = Settings().to_c()
        self.nb_c = (ctypes.c_int32 * self.B)(*self.nb.tolist())
        self.nt_c = (ctypes.c_int32 * self.B)(*self.nt.tolist())
        self.prec = precision
        self.graph = 1 if graph else 0
        self.ctx("cms_set_graph", self.graph)

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


def kernel_models(counts, sc, nb, nt, fp32):
    """Algorithmic work of one launch of each modelled kernel for the loaded
    batch: name -> (bound, work, unit, peak, definition)."""
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
        "attention": ("mfma", float(np.sum(8.0 * N0 * N1)), "TFLOP/s", PEAK_FP64_TFLOPS,
                      "8 flops per binder-target atom pair (dense minimum distance)"),
        "neighbors": ("hbm", float(np.sum(34.0 * R + 12.0 * En + 4.0 * Eb)), "GB/s", PEAK_HBM_GBS,
                      "34 B per row atom + 12 B per neighbour entry + 4 B per buried entry"),
        "probes_count": ("hbm", float(np.sum(36.0 * En + 108.0 * Np)), "GB/s", PEAK_HBM_GBS,
                         "36 B per neighbour entry (index + coordinates, rows read once) + 108 B per torus "
                         "pair (74 B record in, 34 B counts, flags and masks out)"),
    }
    # two searches per step share the name; the SC trim (buried vs the same
    # side's accessible dots within the band) is a cell-grid sweep, modelled by
    # the dense search it replaces
    if fp32:
        m["nearest32"] = ("mfma", float(np.sum(search)) / 2.0, "TFLOP/s", PEAK_FP32_TFLOPS,
                          "8 flops per (query, reference) dot pair over the two searches of a step")
    else:
        m["nearest"] = ("mfma", float(np.sum(search)) / 2.0, "TFLOP/s", PEAK_FP64_TFLOPS,
                        "8 flops per (query, reference) dot pair over the two searches of a step")
    m["trim"] = ("mfma", float(np.sum(trim)), "TFLOP/s", PEAK_FP64_TFLOPS,
                 "8 flops per (buried, accessible) dot pair of a side (the dense search the grid trim replaces)")
    return m


def cpu_baseline(cs, budget, gpu_tot, gpu_sc, log_fn):
    """NumPy oracle on a size-stratified sample of the workload, 1 core."""
    from oracle.cms_oracle import OracleCalculator
    sizes = np.array([len(c["binder_radii"]) + len(c["target_radii"]) for c in cs])
    order = np.argsort(sizes, kind="stable")
    # visit the sorted batch in a stride that covers every size band early
    stride = max(1, len(cs) // 16)
    seq = [int(order[(k * stride + k // 16) % len(cs)]) for k in range(len(cs))]
    done, errs = 0, []
    t0 = time.perf_counter()
    for idx in seq:
        c = cs[idx]
        o = OracleCalculator()
        o.add_binder_and_target(c["binder_xyz"], c["binder_radii"], c["target_xyz"], c["target_radii"])
        ot, _ = o.contact_ms(True)
        ob, _ = o.contact_ms(False)
        osc = o.shape_complementarity()
        got = [gpu_tot[idx, 1], gpu_tot[idx, 0], gpu_sc[idx, 0], gpu_sc[idx, 1], gpu_sc[idx, 2]]
        ref = [ot, ob, osc[0], osc[1], osc[2]]
        errs += [abs(a - b) / max(1.0, abs(a), abs(b)) for a, b in zip(got, ref)]
        done += 1
        if time.perf_counter() - t0 >= budget:
            break
    el = time.perf_counter() - t0
    mean_atoms = float(np.mean([sizes[i] for i in seq[:done]]))
    log_fn("CPU oracle: %d complexes in %.1f s" % (done, el))
    return done / el, done, el, mean_atoms, max(errs)


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
    ap.add_argument("--no-grad", action="store_true", help="skip the c5 forward + backward leg")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of CPU baseline work")
    ap.add_argument("--kernel", default="auto", help="kernel timed for the roofline (auto: the slowest modelled)")
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
    graph = not args.no_graph

    log("generating the %s workload" % args.workload)
    t0 = time.perf_counter()
    cs = make_workload(args.workload, args.batch, args.seed)
    log("%d complexes in %.1f s" % (len(cs), time.perf_counter() - t0))
    main_leg = Leg(cs, prec, graph, grad=(args.workload == "c5"))
    for _ in range(args.warmup):
        main_leg.step()
    torch.cuda.synchronize()
    breakdown = main_leg.breakdown()
    log("per-kernel us in one step: %s" % json.dumps(breakdown))
    log("timing %d steps (graph %s)" % (args.steps, "on" if graph else "off"))
    value, ms_per_step = main_leg.timed(args.steps, 2)
    log("%.1f complexes/s, %.3f ms/step" % (value, ms_per_step))
    gstats = main_leg.stats()
    tot, per, sc = main_leg.results()
    counts = main_leg.counts()
    models = kernel_models(counts, sc, main_leg.nb, main_leg.nt, prec_name == "fp32")
    kname = args.kernel
    if kname == "auto":
        cand = [k for k in models if k in breakdown]
        kname = max(cand, key=lambda k: breakdown[k]) if cand else "attention"
    # HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE on
    # this workload (scaled by a copy kernel of known size), when recorded
    traffic = {}
    tpath = os.path.join(ROOT, "profiles", "traffic_%s.json" % args.workload)
    if os.path.exists(tpath):
        with open(tpath) as fh:
            tr = json.load(fh)
        if tr.get("precision") == prec_name:
            traffic = tr.get("bytes_per_launch", {})
    roofline = None
    by_kernel = {}
    for k in models:
        if k not in breakdown:
            continue
        avg_s, launches = main_leg.kernel_time(k, args.steps)
        bound, work, unit, peak, definition = models[k]
        scale = 1e12 if unit == "TFLOP/s" else 1e9
        ach = work / avg_s / scale
        rec = {"bound": bound, "achieved": ach, "peak": peak, "unit": unit, "frac": ach / peak,
               "traffic": traffic.get(k), "avg_launch_us": avg_s * 1e6, "launches": launches,
               "work_per_launch": work, "work": definition}
        by_kernel[k] = rec
        if k == kname:
            roofline = dict(rec, kernel=k)

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

    with_grad = None
    if args.workload != "c5" and batched and not args.no_grad:
        log("c5 leg: FP32-search forward + backward")
        gcs = make_workload("c5", 0, args.seed)
        gl = Leg(gcs, _lib.PREC_FP32_RESCORE, graph, grad=True)
        gv, gms = gl.timed(args.steps, args.warmup)
        bd = gl.breakdown()
        with_grad = {"workload": WORKLOAD_TEXT["c5"], "value": gv, "unit": "complexes/s", "ms_per_step": gms,
                     "precision": "fp32-search+f64",
                     "backward_us": round(bd.get("cms_adjoint", 0) + bd.get("dot_grad", 0), 1)}
        gl.close()
        log("with grad: %.1f complexes/s" % gv)

    cpu = None
    max_err = None
    if not args.no_cpu:
        log("CPU baseline (oracle, 1 core) for up to %.0f s" % args.cpu_budget)
        v, done, el, mean_atoms, max_err = cpu_baseline(cs, args.cpu_budget, tot, sc, log)
        cpu = {"value": v, "unit": "complexes/s", "cores": 1, "kind": "port",
               "sample": "%d of the %d %s complexes (size-stratified, mean %.0f atoms), CMS both sides + SC "
                         "through the NumPy oracle in %.1f s%s"
                         % (done, len(cs), args.workload, mean_atoms, el,
                            "; no gradient on the CPU" if args.workload == "c5" else "")}

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
        "plain_fp64": plain,
        "speedup_vs_plain_fp64": (value / plain["value"]) if plain else None,
        "with_grad": with_grad,
        "fp32_vs_fp64": fp32_vs_fp64,
        "gpu_vs_oracle_max_rel_err": max_err,
        "kernel_us_per_step": breakdown,
        "roofline_by_kernel": by_kernel,
        "graph": gstats,
        "counts_total": counts.sum(axis=0).tolist(),
    }
    print(json.dumps(out), flush=True)
    main_leg.close()


if __name__ == "__main__":
    main()
