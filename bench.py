#!/usr/bin/env python
"""Benchmark: complexes/s of CMS (both sides) + SC on one MI355X.

A step is one pass of the hot path over one batch of synthetic complexes
whose coordinates and radii are already resident in HBM:
load (pack) -> attention -> neighbours -> probes -> convex / toroidal /
concave dots with burial -> CMS both sides -> SC.  Default workload: the
antibody-antigen-scale complex (1,800-atom scFv-like binder + 4,000-atom
target, gap 4.0 A, probe 1.7 A, 15 dots/A^2), plain FP64, one complex per
step.  --batch B runs B seeded complexes of that size per step.

Prints one JSON line (see the roofline / cpu_baseline fields).
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


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def make_workload(name, batch, seed0):
    from cms_cuda import synth
    if name == "mixed":
        workers = max(1, min(16, len(os.sched_getaffinity(0))))
        return synth.mixed_batch(batch, seed=seed0, workers=workers)
    cs = []
    for b in range(batch):
        if name == "c1":
            cs.append(synth.config1(seed=seed0 + b))
        else:
            cs.append(synth.config2(seed=seed0 + b))
    return cs


def kernel_work(name, counts, sizes):
    """Algorithmic work of one launch of `name` for the loaded batch:
    ("mfma", flops) for FP64 pair sweeps, ("hbm", bytes) otherwise."""
    if name == "attention":
        # every unordered atom pair of a complex once: 3 sub + 3 mul + 2 add
        return "mfma", sum(8.0 * n * (n - 1) / 2 for n in sizes)
    raise KeyError(name)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "mixed"])
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--batch", type=int, default=1)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of CPU baseline work")
    ap.add_argument("--kernel", default="attention", help="kernel timed for the roofline")
    ap.add_argument("--no-graph", action="store_true", help="enqueue every step instead of replaying a HIP graph")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU; the CMS/SC path shards by complex across GPUs "
              "but no multi-GPU figure is produced here", file=sys.stderr)
        sys.exit(2)

    import torch
    from cms_cuda import _lib

    assert torch.cuda.is_available(), "no GPU"
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    log("generating %d synthetic %s complex(es)" % (args.batch, args.workload))
    cs = make_workload(args.workload, args.batch, args.seed)
    nb = np.array([len(c["binder_radii"]) for c in cs], dtype=np.int32)
    nt = np.array([len(c["target_radii"]) for c in cs], dtype=np.int32)
    xyz = np.concatenate([np.concatenate([c["binder_xyz"], c["target_xyz"]]) for c in cs])
    rad = np.concatenate([np.concatenate([c["binder_radii"], c["target_radii"]]) for c in cs])
    d_xyz = torch.from_numpy(np.ascontiguousarray(xyz)).to(dev)
    d_rad = torch.from_numpy(np.ascontiguousarray(rad)).to(dev)
    ntot = len(rad)
    B = len(cs)
    stream = torch.cuda.current_stream(dev)
    ctx = _lib.Context(0, stream.cuda_stream)
    from cms_cuda._calc import Settings
    st = Settings().to_c()
    per_atom = torch.empty(ntot, dtype=torch.float64, device=dev)
    totals = torch.empty(2 * B, dtype=torch.float64, device=dev)
    scres = torch.empty(B * _lib.SC_NRES, dtype=torch.float64, device=dev)
    nb_c = (ctypes.c_int32 * B)(*nb.tolist())
    nt_c = (ctypes.c_int32 * B)(*nt.tolist())
    P = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731

    prec = _lib.PREC_FP64 if args.precision == "fp64" else _lib.PREC_FP32_RESCORE

    def step():
        # pack the resident coordinates, then surface + CMS (both sides) + SC
        # in one call: one host wait per step, a graph replay once warm
        ctx("cms_load", B, nb_c, nt_c, P(d_xyz), P(d_rad), None, ctypes.byref(st))
        ctx("cms_run", prec, P(per_atom), P(totals), P(scres))

    use_graph = 0 if args.no_graph else 1
    ctx("cms_set_graph", use_graph)
    log("warmup %d" % args.warmup)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    # which kernels take the time (one profiled step, enqueued eagerly: a
    # graph replay records no events)
    ctx("cms_set_graph", 0)
    ctx("cms_profile", b"*")
    step()
    torch.cuda.synchronize()
    ctx("cms_profile", b"")
    ctx("cms_set_graph", use_graph)
    for _ in range(2):
        step()
    names = ["attention", "region", "neighbors", "pairs_count", "pairs_emit", "probes_count",
             "probes_emit", "near", "convex_count", "convex_emit", "toroid_count", "toroid_emit",
             "concave_count", "concave_emit", "nearest", "nearest32", "owner_sums", "sc_sides"]
    breakdown = {}
    for nm in names:
        ms = ctypes.c_double()
        cnt = ctypes.c_int64()
        ctx("cms_profile_read", nm.encode(), ctypes.byref(ms), ctypes.byref(cnt))
        if cnt.value:
            breakdown[nm] = round(ms.value * 1e3, 2)
    log("per-kernel us in one step: %s" % json.dumps(breakdown))

    # timed region
    torch.cuda.synchronize()
    log("timing %d steps (graph %s)" % (args.steps, "on" if use_graph else "off"))
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    stats = (ctypes.c_int64 * 4)()
    ctx("cms_run_stats", stats)
    # the roofline kernel's launch duration: HIP events around that kernel
    # on the context's stream, over the same number of eagerly enqueued
    # steps (events cannot be recorded inside a graph replay)
    ctx("cms_set_graph", 0)
    ctx("cms_profile", args.kernel.encode())
    for _ in range(args.steps):
        step()
    ms = ctypes.c_double()
    cnt = ctypes.c_int64()
    ctx("cms_profile_read", args.kernel.encode(), ctypes.byref(ms), ctypes.byref(cnt))
    ctx("cms_profile", b"")
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = B * args.steps / (t1 - t0)
    counts = np.zeros(B * _lib.NCOUNTS, dtype=np.int64)
    ctx("cms_counts", ctypes.c_void_p(counts.ctypes.data))
    counts = counts.reshape(B, _lib.NCOUNTS)
    sizes = (nb + nt).tolist()
    bound, work = kernel_work(args.kernel, counts, sizes)
    avg_s = ms.value / max(cnt.value, 1) / 1e3
    if bound == "mfma":
        achieved = work / avg_s / 1e12
        peak = PEAK_FP64_TFLOPS
        unit = "TFLOP/s"
    else:
        achieved = work / avg_s / 1e9
        peak = PEAK_HBM_GBS
        unit = "GB/s"
    gpu_tot = totals.cpu().numpy()
    gpu_sc = scres.cpu().numpy().reshape(B, _lib.SC_NRES)

    cpu = None
    max_err = None
    if not args.no_cpu:
        from oracle.cms_oracle import OracleCalculator
        log("CPU baseline (oracle, 1 core) for up to %.0f s" % args.cpu_budget)
        done = 0
        t0 = time.perf_counter()
        errs = []
        while True:
            c = cs[done % B]
            o = OracleCalculator()
            o.add_binder_and_target(c["binder_xyz"], c["binder_radii"], c["target_xyz"], c["target_radii"])
            ocms_t, _ = o.contact_ms(True)
            ocms_b, _ = o.contact_ms(False)
            osc = o.shape_complementarity()
            k = done % B
            got = [gpu_tot[2 * k + 1], gpu_tot[2 * k], gpu_sc[k, 0], gpu_sc[k, 1], gpu_sc[k, 2]]
            ref = [ocms_t, ocms_b, osc[0], osc[1], osc[2]]
            errs += [abs(a - b) / max(1.0, abs(a), abs(b)) for a, b in zip(got, ref)]
            done += 1
            if time.perf_counter() - t0 >= args.cpu_budget:
                break
        el = time.perf_counter() - t0
        cpu = {"value": done / el, "unit": "complexes/s", "cores": 1, "kind": "port",
               "sample": "%d x %s complex (CMS both sides + SC) through the NumPy oracle in %.1f s"
                         % (done, args.workload, el)}
        max_err = max(errs)

    out = {
        "metric": "complexes/s CMS+SC(+grad): batched-FP32 vs plain FP64, 1 MI355X @≤1e-6 rel",
        "value": value,
        "unit": "complexes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32-search+f64",
        "data": "synthetic",
        "config": {"workload": "%s %s CMS(both sides)+SC, %d complex(es)/step, binder %d-%d + target %d-%d atoms"
                               % (args.workload, args.precision, B, nb.min(), nb.max(), nt.min(), nt.max()),
                   "precision": args.precision, "batch": B},
        "roofline": {"bound": bound, "kernel": args.kernel, "achieved": achieved, "peak": peak,
                     "unit": unit, "frac": achieved / peak, "traffic": None,
                     "avg_launch_us": avg_s * 1e6, "launches": cnt.value, "work_per_launch": work},
        "cpu_baseline": cpu,
        "gpu_vs_oracle_max_rel_err": max_err,
        "kernel_us_per_step": breakdown,
        "graph": {"enabled": bool(use_graph), "captures": stats[0], "replays": stats[1],
                  "exact_runs": stats[2], "overflow_reruns": stats[3]},
        "counts_per_complex": counts[0].tolist(),
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
