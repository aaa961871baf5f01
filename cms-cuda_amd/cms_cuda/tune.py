"""Autotuner of the launch shapes, one table entry per batch-size bucket.

    python -m cms_cuda.tune [--out tuning.json] [--buckets single,small,medium,large]

For each bucket it builds a representative seeded workload (single: one
1,800 + 4,000-atom complex in FP64, the plain per-complex path; small: 64
complexes of 400 + 1,000 atoms; medium: 256 complexes with 1,500-atom
targets and 80-600-atom binders; large: 2,048 mixed complexes plus 64 of
400 + 1,000 atoms; the batches with the FP32 search and graph replay), and
runs a coordinate descent over the launch-shape fields (two passes; a field
value is kept when a step gets at least 1 % faster), timing each candidate
as the best of three runs of several steps.  Every candidate's results
(CMS totals, per-atom CMS, SC records) must equal the built-in rules'
results bit for bit, or the tuner stops with an error.  The table records
the choice, the default and tuned step times and every trial.
"""

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

CANDIDATES = (
    ("probe_count", (1, 2)),
    ("probe_slots", (2, 4, 8)),
    ("trim_lanes", (1, 4, 32)),
    ("cell_sweep", (1, 2)),
    ("search_split", (512, 1024, 2048, 4096, 8192)),
)


def workloads():
    from . import synth
    workers = max(1, min(16, len(os.sched_getaffinity(0))))
    return {
        "single": ("one 1,800 + 4,000-atom complex, FP64, no graph", lambda: [synth.config2(seed=2)], 0, False, 20),
        "small": ("64 complexes of 400 + 1,000 atoms, FP32 search, graph",
                  lambda: synth.generate([(1000, 400, 9100 + i, 4.0) for i in range(64)], workers), 1, True, 10),
        "medium": ("256 complexes, 1,500-atom targets, binders 80-600 atoms, FP32 search, graph",
                   lambda: synth.c5_batch(5, workers), 1, True, 10),
        "large": ("2,048 mixed complexes + 64 of 400 + 1,000 atoms, FP32 search, graph",
                  lambda: synth.c4_batch(3, workers), 1, True, 5),
    }


class Runner:
    def __init__(self, cs, precision, graph):
        import torch
        from . import _lib
        from ._calc import Settings
        self.torch = torch
        dev = torch.device("cuda", 0)
        self.B = len(cs)
        nb = [len(c["binder_radii"]) for c in cs]
        nt = [len(c["target_radii"]) for c in cs]
        xyz = np.concatenate([np.concatenate([c["binder_xyz"], c["target_xyz"]]) for c in cs])
        rad = np.concatenate([np.concatenate([c["binder_radii"], c["target_radii"]]) for c in cs])
        self.ntot = len(rad)
        self.x = torch.from_numpy(xyz).to(dev)
        self.r = torch.from_numpy(rad).to(dev)
        self.pa = torch.empty(self.ntot, dtype=torch.float64, device=dev)
        self.tot = torch.empty(2 * self.B, dtype=torch.float64, device=dev)
        self.sc = torch.empty(self.B * _lib.SC_NRES, dtype=torch.float64, device=dev)
        self.ctx = _lib.Context(0, torch.cuda.current_stream(dev).cuda_stream)
        self.st = Settings().to_c()
        self.nb = (ctypes.c_int32 * self.B)(*nb)
        self.nt = (ctypes.c_int32 * self.B)(*nt)
        self.prec = precision
        self.ctx("cms_set_graph", 1 if graph else 0)
        self.ctx("cms_set_keep_descriptors", 0)

    def step(self):
        P = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
        self.ctx("cms_load", self.B, self.nb, self.nt, P(self.x), P(self.r), None, ctypes.byref(self.st))
        self.ctx("cms_run", self.prec, P(self.pa), P(self.tot), P(self.sc))

    def measure(self, ch, steps):
        from . import tuning
        tuning.apply(self.ctx, ch)
        for _ in range(3):
            self.step()
        self.torch.cuda.synchronize()
        best = float("inf")
        for _ in range(3):
            t0 = time.perf_counter()
            for _ in range(steps):
                self.step()
            self.torch.cuda.synchronize()
            best = min(best, (time.perf_counter() - t0) / steps)
        res = (self.tot.cpu().numpy().copy(), self.pa.cpu().numpy().copy(), self.sc.cpu().numpy().copy())
        return best * 1e3, res

    def close(self):
        self.ctx.close()


def tune_bucket(name, desc, make, precision, graph, steps, log):
    from . import tuning
    cs = make()
    run = Runner(cs, precision, graph)
    cur = tuning.default()
    t_def, ref = run.measure(cur, steps)
    trials = [{"choice": dict(cur), "ms": t_def}]
    best = t_def
    log("%s: %d complexes, %d atoms; built-in rules %.3f ms/step" % (name, len(cs), run.ntot, t_def))
    for _ in range(2):
        for field, values in CANDIDATES:
            for v in values:
                cand = dict(cur, **{field: v})
                if cand == cur:
                    continue
                t, res = run.measure(cand, steps)
                if not all(np.array_equal(a, b) for a, b in zip(res, ref)):
                    raise RuntimeError("%s: launch shapes %r changed the results" % (name, cand))
                trials.append({"choice": dict(cand), "ms": t})
                if t < best * 0.99:
                    cur, best = cand, t
                    log("  %s = %d: %.3f ms/step" % (field, v, t))
    run.close()
    return {"workload": desc, "atoms": run.ntot, "complexes": len(cs), "choice": cur, "default_ms": t_def,
            "tuned_ms": best, "speedup": t_def / best, "trials": trials}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--out", default=None, help="table to write (default: the package's tuning.json)")
    ap.add_argument("--buckets", default="single,small,medium,large")
    args = ap.parse_args(argv)
    import torch
    from . import tuning

    def log(msg):
        print("[tune] " + msg, file=sys.stderr, flush=True)

    assert torch.cuda.is_available(), "the tuner needs a GPU"
    wl = workloads()
    out = {"device": torch.cuda.get_device_name(0), "buckets": {},
           "bucket_bounds": {nm: lo for nm, lo in tuning.BUCKETS}}
    for name in args.buckets.split(","):
        desc, make, prec, graph, steps = wl[name]
        out["buckets"][name] = tune_bucket(name, desc, make, prec, graph, steps, log)
        torch.cuda.empty_cache()
    path = args.out or tuning.TABLE_PATH
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
    log("wrote %s" % path)


if __name__ == "__main__":
    main()
