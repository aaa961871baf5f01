#!/usr/bin/env python
"""Write reference vectors with the reference's own NumPy calculator.

Runs cms-cuda/_core.py (loaded by file path from the reference checkout,
default /root/reference, or $CMS_REFERENCE) in the development container on
inputs the known-answer complex does not cover, and stores inputs and
outputs as .npz fixtures next to this script.  The reference itself is not
copied anywhere; only these vectors are committed.

    python tests/golden/make_golden.py

Fixtures (every array float64 unless noted):
  ref_small.npz   a 260 + 80-atom synthetic complex (gap 0.5 A): inputs, CMS of both
                  sides with per-atom arrays, SC record, every dot of both
                  molecules (coordinates, normals, areas, buried, type,
                  owner), probes
  ref_golden.npz  the known-answer complex: per-atom CMS of both sides, SC
                  side statistics, per-molecule dot checksums and 600
                  sampled dots
  ref_radius0.npz the known-answer complex with every 4th binder radius 0
  ref_maxcms.npz  maximum possible CMS of the known-answer binder alone
  ref_c1.npz      a 400 + 1,000-atom complex (gap 4 A): CMS both sides,
                  per-atom arrays, SC record, dot counts
  ref_examples.npz the reference's five example complexes
                  (example_*.pdb, selections as the reference's README
                  runs them), inputs as cms_cuda.structure.load_atoms reads
                  them: CMS of both sides with per-atom arrays, SC record,
                  counts, every dot's owner / burial / type in order, dot
                  checksums and 400 sampled dots per molecule; and the
                  maximum possible CMS of the 4,167-atom target (chain A)
                  alone with every dot's owner in order

    python tests/golden/make_golden.py [--only examples]
"""

import contextlib
import importlib.util
import io
import os
import platform
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "cms-cuda_amd"))


def load_reference():
    path = os.path.join(os.environ.get("CMS_REFERENCE", "/root/reference"), "cms-cuda", "_core.py")
    spec = importlib.util.spec_from_file_location("cms_core_reference", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def runtime_info():
    buf = io.StringIO()
    with contextlib.redirect_stdout(buf):
        try:
            np.show_runtime()
        except Exception as e:  # noqa: BLE001
            print("show_runtime failed: %r" % e)
    import scipy
    return "numpy %s scipy %s python %s %s\n%s" % (np.__version__, scipy.__version__, platform.python_version(),
                                                  platform.machine(), buf.getvalue())


def dots_of(calc, m):
    d = calc.run.dots[m]
    n = len(d)
    return dict(coor=d.coor_xyz[:n].copy(), nml=d.outnml_xyz[:n].copy(), area=d.area[:n].copy(),
                buried=d.buried[:n].astype(np.int8), type=d.type_[:n].astype(np.int8),
                atom=d.atom_idx[:n].astype(np.int32))


def run_complex(core, bx, br, tx, tr):
    """CMS of both sides, SC, the calculator."""
    calc = core.MolecularSurfaceCalculator()
    calc.add_binder_and_target(bx, br, tx, tr)
    cms_t, per_t = calc.CalcLoaded()
    cms_b, per_b = calc.calc_contact_molecular_surface(target_side=False)
    sc = calc.CalcLoadedSC()
    res = calc.run.results
    side = []
    for m in (0, 1):
        s = res.surface[m]
        side += [s.d_mean, s.d_median, s.s_mean, s.s_median, s.trimmedArea, len(calc.run.trimmed_dots[m])]
    return calc, dict(cms_target=cms_t, cms_binder=cms_b, per_target=np.asarray(per_t, np.float64),
                      per_binder=np.asarray(per_b, np.float64), sc=np.array(sc, np.float64),
                      side=np.array(side, np.float64))


def counts(calc):
    out = []
    for m in (0, 1):
        d = dots_of(calc, m)
        out += [len(d["area"]), int(d["buried"].sum())] + [int((d["type"] == t).sum()) for t in (1, 2, 3)]
    out.append(len(calc.run.probes))
    return np.array(out, np.int64)


EXAMPLES = [
    # file, binder selection, target selection
    ("linearpeptide", dict(chains="B"), dict(chains="A")),
    ("Nanobody", dict(chains="B"), dict(chains="A")),
    ("binder", dict(chains="B"), dict(chains="A")),
    ("scFv", dict(chains="B"), dict(chains="A")),
    ("7jkv-smallmol", dict(chains="A", exclude_resnames={"V7G"}), dict(resnames={"V7G"})),
]


def example_inputs(name, bsel, tsel):
    from cms_cuda.structure import load_atoms
    path = os.path.join(HERE, "example_%s.pdb" % name)
    bx, br, _ = load_atoms(path, **bsel)
    tx, tr, _ = load_atoms(path, **tsel)
    return bx, br, tx, tr


def examples(core, info):
    out = {}
    rng = np.random.default_rng(9)
    for name, bsel, tsel in EXAMPLES:
        t0 = time.time()
        key = name.replace("-", "_")
        bx, br, tx, tr = example_inputs(name, bsel, tsel)
        calc, r = run_complex(core, bx, br, tx, tr)
        out.update({"%s_%s" % (key, k): v for k, v in dict(binder_xyz=bx, binder_radii=br, target_xyz=tx,
                                                          target_radii=tr, counts=counts(calc), **r).items()})
        for m in (0, 1):
            d = dots_of(calc, m)
            for k in ("atom", "buried", "type"):
                out["%s_dots%d_%s" % (key, m, k)] = d[k]
            idx = np.sort(rng.choice(len(d["area"]), size=min(400, len(d["area"])), replace=False))
            out["%s_sample%d_idx" % (key, m)] = idx.astype(np.int64)
            for k in ("coor", "nml", "area"):
                out["%s_sample%d_%s" % (key, m, k)] = d[k][idx]
            out["%s_sum%d" % (key, m)] = np.array([d["coor"].sum(), d["nml"].sum(), d["area"].sum()])
        print("%s: %d + %d atoms, cms %.16g / %.16g, sc %s, %d + %d dots (%.1f s)"
              % (name, len(br), len(tr), r["cms_target"], r["cms_binder"], r["sc"],
                 len(calc.run.dots[0]), len(calc.run.dots[1]), time.time() - t0), flush=True)
    # whole-molecule surface of the 4,167-atom target
    t0 = time.time()
    _, _, tx, tr = example_inputs("scFv", dict(chains="B"), dict(chains="A"))
    mx, per, calc = core.calculate_maximum_possible_contact_ms(tx, tr)
    d = dots_of(calc, 0)
    out.update(maxcms_xyz=tx, maxcms_radii=tr, maxcms_total=np.float64(mx),
               maxcms_per_atom=np.asarray(per, np.float64), maxcms_atom=d["atom"],
               maxcms_sum=np.array([d["coor"].sum(), d["area"].sum()]),
               maxcms_counts=np.array([len(d["area"])] + [int((d["type"] == t).sum()) for t in (1, 2, 3)]))
    print("max-cms of the target: %.16g from %d dots (%.1f s)" % (mx, len(d["area"]), time.time() - t0))
    np.savez_compressed(os.path.join(HERE, "ref_examples.npz"), **out, runtime=np.array(info))


def main():
    from cms_cuda import synth
    from cms_cuda import get_radii_from_names
    from cms_cuda.structure import read_pdb_atoms, split_chains
    core = load_reference()
    info = runtime_info()
    t0 = time.time()
    if "--only" in sys.argv:
        if sys.argv[sys.argv.index("--only") + 1] == "examples":
            examples(core, info)
        return

    # small synthetic complex, everything
    c = synth.make_complex(260, 80, 21, gap=0.5)
    bx, br, tx, tr = c["binder_xyz"], c["binder_radii"], c["target_xyz"], c["target_radii"]
    calc, r = run_complex(core, bx, br, tx, tr)
    d0, d1 = dots_of(calc, 0), dots_of(calc, 1)
    pr = calc.run.probes
    n = len(pr)
    probes = np.stack([pr.atom_idx_0[:n], pr.atom_idx_1[:n], pr.atom_idx_2[:n]], 1).astype(np.int32) \
        if n else np.zeros((0, 3), np.int32)
    np.savez_compressed(os.path.join(HERE, "ref_small.npz"), binder_xyz=bx, binder_radii=br, target_xyz=tx,
                        target_radii=tr, counts=counts(calc), probes=probes,
                        **r, **{"dots0_" + k: v for k, v in d0.items()}, **{"dots1_" + k: v for k, v in d1.items()},
                        runtime=np.array(info))
    print("small: %d + %d dots, %d probes, cms %.12g" % (len(d0["area"]), len(d1["area"]), n, r["cms_target"]))

    # known-answer complex: checksums and a sample of dots
    atoms = read_pdb_atoms(os.path.join(HERE, "sc_example.pdb"))
    (gbx, gbrn, gban), (gtx, gtrn, gtan) = split_chains(atoms, "A")
    gbr, gtr = get_radii_from_names(gbrn, gban), get_radii_from_names(gtrn, gtan)
    calc, r = run_complex(core, gbx, gbr, gtx, gtr)
    extra = {}
    rng = np.random.default_rng(5)
    for m in (0, 1):
        d = dots_of(calc, m)
        idx = np.sort(rng.choice(len(d["area"]), size=300, replace=False))
        extra["sample%d_idx" % m] = idx.astype(np.int64)
        for k in ("coor", "nml", "area", "buried", "type", "atom"):
            extra["sample%d_%s" % (m, k)] = d[k][idx]
        extra["sum%d" % m] = np.array([d["coor"].sum(), d["nml"].sum(), d["area"].sum(),
                                       float(d["atom"].sum()), float(d["buried"].sum())])
    np.savez_compressed(os.path.join(HERE, "ref_golden.npz"), counts=counts(calc), **r, **extra,
                        runtime=np.array(info))
    print("golden: cms %.16g sc %s" % (r["cms_target"], r["sc"]))

    # radius-0 variant of the known-answer complex
    r0 = gbr.copy()
    r0[::4] = 0.0
    calc, r = run_complex(core, gbx, r0, gtx, gtr)
    np.savez_compressed(os.path.join(HERE, "ref_radius0.npz"), binder_radii=r0, counts=counts(calc), **r,
                        runtime=np.array(info))
    print("radius0: cms %.16g" % r["cms_target"])

    # maximum possible CMS of the known-answer binder
    mx, per, calc = core.calculate_maximum_possible_contact_ms(gbx, gbr)
    np.savez_compressed(os.path.join(HERE, "ref_maxcms.npz"), total=np.float64(mx),
                        per_atom=np.asarray(per, np.float64), ndots=np.int64(len(calc.run.dots[0])),
                        runtime=np.array(info))
    print("max-cms: %.16g" % mx)

    # c1-size complex
    c = synth.config1(seed=1)
    bx, br, tx, tr = c["binder_xyz"], c["binder_radii"], c["target_xyz"], c["target_radii"]
    calc, r = run_complex(core, bx, br, tx, tr)
    np.savez_compressed(os.path.join(HERE, "ref_c1.npz"), binder_xyz=bx, binder_radii=br, target_xyz=tx,
                        target_radii=tr, counts=counts(calc), **r, runtime=np.array(info))
    print("c1: cms %.16g sc %s (%.1f s total)" % (r["cms_target"], r["sc"], time.time() - t0))


if __name__ == "__main__":
    main()
