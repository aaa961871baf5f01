"""Command line: CMS and SC of complex structures (PDB or mmCIF).

    python -m cms_cuda.cli complex.pdb --binder-chains B --target-chains A
    python -m cms_cuda.cli fab.cif --binder-chains H,L --target-chains A
    python -m cms_cuda.cli 1STP.pdb --binder-chains A --binder-exclude BTN --target-resnames BTN --max-cms
    python -m cms_cuda.cli designs/*.pdb --binder-chains B --target-chains A --csv cms_sc.csv --per-residue out/

The arguments, selections, printed rows, CSV columns and per-residue tables
are those of the reference's examples/calculate_cms_sc.py (lines 43-212).
Each structure is split into a binder and a target group of heavy atoms by
chain ID and / or residue name (cms_cuda.structure.load_atoms: first model,
one alternate location per atom, waters and single-atom ions dropped unless
asked for); CMS is reported on the target side and on the binder side, SC
is symmetric.  All structures of one invocation run as one batch on the GPU
(per-structure results equal single-structure runs bit for bit; a structure
that fails -- coincident atoms, an empty selection -- gets an error row and
the others are unaffected); --max-cms runs every selected molecule's whole
surface as a second batch.  "seconds" is the batch time divided by the
number of structures.
"""

import argparse
import csv
import ctypes
import os
import sys
import time

import numpy as np

from .structure import load_atoms, parse_structure


def unmatched_atoms(info, radii):
    """(residue name, atom name) pairs that got radius 0 from the radius table."""
    return sorted({(i[2], i[3]) for i, r in zip(info, radii) if r == 0.0})


def per_residue(info, per_atom):
    """Sum a per-atom array over residues -> list of (chain, resnum, resname, value), in first-seen order."""
    sums, order = {}, []
    for (chain, resnum, resname, _, _), v in zip(info, per_atom):
        key = (chain, resnum, resname)
        if key not in sums:
            sums[key] = 0.0
            order.append(key)
        sums[key] += float(v)
    return [(*k, sums[k]) for k in order]


def split(s):
    return [x for x in s.split(",") if x] if s else None


def parser():
    ap = argparse.ArgumentParser(prog="python -m cms_cuda.cli", description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("structures", nargs="+", help="PDB / mmCIF files of the complex")
    for side in ("binder", "target"):
        ap.add_argument(f"--{side}-chains", help=f"comma-separated chain IDs of the {side} (default: all)")
        ap.add_argument(f"--{side}-resnames", help=f"only these residue names in the {side}, e.g. BTN")
        ap.add_argument(f"--{side}-exclude", help=f"residue names to drop from the {side}, e.g. BTN")
    ap.add_argument("--keep-water", action="store_true", help="keep water molecules")
    ap.add_argument("--keep-ions", action="store_true", help="keep single-atom ions (most get radius 0)")
    ap.add_argument("--device", choices=("auto", "cpu", "cuda"), default="auto")
    ap.add_argument("--no-sc", action="store_true", help="skip shape complementarity")
    ap.add_argument("--max-cms", action="store_true",
                    help="also compute the maximum possible CMS (total molecular surface) of each side")
    ap.add_argument("--csv", help="write one row per structure to this CSV file")
    ap.add_argument("--per-residue", metavar="DIR", help="write per-residue CMS tables to this directory")
    return ap


def selection(args, side):
    return dict(chains=split(getattr(args, f"{side}_chains")),
                resnames=set(split(getattr(args, f"{side}_resnames")) or []) or None,
                exclude_resnames=set(split(getattr(args, f"{side}_exclude")) or []),
                keep_water=args.keep_water, keep_ions=args.keep_ions)


def run(args, out=sys.stdout, err=sys.stderr):
    """Rows (dicts) for the structures of args, in order."""
    from . import gpu_available, _lib
    from .batch import BatchCalculator
    if args.device == "cpu":
        raise SystemExit("--device cpu: this MI355X build has no CPU calculator; use --device auto or cuda")
    if not gpu_available():
        raise SystemExit("no usable ROCm GPU / libcms_hip.so was found")
    names = [os.path.basename(p) for p in args.structures]
    rows = [None] * len(names)
    loaded = []                      # (row index, bx, br, binfo, tx, tr, tinfo)
    for k, path in enumerate(args.structures):
        try:
            model = parse_structure(path)
            bx, br, binfo = load_atoms(path, model=model, **selection(args, "binder"))
            tx, tr, tinfo = load_atoms(path, model=model, **selection(args, "target"))
            if len(bx) == 0 or len(tx) == 0:
                raise ValueError(f"empty selection (binder {len(bx)} atoms, target {len(tx)} atoms)")
            for side, info, radii in (("binder", binfo, br), ("target", tinfo, tr)):
                bad = unmatched_atoms(info, radii)
                if bad:
                    print(f"  warning: {names[k]} {side}: {int((radii == 0).sum())} atoms have radius 0 "
                          f"(no entry in the radius table, they are invisible): {bad[:8]}", file=err)
            loaded.append((k, bx, br, binfo, tx, tr, tinfo))
        except Exception as exc:            # keep going through a batch
            rows[k] = dict(structure=names[k], error=str(exc))
    if loaded:
        t0 = time.perf_counter()
        bc = BatchCalculator()
        try:
            bc.load([(bx, br, tx, tr) for _, bx, br, _, tx, tr, _ in loaded])
            res = bc.run(_lib.PREC_FP32_RESCORE, sc=not args.no_sc)
            status = bc.status()
            coinc = []
            for c in range(len(loaded)):
                if status[c] & 1:
                    i, j = ctypes.c_int32(), ctypes.c_int32()
                    bc.ctx("cms_coincident", c, ctypes.byref(i), ctypes.byref(j))
                    coinc.append((c, i.value, j.value))
        except RuntimeError as exc:
            # a one-structure batch with coincident atoms (or every structure failing)
            for k, *_ in loaded:
                rows[k] = dict(structure=names[k], error=str(exc))
            loaded, res = [], None
        finally:
            bc.close()
        mx = None
        if loaded and args.max_cms:
            mols = []
            for _, bx, br, _, tx, tr, _ in loaded:
                mols += [(tx, tr, np.zeros((0, 3)), np.zeros(0)), (bx, br, np.zeros((0, 3)), np.zeros(0))]
            mc = BatchCalculator()
            try:
                mc.load(mols)
                mx = mc.max_contact_ms()["total"].reshape(len(loaded), 2)
            finally:
                mc.close()
        seconds = (time.perf_counter() - t0) / max(len(loaded), 1)
        per = bc.per_atom_split(res["per_atom"]) if loaded else []
        cmsg = {c: "Coincident atoms: %d:: == %d::" % (i, j) for c, i, j in (coinc if loaded else [])}
        for c, (k, bx, br, binfo, tx, tr, tinfo) in enumerate(loaded):
            if status[c]:
                rows[k] = dict(structure=names[k], error=cmsg.get(c, "row capacity exceeded (status %d)" % status[c]))
                continue
            cms, cms_b = float(res["target"][c]), float(res["binder"][c])
            row = dict(structure=names[k], device="cuda", binder_atoms=len(bx), target_atoms=len(tx),
                       cms=round(cms, 4), cms_binder=round(cms_b, 4))
            if not args.no_sc:
                sc = res["sc"][c]
                row.update(sc=round(float(sc[0]), 4), sc_area=round(float(sc[1]), 4),
                           sc_median_dist=round(float(sc[2]), 4))
            if mx is not None:
                max_t, max_b = float(mx[c, 0]), float(mx[c, 1])
                row.update(max_cms_target=round(max_t, 4), max_cms_binder=round(max_b, 4),
                           cms_fraction_target=round(cms / max_t, 4) if max_t else 0.0,
                           cms_fraction_binder=round(cms_b / max_b, 4) if max_b else 0.0)
            row["seconds"] = round(seconds, 3)
            if args.per_residue:
                base = os.path.splitext(names[k])[0]
                pb, pt = per[c]
                for side, info, values in (("target", tinfo, pt), ("binder", binfo, pb)):
                    with open(os.path.join(args.per_residue, f"{base}_{side}_cms.csv"), "w", newline="") as f:
                        w = csv.writer(f)
                        w.writerow(["chain", "resnum", "resname", "cms"])
                        w.writerows((ch, n, r, round(v, 4)) for ch, n, r, v in per_residue(info, values))
            rows[k] = row
    for row in rows:
        if "error" in row:
            print(f"  error: {row['structure']}: {row['error']}", file=err)
        print("  ".join(f"{k}={v}" for k, v in row.items()), file=out)
    return rows


def main(argv=None):
    ap = parser()
    args = ap.parse_args(argv)
    if not (args.binder_chains or args.binder_resnames) or not (args.target_chains or args.target_resnames):
        ap.error("select both sides, e.g. --binder-chains B --target-chains A")
    if args.per_residue:
        os.makedirs(args.per_residue, exist_ok=True)
    rows = run(args)
    if args.csv:
        keys = []
        for r in rows:
            keys += [k for k in r if k not in keys]
        with open(args.csv, "w", newline="") as f:
            w = csv.DictWriter(f, fieldnames=keys)
            w.writeheader()
            w.writerows(rows)
        print(f"wrote {args.csv}")
    return rows


if __name__ == "__main__":
    main()
