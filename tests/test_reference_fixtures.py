"""The oracle against vectors written by the reference's own calculator
(cms-cuda/_core.py, see tests/golden/make_golden.py).

These pin what the known-answer complex leaves open: every dot (position,
normal, area, burial, type, owner, order) of a second, synthetic complex,
per-atom CMS arrays, the SC side statistics, radius-0 atoms, the max-CMS
per-atom split and a c1-size complex.  Bar: counts, owners, flags, types
and order identical; floats within 1e-12 relative.
"""

import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
TOL = 1e-12


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.size == 0 and b.size == 0:
        return 0.0
    return float(np.max(np.abs(a - b) / np.maximum(1.0, np.maximum(np.abs(a), np.abs(b)))))


def load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def oracle_run(bx, br, tx, tr):
    from oracle.cms_oracle import OracleCalculator
    o = OracleCalculator()
    o.add_binder_and_target(bx, br, tx, tr)
    cms_t, per_t = o.contact_ms(True)
    cms_b, per_b = o.contact_ms(False)
    sc = o.shape_complementarity()
    return o, cms_t, per_t, cms_b, per_b, sc


def counts(o):
    out = []
    for m in (0, 1):
        d = o.dots[m]
        out += [len(d), int(d.buried.sum())] + [int((d.type == t).sum()) for t in (1, 2, 3)]
    out.append(len(o.pr_h))
    return np.array(out)


def side_stats(o):
    out = []
    r = o.sc_results
    for m in (0, 1):
        s = r["sides"][m] or dict(d_mean=0.0, d_median=0.0, s_mean=0.0, s_median=0.0)
        out += [s["d_mean"], s["d_median"], s["s_mean"], s["s_median"], r["trimmed_area"][m], r["n_trimmed"][m]]
    return np.array(out)


def check_scalars(f, cms_t, per_t, cms_b, per_b, sc):
    assert rel(cms_t, f["cms_target"]) <= TOL
    assert rel(cms_b, f["cms_binder"]) <= TOL
    assert rel(per_t, f["per_target"]) <= TOL and len(per_t) == len(f["per_target"])
    assert rel(per_b, f["per_binder"]) <= TOL and len(per_b) == len(f["per_binder"])
    assert rel(sc, f["sc"]) <= TOL


def test_small_complex_every_dot():
    f = load("ref_small.npz")
    o, *out = oracle_run(f["binder_xyz"], f["binder_radii"], f["target_xyz"], f["target_radii"])
    assert np.array_equal(counts(o), f["counts"])
    check_scalars(f, *out)
    assert rel(side_stats(o), f["side"]) <= TOL
    for m in (0, 1):
        d = o.dots[m]
        assert np.array_equal(d.atom, f["dots%d_atom" % m])
        assert np.array_equal(d.buried, f["dots%d_buried" % m])
        assert np.array_equal(d.type, f["dots%d_type" % m])
        for k, fk in (("coor", "coor"), ("nml", "nml"), ("area", "area")):
            assert rel(getattr(d, k), f["dots%d_%s" % (m, fk)]) <= TOL, (m, k)
    assert np.array_equal(o.pr_a, f["probes"])


def test_known_answer_complex_dots(golden_complex):
    f = load("ref_golden.npz")
    c = golden_complex
    o, *out = oracle_run(c["binder_xyz"], c["binder_radii"], c["target_xyz"], c["target_radii"])
    assert np.array_equal(counts(o), f["counts"])
    check_scalars(f, *out)
    assert rel(side_stats(o), f["side"]) <= TOL
    for m in (0, 1):
        d = o.dots[m]
        idx = f["sample%d_idx" % m]
        assert np.array_equal(d.atom[idx], f["sample%d_atom" % m])
        assert np.array_equal(d.buried[idx], f["sample%d_buried" % m])
        assert np.array_equal(d.type[idx], f["sample%d_type" % m])
        assert rel(d.coor[idx], f["sample%d_coor" % m]) <= TOL
        assert rel(d.nml[idx], f["sample%d_nml" % m]) <= TOL
        assert rel(d.area[idx], f["sample%d_area" % m]) <= TOL
        sums = np.array([d.coor.sum(), d.nml.sum(), d.area.sum(), float(d.atom.sum()), float(d.buried.sum())])
        assert rel(sums, f["sum%d" % m]) <= 1e-10


def test_radius_zero_atoms(golden_complex):
    f = load("ref_radius0.npz")
    c = golden_complex
    o, *out = oracle_run(c["binder_xyz"], f["binder_radii"], c["target_xyz"], c["target_radii"])
    assert np.array_equal(counts(o), f["counts"])
    check_scalars(f, *out)


def test_max_cms_per_atom(golden_complex):
    from oracle.cms_oracle import calculate_maximum_possible_contact_ms
    f = load("ref_maxcms.npz")
    c = golden_complex
    tot, per, o = calculate_maximum_possible_contact_ms(c["binder_xyz"], c["binder_radii"])
    assert len(o.dots[0]) == int(f["ndots"])
    assert rel(tot, f["total"]) <= TOL
    assert rel(per, f["per_atom"]) <= TOL


@pytest.mark.slow
def test_c1_complex():
    f = load("ref_c1.npz")
    o, *out = oracle_run(f["binder_xyz"], f["binder_radii"], f["target_xyz"], f["target_radii"])
    assert np.array_equal(counts(o), f["counts"])
    check_scalars(f, *out)
    assert rel(side_stats(o), f["side"]) <= TOL


EXAMPLES = [
    ("linearpeptide", dict(chains="B"), dict(chains="A")),
    ("Nanobody", dict(chains="B"), dict(chains="A")),
    ("binder", dict(chains="B"), dict(chains="A")),
    ("scFv", dict(chains="B"), dict(chains="A")),
    ("7jkv-smallmol", dict(chains="A", exclude_resnames={"V7G"}), dict(resnames={"V7G"})),
]


@pytest.mark.parametrize("name,bsel,tsel", EXAMPLES)
def test_example_inputs_from_structure_files(name, bsel, tsel):
    """cms_cuda.structure.load_atoms reads the example files into exactly the
    inputs the reference vectors were computed on (float32 coordinates, as
    Bio.PDB stores them; radii from the table, none unmatched)."""
    from cms_cuda.structure import load_atoms
    z = load("ref_examples.npz")
    key = name.replace("-", "_")
    path = os.path.join(GOLD, "example_%s.pdb" % name)
    bx, br, binfo = load_atoms(path, **bsel)
    tx, tr, tinfo = load_atoms(path, **tsel)
    assert np.array_equal(bx, z[key + "_binder_xyz"]) and np.array_equal(br, z[key + "_binder_radii"])
    assert np.array_equal(tx, z[key + "_target_xyz"]) and np.array_equal(tr, z[key + "_target_radii"])
    assert np.array_equal(bx, bx.astype(np.float32).astype(np.float64))
    assert (br > 0).all() and (tr > 0).all()
    assert len(binfo) == len(bx) and len(tinfo) == len(tx)


@pytest.mark.slow
def test_example_ligand_complex_oracle():
    """The oracle on the example protein-ligand complex (2,329 + 41 atoms)
    against the reference's vectors: every dot's owner, burial and type in
    order, sampled dots, CMS of both sides with per-atom arrays, SC and the
    per-side statistics."""
    z = load("ref_examples.npz")
    k = "7jkv_smallmol_"
    o, cms_t, per_t, cms_b, per_b, sc = oracle_run(z[k + "binder_xyz"], z[k + "binder_radii"],
                                                   z[k + "target_xyz"], z[k + "target_radii"])
    assert np.array_equal(counts(o), z[k + "counts"])
    for m in (0, 1):
        d = o.dots[m]
        assert np.array_equal(d.atom, z[k + "dots%d_atom" % m])
        assert np.array_equal(d.buried, z[k + "dots%d_buried" % m])
        assert np.array_equal(d.type, z[k + "dots%d_type" % m])
        idx = z[k + "sample%d_idx" % m]
        assert rel(d.coor[idx], z[k + "sample%d_coor" % m]) <= TOL
        assert rel(d.nml[idx], z[k + "sample%d_nml" % m]) <= TOL
        assert rel(d.area[idx], z[k + "sample%d_area" % m]) <= TOL
    assert rel(cms_t, z[k + "cms_target"]) <= TOL and rel(per_t, z[k + "per_target"]) <= TOL
    assert rel(cms_b, z[k + "cms_binder"]) <= TOL and rel(per_b, z[k + "per_binder"]) <= TOL
    assert rel(sc, z[k + "sc"]) <= TOL
    assert rel(side_stats(o), z[k + "side"]) <= TOL
