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
