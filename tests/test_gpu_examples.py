"""The reference's five example complexes (examples/*.pdb, copied to
tests/golden/example_*.pdb) through the drop-in calculator, against the
vectors the reference's own NumPy calculator wrote for them
(tests/golden/ref_examples.npz, tests/golden/make_golden.py):

* binders of 136 (linear peptide), 1,022 (nanobody), 1,105 (designed
  binder), 1,875 (scFv) atoms against the 4,167-atom target, and a
  2,329-atom protein against its 41-atom ligand;
* every dot's owner, burial and type identical and in order; dot counts by
  type, probe counts; 400 sampled dots per molecule and dot sums within
  1e-9; CMS of both sides with per-atom arrays, SC, area, distance and the
  per-side statistics within 1e-10 relative;
* the maximum possible CMS of the 4,167-atom target (301,660 dots, the
  whole-molecule surface where neighbour rows and near-probe lists are
  longest): total, per-atom values and every dot's owner;
* the five complexes as one batch equal the single-complex runs bit for bit.
"""

import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
NAMES = ["linearpeptide", "Nanobody", "binder", "scFv", "7jkv_smallmol"]
TOL_DOT = 1e-9
TOL_OUT = 1e-10


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.size == 0:
        return 0.0
    return float(np.max(np.abs(a - b) / np.maximum(1.0, np.maximum(np.abs(a), np.abs(b)))))


@pytest.fixture(scope="module")
def ref():
    with np.load(os.path.join(HERE, "golden", "ref_examples.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def inputs(ref, name):
    return tuple(ref["%s_%s" % (name, k)] for k in ("binder_xyz", "binder_radii", "target_xyz", "target_radii"))


@pytest.mark.parametrize("name", NAMES)
def test_example_complex(ref, name):
    import cms_cuda
    bx, br, tx, tr = inputs(ref, name)
    say("%s: %d + %d atoms" % (name, len(br), len(tr)))
    calc = cms_cuda.MolecularSurfaceCalculator()
    calc.add_binder_and_target(bx, br, tx, tr)
    cms, per_t = calc.CalcLoaded()
    cms_b, per_b = calc.calc_contact_molecular_surface(target_side=False)
    sc, area, dist = calc.CalcLoadedSC()
    g = lambda k: ref["%s_%s" % (name, k)]  # noqa: E731
    assert rel_err(cms, g("cms_target")) <= TOL_OUT
    assert rel_err(per_t, g("per_target")) <= TOL_OUT
    assert rel_err(cms_b, g("cms_binder")) <= TOL_OUT
    assert rel_err(per_b, g("per_binder")) <= TOL_OUT
    assert rel_err([sc, area, dist], g("sc")) <= TOL_OUT
    side = g("side")
    for m in (0, 1):
        s = calc.run.results.surface[m]
        assert rel_err([s.d_mean, s.d_median, s.s_mean, s.s_median, s.trimmedArea], side[6 * m:6 * m + 5]) <= TOL_OUT
        assert s.nTrimmedDots == int(side[6 * m + 5])
    cnt = g("counts")
    res = calc.run.results
    for m in (0, 1):
        d = calc.run.dots[m]
        assert len(d) == cnt[5 * m]
        assert int(d.buried.sum()) == cnt[5 * m + 1]
        assert [int((d.type_ == t).sum()) for t in (1, 2, 3)] == list(cnt[5 * m + 2:5 * m + 5])
        assert np.array_equal(d.atom_idx, g("dots%d_atom" % m))
        assert np.array_equal(d.buried, g("dots%d_buried" % m))
        assert np.array_equal(d.type_, g("dots%d_type" % m))
        idx = g("sample%d_idx" % m)
        assert rel_err(d.coor_xyz[idx], g("sample%d_coor" % m)) <= TOL_DOT
        assert rel_err(d.outnml_xyz[idx], g("sample%d_nml" % m)) <= TOL_DOT
        assert rel_err(d.area[idx], g("sample%d_area" % m)) <= TOL_DOT
        s = g("sum%d" % m)
        assert rel_err([d.coor_xyz.sum(), d.outnml_xyz.sum(), d.area.sum()], s) <= TOL_DOT
    assert len(calc.run.probes) == cnt[10]
    assert (res.dots.convex, res.dots.toroidal, res.dots.concave) == (
        cnt[2] + cnt[7], cnt[3] + cnt[8], cnt[4] + cnt[9])
    say("  cms %.10f sc %.10f, %d trimmed dots" % (cms, sc, res.surface[0].nTrimmedDots + res.surface[1].nTrimmedDots))


def test_example_target_max_cms(ref):
    import cms_cuda
    say("max-CMS of the 4,167-atom target")
    mx, per, calc = cms_cuda.calculate_maximum_possible_contact_ms(ref["maxcms_xyz"], ref["maxcms_radii"])
    assert rel_err(mx, ref["maxcms_total"]) <= TOL_OUT
    assert rel_err(per, ref["maxcms_per_atom"]) <= TOL_OUT
    d = calc.run.dots[0]
    cnt = ref["maxcms_counts"]
    assert len(d) == cnt[0]
    assert [int((d.type_ == t).sum()) for t in (1, 2, 3)] == list(cnt[1:4])
    assert np.array_equal(d.atom_idx, ref["maxcms_atom"])
    assert rel_err([d.coor_xyz.sum(), d.area.sum()], ref["maxcms_sum"]) <= TOL_DOT


def test_examples_as_one_batch(ref):
    from cms_cuda.batch import BatchCalculator
    from cms_cuda import _lib
    cs = [inputs(ref, n) for n in NAMES]
    say("five examples as one batch")
    bc = BatchCalculator()
    bc.load(cs)
    both = bc.run(_lib.PREC_FP32_RESCORE)
    bc.close()
    per = bc.per_atom_split(both["per_atom"])
    for c, n in enumerate(NAMES):
        g = lambda k: ref["%s_%s" % (n, k)]  # noqa: E731
        assert rel_err(both["target"][c], g("cms_target")) <= TOL_OUT
        assert rel_err(both["binder"][c], g("cms_binder")) <= TOL_OUT
        assert rel_err(per[c][1], g("per_target")) <= TOL_OUT
        assert rel_err(per[c][0], g("per_binder")) <= TOL_OUT
        assert rel_err(both["sc"][c][:3], g("sc")) <= TOL_OUT
        bc1 = BatchCalculator()
        bc1.load([cs[c]])
        one = bc1.run(_lib.PREC_FP64)
        bc1.close()
        assert one["target"][0] == both["target"][c] and one["binder"][0] == both["binder"][c]
        assert np.array_equal(one["sc"][0], both["sc"][c])
        assert np.array_equal(one["per_atom"], np.concatenate(per[c]))
