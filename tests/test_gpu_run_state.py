"""calc.run of the drop-in calculator against the oracle: the fields the
reference's full-state parity harness compares (test/test_gpu_parity.py:
153-204, compare_calcs) -- radmax, result counters and scalars, per-atom
state, neighbour and buried tables, dots, trimmed dots, probes and the
toroid queue -- on the known-answer complex and a 260 + 80-atom synthetic
complex; neighbour rows with exact distance ties (a lattice complex) and
what is built on them; and the CMS / max-CMS readouts on a surface generated in the other
all_atoms mode (the reference reads whatever dots exist,
cms-cuda/_core.py:946-1034).

Bar: integers, flags, indices and orders identical; floats within 1e-9
(dots, probes, queue geometry) or 1e-10 relative (scalars).
"""

import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
TOL_DOT = 1e-9
TOL_OUT = 1e-10


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.size == 0 and b.size == 0:
        return 0.0
    return float(np.max(np.abs(a - b) / np.maximum(1.0, np.maximum(np.abs(a), np.abs(b)))))


def small_complex():
    with np.load(os.path.join(HERE, "golden", "ref_small.npz"), allow_pickle=False) as z:
        return dict(binder_xyz=z["binder_xyz"], binder_radii=z["binder_radii"], target_xyz=z["target_xyz"],
                    target_radii=z["target_radii"])


def compare_dots(g, o):
    assert len(g) == len(o.area)
    assert np.array_equal(g.atom_idx, o.atom)
    assert np.array_equal(g.buried, o.buried)
    assert np.array_equal(g.type_, o.type)
    assert rel_err(g.coor_xyz, o.coor) <= TOL_DOT
    assert rel_err(g.outnml_xyz, o.nml) <= TOL_DOT
    assert rel_err(g.area, o.area) <= TOL_DOT


def compare_rows(arr, rows, counts, xyz, rad):
    assert np.array_equal(arr.nneighbors, counts)
    assert np.array_equal(arr.natom, rows)
    m = rows >= 0
    exp_xyz = np.full(rows.shape + (3,), np.nan)
    exp_xyz[m] = xyz[rows[m]]
    exp_rad = np.full(rows.shape, np.nan)
    exp_rad[m] = rad[rows[m]]
    assert np.array_equal(arr.xyz, exp_xyz, equal_nan=True)
    assert np.array_equal(arr.radius, exp_rad, equal_nan=True)


@pytest.mark.parametrize("which", ["golden", "small"])
def test_run_state_matches_oracle(which, golden_complex):
    import cms_cuda
    from oracle.cms_oracle import OracleCalculator
    c = golden_complex if which == "golden" else small_complex()
    bx, br, tx, tr = c["binder_xyz"], c["binder_radii"], c["target_xyz"], c["target_radii"]
    say("run state, %s complex (%d + %d atoms)" % (which, len(br), len(tr)))
    o = OracleCalculator()
    o.add_binder_and_target(bx, br, tx, tr)
    ocms = o.contact_ms(True)
    osc = o.shape_complementarity()
    calc = cms_cuda.MolecularSurfaceCalculator()
    calc.add_binder_and_target(bx, br, tx, tr)
    cms, _ = calc.CalcLoaded()
    sc = calc.CalcLoadedSC()
    rg = calc.run
    assert rel_err(cms, ocms[0]) <= TOL_OUT and rel_err(sc, osc) <= TOL_OUT
    assert rg.radmax == float(np.concatenate([br, tr]).max())
    assert (rg.surfaces_generated, rg.surfaces_generated_all_atoms) == (True, False)
    res = rg.results
    n = len(br) + len(tr)
    assert res.nAtoms == n and res.valid == 1
    assert (res.dots.convex, res.dots.toroidal, res.dots.concave) == (
        o.counts["convex"], o.counts["toroidal"], o.counts["concave"])
    r = o.sc_results
    for m in (0, 1):
        s = res.surface[m]
        mol = o.mol == m
        assert s.nAtoms == int(mol.sum())
        assert s.nBlockedAtoms == int((o.atten[mol] == 1).sum())
        assert s.nBuriedAtoms == int(mol.sum()) - s.nBlockedAtoms
        assert s.nAllDots == len(o.dots[m]) and s.nTrimmedDots == r["n_trimmed"][m]
        assert (s.nBuriedDots, s.nAccessibleDots) == (0, 0)      # never set by the reference either
        assert rel_err(s.trimmedArea, r["trimmed_area"][m]) <= TOL_OUT
    # per-atom state over every atom
    at = rg.atoms
    assert np.array_equal(at.natom, np.arange(n))
    assert np.array_equal(at.molecule, o.mol)
    assert np.array_equal(at.atten, o.atten) and np.array_equal(at.access, o.access)
    assert np.array_equal(at.radius, o.radius) and np.array_equal(at.xyz, o.xyz)
    assert np.array_equal(at.density, o.dens)
    # neighbour and buried tables
    compare_rows(rg.neighbor_array, o.nbr, o.nn, o.xyz, o.radius)
    compare_rows(rg.buried_array, o.bur, o.nb, o.xyz, o.radius)
    # dots and trimmed dots
    for m in (0, 1):
        compare_dots(rg.dots[m], o.dots[m])
        compare_dots(rg.trimmed_dots[m], r["trimmed"][m])
    assert len(rg.trimmed_dots[0]) + len(rg.trimmed_dots[1]) > 100
    # probes in the reference's order within each molecule
    pr = rg.probes
    assert len(pr) == len(o.pr_h)
    # toroid queue
    q = rg.toroid_queue
    tor = np.nonzero(o.p_tor)[0]
    assert len(q) == len(tor) > 100
    assert [(t[0], t[1], t[5]) for t in q] == [(int(o.p_i[k]), int(o.p_j[k]), bool(o.p_between[k])) for k in tor]
    assert rel_err(np.array([t[2] for t in q]), o.p_u[tor]) <= TOL_DOT
    assert rel_err(np.array([t[3] for t in q]), o.p_t[tor]) <= TOL_DOT
    assert rel_err(np.array([t[4] for t in q]), o.p_r[tor]) <= TOL_DOT
    say("  %d toroid-queue pairs, %d + %d trimmed dots, rows up to %d / %d"
        % (len(q), len(rg.trimmed_dots[0]), len(rg.trimmed_dots[1]), rg.neighbor_array.natom.shape[1],
           rg.buried_array.natom.shape[1]))


def lattice_complex():
    """Two slabs of a cubic lattice with spacing 3.5 A (exact in binary, so
    every lattice distance is computed exactly and neighbour rows hold many
    exact d^2 ties: 6 at 3.5 A, 12 at 3.5 sqrt 2, 8 at 3.5 sqrt 3); radii
    alternate so that the bridge distances differ between rows."""
    k = np.arange(6, dtype=np.float64)
    g = np.stack(np.meshgrid(k, k, k, indexing="ij"), -1).reshape(-1, 3) * 3.5
    rad = np.array([1.6, 1.8, 1.7, 1.9])
    b = g[g[:, 2] <= 3.5]
    t = g[g[:, 2] >= 7.0] + np.array([0.25, 0.5, 0.5])
    return dict(binder_xyz=b, binder_radii=rad[np.arange(len(b)) % 4],
                target_xyz=t, target_radii=rad[(np.arange(len(t)) + 1) % 4])


def test_lattice_ties_match_oracle():
    """Neighbour rows with exact d^2 ties: the (d^2, j) order, and everything
    built on it (pairs, probes, dots, CMS, SC), equal the oracle's."""
    import cms_cuda
    from oracle.cms_oracle import OracleCalculator
    c = lattice_complex()
    bx, br, tx, tr = c["binder_xyz"], c["binder_radii"], c["target_xyz"], c["target_radii"]
    say("lattice complex with exact distance ties (%d + %d atoms)" % (len(br), len(tr)))
    o = OracleCalculator()
    o.add_binder_and_target(bx, br, tx, tr)
    ocms = o.contact_ms(True)
    osc = o.shape_complementarity()
    calc = cms_cuda.MolecularSurfaceCalculator()
    calc.add_binder_and_target(bx, br, tx, tr)
    cms, per = calc.CalcLoaded()
    sc = calc.CalcLoadedSC()
    rg = calc.run
    # the rows do hold ties: some row has two entries at the same distance
    d2 = np.where(o.nbr >= 0, ((o.xyz[np.maximum(o.nbr, 0)] - o.xyz[:, None, :]) ** 2).sum(-1), np.nan)
    assert any(len(np.unique(r[~np.isnan(r)])) < int((~np.isnan(r)).sum()) for r in d2)
    compare_rows(rg.neighbor_array, o.nbr, o.nn, o.xyz, o.radius)
    compare_rows(rg.buried_array, o.bur, o.nb, o.xyz, o.radius)
    for m in (0, 1):
        compare_dots(rg.dots[m], o.dots[m])
    assert len(rg.probes) == len(o.pr_h)
    assert rel_err(cms, ocms[0]) <= TOL_OUT and rel_err(per, ocms[1]) <= TOL_OUT
    assert rel_err(sc, osc) <= TOL_OUT
    assert len(rg.dots[0]) > 0 and len(rg.dots[1]) > 0


def test_run_state_before_generation(golden_complex):
    import cms_cuda
    c = golden_complex
    calc = cms_cuda.MolecularSurfaceCalculator()
    calc.add_binder_and_target(c["binder_xyz"], c["binder_radii"], c["target_xyz"], c["target_radii"])
    rg = calc.run
    assert rg.neighbor_array is None and rg.buried_array is None
    assert rg.toroid_queue == [] and len(rg.trimmed_dots[0]) == 0
    assert calc.calc_contact_molecular_surface(True)[0] == 0.0
    assert calc.calc_max_possible_contact_molecular_surface()[0] == 0.0
    calc.CalcLoaded()
    assert len(calc.run.trimmed_dots[1]) == 0            # no SC yet


def test_cross_mode_readouts(golden_complex):
    """CMS on the all-atoms surface and max-CMS on the CMS-mode surface."""
    import cms_cuda
    from oracle.cms_oracle import OracleCalculator
    c = golden_complex
    args = (c["binder_xyz"], c["binder_radii"], c["target_xyz"], c["target_radii"])
    say("cross-mode readouts")
    calc = cms_cuda.MolecularSurfaceCalculator()
    calc.add_binder_and_target(*args)
    mx, per = calc.CalcLoadedMaxPossibleCMS()
    o = OracleCalculator()
    o.add_binder_and_target(*args)
    o.generate(True)
    omx = o.max_contact_ms_of_surface()
    assert rel_err(mx, omx[0]) <= TOL_OUT and rel_err(per, omx[1]) <= TOL_OUT
    for side in (True, False):
        g = calc.calc_contact_molecular_surface(target_side=side)
        e = o.contact_ms_of_surface(side)
        assert rel_err(g[0], e[0]) <= TOL_OUT and rel_err(g[1], e[1]) <= TOL_OUT
        assert g[0] > 0
    calc = cms_cuda.MolecularSurfaceCalculator()
    calc.add_binder_and_target(*args)
    calc.CalcLoaded()
    o = OracleCalculator()
    o.add_binder_and_target(*args)
    o.generate(False)
    g = calc.calc_max_possible_contact_molecular_surface()
    e = o.max_contact_ms_of_surface()
    assert rel_err(g[0], e[0]) <= TOL_OUT and rel_err(g[1], e[1]) <= TOL_OUT and g[0] > 0
