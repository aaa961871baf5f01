"""GPU parity: the HIP path (through the C ABI) against the oracle and the
reference's golden numbers.

Bar (integer / index / flag work bit-exact; floats within tolerance):
* dot counts, dot owners (atom_idx), buried flags and dot types: identical,
  in the same order;
* dot coordinates, normals, areas: |a-b| <= 1e-9 * max(1, |a|, |b|) (the
  reference's own CPU/GPU parity tolerance, test/test_gpu_parity.py:47);
  only sin/cos/atan2/acos/exp may differ from NumPy in the last ulp;
* per-atom and total CMS, SC, area, distance: <= 1e-10 relative.
"""

import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL_DOT = 1e-9
TOL_OUT = 1e-10


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.size == 0:
        return 0.0
    return float(np.max(np.abs(a - b) / np.maximum(1.0, np.maximum(np.abs(a), np.abs(b)))))


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def cms_cuda():
    import cms_cuda
    assert cms_cuda.gpu_available(), "HIP library or GPU missing"
    return cms_cuda


def compare_dots(gpu_calc, oracle_calc):
    for m in (0, 1):
        g = gpu_calc.run.dots[m]
        o = oracle_calc.dots[m]
        assert len(g) == len(o), (m, len(g), len(o))
        assert np.array_equal(g.atom_idx, o.atom), m
        assert np.array_equal(g.buried, o.buried), m
        assert np.array_equal(g.type_, o.type), m
        assert rel_err(g.coor_xyz, o.coor) <= TOL_DOT
        assert rel_err(g.outnml_xyz, o.nml) <= TOL_DOT
        assert rel_err(g.area, o.area) <= TOL_DOT


def test_nearest_bruteforce(cms_cuda):
    """Nearest-point primitive: exact squared distances (same summation
    order as cdist) and first-index ties, including duplicated refs."""
    import ctypes
    import torch
    from cms_cuda import _lib
    from scipy.spatial.distance import cdist
    say("nearest: random 3000 x 2500 with duplicates")
    rng = np.random.default_rng(11)
    q = rng.uniform(-20, 20, size=(3000, 3))
    r = rng.uniform(-20, 20, size=(2500, 3))
    r[1500:1600] = r[100:200]          # exact duplicates -> ties
    q[:50] = r[100:150]                # zero distances
    ctx = _lib.Context(0, torch.cuda.current_stream().cuda_stream)
    dq = torch.from_numpy(q).cuda()
    dr = torch.from_numpy(r).cuda()
    d2 = torch.empty(len(q), dtype=torch.float64, device="cuda")
    idx = torch.empty(len(q), dtype=torch.int32, device="cuda")
    ctx("cms_nearest", 0, ctypes.c_void_p(dq.data_ptr()), len(q), ctypes.c_void_p(dr.data_ptr()), len(r),
        ctypes.c_void_p(d2.data_ptr()), ctypes.c_void_p(idx.data_ptr()))
    D = cdist(r, q, metric="sqeuclidean")
    ref_idx = np.argmin(D, axis=0)
    ref_d2 = D[ref_idx, np.arange(len(q))]
    assert np.array_equal(d2.cpu().numpy(), ref_d2)
    assert np.array_equal(idx.cpu().numpy(), ref_idx)
    ctx.close()


def test_golden_complex(cms_cuda, golden_complex, golden_oracle):
    from test_oracle import GOLDEN
    c = golden_complex
    say("golden complex: CMS")
    cms, per_atom, calc = cms_cuda.calculate_contact_ms(c["binder_xyz"], c["binder_radii"],
                                                        c["target_xyz"], c["target_radii"])
    assert isinstance(cms, float) and per_atom.dtype == np.float64 and per_atom.shape == (242,)
    assert rel_err(cms, GOLDEN["cms"]) <= 1e-9
    ocms, oper = golden_oracle["cms"]
    assert rel_err(cms, ocms) <= TOL_OUT
    assert rel_err(per_atom, oper) <= TOL_OUT
    res = calc.run.results
    assert (res.dots.convex, res.dots.toroidal, res.dots.concave) == (
        GOLDEN["convex"], GOLDEN["toroidal"], GOLDEN["concave"])
    compare_dots(calc, golden_oracle["calc"])
    say("golden complex: binder side")
    bcms, bper = calc.calc_contact_molecular_surface(target_side=False)
    assert rel_err(bcms, golden_oracle["cms_binder"][0]) <= TOL_OUT
    assert rel_err(bper, golden_oracle["cms_binder"][1]) <= TOL_OUT
    say("golden complex: SC on the same calculator")
    sc, area, dist = calc.CalcLoadedSC()
    assert rel_err(sc, GOLDEN["sc"]) <= 1e-9
    assert rel_err(area, GOLDEN["area"]) <= 1e-9
    assert rel_err(dist, GOLDEN["distance"]) <= 1e-9
    osc = golden_oracle["sc"]
    assert rel_err([sc, area, dist], list(osc)) <= TOL_OUT
    s = calc.run.results.surface
    assert (s[0].nTrimmedDots, s[1].nTrimmedDots) == GOLDEN["trimmed"]
    atoms = calc.run.atoms
    assert int((atoms.atten == 5).sum()) == GOLDEN["atten5"]
    assert np.array_equal(atoms.atten, golden_oracle["calc"].atten)


def test_golden_probes(cms_cuda, golden_complex, golden_oracle):
    c = golden_complex
    say("golden complex: probes")
    calc = cms_cuda.MolecularSurfaceCalculator()
    calc.add_binder_and_target(c["binder_xyz"], c["binder_radii"], c["target_xyz"], c["target_radii"])
    calc.CalcLoaded()
    pr = calc.run.probes
    o = golden_oracle["calc"]
    assert len(pr) == len(o.pr_h)
    # within each molecule the GPU keeps the reference order; molecules are
    # stored one after the other
    mol = (np.asarray(pr.atom_idx_0) >= len(c["binder_radii"])).astype(int)
    omol = (o.pr_a[:, 0] >= len(c["binder_radii"])).astype(int)
    for m in (0, 1):
        g = np.stack([pr.atom_idx_0, pr.atom_idx_1, pr.atom_idx_2], 1)[mol == m]
        assert np.array_equal(g, o.pr_a[omol == m])
        assert rel_err(pr.point_xyz[mol == m], o.pr_p[omol == m]) <= TOL_DOT


@pytest.mark.parametrize("which", ["config1", "config2"])
def test_synthetic_configs(cms_cuda, which):
    """400 + 1,000 and 1,800 + 4,000 atom synthetic complexes against the
    oracle."""
    from cms_cuda import synth
    from oracle.cms_oracle import OracleCalculator
    say("synthetic " + which)
    c = getattr(synth, which)()
    args = (c["binder_xyz"], c["binder_radii"], c["target_xyz"], c["target_radii"])
    o = OracleCalculator()
    o.add_binder_and_target(*args)
    ocms, oper = o.contact_ms(True)
    osc = o.shape_complementarity()
    calc = cms_cuda.MolecularSurfaceCalculator()
    calc.add_binder_and_target(*args)
    cms, per = calc.CalcLoaded()
    sc = calc.CalcLoadedSC()
    compare_dots(calc, o)
    assert rel_err(cms, ocms) <= TOL_OUT
    assert rel_err(per, oper) <= TOL_OUT
    assert rel_err(list(sc), list(osc)) <= TOL_OUT


def test_max_cms_golden_molecules(cms_cuda, golden_complex):
    from oracle.cms_oracle import calculate_maximum_possible_contact_ms as omax
    for key in ("binder", "target"):
        say("max-CMS " + key)
        xyz, r = golden_complex[key + "_xyz"], golden_complex[key + "_radii"]
        cms, per, calc = cms_cuda.calculate_maximum_possible_contact_ms(xyz, r)
        ocms, oper, ocalc = omax(xyz, r)
        assert rel_err(cms, ocms) <= TOL_OUT
        assert rel_err(per, oper) <= TOL_OUT
        assert len(calc.run.dots[0]) == len(ocalc.dots[0])
        assert np.array_equal(calc.run.dots[0].atom_idx, ocalc.dots[0].atom)


def test_two_atom_molecule(cms_cuda):
    from test_oracle import two_atom_area
    say("two-atom molecule")
    xyz = np.array([[0.0, 0.0, 0.0], [3.0, 0.0, 0.0]])
    cms, per, calc = cms_cuda.calculate_maximum_possible_contact_ms(xyz, np.array([1.8, 1.6]))
    assert abs(cms - two_atom_area(3.0, 1.8, 1.6)) <= 1e-9 * cms


def test_radius_zero_variant(cms_cuda, golden_complex):
    from oracle.cms_oracle import OracleCalculator
    say("radius-0 variant")
    c = golden_complex
    br = c["binder_radii"].copy()
    tr = c["target_radii"].copy()
    br[::7] = 0.0
    tr[3::5] = 0.0
    args = (c["binder_xyz"], br, c["target_xyz"], tr)
    o = OracleCalculator()
    o.add_binder_and_target(*args)
    ocms, oper = o.contact_ms(True)
    osc = o.shape_complementarity()
    calc = cms_cuda.MolecularSurfaceCalculator()
    calc.add_binder_and_target(*args)
    cms, per = calc.CalcLoaded()
    sc = calc.CalcLoadedSC()
    compare_dots(calc, o)
    assert rel_err(cms, ocms) <= TOL_OUT and rel_err(per, oper) <= TOL_OUT
    assert rel_err(list(sc), list(osc)) <= TOL_OUT


def test_far_apart_and_errors(cms_cuda):
    say("far apart / coincident / cache mode")
    rng = np.random.default_rng(2)
    a = rng.uniform(-3, 3, size=(12, 3))
    r = np.full(12, 1.8)
    cms, per, calc = cms_cuda.calculate_contact_ms(a, r, a + 40.0, r)
    assert cms == 0.0 and per.shape == (12,) and not per.any()
    assert calc.CalcLoadedSC() == (0.0, 0.0, 0.0)
    with pytest.raises(RuntimeError, match="reuse it with all_atoms=True"):
        calc.CalcLoadedMaxPossibleCMS()
    xyz = np.array([[0.0, 0, 0], [1.5, 0, 0], [1.5, 0.005, 0]])
    with pytest.raises(RuntimeError, match="Coincident atoms: 1:: == 2::"):
        cms_cuda.calculate_maximum_possible_contact_ms(xyz, np.full(3, 1.8))


def test_deterministic(cms_cuda, golden_complex):
    say("determinism")
    c = golden_complex
    args = (c["binder_xyz"], c["binder_radii"], c["target_xyz"], c["target_radii"])
    r1 = cms_cuda.calculate_contact_ms(*args)
    r2 = cms_cuda.calculate_contact_ms(*args)
    assert r1[0] == r2[0] and np.array_equal(r1[1], r2[1])
    d1, d2 = r1[2].run.dots[1], r2[2].run.dots[1]
    assert np.array_equal(d1.coor_xyz, d2.coor_xyz) and np.array_equal(d1.area, d2.area)


def test_batch_matches_single(cms_cuda, golden_complex):
    """A batch of uneven complexes (ligand, small, golden, radius-0) gives
    results bit-identical to running each complex alone."""
    from cms_cuda import synth
    from cms_cuda.batch import BatchCalculator
    say("batch vs single")
    g = golden_complex
    cs = [(g["binder_xyz"], g["binder_radii"], g["target_xyz"], g["target_radii"])]
    for seed, (nt, nb) in enumerate([(300, 30), (500, 120), (250, 250)]):
        c = synth.make_complex(nt, nb, 100 + seed, gap=3.5)
        cs.append((c["binder_xyz"], c["binder_radii"], c["target_xyz"], c["target_radii"]))
    r0 = g["binder_radii"].copy()
    r0[::4] = 0.0
    cs.append((g["binder_xyz"], r0, g["target_xyz"], g["target_radii"]))
    bc = BatchCalculator()
    bc.load(cs)
    out = bc.contact_ms()
    sc = bc.shape_complementarity()
    per = bc.per_atom_split(out["per_atom"])
    for k, c in enumerate(cs):
        calc = cms_cuda.MolecularSurfaceCalculator()
        calc.add_binder_and_target(*c)
        cms_t, per_t = calc.CalcLoaded()
        cms_b, per_b = calc.calc_contact_molecular_surface(target_side=False)
        s = calc.CalcLoadedSC()
        assert out["target"][k] == cms_t and out["binder"][k] == cms_b, k
        assert np.array_equal(per[k][1], per_t) and np.array_equal(per[k][0], per_b), k
        assert tuple(sc[k, :3]) == s, k
    bc.close()


def test_fp32_search_matches_fp64(cms_cuda):
    """FP32 search with FP64 rescoring returns the FP64 distances and the
    FP64 first-index partners exactly, including exact ties and near-ties
    closer than float resolution."""
    import ctypes
    import torch
    from cms_cuda import _lib
    say("fp32 vs fp64 nearest")
    rng = np.random.default_rng(7)
    q = rng.uniform(-60, 60, size=(4000, 3)) + np.array([30.0, -20.0, 10.0])
    r = rng.uniform(-60, 60, size=(3000, 3)) + np.array([30.0, -20.0, 10.0])
    r[2000:2100] = r[0:100]                       # exact ties
    r[2100:2200] = r[100:200] + 1e-12             # near-ties below float resolution
    q[:200] = r[:200] + rng.normal(scale=1e-3, size=(200, 3))
    ctx = _lib.Context(0, torch.cuda.current_stream().cuda_stream)
    dq = torch.from_numpy(q).cuda()
    dr = torch.from_numpy(r).cuda()
    res = {}
    for prec in (_lib.PREC_FP64, _lib.PREC_FP32_RESCORE):
        d2 = torch.empty(len(q), dtype=torch.float64, device="cuda")
        idx = torch.empty(len(q), dtype=torch.int32, device="cuda")
        ctx("cms_nearest", prec, ctypes.c_void_p(dq.data_ptr()), len(q), ctypes.c_void_p(dr.data_ptr()), len(r),
            ctypes.c_void_p(d2.data_ptr()), ctypes.c_void_p(idx.data_ptr()))
        res[prec] = (d2.cpu().numpy(), idx.cpu().numpy())
    n = ctypes.c_uint64()
    ctx("cms_search_stats", ctypes.byref(n))
    assert np.array_equal(res[0][0], res[1][0])
    assert np.array_equal(res[0][1], res[1][1])
    assert n.value < 200 * len(q)   # rescoring stays a small fraction of all pairs
    ctx.close()


def test_fp32_search_equidistant_shell(cms_cuda):
    """References on a sphere of radius 55 A around the queries, radii
    spread by at most 1e-5 A: the float filter cannot order them (its error
    at this frame size is ~1e-2 A^2), so only a rescoring band as wide as the
    filter's error bound finds the FP64 nearest reference and its
    first-index ties."""
    import ctypes
    import torch
    from cms_cuda import _lib
    say("fp32 vs fp64 nearest, equidistant shell")
    rng = np.random.default_rng(11)
    c0 = np.array([12.5, -7.25, 3.0])
    u = rng.normal(size=(3000, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    r = c0 + (55.0 + rng.uniform(-1e-5, 1e-5, size=(3000, 1))) * u
    r[2500:2600] = r[1000:1100]                   # exact ties
    q = c0 + rng.normal(scale=1e-7, size=(64, 3))
    q[:8] = c0
    ctx = _lib.Context(0, torch.cuda.current_stream().cuda_stream)
    dq = torch.from_numpy(q).cuda()
    dr = torch.from_numpy(r).cuda()
    res = {}
    for prec in (_lib.PREC_FP64, _lib.PREC_FP32_RESCORE):
        d2 = torch.empty(len(q), dtype=torch.float64, device="cuda")
        idx = torch.empty(len(q), dtype=torch.int32, device="cuda")
        ctx("cms_nearest", prec, ctypes.c_void_p(dq.data_ptr()), len(q), ctypes.c_void_p(dr.data_ptr()), len(r),
            ctypes.c_void_p(d2.data_ptr()), ctypes.c_void_p(idx.data_ptr()))
        res[prec] = (d2.cpu().numpy(), idx.cpu().numpy())
    # FP64 partners against numpy's squared distances (x^2 + y^2 first)
    e = r[None, :, :] - q[:, None, :]
    dn = (e[..., 0] * e[..., 0] + e[..., 1] * e[..., 1]) + e[..., 2] * e[..., 2]
    assert np.array_equal(res[0][0], dn.min(axis=1))
    assert np.array_equal(res[0][1], dn.argmin(axis=1))
    assert np.array_equal(res[0][0], res[1][0])
    assert np.array_equal(res[0][1], res[1][1])
    ctx.close()


def test_fp32_mode_batch_identical(cms_cuda, golden_complex):
    """Batched CMS / SC in FP32-search mode equal the FP64 mode bit for bit."""
    from cms_cuda import synth
    from cms_cuda import _lib
    from cms_cuda.batch import BatchCalculator
    say("fp32 mode batch")
    g = golden_complex
    cs = [(g["binder_xyz"], g["binder_radii"], g["target_xyz"], g["target_radii"])]
    for seed, (nt, nb) in enumerate([(800, 40), (1000, 400), (1200, 900)]):
        c = synth.make_complex(nt, nb, 200 + seed, gap=3.5)
        cs.append((c["binder_xyz"], c["binder_radii"], c["target_xyz"], c["target_radii"]))
    bc = BatchCalculator()
    bc.load(cs)
    a64 = bc.contact_ms(_lib.PREC_FP64)
    s64 = bc.shape_complementarity(_lib.PREC_FP64)
    a32 = bc.contact_ms(_lib.PREC_FP32_RESCORE)
    s32 = bc.shape_complementarity(_lib.PREC_FP32_RESCORE)
    assert np.array_equal(a64["target"], a32["target"]) and np.array_equal(a64["binder"], a32["binder"])
    assert np.array_equal(a64["per_atom"], a32["per_atom"])
    assert np.array_equal(s64, s32)
    bc.close()


# ---------------------------------------------------------------- vectors
# written by the reference's own calculator (tests/golden/make_golden.py)

def _fixture(name):
    return np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", name), allow_pickle=False)


def _gpu_run(cms_cuda, bx, br, tx, tr):
    calc = cms_cuda.MolecularSurfaceCalculator()
    calc.add_binder_and_target(bx, br, tx, tr)
    cms_t, per_t = calc.CalcLoaded()
    cms_b, per_b = calc.calc_contact_molecular_surface(target_side=False)
    sc = calc.CalcLoadedSC()
    return calc, cms_t, per_t, cms_b, per_b, sc


def _check_scalars(f, cms_t, per_t, cms_b, per_b, sc):
    assert rel_err(cms_t, f["cms_target"]) <= TOL_OUT
    assert rel_err(cms_b, f["cms_binder"]) <= TOL_OUT
    assert len(per_t) == len(f["per_target"]) and rel_err(per_t, f["per_target"]) <= TOL_OUT
    assert len(per_b) == len(f["per_binder"]) and rel_err(per_b, f["per_binder"]) <= TOL_OUT
    assert rel_err(list(sc), f["sc"]) <= TOL_OUT


def test_reference_vectors_small_complex(cms_cuda):
    """Every dot of a synthetic complex in the reference's order."""
    say("reference vectors: small complex")
    f = _fixture("ref_small.npz")
    calc, *out = _gpu_run(cms_cuda, f["binder_xyz"], f["binder_radii"], f["target_xyz"], f["target_radii"])
    _check_scalars(f, *out)
    for m in (0, 1):
        g = calc.run.dots[m]
        assert len(g) == len(f["dots%d_area" % m])
        assert np.array_equal(g.atom_idx, f["dots%d_atom" % m])
        assert np.array_equal(g.buried, f["dots%d_buried" % m])
        assert np.array_equal(g.type_, f["dots%d_type" % m])
        assert rel_err(g.coor_xyz, f["dots%d_coor" % m]) <= TOL_DOT
        assert rel_err(g.outnml_xyz, f["dots%d_nml" % m]) <= TOL_DOT
        assert rel_err(g.area, f["dots%d_area" % m]) <= TOL_DOT


def test_reference_vectors_radius0_maxcms_c1(cms_cuda, golden_complex):
    say("reference vectors: radius 0, max-CMS, c1")
    g = golden_complex
    f = _fixture("ref_radius0.npz")
    _, *out = _gpu_run(cms_cuda, g["binder_xyz"], f["binder_radii"], g["target_xyz"], g["target_radii"])
    _check_scalars(f, *out)
    f = _fixture("ref_maxcms.npz")
    tot, per, calc = cms_cuda.calculate_maximum_possible_contact_ms(g["binder_xyz"], g["binder_radii"])
    assert rel_err(tot, f["total"]) <= TOL_OUT and rel_err(per, f["per_atom"]) <= TOL_OUT
    assert len(calc.run.dots[0]) == int(f["ndots"])
    f = _fixture("ref_c1.npz")
    _, *out = _gpu_run(cms_cuda, f["binder_xyz"], f["binder_radii"], f["target_xyz"], f["target_radii"])
    _check_scalars(f, *out)


def test_atom_state_matches_oracle(cms_cuda, golden_complex, golden_oracle):
    """calc.run.atoms.atten / .access for every atom, including the
    reference's last-write-wins toroidal access merge (golden tally: 258
    atoms with access)."""
    from oracle.cms_oracle import OracleCalculator
    say("atom state")
    g = golden_complex
    cases = [(g["binder_xyz"], g["binder_radii"], g["target_xyz"], g["target_radii"], golden_oracle["calc"])]
    f = _fixture("ref_small.npz")
    o = OracleCalculator()
    o.add_binder_and_target(f["binder_xyz"], f["binder_radii"], f["target_xyz"], f["target_radii"])
    o.generate(False)
    cases.append((f["binder_xyz"], f["binder_radii"], f["target_xyz"], f["target_radii"], o))
    for k, (bx, br, tx, tr, oc) in enumerate(cases):
        calc = cms_cuda.MolecularSurfaceCalculator()
        calc.add_binder_and_target(bx, br, tx, tr)
        calc.CalcLoaded()
        at = calc.run.atoms
        assert np.array_equal(at.atten, oc.atten), k
        assert np.array_equal(at.access, oc.access), (k, int(at.access.sum()), int(oc.access.sum()))
        if k == 0:
            assert int(at.access.sum()) == 258
