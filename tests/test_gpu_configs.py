"""The benchmark workloads as GPU tests (through the C ABI).

* c4: 2,048 mixed complexes (binders log-uniform 20-1,800 atoms, the
  20-60-atom ones ligand-like; targets 800-6,000; gaps 3.5-5.0 A) plus 64
  complexes of 400 + 1,000 atoms.  FP32 search with FP64 rescoring must give
  the FP64 results bit for bit on every complex, with identical counts; a
  size-stratified sample of the batch is checked against the NumPy oracle
  (identical dot counts, owners, burial and types in order; coordinates,
  normals and areas within 1e-9; CMS, per-atom CMS, SC, area, distance and
  the per-side statistics within 1e-10 relative).
* c3: the first 2,048 complexes of that batch in FP64 as one graph-replayed
  step, equal bit for bit to the same complexes run in chunks of 128.
* c5: 256 complexes (1,500-atom targets, binders log-uniform 80-600 atoms)
  with the FP32-search forward; the hand-written backward against central
  finite differences of the frozen-topology CMS on 64 atoms per complex (up
  to 48 of them with a nonzero gradient)
  (h = 1e-4 A, FP64, one perturbed atom in every complex per evaluation),
  and against torch autograd through the FP64 torch re-implementation on 8
  complexes.
"""

import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

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


def as_tuples(cs):
    return [(c["binder_xyz"], c["binder_radii"], c["target_xyz"], c["target_radii"]) for c in cs]


def same_results(a, b):
    return all(np.array_equal(a[k], b[k]) for k in ("target", "binder", "per_atom", "sc"))


@pytest.fixture(scope="module")
def c4_batch():
    from cms_cuda import synth
    from oracle_jobs import pool_workers
    say("c4: generating 2,112 complexes")
    return synth.c4_batch(3, pool_workers())


@pytest.fixture(scope="module")
def c4_fp64(c4_batch):
    """One FP64 step over the whole c4 batch: results, counts, and the dots
    of the size-stratified sample."""
    from cms_cuda.batch import BatchCalculator
    from cms_cuda import _lib
    say("c4: FP64 step over the batch")
    bc = BatchCalculator()
    bc.load(as_tuples(c4_batch))
    out = bc.run(_lib.PREC_FP64)
    out["counts"] = bc.counts()
    sample = stratified_sample(c4_batch)
    out["sample"] = sample
    out["dots"] = {c: [copy_dots(bc, c, m) for m in (0, 1)] for c in sample}
    bc.close()
    return out


def stratified_sample(cs, n=32):
    """n complexes spread over the binder-size order (ligands to the largest
    binders) plus the largest complex and two of the 400 + 1,000-atom ones."""
    nb = np.array([len(c["binder_radii"]) for c in cs[:2048]])
    order = np.argsort(nb, kind="stable")
    pick = [int(order[int(round(q))]) for q in np.linspace(0, len(order) - 1, n)]
    tot = np.array([len(c["binder_radii"]) + len(c["target_radii"]) for c in cs[:2048]])
    pick.append(int(np.argmax(tot)))
    pick += [2048, len(cs) - 1]
    return sorted(set(pick))


def copy_dots(bc, c, m):
    import torch
    n = ctypes.c_int64()
    bc.ctx("cms_dot_count", c, m, ctypes.byref(n))
    n = n.value
    k = max(n, 1)
    coor = torch.empty((k, 3), dtype=torch.float64, device="cuda")
    nml = torch.empty((k, 3), dtype=torch.float64, device="cuda")
    area = torch.empty(k, dtype=torch.float64, device="cuda")
    bur = torch.empty(k, dtype=torch.int8, device="cuda")
    typ = torch.empty(k, dtype=torch.int8, device="cuda")
    atom = torch.empty(k, dtype=torch.int32, device="cuda")
    P = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    if n:
        bc.ctx("cms_copy_dots", c, m, P(coor), P(nml), P(area), P(bur), P(typ), P(atom))
    return dict(coor=coor[:n].cpu().numpy(), nml=nml[:n].cpu().numpy(), area=area[:n].cpu().numpy(),
                buried=bur[:n].cpu().numpy(), type=typ[:n].cpu().numpy(), atom=atom[:n].cpu().numpy())


def test_c4_fp32_search_equals_fp64(c4_batch, c4_fp64):
    from cms_cuda.batch import BatchCalculator
    from cms_cuda import _lib
    say("c4: FP32-search step over the batch")
    bc = BatchCalculator()
    bc.load(as_tuples(c4_batch))
    out = bc.run(_lib.PREC_FP32_RESCORE)
    counts = bc.counts()
    bc.close()
    assert np.array_equal(counts, c4_fp64["counts"])
    assert same_results(out, c4_fp64)
    # the batch is the workload it claims to be
    nb = np.array([len(c["binder_radii"]) for c in c4_batch])
    assert len(c4_batch) == 2112 and nb[:2048].min() <= 25 and nb[:2048].max() >= 1700
    assert (c4_fp64["target"] > 0).sum() > 2000


def test_c4_sample_matches_oracle(c4_batch, c4_fp64):
    from oracle_jobs import oracle_many
    sample = c4_fp64["sample"]
    say("c4: oracle on %d sampled complexes" % len(sample))
    refs = oracle_many([{k: c4_batch[c][k] for k in ("binder_xyz", "binder_radii", "target_xyz", "target_radii")}
                        for c in sample])
    off = np.concatenate([[0], np.cumsum([len(c["binder_radii"]) + len(c["target_radii"]) for c in c4_batch])])
    sizes = []
    for c, ref in zip(sample, refs):
        nb = len(c4_batch[c]["binder_radii"])
        nt = len(c4_batch[c]["target_radii"])
        sizes.append(nb)
        for m in (0, 1):
            g, o = c4_fp64["dots"][c][m], ref["dots"][m]
            assert len(g["area"]) == len(o["area"]), (c, m)
            assert np.array_equal(g["atom"], o["atom"]), (c, m)
            assert np.array_equal(g["buried"], o["buried"]), (c, m)
            assert np.array_equal(g["type"], o["type"]), (c, m)
            for f in ("coor", "nml", "area"):
                assert rel_err(g[f], o[f]) <= TOL_DOT, (c, m, f)
        cnt = c4_fp64["counts"][c]
        assert (cnt[4], cnt[5], cnt[6]) == (ref["counts"]["convex"], ref["counts"]["toroidal"],
                                            ref["counts"]["concave"]), c
        per = c4_fp64["per_atom"][off[c]:off[c + 1]]
        assert rel_err(c4_fp64["target"][c], ref["target"][0]) <= TOL_OUT, c
        assert rel_err(per[nb:], ref["target"][1]) <= TOL_OUT, c
        assert rel_err(c4_fp64["binder"][c], ref["binder"][0]) <= TOL_OUT, c
        assert rel_err(per[:nb], ref["binder"][1]) <= TOL_OUT, c
        rec = c4_fp64["sc"][c]
        assert rel_err(rec[:3], ref["sc"]) <= TOL_OUT, c
        for m in (0, 1):
            side = rec[8 + 8 * m: 16 + 8 * m]
            assert int(side[5]) == ref["n_trimmed"][m], (c, m)
            assert rel_err(side[4], ref["trimmed_area"][m]) <= TOL_OUT, (c, m)
            s = ref["sides"][m]
            if s is not None:
                assert rel_err(side[:4], [s["d_mean"], s["d_median"], s["s_mean"], s["s_median"]]) <= TOL_OUT
        assert nt >= 800
    trimmed = [int(c4_fp64["sc"][c][13] + c4_fp64["sc"][c][21]) for c in sample]
    say("  sample binder sizes %d..%d, trimmed dots up to %d" % (min(sizes), max(sizes), max(trimmed)))
    assert min(sizes) <= 60 and max(sizes) >= 1700 and max(trimmed) > 1000


def test_c3_graph_step_equals_chunks(c4_batch):
    """c3 (the first 2,048 complexes) as one graph-captured FP64 step, replayed,
    equal to the same complexes in chunks of 128."""
    import torch
    from cms_cuda.batch import BatchCalculator
    from cms_cuda import _lib
    cs = as_tuples(c4_batch[:2048])
    say("c3: chunks of 128")
    parts = []
    bc = BatchCalculator()
    for k in range(0, len(cs), 128):
        bc.load(cs[k:k + 128])
        parts.append(bc.run(_lib.PREC_FP64))
    bc.close()
    ref = {key: np.concatenate([p[key] for p in parts]) for key in ("target", "binder", "per_atom", "sc")}
    say("c3: one batch, graph replay")
    bc = BatchCalculator()
    bc.load(cs)
    ntot = bc.ntot
    per_atom = torch.empty(ntot, dtype=torch.float64, device="cuda")
    totals = torch.empty(2 * len(cs), dtype=torch.float64, device="cuda")
    scr = torch.empty(len(cs) * _lib.SC_NRES, dtype=torch.float64, device="cuda")
    for step in range(4):
        bc.load_device(bc.nb, bc.nt, bc.d_xyz, bc.d_rad)
        bc.run(_lib.PREC_FP64, graph=True, out=(per_atom, totals, scr))
        torch.cuda.synchronize()
        t = totals.cpu().numpy().reshape(len(cs), 2)
        got = dict(target=t[:, 1], binder=t[:, 0], per_atom=per_atom.cpu().numpy(),
                   sc=scr.cpu().numpy().reshape(len(cs), _lib.SC_NRES))
        assert same_results(got, ref), step
    cap, rep, exact, ovf = bc.run_stats()
    say("  captures %d replays %d exact %d overflow %d" % (cap, rep, exact, ovf))
    assert cap >= 1 and rep >= 1
    bc.close()


@pytest.fixture(scope="module")
def c5_state():
    import torch
    from cms_cuda import synth, _lib
    from cms_cuda.autograd import _ContactMS
    from oracle_jobs import pool_workers
    say("c5: generating 256 complexes")
    cs = synth.c5_batch(5, pool_workers())
    nb = [len(c["binder_radii"]) for c in cs]
    nt = [len(c["target_radii"]) for c in cs]
    xyz = np.concatenate([np.concatenate([c["binder_xyz"], c["target_xyz"]]) for c in cs])
    rad = np.concatenate([np.concatenate([c["binder_radii"], c["target_radii"]]) for c in cs])
    st = _ContactMS(torch.from_numpy(xyz).cuda(), torch.from_numpy(rad).cuda(), nb, nt, None,
                    _lib.PREC_FP32_RESCORE)
    yield dict(cs=cs, nb=np.array(nb), nt=np.array(nt), xyz=xyz, rad=rad, st=st)
    st.close()


def test_c5_gradient_matches_finite_differences(c5_state):
    """dCMS/dxyz of the sum of both sides of every complex against central
    differences at frozen topology, h = 1e-4 A.  The complexes are
    independent, so one evaluation perturbs one atom in each of the 256
    complexes.  Bars: per component, |g - FD| <= 1e-5 * s against the
    Richardson extrapolation (4 FD(h/2) - FD(h)) / 3 and <= 1e-3 * s against
    the plain step-h difference, s = max(|g|, |FD|, 1e-3); per complex, the
    plain difference within 1e-5 of g in the 2-norm over the checked
    components."""
    import torch
    s = c5_state
    st = s["st"]
    B = len(s["nb"])
    g_tot = torch.ones(2 * B, dtype=torch.float64, device="cuda")
    grad = st.backward(g_tot, None).cpu().numpy()
    off = np.concatenate([[0], np.cumsum(s["nb"] + s["nt"])])
    rng = np.random.default_rng(2024)
    picks = np.zeros((B, 64), dtype=np.int64)
    n_active = 0
    for c in range(B):
        lo, hi = off[c], off[c + 1]
        active = lo + np.nonzero(np.abs(grad[lo:hi]).sum(1) > 0)[0]
        n_active += min(48, len(active))
        # up to 48 atoms with a nonzero gradient, the rest drawn from all atoms
        a = rng.choice(active, size=min(48, len(active)), replace=False)
        rest = np.setdiff1d(np.arange(lo, hi), a)
        picks[c] = np.concatenate([a, rng.choice(rest, size=64 - len(a), replace=False)])
    say("c5: %d checked atoms with a nonzero gradient over %d complexes" % (n_active, B))
    assert n_active >= B
    x0 = torch.from_numpy(s["xyz"]).cuda()
    h = 1e-4

    def total(x):
        t = st.frozen_totals(x).cpu().numpy().reshape(B, 2)
        return t[:, 0] + t[:, 1]

    say("c5: %d frozen-topology evaluations" % (64 * 3 * 4))
    fd = np.zeros((B, 64, 3))
    rich = np.zeros((B, 64, 3))
    for k in range(64):
        idx = torch.from_numpy(picks[:, k]).cuda()
        for d in range(3):
            vals = []
            for hh in (h, h / 2):
                xp = x0.clone()
                xp[idx, d] += hh
                xm = x0.clone()
                xm[idx, d] -= hh
                vals.append((total(xp) - total(xm)) / (2 * hh))
            fd[:, k, d] = vals[0]
            rich[:, k, d] = (4 * vals[1] - vals[0]) / 3
    g = grad[picks]                                   # (B, 64, 3)
    scale = np.maximum(np.maximum(np.abs(g), np.abs(rich)), 1e-3)
    e_rich = np.abs(g - rich) / scale
    e_plain = np.abs(g - fd) / scale
    gn = np.linalg.norm(g.reshape(B, -1), axis=1)
    live = gn > 0
    # complexes without an interface: zero gradient, and the difference
    # quotients vanish too
    assert np.abs(fd[~live]).max(initial=0.0) <= 1e-8
    e_norm = np.linalg.norm((g - fd).reshape(B, -1)[live], axis=1) / gn[live]
    say("  per component: Richardson %.2e, plain %.2e; per complex 2-norm, plain %.2e (%d complexes without "
        "an interface)" % (e_rich.max(), e_plain.max(), e_norm.max(), int((~live).sum())))
    assert e_rich.max() <= 1e-5
    assert e_plain.max() <= 1e-3
    assert e_norm.max() <= 1e-5


def test_c5_gradient_matches_torch_reference(c5_state):
    """Eight complexes of the batch (spread over the binder sizes): the HIP
    gradient of each complex's target-side CMS against torch autograd
    through the FP64 torch re-implementation of the frozen-topology CMS."""
    import torch
    from scipy.spatial import cKDTree
    import torch_frozen as tf
    s = c5_state
    st = s["st"]
    B = len(s["nb"])
    order = np.argsort(s["nb"], kind="stable")
    sel = [int(order[int(round(q))]) for q in np.linspace(0, B - 1, 8)]
    X = torch.from_numpy(s["xyz"].copy()).requires_grad_(True)
    R = torch.from_numpy(s["rad"].copy())
    D = torch.full((len(s["rad"]),), 15.0, dtype=torch.float64)
    tot = st.totals.cpu().numpy()
    off = np.concatenate([[0], np.cumsum(s["nb"] + s["nt"])])
    for c in sel:
        per_mol = []
        for mol in (0, 1):
            n = ctypes.c_int64()
            st.ctx("cms_dot_count", c, mol, ctypes.byref(n))
            n = n.value
            k = max(n, 1)
            coor = torch.empty((k, 3), dtype=torch.float64, device="cuda")
            bur = torch.empty(k, dtype=torch.int8, device="cuda")
            d1 = torch.empty((k, 4), dtype=torch.int32, device="cuda")
            d2 = torch.empty((k, 4), dtype=torch.int32, device="cuda")
            P = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
            st.ctx("cms_copy_dots", c, mol, P(coor), None, None, P(bur), None, None)
            st.ctx("cms_copy_dot_desc", c, mol, P(d1), P(d2))
            b = torch.from_numpy(bur[:n].cpu().numpy().astype(bool))
            per_mol.append(dict(coor=coor[:n].cpu().numpy()[b.numpy()], d1=d1[:n].cpu().long()[b],
                                d2=d2[:n].cpu().long()[b]))
        q, r = per_mol[1], per_mol[0]
        assert len(q["coor"]) > 50 and len(r["coor"]) > 50, c
        partner = torch.from_numpy(cKDTree(r["coor"]).query(q["coor"])[1])
        cms = tf.frozen_cms(X, R, D, 1.7, 0.5, (q["d1"], q["d2"]), (r["d1"], r["d2"]), partner)
        assert abs(cms.item() - tot[2 * c + 1]) <= 1e-10 * max(1.0, abs(tot[2 * c + 1])), c
        (g_ref,) = torch.autograd.grad(cms, X)
        g_tot = torch.zeros(2 * B, dtype=torch.float64, device="cuda")
        g_tot[2 * c + 1] = 1.0
        g_hip = st.backward(g_tot, None).cpu().numpy()
        g_ref = g_ref.numpy()
        sc = np.max(np.abs(g_ref))
        err = np.max(np.abs(g_hip - g_ref)) / sc
        outside = np.abs(np.delete(g_hip, np.arange(off[c], off[c + 1]), axis=0)).max()
        say("  complex %d (binder %d atoms): max |grad| %.3e, rel err %.2e" % (c, s["nb"][c], sc, err))
        assert sc > 0 and err < 1e-9 and outside == 0.0, c
