"""GPU: the one-step entry point cms_run (surface + CMS both sides + SC with
one host wait), its learned buffer sizes, the exact rerun after an
overflow, and HIP-graph capture / replay.

Bar: every result bit-identical to the separate entry points
(cms_generate + cms_contact_ms + cms_shape_complementarity) on a fresh
context, for every step and every replay.
"""

import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _batch(golden, sizes, seed0):
    from cms_cuda import synth
    g = golden
    cs = [(g["binder_xyz"], g["binder_radii"], g["target_xyz"], g["target_radii"])]
    for k, (nt, nb) in enumerate(sizes):
        c = synth.make_complex(nt, nb, seed0 + k, gap=3.5)
        cs.append((c["binder_xyz"], c["binder_radii"], c["target_xyz"], c["target_radii"]))
    return cs


def _reference(cs, precision=0):
    """Separate entry points on a fresh context (exact sizing)."""
    from cms_cuda.batch import BatchCalculator
    bc = BatchCalculator()
    bc.load(cs)
    out = bc.contact_ms(precision)
    out["sc"] = bc.shape_complementarity(precision)
    bc.close()
    return out


def _same(a, b):
    return (np.array_equal(a["target"], b["target"]) and np.array_equal(a["binder"], b["binder"])
            and np.array_equal(a["per_atom"], b["per_atom"]) and np.array_equal(a["sc"], b["sc"]))


def _flat(cs):
    nb = np.array([len(c[1]) for c in cs], dtype=np.int32)
    nt = np.array([len(c[3]) for c in cs], dtype=np.int32)
    xyz = np.concatenate([np.concatenate([c[0], c[2]]) for c in cs])
    rad = np.concatenate([np.concatenate([c[1], c[3]]) for c in cs])
    return nb, nt, xyz, rad


def test_run_matches_separate_calls(golden_complex):
    from cms_cuda.batch import BatchCalculator
    cs = _batch(golden_complex, [(400, 40), (700, 200), (300, 300)], 300)
    say("cms_run vs separate entry points, %d complexes" % len(cs))
    ref = _reference(cs)
    bc = BatchCalculator()
    bc.load(cs)
    for step in range(3):
        say("  step %d" % step)
        got = bc.run(graph=False)
        assert _same(got, ref), step
    cap, rep, exact, ovf = bc.run_stats()
    assert exact == 1 and ovf == 0 and cap == 0 and rep == 0
    bc.close()


def test_run_fp32_search_identical(golden_complex):
    from cms_cuda import _lib
    from cms_cuda.batch import BatchCalculator
    cs = _batch(golden_complex, [(900, 60), (1100, 500)], 320)
    say("cms_run fp32 search vs fp64")
    bc = BatchCalculator()
    bc.load(cs)
    a = bc.run(_lib.PREC_FP64)
    b = bc.run(_lib.PREC_FP32_RESCORE)
    c = bc.run(_lib.PREC_FP32_RESCORE, graph=True)
    d = bc.run(_lib.PREC_FP32_RESCORE, graph=True)
    assert _same(a, b) and _same(a, c) and _same(a, d)
    bc.close()


def test_overflow_repeats_exactly(golden_complex):
    """A context that learned from a batch without any surface (molecules
    40 A apart) under-sizes every buffer for the next batch: the learned
    run overflows, is detected on the device and repeated exactly."""
    from cms_cuda.batch import BatchCalculator
    g = golden_complex
    far = [(g["binder_xyz"], g["binder_radii"], g["target_xyz"] + 40.0, g["target_radii"])]
    cs = _batch(golden_complex, [(500, 100)], 340)
    say("overflow rerun")
    ref = _reference(cs)
    bc = BatchCalculator()
    bc.load(far)
    z = bc.run()
    assert not z["target"].any() and not z["binder"].any()
    bc.load(cs)
    got = bc.run()
    assert _same(got, ref)
    cap, rep, exact, ovf = bc.run_stats()
    assert ovf >= 1 and exact == 2
    # the learned sizes now fit: no further overflow
    got = bc.run()
    assert _same(got, ref) and bc.run_stats()[3] == ovf
    bc.close()


def test_graph_replays_match_eager(golden_complex):
    """Graph capture / replay on one layout while the coordinates change in
    place between steps; eager torch work and synchronisations between the
    replays; every replay checked against a fresh eager run on the same
    coordinates."""
    import torch
    from cms_cuda.batch import BatchCalculator
    cs = _batch(golden_complex, [(600, 80), (800, 300), (350, 350)], 360)
    nb, nt, xyz0, rad = _flat(cs)
    rng = np.random.default_rng(5)
    xyz1 = xyz0 + rng.normal(scale=0.02, size=xyz0.shape)
    dev = torch.device("cuda", 0)
    d_xyz = torch.from_numpy(xyz0.copy()).to(dev)
    d_rad = torch.from_numpy(rad.copy()).to(dev)
    refs = {}
    for key, x in (("a", xyz0), ("b", xyz1)):
        off = np.concatenate([[0], np.cumsum(nb + nt)])
        cs_k = []
        for c in range(len(cs)):
            xc = x[off[c]:off[c + 1]]
            rc = rad[off[c]:off[c + 1]]
            cs_k.append((xc[:nb[c]], rc[:nb[c]], xc[nb[c]:], rc[nb[c]:]))
        refs[key] = _reference(cs_k)
    bc = BatchCalculator()
    per_atom = torch.empty(len(rad), dtype=torch.float64, device=dev)
    totals = torch.empty(2 * len(cs), dtype=torch.float64, device=dev)
    scr = torch.empty(len(cs) * 24, dtype=torch.float64, device=dev)
    seq = ["a", "a", "a", "b", "a", "b", "b", "a"]
    for step, key in enumerate(seq):
        say("  graph step %d coords %s" % (step, key))
        d_xyz.copy_(torch.from_numpy(xyz0 if key == "a" else xyz1))
        bc.load_device(nb, nt, d_xyz, d_rad)
        bc.run(graph=True, out=(per_atom, totals, scr))
        # eager work of the program between replays
        junk = torch.randn(4096, device=dev).sum()
        torch.cuda.synchronize()
        t = totals.cpu().numpy().reshape(len(cs), 2)
        got = dict(target=t[:, 1], binder=t[:, 0], per_atom=per_atom.cpu().numpy(),
                   sc=scr.cpu().numpy().reshape(len(cs), 24))
        assert _same(got, refs[key]), (step, key)
        del junk
    cap, rep, exact, ovf = bc.run_stats()
    say("  captures %d replays %d exact %d overflow %d" % (cap, rep, exact, ovf))
    assert cap >= 1 and rep >= 3
    bc.close()


def _probe_lists(bc, counts):
    """Every complex's probes (atoms, height, position, normal) via the C ABI."""
    import ctypes
    import torch
    out = []
    for c in range(bc.B):
        nn = ctypes.c_int64()
        bc.ctx("cms_probe_count", c, ctypes.byref(nn))
        n = nn.value
        assert n == counts[c, 7]
        atoms = torch.empty((max(n, 1), 3), dtype=torch.int32, device="cuda")
        h = torch.empty(max(n, 1), dtype=torch.float64, device="cuda")
        p = torch.empty((max(n, 1), 3), dtype=torch.float64, device="cuda")
        a = torch.empty((max(n, 1), 3), dtype=torch.float64, device="cuda")
        if n:
            bc.ctx("cms_copy_probes", c, ctypes.c_void_p(atoms.data_ptr()), ctypes.c_void_p(h.data_ptr()),
                   ctypes.c_void_p(p.data_ptr()), ctypes.c_void_p(a.data_ptr()))
        out.append(tuple(t[:n].cpu().numpy() for t in (atoms, h, p, a)))
    return out


def test_probe_count_kernels_agree(golden_complex, monkeypatch):
    """The probes count pass has a per-pair and a per-row form (chosen by
    batch size); the per-row form stages up to 8 surviving probes per row
    for placement after the scan and leaves pairs past them to the emit
    pass.  Per pair, per row, and per row with one slot (staged and emitted
    probes mixed) must give identical counts, probe lists and results."""
    from cms_cuda.batch import BatchCalculator
    cs = _batch(golden_complex, [(400, 40), (700, 200), (300, 300), (900, 500)], 700)
    outs = []
    for mode, slots in (("pairs", None), ("rows", None), ("rows", "1")):
        monkeypatch.setenv("CMS_PROBE_COUNT", mode)
        if slots is None:
            monkeypatch.delenv("CMS_PROBE_SLOTS", raising=False)
        else:
            monkeypatch.setenv("CMS_PROBE_SLOTS", slots)
        say("probes count pass per %s (slots %s), %d complexes" % (mode[:-1], slots, len(cs)))
        bc = BatchCalculator()
        bc.load(cs)
        o = bc.contact_ms(0)
        o["sc"] = bc.shape_complementarity(0)
        o["counts"] = bc.counts()
        o["probes"] = _probe_lists(bc, o["counts"])
        bc.close()
        outs.append(o)
    a = outs[0]
    assert sum(len(pl[1]) for pl in a["probes"]) > 100
    for b in outs[1:]:
        assert np.array_equal(a["counts"], b["counts"])
        for pa, pb in zip(a["probes"], b["probes"]):
            for x, y in zip(pa, pb):
                assert np.array_equal(x, y)
        assert _same({k: v for k, v in a.items() if k != "probes"}, {k: v for k, v in b.items() if k != "probes"})


def test_group_totals_numpy_order(golden_complex):
    """Every molecule's CMS total is np.sum of its per-atom values, bit for
    bit: NumPy's pairwise tree (leaves of at most 128 values with eight
    interleaved accumulators) evaluated on the GPU, for groups from under
    128 atoms (one leaf) to thousands (several tree levels)."""
    from cms_cuda.batch import BatchCalculator
    cs = _batch(golden_complex, [(60, 20), (3000, 900), (1500, 130), (4200, 1700)], 910)
    say("group totals vs np.sum, %d complexes" % len(cs))
    bc = BatchCalculator()
    bc.load(cs)
    o = bc.contact_ms(0)
    bc.close()
    per = o["per_atom"]
    off = 0
    sizes = []
    for c, (bx, br, tx, tr) in enumerate(cs):
        nb, nt = len(br), len(tr)
        b_seg = per[off:off + nb]
        t_seg = per[off + nb:off + nb + nt]
        assert o["binder"][c] == np.sum(b_seg), c
        assert o["target"][c] == np.sum(t_seg), c
        sizes += [nb, nt]
        off += nb + nt
    assert min(sizes) < 128 and max(sizes) > 4000


def test_large_batch_kernels_match_small(golden_complex):
    """Batches of at least 200,000 atoms take the per-complex cell-list
    kernels, the thread-per-atom attention sweep (column runs, early stop,
    cyclic coincidence test) and the region sweep over the attention pass's
    list of non-blockers; smaller batches take the wavefront-per-atom forms.  The same complexes as one large batch and in
    chunks of 20: identical results."""
    cs = _batch(golden_complex, [(1800 + 40 * k, 200 + 53 * (k % 9)) for k in range(80)], 1200)
    n = sum(len(c[1]) + len(c[3]) for c in cs)
    assert n >= 200000
    say("large batch (%d complexes, %d atoms) vs chunks" % (len(cs), n))
    big = _reference(cs)
    parts = [_reference(cs[k:k + 20]) for k in range(0, len(cs), 20)]
    assert max(sum(len(c[1]) + len(c[3]) for c in cs[k:k + 20]) for k in range(0, len(cs), 20)) < 200000
    merged = {key: np.concatenate([p[key] for p in parts]) for key in ("target", "binder", "per_atom", "sc")}
    assert _same(big, merged)
    assert (big["target"] > 0).sum() > 60


def test_trim_lanes_agree(golden_complex, monkeypatch):
    """The SC trim reads each buried dot's candidate cells with 1, 4 or 32
    lanes (chosen by batch size, overridable with CMS_TRIM_LANES): the same
    trimmed dots and SC records on a mid-size batch."""
    from cms_cuda.batch import BatchCalculator
    cs = _batch(golden_complex, [(1800, 400), (900, 120), (3000, 900), (600, 60)], 1400)
    outs = []
    for lanes in ("1", "4", "32"):
        monkeypatch.setenv("CMS_TRIM_LANES", lanes)
        say("SC trim with %s lanes per buried dot" % lanes)
        bc = BatchCalculator()
        bc.load(cs)
        outs.append(bc.run(0))
        bc.close()
    a = outs[0]
    assert (a["sc"][:, 13] + a["sc"][:, 21]).sum() > 500
    for b in outs[1:]:
        assert _same(a, b)


def test_descriptors_optional(golden_complex):
    """Generation with and without dot descriptors (cms_set_keep_descriptors):
    identical results; the entry points that read descriptors refuse a
    surface generated without them and accept one generated with them."""
    import ctypes
    import torch
    from cms_cuda.batch import BatchCalculator
    from cms_cuda import _lib
    cs = _batch(golden_complex, [(700, 200), (1200, 90)], 1500)
    outs = []
    for keep in (True, False):
        say("dot descriptors kept: %s" % keep)
        bc = BatchCalculator(keep_descriptors=keep)
        bc.load(cs)
        outs.append(bc.run(_lib.PREC_FP32_RESCORE))
        n = ctypes.c_int64()
        bc.ctx("cms_dot_count", 1, 0, ctypes.byref(n))
        assert n.value > 100
        d1 = torch.empty((n.value, 4), dtype=torch.int32, device="cuda")
        d2 = torch.empty((n.value, 4), dtype=torch.int32, device="cuda")
        P = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
        bad, dev = ctypes.c_int64(), ctypes.c_double()
        if keep:
            bc.ctx("cms_copy_dot_desc", 1, 0, P(d1), P(d2))
            bc.ctx("cms_check_dots", ctypes.byref(bad), ctypes.byref(dev))
            assert bad.value == 0
            assert (d1[:, 0].cpu() & 15).min() >= 1
        else:
            with pytest.raises(_lib.CmsError, match="without dot descriptors"):
                bc.ctx("cms_copy_dot_desc", 1, 0, P(d1), P(d2))
            with pytest.raises(_lib.CmsError, match="without dot descriptors"):
                bc.ctx("cms_check_dots", ctypes.byref(bad), ctypes.byref(dev))
        bc.close()
    assert _same(outs[0], outs[1])
