"""A failing complex fails alone (through the C ABI and BatchCalculator).

A batch holds two normal complexes, one with two coincident target atoms
(the reference raises for it, cms-cuda/_core.py:1285-1293) and one whose
neighbour row exceeds the row capacity (an atom inside a shell of 300
same-molecule atoms).  The failed complexes get their status word and zero
results in every readout (CMS, SC, max-CMS, frozen CMS, gradient); the
normal ones equal their single-complex runs bit for bit, eagerly and from
graph replays.  A batch in which every complex fails raises, and a
one-complex batch with coincident atoms raises the reference's message.
"""

import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def complexes():
    from cms_cuda import synth
    a = synth.make_complex(700, 150, 4101, gap=3.5)
    b = synth.make_complex(500, 90, 4102, gap=3.6)
    c = synth.make_complex(650, 120, 4103, gap=3.5)
    ok0 = (a["binder_xyz"], a["binder_radii"], a["target_xyz"], a["target_radii"])
    ok1 = (c["binder_xyz"], c["binder_radii"], c["target_xyz"], c["target_radii"])
    tx = b["target_xyz"].copy()
    tx[7] = tx[3]                                      # coincident target atoms 3 and 7
    coinc = (b["binder_xyz"], b["binder_radii"], tx, b["target_radii"])
    # an atom at the centre of a 5 A shell of 300 atoms (Fibonacci sphere):
    # its neighbour row holds 300 entries; a small binder 4 A off the shell
    k = np.arange(300) + 0.5
    phi = np.arccos(1 - 2 * k / 300)
    th = np.pi * (1 + 5 ** 0.5) * k
    shell = 5.0 * np.stack([np.cos(th) * np.sin(phi), np.sin(th) * np.sin(phi), np.cos(phi)], 1)
    t_xyz = np.concatenate([[[0.0, 0.0, 0.0]], shell])
    b_xyz = np.array([[9.0, 0.0, 0.0], [10.5, 0.8, 0.0], [9.5, -1.2, 0.7]])
    cap = (b_xyz, np.full(3, 1.8), t_xyz, np.full(301, 1.8))
    return [ok0, coinc, cap, ok1]


def solo(c, precision):
    from cms_cuda.batch import BatchCalculator
    bc = BatchCalculator()
    bc.load([c])
    out = bc.run(precision)
    bc.close()
    return out


def test_failed_complexes_are_isolated():
    from cms_cuda.batch import BatchCalculator
    from cms_cuda import _lib
    cs = complexes()
    off = np.concatenate([[0], np.cumsum([len(c[1]) + len(c[3]) for c in cs])])
    for prec in (_lib.PREC_FP64, _lib.PREC_FP32_RESCORE):
        say("batch of 4 with a coincident and an over-capacity complex, precision %d" % prec)
        bc = BatchCalculator()
        bc.load(cs)
        outs = [bc.run(prec, graph=True) for _ in range(3)]
        st = bc.status()
        assert bc.failed_count() == 2
        cap, rep, _, _ = bc.run_stats()
        say("  status %s, captures %d replays %d" % (st.tolist(), cap, rep))
        bc.close()
        assert st[0] == 0 and st[3] == 0
        assert st[1] & 1
        assert st[2] & 2
        for out in outs:
            for c in (1, 2):
                assert out["target"][c] == 0.0 and out["binder"][c] == 0.0
                assert not out["per_atom"][off[c]:off[c + 1]].any()
                assert not out["sc"][c].any()
            for c in (0, 3):
                ref = solo(cs[c], prec)
                assert ref["target"][0] == out["target"][c] and ref["binder"][0] == out["binder"][c]
                assert np.array_equal(ref["per_atom"], out["per_atom"][off[c]:off[c + 1]])
                assert np.array_equal(ref["sc"][0], out["sc"][c])
                assert out["target"][c] > 0
        assert rep >= 1


def test_failed_complexes_in_max_cms_and_backward():
    import torch
    from cms_cuda.batch import BatchCalculator
    from cms_cuda.autograd import _ContactMS
    from cms_cuda import _lib
    cs = complexes()
    off = np.concatenate([[0], np.cumsum([len(c[1]) + len(c[3]) for c in cs])])
    say("max-CMS of the batch")
    bc = BatchCalculator()
    bc.load(cs)
    mx = bc.max_contact_ms()
    st = bc.status()
    bc.close()
    assert st[0] == 0 and st[3] == 0 and st[1] & 1 and st[2] & 2
    for c in (1, 2):
        assert mx["total"][c] == 0.0 and not mx["per_atom"][off[c]:off[c + 1]].any()
    for c in (0, 3):
        one = BatchCalculator()
        one.load([cs[c]])
        m1 = one.max_contact_ms()
        one.close()
        assert m1["total"][0] == mx["total"][c] > 0
        assert np.array_equal(m1["per_atom"], mx["per_atom"][off[c]:off[c + 1]])
    say("backward of the batch")
    xyz = np.concatenate([np.concatenate([c[0], c[2]]) for c in cs])
    rad = np.concatenate([np.concatenate([c[1], c[3]]) for c in cs])
    s = _ContactMS(torch.from_numpy(xyz).cuda(), torch.from_numpy(rad).cuda(), [len(c[1]) for c in cs],
                   [len(c[3]) for c in cs], None, _lib.PREC_FP64)
    g = s.backward(torch.ones(8, dtype=torch.float64, device="cuda"), None).cpu().numpy()
    fz = s.frozen_totals(torch.from_numpy(xyz).cuda()).cpu().numpy().reshape(4, 2)
    tot = s.totals.cpu().numpy().reshape(4, 2)
    s.close()
    for c in (1, 2):
        assert not g[off[c]:off[c + 1]].any() and not fz[c].any() and not tot[c].any()
    for c in (0, 3):
        assert np.abs(g[off[c]:off[c + 1]]).sum() > 0
        assert np.allclose(fz[c], tot[c], rtol=1e-12, atol=0)


def test_all_failed_and_single_complex_errors():
    from cms_cuda.batch import BatchCalculator
    from cms_cuda import _lib
    cs = complexes()
    say("a batch in which every complex fails")
    bc = BatchCalculator()
    bc.load([cs[1], cs[2]])
    with pytest.raises(_lib.CmsError):
        bc.run(_lib.PREC_FP64)
    bc.close()
    say("one complex with coincident atoms")
    bc = BatchCalculator()
    bc.load([cs[1]])
    with pytest.raises(RuntimeError, match=r"^Coincident atoms: 93:: == 97::$"):
        bc.run(_lib.PREC_FP64)
    bc.close()


def test_coincident_pair_order_in_both_sweeps():
    """Several coincident pairs: the message names np.argwhere's first pair
    (cms-cuda/_core.py:1285-1293) in the wavefront-per-atom and the
    thread-per-atom attention sweeps alike."""
    from cms_cuda.batch import BatchCalculator
    from cms_cuda import _lib, synth, tuning
    b = synth.make_complex(500, 90, 4102, gap=3.6)
    tx = b["target_xyz"].copy()
    tx[40] = tx[17]
    tx[9] = tx[30]
    tx[11] = tx[30]
    c = (b["binder_xyz"], b["binder_radii"], tx, b["target_radii"])
    # target pairs (9, 11), (9, 30), (11, 30), (17, 40); binder atoms first
    for sweep in (1, 2):
        say("coincident pairs, cell sweep %d" % sweep)
        bc = BatchCalculator()
        bc.load([c])
        tuning.apply(bc.ctx, dict(cell_sweep=sweep))
        with pytest.raises(RuntimeError, match=r"^Coincident atoms: 99:: == 101::$"):
            bc.run(_lib.PREC_FP64)
        bc.close()
