"""Backward pass of the CMS at frozen topology.

* every dot re-evaluated from its descriptor equals the forward dot bit for
  bit (the evaluator the backward differentiates is the forward pipeline);
* the frozen-topology CMS at the forward coordinates equals the CMS;
* the analytic gradient agrees with central finite differences of the
  frozen-topology CMS (FP64, h = 1e-4 A): to 1e-5 relative against the
  Richardson-extrapolated difference of steps h and h/2 (O(h^4)), and to
  1e-3 against the plain step-h difference, whose O(h^2) truncation reaches
  ~2e-4 on a few components with large third derivatives.
"""

import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def batch_inputs(cs):
    nb = [len(c[1]) for c in cs]
    nt = [len(c[3]) for c in cs]
    xyz = np.concatenate([np.concatenate([c[0], c[2]]) for c in cs])
    rad = np.concatenate([np.concatenate([c[1], c[3]]) for c in cs])
    return xyz, rad, nb, nt


@pytest.fixture(scope="module")
def complexes(golden_complex):
    from cms_cuda import synth
    g = golden_complex
    cs = [(g["binder_xyz"], g["binder_radii"], g["target_xyz"], g["target_radii"])]
    c = synth.make_complex(900, 150, 301, gap=3.5)
    cs.append((c["binder_xyz"], c["binder_radii"], c["target_xyz"], c["target_radii"]))
    return cs


def test_descriptor_reevaluation_is_exact(complexes):
    import torch
    from cms_cuda import _lib
    from cms_cuda.autograd import _ContactMS
    say("descriptor self-check")
    xyz, rad, nb, nt = batch_inputs(complexes)
    st = _ContactMS(torch.from_numpy(xyz).cuda(), torch.from_numpy(rad).cuda(), nb, nt, None, _lib.PREC_FP64)
    bad = ctypes.c_int64()
    dev = ctypes.c_double()
    st.ctx("cms_check_dots", ctypes.byref(bad), ctypes.byref(dev))
    assert bad.value == 0, (bad.value, dev.value)
    fz = st.frozen_totals(torch.from_numpy(xyz).cuda()).cpu().numpy()
    tot = st.totals.cpu().numpy()
    assert np.allclose(fz, tot, rtol=1e-12, atol=1e-12)
    st.close()


@pytest.mark.parametrize("side", [1, 0])
def test_gradient_matches_finite_differences(complexes, side):
    import torch
    from cms_cuda import _lib
    from cms_cuda.autograd import _ContactMS
    say("FD check side %d" % side)
    xyz, rad, nb, nt = batch_inputs(complexes)
    B = len(nb)
    dx = torch.from_numpy(xyz).cuda()
    st = _ContactMS(dx, torch.from_numpy(rad).cuda(), nb, nt, None, _lib.PREC_FP64)
    g_tot = torch.zeros(2 * B, dtype=torch.float64, device="cuda")
    g_tot[side::2] = 1.0
    grad = st.backward(g_tot, None).cpu().numpy()
    rng = np.random.default_rng(17 + side)
    offs = np.concatenate([[0], np.cumsum(np.array(nb) + np.array(nt))])
    h = 1e-4
    worst = 0.0
    checked = 0
    for c in range(B):
        lo, hi = offs[c], offs[c + 1]
        active = lo + np.nonzero(np.abs(grad[lo:hi]).sum(1) > 0)[0]
        others = rng.integers(lo, hi, size=16)
        pick = np.unique(np.concatenate([rng.choice(active, size=min(48, len(active)), replace=False), others]))
        for a in pick:
            for k in range(3):
                xp = xyz.copy(); xp[a, k] += h
                xm = xyz.copy(); xm[a, k] -= h
                fds = []
                for hh in (h, h / 2):
                    xp = xyz.copy(); xp[a, k] += hh
                    xm = xyz.copy(); xm[a, k] -= hh
                    fp = st.frozen_totals(torch.from_numpy(xp).cuda()).cpu().numpy()[2 * c + side]
                    fm = st.frozen_totals(torch.from_numpy(xm).cuda()).cpu().numpy()[2 * c + side]
                    fds.append((fp - fm) / (2 * hh))
                fd = fds[0]
                rich = (4 * fds[1] - fds[0]) / 3
                g = grad[a, k]
                scale = max(abs(g), abs(rich), 1e-3)
                err = abs(g - rich) / scale
                worst = max(worst, err)
                checked += 1
                assert err <= 1e-5, (c, a, k, g, rich, fd)
                assert abs(g - fd) / scale <= 1e-3, (c, a, k, g, fd)
    say("checked %d components, worst rel err %.2e" % (checked, worst))
    st.close()


def test_autograd_function(complexes):
    import torch
    from cms_cuda.autograd import contact_ms_pair
    say("autograd function")
    bx, br, tx, tr = complexes[0]
    b = torch.from_numpy(bx).cuda().requires_grad_(True)
    t = torch.from_numpy(tx).cuda().requires_grad_(True)
    cms = contact_ms_pair(b, torch.from_numpy(br).cuda(), t, torch.from_numpy(tr).cuda())
    assert abs(cms.item() - 275.5999647907913) < 1e-9 * 275.6
    cms.backward()
    assert b.grad is not None and t.grad is not None
    assert torch.isfinite(b.grad).all() and torch.isfinite(t.grad).all()
    assert b.grad.abs().sum() > 0 and t.grad.abs().sum() > 0


@pytest.mark.parametrize("which", [0, 1])
def test_gradient_matches_torch_fp64_reference(complexes, which):
    """The HIP backward (hand-written adjoint + dual-number VJP) against
    torch autograd through an independent FP64 torch re-implementation of
    the frozen-topology CMS (tests/torch_frozen.py): dot positions and areas
    from the exported descriptors, partners from the forward dots."""
    import torch
    from scipy.spatial import cKDTree
    from cms_cuda import _lib
    from cms_cuda.autograd import _ContactMS
    import torch_frozen as tf
    say("torch FP64 reference gradient, complex %d" % which)
    c = complexes[which]
    xyz, rad, nb, nt = batch_inputs([c])
    st = _ContactMS(torch.from_numpy(xyz).cuda(), torch.from_numpy(rad).cuda(), nb, nt, None, _lib.PREC_FP64)
    per_mol = []
    for mol in (0, 1):
        n = ctypes.c_int64()
        st.ctx("cms_dot_count", 0, mol, ctypes.byref(n))
        n = n.value
        coor = torch.empty((max(n, 1), 3), dtype=torch.float64, device="cuda")
        area = torch.empty(max(n, 1), dtype=torch.float64, device="cuda")
        bur = torch.empty(max(n, 1), dtype=torch.int8, device="cuda")
        d1 = torch.empty((max(n, 1), 4), dtype=torch.int32, device="cuda")
        d2 = torch.empty((max(n, 1), 4), dtype=torch.int32, device="cuda")
        P = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
        st.ctx("cms_copy_dots", 0, mol, P(coor), None, P(area), P(bur), None, None)
        st.ctx("cms_copy_dot_desc", 0, mol, P(d1), P(d2))
        b = bur[:n].cpu().numpy().astype(bool)
        per_mol.append(dict(coor=coor[:n].cpu().numpy()[b], area=area[:n].cpu().numpy()[b],
                            d1=d1[:n].cpu().long()[torch.from_numpy(b)], d2=d2[:n].cpu().long()[torch.from_numpy(b)]))
    assert len(per_mol[0]["area"]) > 100 and len(per_mol[1]["area"]) > 100
    X = torch.from_numpy(xyz.copy()).requires_grad_(True)
    R = torch.from_numpy(rad.copy())
    D = torch.full((len(rad),), 15.0, dtype=torch.float64)
    # the torch evaluator reproduces the forward dots
    for m in (0, 1):
        Pq, Aq = tf.dots(X.detach(), R, D, 1.7, per_mol[m]["d1"], per_mol[m]["d2"])
        assert np.max(np.abs(Pq.numpy() - per_mol[m]["coor"])) < 1e-11
        assert np.max(np.abs(Aq.numpy() - per_mol[m]["area"])) < 1e-11
    # target side: query = buried dots of molecule 1, partners among molecule 0's
    q, r = per_mol[1], per_mol[0]
    partner = torch.from_numpy(cKDTree(r["coor"]).query(q["coor"])[1])
    cms = tf.frozen_cms(X, R, D, 1.7, 0.5, (q["d1"], q["d2"]), (r["d1"], r["d2"]), partner)
    tot = st.totals.cpu().numpy()
    assert abs(cms.item() - tot[1]) <= 1e-10 * max(1.0, abs(tot[1]))
    (g_ref,) = torch.autograd.grad(cms, X)
    g_tot = torch.tensor([0.0, 1.0], dtype=torch.float64, device="cuda")
    g_hip = st.backward(g_tot, None).cpu().numpy()
    g_ref = g_ref.numpy()
    scale = np.max(np.abs(g_ref))
    err = np.max(np.abs(g_hip - g_ref)) / scale
    say("  max |grad| %.3e, max abs diff / max |grad| %.2e" % (scale, err))
    assert scale > 0 and err < 1e-9
    st.close()


def test_autograd_retain_graph(complexes):
    """A graph kept with retain_graph=True is differentiated twice with the
    same result (to rounding: per-atom gradients are accumulated with FP64
    atomics, whose order varies between calls)."""
    import torch
    from cms_cuda.autograd import contact_ms_pair
    bx, br, tx, tr = complexes[1]
    b = torch.from_numpy(bx).cuda().requires_grad_(True)
    t = torch.from_numpy(tx).cuda().requires_grad_(True)
    cms = contact_ms_pair(b, torch.from_numpy(br).cuda(), t, torch.from_numpy(tr).cuda())
    (g1,) = torch.autograd.grad(cms, b, retain_graph=True)
    (g2,) = torch.autograd.grad(cms, b)
    assert torch.allclose(g1, g2, rtol=1e-12, atol=1e-15 * float(g1.abs().max())) and g1.abs().sum() > 0
