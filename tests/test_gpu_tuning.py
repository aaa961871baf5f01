"""Launch shapes change speed only: a mixed batch gives identical results
under extreme choices of every cms_tuning field, out-of-range values are
refused, and BatchCalculator applies the table's choice for its size."""

import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_launch_shapes_give_identical_results(golden_complex):
    from cms_cuda import synth, _lib, tuning
    from cms_cuda.batch import BatchCalculator
    g = golden_complex
    cs = [(g["binder_xyz"], g["binder_radii"], g["target_xyz"], g["target_radii"])]
    for k, (nt, nb) in enumerate([(2400, 30), (1800, 600), (900, 150), (3100, 1200), (700, 55)]):
        c = synth.make_complex(nt, nb, 5200 + k, gap=3.6)
        cs.append((c["binder_xyz"], c["binder_radii"], c["target_xyz"], c["target_radii"]))
    combos = [tuning.default(),
              dict(trim_lanes=1, probe_count=1, probe_slots=0, cell_sweep=1, search_split=512),
              dict(trim_lanes=32, probe_count=2, probe_slots=1, cell_sweep=2, search_split=8192),
              dict(trim_lanes=4, probe_count=2, probe_slots=8, cell_sweep=1, search_split=1)]
    outs = []
    for ch in combos:
        say("launch shapes %s" % ch)
        bc = BatchCalculator(tuned=False)
        bc.load(cs)
        tuning.apply(bc.ctx, ch)
        o = bc.run(_lib.PREC_FP32_RESCORE, graph=True)
        o2 = bc.run(_lib.PREC_FP32_RESCORE, graph=True)
        o3 = bc.run(_lib.PREC_FP64)
        bc.close()
        for key in ("target", "binder", "per_atom", "sc"):
            assert np.array_equal(o[key], o2[key]) and np.array_equal(o[key], o3[key])
        outs.append(o)
    assert (outs[0]["sc"][:, 13] + outs[0]["sc"][:, 21]).sum() > 100
    for o in outs[1:]:
        for key in ("target", "binder", "per_atom", "sc"):
            assert np.array_equal(outs[0][key], o[key]), key


def test_tuning_values_checked_and_applied():
    import torch
    from cms_cuda import _lib, tuning, synth
    from cms_cuda.batch import BatchCalculator
    ctx = _lib.Context(0, torch.cuda.current_stream().cuda_stream)
    for bad in (dict(trim_lanes=3), dict(probe_count=5), dict(probe_slots=9), dict(cell_sweep=-1),
                dict(search_split=70000)):
        with pytest.raises(ValueError):
            tuning.apply(ctx, dict(tuning.default(), **bad))
    ctx.close()
    c = synth.config1(seed=1)
    bc = BatchCalculator()
    bc.load([(c["binder_xyz"], c["binder_radii"], c["target_xyz"], c["target_radii"])])
    got = tuning.current(bc.ctx)
    want = tuning.choice(1400)
    for f in tuning.FIELDS:
        if f == "probe_slots":
            assert got[f] == (want[f] or 8)
        else:
            assert got[f] == want[f]
    bc.close()
