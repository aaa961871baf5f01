"""Batched CMS / SC: many complexes of uneven size in one pass.

All complexes of a batch go through every stage together (one launch per
stage for the whole batch, per-complex offsets on the device), which is what
makes small complexes cheap: a single complex leaves most of the GPU idle.
Results per complex are bit-identical to running it alone.

    calc = BatchCalculator()
    calc.load([(binder_xyz, binder_radii, target_xyz, target_radii), ...])
    cms = calc.contact_ms()          # dict: target, binder totals + per-atom
    sc = calc.shape_complementarity() # (B, 24) records (see include/cms_hip.h)

A complex with coincident atoms or a neighbour row over capacity fails
alone: ``calc.status()`` flags it, its results are zero, and the other
complexes are unaffected.  Only a batch in which every complex fails raises
(for one complex, the reference's ``RuntimeError``).
"""

import ctypes

import numpy as np

from . import _lib, tuning
from ._calc import Settings, _device, _ptr


class BatchCalculator:
    def __init__(self, settings=None, device=None, stream=None, tuned=True, keep_descriptors=False):
        """tuned: launch shapes from the tuning table for the loaded batch's
        size (cms_cuda.tuning); False keeps the library's built-in rules.
        keep_descriptors: write each dot's descriptors during generation
        (what cms_copy_dot_desc / the backward pass read; forward-only
        batches skip them)."""
        import torch
        self.settings = settings or Settings()
        self.dev = device or _device()
        self.stream = stream if stream is not None else torch.cuda.current_stream(self.dev)
        self.ctx = _lib.Context(self.dev.index, self.stream.cuda_stream)
        self.ctx("cms_set_keep_descriptors", 1 if keep_descriptors else 0)
        self.B = 0
        self.mode = None
        self.tuned = tuned

    def close(self):
        if self.ctx is not None:
            self.ctx.close()
            self.ctx = None

    # ------------------------------------------------------------ inputs
    def load(self, complexes):
        """complexes: iterable of (binder_xyz, binder_radii, target_xyz,
        target_radii).  Host arrays are copied to the GPU once."""
        import torch
        cs = list(complexes)
        if not cs:
            raise ValueError("empty batch")
        nb = np.array([len(np.asarray(c[1]).reshape(-1)) for c in cs], dtype=np.int32)
        nt = np.array([len(np.asarray(c[3]).reshape(-1)) for c in cs], dtype=np.int32)
        xyz = np.concatenate([np.concatenate([np.asarray(c[0], np.float64).reshape(-1, 3),
                                              np.asarray(c[2], np.float64).reshape(-1, 3)]) for c in cs])
        rad = np.concatenate([np.concatenate([np.asarray(c[1], np.float64).reshape(-1),
                                              np.asarray(c[3], np.float64).reshape(-1)]) for c in cs])
        self.d_xyz = torch.from_numpy(np.ascontiguousarray(xyz)).to(self.dev)
        self.d_rad = torch.from_numpy(np.ascontiguousarray(rad)).to(self.dev)
        self.load_device(nb, nt, self.d_xyz, self.d_rad)

    def load_device(self, n_binder, n_target, d_xyz, d_rad, d_dens=None):
        """Inputs already in HBM: d_xyz (ntot, 3) and d_rad (ntot,) float64
        torch tensors, complexes back to back, binder atoms first."""
        self.nb = np.asarray(n_binder, dtype=np.int32)
        self.nt = np.asarray(n_target, dtype=np.int32)
        self.B = len(self.nb)
        self.off = np.concatenate([[0], np.cumsum(self.nb + self.nt)]).astype(np.int64)
        self.ntot = int(self.off[-1])
        self._keep = (d_xyz, d_rad, d_dens)
        self._nb_c = (ctypes.c_int32 * self.B)(*self.nb.tolist())
        self._nt_c = (ctypes.c_int32 * self.B)(*self.nt.tolist())
        self._st = self.settings.to_c()
        tuning.apply(self.ctx, tuning.choice(self.ntot) if self.tuned else tuning.default())
        self.ctx("cms_load", self.B, self._nb_c, self._nt_c, _ptr(d_xyz), _ptr(d_rad), _ptr(d_dens),
                 ctypes.byref(self._st))
        self.mode = None

    # ------------------------------------------------------------ stages
    def _surface_call(self, name, *args):
        """A library call that generates the surface; a one-complex batch
        with coincident atoms raises the reference's RuntimeError
        (cms-cuda/_core.py:1285-1293)."""
        try:
            self.ctx(name, *args)
        except _lib.CmsError as e:
            if e.code == _lib.CMS_ERR_COINCIDENT and self.B == 1:
                i, j = ctypes.c_int32(), ctypes.c_int32()
                self.ctx("cms_coincident", 0, ctypes.byref(i), ctypes.byref(j))
                raise RuntimeError("Coincident atoms: %d:: == %d::" % (i.value, j.value)) from None
            raise

    def generate(self, all_atoms=False, full_state=False):
        self._surface_call("cms_generate", 1 if all_atoms else 0, 1 if full_state else 0)
        self.mode = bool(all_atoms)

    def status(self):
        """Per-complex status words of the last generation (0: ok; bit 1
        coincident atoms, 2 / 4 a neighbour / buried row over capacity, 8 a
        near-probe list over capacity).  Complexes with a non-zero word have
        zero results."""
        out = np.zeros(max(self.B, 1), dtype=np.int32)
        self.ctx("cms_status", ctypes.c_void_p(out.ctypes.data))
        return out[:self.B]

    def failed_count(self):
        n = ctypes.c_int32()
        self.ctx("cms_failed_count", ctypes.byref(n))
        return n.value

    def contact_ms(self, precision=_lib.PREC_FP64, out=None):
        """CMS of every complex, both sides.  Returns dict with numpy arrays
        target (B,), binder (B,) and per_atom (ntot,), unless out=(per_atom,
        totals) device tensors are given (then nothing is copied back)."""
        import torch
        if self.mode is None:
            self.generate(False)
        if out is not None:
            per_atom, totals = out
        else:
            per_atom = torch.empty(max(self.ntot, 1), dtype=torch.float64, device=self.dev)
            totals = torch.empty(2 * self.B, dtype=torch.float64, device=self.dev)
        self.ctx("cms_contact_ms", int(precision), _ptr(per_atom), _ptr(totals))
        if out is not None:
            return None
        t = totals.cpu().numpy().reshape(self.B, 2)
        return dict(target=t[:, 1].copy(), binder=t[:, 0].copy(), per_atom=per_atom[:self.ntot].cpu().numpy())

    def run(self, precision=_lib.PREC_FP64, sc=True, graph=False, out=None):
        """One step of the hot path on the loaded batch (cms_run): surface,
        CMS of both sides and, with sc=True, SC, with one host wait; with
        graph=True repeated steps on the same layout replay a captured HIP
        graph.  Returns dict(target, binder, per_atom, sc) as numpy arrays,
        or None when out=(per_atom, totals, sc_records) device tensors are
        given (sc_records may be None)."""
        import torch
        self.ctx("cms_set_graph", 1 if graph else 0)
        if out is not None:
            per_atom, totals, res = out
        else:
            per_atom = torch.empty(max(self.ntot, 1), dtype=torch.float64, device=self.dev)
            totals = torch.empty(2 * self.B, dtype=torch.float64, device=self.dev)
            res = torch.empty(self.B * _lib.SC_NRES, dtype=torch.float64, device=self.dev) if sc else None
        self._surface_call("cms_run", int(precision), _ptr(per_atom), _ptr(totals), _ptr(res))
        self.mode = False
        if out is not None:
            return None
        t = totals.cpu().numpy().reshape(self.B, 2)
        return dict(target=t[:, 1].copy(), binder=t[:, 0].copy(), per_atom=per_atom[:self.ntot].cpu().numpy(),
                    sc=res.cpu().numpy().reshape(self.B, _lib.SC_NRES) if res is not None else None)

    def run_stats(self):
        """(graphs captured, graph replays, exact runs, overflow reruns)."""
        out = (ctypes.c_int64 * 4)()
        self.ctx("cms_run_stats", out)
        return tuple(out)

    def per_atom_split(self, per_atom):
        """Split a batch per-atom array into [(binder (N0,), target (N1,))]."""
        out = []
        for c in range(self.B):
            a = per_atom[self.off[c]:self.off[c + 1]]
            out.append((a[:self.nb[c]], a[self.nb[c]:]))
        return out

    def shape_complementarity(self, precision=_lib.PREC_FP64, out=None):
        import torch
        if self.mode is None:
            self.generate(False)
        res = out if out is not None else torch.empty(self.B * _lib.SC_NRES, dtype=torch.float64, device=self.dev)
        self.ctx("cms_shape_complementarity", int(precision), _ptr(res))
        if out is not None:
            return None
        return res.cpu().numpy().reshape(self.B, _lib.SC_NRES)

    def max_contact_ms(self):
        import torch
        if self.mode is None:
            self.generate(True)
        per_atom = torch.empty(max(self.ntot, 1), dtype=torch.float64, device=self.dev)
        totals = torch.empty(self.B, dtype=torch.float64, device=self.dev)
        self.ctx("cms_max_contact_ms", _ptr(per_atom), _ptr(totals))
        return dict(total=totals.cpu().numpy(), per_atom=per_atom[:self.ntot].cpu().numpy())

    def counts(self):
        out = np.zeros(self.B * _lib.NCOUNTS, dtype=np.int64)
        self.ctx("cms_counts", ctypes.c_void_p(out.ctypes.data))
        return out.reshape(self.B, _lib.NCOUNTS)


def calculate_batch(complexes, sc=True, precision=_lib.PREC_FP64):
    """Convenience: CMS (both sides) and optionally SC of every complex."""
    calc = BatchCalculator()
    calc.load(complexes)
    out = calc.contact_ms(precision)
    if sc:
        out["sc"] = calc.shape_complementarity(precision)
    out["per_atom"] = calc.per_atom_split(out["per_atom"])
    calc.close()
    return out
