"""Differentiable CMS: dCMS/dxyz behind torch.autograd.Function.

The forward pass runs the HIP pipeline (surface generation, nearest-dot
search, CMS of both sides of every complex); the backward pass
(cms_backward_cms) differentiates the CMS at frozen topology -- dot
existence, burial, nearest partners and subdivision counts fixed by the
forward pass, dot positions and areas differentiated through the atoms
that generate each dot, recomputed on the GPU from compact per-dot
descriptors.
"""

import ctypes

import numpy as np

from . import _lib
from ._calc import Settings, _ptr


def _torch():
    import torch
    return torch


class _ContactMS:
    """Holds one library context per forward call (the backward pass reads
    its surface and search state)."""

    def __init__(self, xyz, radii, n_binder, n_target, settings, precision):
        torch = _torch()
        self.dev = xyz.device
        self.stream = torch.cuda.current_stream(self.dev)
        self.ctx = _lib.Context(self.dev.index, self.stream.cuda_stream)
        self.ctx("cms_set_keep_descriptors", 1)      # what the backward pass re-evaluates dots from
        self.B = len(n_binder)
        self.ntot = int(xyz.shape[0])
        self.nb = (ctypes.c_int32 * self.B)(*[int(v) for v in n_binder])
        self.nt = (ctypes.c_int32 * self.B)(*[int(v) for v in n_target])
        self.st = (settings or Settings()).to_c()
        self.xyz = xyz.detach().contiguous()
        self.radii = radii.detach().contiguous()
        self.ctx("cms_load", self.B, self.nb, self.nt, _ptr(self.xyz), _ptr(self.radii), None,
                 ctypes.byref(self.st))
        self.ctx("cms_generate", 0, 0)
        self.per_atom = torch.empty(max(self.ntot, 1), dtype=torch.float64, device=self.dev)
        self.totals = torch.empty(2 * self.B, dtype=torch.float64, device=self.dev)
        self.ctx("cms_contact_ms", int(precision), _ptr(self.per_atom), _ptr(self.totals))

    def backward(self, g_totals, g_per_atom):
        torch = _torch()
        grad = torch.empty((self.ntot, 3), dtype=torch.float64, device=self.dev)
        gt = g_totals.detach().contiguous() if g_totals is not None else None
        ga = g_per_atom.detach().contiguous() if g_per_atom is not None else None
        self.ctx("cms_backward_cms", _ptr(gt), _ptr(ga), _ptr(grad))
        return grad

    def frozen_totals(self, xyz_eval):
        torch = _torch()
        out = torch.empty(2 * self.B, dtype=torch.float64, device=self.dev)
        self.ctx("cms_frozen_cms", _ptr(xyz_eval.detach().contiguous()), _ptr(out))
        return out

    def close(self):
        if self.ctx is not None:
            self.ctx.close()
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _function():
    torch = _torch()

    class ContactMSFunction(torch.autograd.Function):
        @staticmethod
        def forward(fctx, xyz, radii, n_binder, n_target, settings, precision):
            state = _ContactMS(xyz, radii, n_binder, n_target, settings, precision)
            fctx.state = state
            B = state.B
            return state.totals.view(B, 2).clone(), state.per_atom[:state.ntot].clone()

        @staticmethod
        def backward(fctx, g_totals, g_per_atom):
            # the state (and its library context) lives as long as the graph
            # node, so a graph kept with retain_graph=True can be
            # differentiated again
            state = fctx.state
            grad = state.backward(g_totals.reshape(-1) if g_totals is not None else None,
                                  g_per_atom)
            return grad, None, None, None, None, None

    return ContactMSFunction


_FN = None


def contact_ms(xyz, radii, n_binder, n_target, settings=None, precision=_lib.PREC_FP64):
    """Differentiable CMS of a batch.

    xyz: (ntot, 3) float64 CUDA tensor (complexes back to back, binder atoms
    first), radii: (ntot,) float64 CUDA tensor, n_binder / n_target: per
    complex atom counts.  Returns (totals (B, 2): [binder side, target
    side], per_atom (ntot,)); both are differentiable with respect to xyz.
    """
    global _FN
    if _FN is None:
        _FN = _function()
    return _FN.apply(xyz, radii, list(n_binder), list(n_target), settings, precision)


def contact_ms_pair(binder_xyz, binder_radii, target_xyz, target_radii, settings=None):
    """Target-side CMS of one complex as a differentiable scalar (the value
    calculate_contact_ms returns), for torch tensors of coordinates."""
    torch = _torch()
    xyz = torch.cat([binder_xyz, target_xyz], 0)
    radii = torch.cat([binder_radii, target_radii], 0)
    totals, _ = contact_ms(xyz, radii, [binder_xyz.shape[0]], [target_xyz.shape[0]], settings)
    return totals[0, 1]
