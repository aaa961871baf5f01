"""Public entry points (same names, arguments and returns as cms-cuda/_api.py).

This package is the MI355X implementation: every entry point runs on the
GPU through libcms_hip.so.  ``device="auto"``, ``"cuda"`` and ``"gpu"`` all
select it; ``device="cpu"`` is not served by this package (the NumPy CPU
calculator stays the reference's ``_core.py``) and raises, as does any
request when no ROCm GPU is usable -- there is no silent CPU fallback.
"""

import functools

from ._calc import MolecularSurfaceCalculator


@functools.lru_cache(maxsize=1)
def gpu_available():
    """True when the HIP library loads and a ROCm GPU is visible."""
    try:
        import torch
        if not torch.cuda.is_available() or torch.cuda.device_count() == 0:
            return False
        from . import _lib
        _lib.load()
        return True
    except Exception:
        return False


def _check_device(device):
    d = "auto" if device is None else str(device).lower()
    if d == "cpu":
        raise RuntimeError("device='cpu' is not provided by this MI355X build (it has no CPU calculator): "
                           "use device='auto' or 'cuda' here, or the NumPy calculator of the original "
                           "cms-cuda package for CPU runs")
    if d not in ("auto", "cuda", "gpu"):
        raise ValueError("device must be 'auto', 'cpu', 'cuda' or 'gpu', got %r" % (device,))
    if not gpu_available():
        raise RuntimeError("device=%r requested but no ROCm GPU is usable (torch.cuda.is_available() "
                           "is False or libcms_hip.so is not built)" % (device,))


def calculate_contact_ms(binder_xyz, binder_radii, target_xyz, target_radii, device="auto"):
    """Contact molecular surface of the binder against the target.

    Returns (cms, per_atom_target_cms (N1,), calc); call
    calc.calc_contact_molecular_surface(target_side=False) for the binder side.
    """
    _check_device(device)
    calc = MolecularSurfaceCalculator()
    calc.add_binder_and_target(binder_xyz, binder_radii, target_xyz, target_radii)
    cms, per_atom = calc.CalcLoaded()
    return cms, per_atom, calc


def calculate_shape_complementarity(binder_xyz, binder_radii, target_xyz, target_radii, device="auto"):
    """Lawrence & Colman shape complementarity: (sc, sc_int_area, median_dist, calc)."""
    _check_device(device)
    calc = MolecularSurfaceCalculator()
    calc.add_binder_and_target(binder_xyz, binder_radii, target_xyz, target_radii)
    sc, area, dist = calc.CalcLoadedSC()
    return sc, area, dist, calc


def calculate_maximum_possible_contact_ms(xyz, radii, device="auto"):
    """Maximum possible CMS (surface area) of one molecule: (cms, per_atom, calc)."""
    _check_device(device)
    calc = MolecularSurfaceCalculator()
    calc.AddMolecule(0, xyz, radii)
    cms, per_atom = calc.CalcLoadedMaxPossibleCMS()
    return cms, per_atom, calc
