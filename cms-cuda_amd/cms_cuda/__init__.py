"""cms_cuda on MI355X: contact molecular surface (CMS) and Lawrence & Colman
shape complementarity (SC) computed by hand-written HIP kernels for gfx950.

Drop-in for the cms-cuda package API (cms-cuda/__init__.py:15-44): the same
entry points, calculator class and return types.  ``cms_cuda.batch`` adds
the batched many-complex path.
"""

__version__ = "0.2.0"

from ._radii import get_radii_from_names
from ._calc import MolecularSurfaceCalculator
from ._api import (
    calculate_contact_ms,
    calculate_shape_complementarity,
    calculate_maximum_possible_contact_ms,
    gpu_available,
)

MolecularSurfaceCalculatorGPU = MolecularSurfaceCalculator

__all__ = [
    "calculate_contact_ms",
    "calculate_shape_complementarity",
    "get_radii_from_names",
    "calculate_maximum_possible_contact_ms",
    "partition_pose",
    "MolecularSurfaceCalculator",
    "MolecularSurfaceCalculatorGPU",
    "gpu_available",
]


def partition_pose(pose, jump_id=1):
    """Split a PyRosetta pose by jump into (xyz_0, radii_0, xyz_1, radii_1),
    dropping atoms without a radius (cms-cuda/_core.py:143-168, 622-676).
    Needs PyRosetta, imported lazily."""
    import numpy as np
    from pyrosetta import rosetta  # noqa: F401  (raises ImportError when absent)

    if jump_id > pose.num_jump():
        raise ValueError("Jump ID out of bounds")
    is_upstream = rosetta.utility.vector1_bool(pose.size())
    if jump_id > 0:
        is_upstream = pose.fold_tree().partition_by_jump(jump_id)
    else:
        for i in range(1, pose.size() + 1):
            is_upstream[i] = True
    names = [[], []]
    atoms = [[], []]
    xyzs = [[], []]
    for i in range(1, pose.size() + 1):
        res = pose.residue(i)
        if res.type().name() == "VRT" or res.type().is_metal():
            continue
        m = 0 if is_upstream[i] else 1
        for j in range(1, res.nheavyatoms() + 1):
            if res.is_virtual(j):
                continue
            v = res.xyz(j)
            names[m].append(res.name3())
            atoms[m].append(res.atom_name(j).strip())
            xyzs[m].append([v.x, v.y, v.z])
    out = []
    for m in (0, 1):
        xyz = np.array(xyzs[m], dtype=np.float64).reshape(-1, 3)
        r = get_radii_from_names(names[m], atoms[m])
        keep = r > 0
        out += [xyz[keep], r[keep]]
    return tuple(out)
