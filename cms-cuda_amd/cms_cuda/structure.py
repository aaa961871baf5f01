"""Minimal PDB heavy-atom reader.

Reads ATOM/HETATM records of the first model, skipping hydrogens (element H
or D, or names starting with H after digits are stripped), waters and common
ions, so a two-chain complex can be fed to ``calculate_contact_ms`` without a
structure library.  The record columns follow the PDB format (name 13-16,
residue 18-20, chain 22, x/y/z 31-54, element 77-78).
"""

import numpy as np

_WATERS = {"HOH", "WAT", "DOD", "H2O", "TIP", "TIP3", "SOL"}
_IONS = {"NA", "CL", "K", "MG", "CA", "ZN", "MN", "FE", "CU", "CO", "NI", "CD",
         "IOD", "BR", "SO4", "PO4"}


def _is_hydrogen(name, element):
    el = element.strip().upper()
    if el in ("H", "D"):
        return True
    if el:
        return False
    n = name.strip().lstrip("0123456789").upper()
    return n.startswith("H") or n.startswith("D")


def read_pdb_atoms(path, keep_hetero=True):
    """Return a list of dicts (chain, resname, resseq, name, xyz) for the
    heavy atoms of the first model, in file order."""
    atoms = []
    with open(path) as fh:
        for line in fh:
            rec = line[:6]
            if rec.startswith("ENDMDL"):
                break
            if not (rec.startswith("ATOM") or (keep_hetero and rec.startswith("HETATM"))):
                continue
            name = line[12:16]
            res = line[17:20].strip()
            element = line[76:78] if len(line) >= 78 else ""
            if _is_hydrogen(name, element):
                continue
            if res in _WATERS:
                continue
            if rec.startswith("HETATM") and res in _IONS:
                continue
            alt = line[16]
            if alt not in (" ", "A", "1"):
                continue
            atoms.append(dict(
                chain=line[21], resname=res, resseq=line[22:27].strip(),
                name=name.strip(),
                xyz=(float(line[30:38]), float(line[38:46]), float(line[46:54])),
            ))
    return atoms


def split_chains(atoms, binder_chains):
    """Split atoms into (binder, target) by chain id; returns for each side
    (xyz (N,3) float64, residue names, atom names)."""
    binder_chains = set(binder_chains)
    out = []
    for want in (True, False):
        sel = [a for a in atoms if (a["chain"] in binder_chains) == want]
        xyz = np.array([a["xyz"] for a in sel], dtype=np.float64).reshape(-1, 3)
        out.append((xyz, [a["resname"] for a in sel], [a["name"] for a in sel]))
    return out[0], out[1]
