"""Heavy-atom loader for PDB and mmCIF files, without a structure library.

``load_atoms`` takes the atoms of one side of an interface with the rules of
the reference's command-line loader (examples/calculate_cms_sc.py:43-97,
which goes through Bio.PDB): first model only; one alternate location per
atom, the one with the highest occupancy (the first of equal ones); waters
(HOH, WAT, ... or a water hetero flag) and single-atom hetero residues
(ions) dropped unless asked for; hydrogens (element H or D, or, without an
element, a name starting with H or D after leading digits) dropped; chain,
residue-name and excluded-residue-name selections.  Coordinates pass
through float32 as Bio.PDB stores them (``atom.coord``), then float64.

``read_pdb_atoms`` / ``split_chains`` keep the plain float64 reader used
for the known-answer complex, whose golden values were computed from
double-precision coordinates.
"""

import shlex

import numpy as np

WATER_NAMES = {"HOH", "WAT", "DOD", "H2O", "TIP", "TIP3", "SOL"}

# element symbols Bio.PDB accepts when it guesses an element from the atom name
_ELEMENTS = {
    "H", "D", "HE", "LI", "BE", "B", "C", "N", "O", "F", "NE", "NA", "MG", "AL", "SI", "P", "S", "CL", "AR",
    "K", "CA", "SC", "TI", "V", "CR", "MN", "FE", "CO", "NI", "CU", "ZN", "GA", "GE", "AS", "SE", "BR", "KR",
    "RB", "SR", "Y", "ZR", "NB", "MO", "TC", "RU", "RH", "PD", "AG", "CD", "IN", "SN", "SB", "TE", "I", "XE",
    "CS", "BA", "LA", "CE", "PR", "ND", "PM", "SM", "EU", "GD", "TB", "DY", "HO", "ER", "TM", "YB", "LU",
    "HF", "TA", "W", "RE", "OS", "IR", "PT", "AU", "HG", "TL", "PB", "BI", "PO", "AT", "RN", "U",
}


def _element(fullname, element):
    """Element from the element field, else guessed from the 4-character
    atom name as Bio.PDB does (the name, or its first non-digit character);
    'X' when nothing fits."""
    el = (element or "").strip().upper()
    if el in _ELEMENTS:
        return el
    name = fullname.strip()
    if fullname[:1].isalpha() and not fullname[2:].strip().isdigit():
        guess = name
    elif fullname[:1].isdigit():
        guess = fullname[1:2]
    else:
        guess = fullname[:1] if fullname[:1].strip() else name[:1]
    guess = guess.upper()
    if guess in _ELEMENTS:
        return guess
    if guess[:1] in _ELEMENTS:
        return guess[:1]
    return "X"


def _is_hydrogen(name, element):
    if element in ("H", "D"):
        return True
    return element in ("", "X") and name.lstrip("0123456789").startswith(("H", "D"))


class _Atom:
    __slots__ = ("name", "element", "xyz", "occ")

    def __init__(self, name, element, xyz, occ):
        self.name = name
        self.element = element
        self.xyz = xyz
        self.occ = occ


class _Residue:
    __slots__ = ("resname", "het", "resseq", "icode", "atoms")

    def __init__(self, resname, het, resseq, icode):
        self.resname = resname
        self.het = het
        self.resseq = resseq
        self.icode = icode
        self.atoms = {}     # atom name -> selected alternate location (insertion order kept)

    def add(self, atom):
        cur = self.atoms.get(atom.name)
        if cur is None or atom.occ > cur.occ:
            self.atoms[atom.name] = atom if cur is None else _replace(cur, atom)

    def __len__(self):
        return len(self.atoms)


def _replace(old, new):
    old.element, old.xyz, old.occ = new.element, new.xyz, new.occ
    return old


def _het_flag(record, resname):
    if record != "HETATM":
        return " "
    return "W" if resname in ("HOH", "WAT") else "H_" + resname


class _Model:
    """Chains in order of appearance, residues in order of appearance."""

    def __init__(self):
        self.chains = {}

    def add(self, chain, record, resname, resseq, icode, name, element, xyz, occ):
        residues = self.chains.setdefault(chain, {})
        het = _het_flag(record, resname)
        key = (het, resseq, icode)
        res = residues.get(key)
        if res is None:
            res = residues[key] = _Residue(resname, het, resseq, icode)
        res.add(_Atom(name, element, xyz, occ))


def _f32(x):
    return float(np.float32(float(x)))


def _parse_pdb(path):
    model = _Model()
    seen_model = False
    with open(path) as fh:
        for line in fh:
            rec = line[:6].strip()
            if rec == "MODEL":
                if seen_model:
                    break
                seen_model = True
                continue
            if rec == "ENDMDL":
                break
            if rec not in ("ATOM", "HETATM"):
                continue
            fullname = line[12:16]
            resname = line[17:20].strip()
            chain = line[21]
            resseq = int(line[22:26])
            icode = line[26].strip()
            try:
                occ = float(line[54:60])
            except ValueError:
                occ = 1.0
            element = _element(fullname, line[76:78] if len(line) >= 77 else "")
            xyz = (_f32(line[30:38]), _f32(line[38:46]), _f32(line[46:54]))
            model.add(chain, rec, resname, resseq, icode, fullname.strip(), element, xyz, occ)
    return model


def _cif_tokens(lines):
    for ln in lines:
        if ln.startswith("#"):
            continue
        if ln.startswith(";"):
            yield ln[1:].rstrip("\n")
            continue
        for t in shlex.split(ln, posix=True):
            yield t


def _parse_cif(path):
    """The _atom_site loop of an mmCIF file, with the columns Bio.PDB's
    MMCIFParser reads: label_atom_id, label_comp_id, auth_asym_id (chain),
    auth_seq_id, pdbx_PDB_ins_code, label_alt_id, occupancy, type_symbol,
    Cartn_x/y/z, group_PDB, pdbx_PDB_model_num."""
    with open(path) as fh:
        lines = fh.readlines()
    i = 0
    cols = []
    while i < len(lines):
        if lines[i].strip() == "loop_" and i + 1 < len(lines) and lines[i + 1].startswith("_atom_site."):
            i += 1
            while i < len(lines) and lines[i].startswith("_atom_site."):
                cols.append(lines[i].strip()[len("_atom_site."):])
                i += 1
            break
        i += 1
    if not cols:
        raise ValueError("%s: no _atom_site loop" % path)
    body = []
    while i < len(lines) and not lines[i].startswith(("loop_", "_", "data_")):
        body.append(lines[i])
        i += 1
    toks = list(_cif_tokens(body))
    n = len(cols)
    col = {c: k for k, c in enumerate(cols)}

    def get(row, *names, default=""):
        for nm in names:
            if nm in col:
                v = row[col[nm]]
                return "" if v in ("?", ".") else v
        return default

    model = _Model()
    first_model = None
    for s in range(0, len(toks) - n + 1, n):
        row = toks[s:s + n]
        mnum = get(row, "pdbx_PDB_model_num", default="1")
        if first_model is None:
            first_model = mnum
        if mnum != first_model:
            break
        rec = get(row, "group_PDB", default="ATOM")
        name = get(row, "label_atom_id", "auth_atom_id")
        resname = get(row, "label_comp_id", "auth_comp_id")
        chain = get(row, "auth_asym_id", "label_asym_id")
        seq = get(row, "auth_seq_id", "label_seq_id")
        try:
            resseq = int(seq)
        except ValueError:
            resseq = 0
        icode = get(row, "pdbx_PDB_ins_code")
        try:
            occ = float(get(row, "occupancy", default="1"))
        except ValueError:
            occ = 1.0
        element = _element(name.ljust(4)[:4], get(row, "type_symbol"))
        xyz = (_f32(get(row, "Cartn_x")), _f32(get(row, "Cartn_y")), _f32(get(row, "Cartn_z")))
        model.add(chain, rec, resname, resseq, icode, name, element, xyz, occ)
    return model


def parse_structure(path):
    """First model of a PDB / mmCIF file (chains -> residues -> atoms)."""
    is_cif = str(path).lower().endswith((".cif", ".mmcif"))
    return _parse_cif(path) if is_cif else _parse_pdb(path)


def load_atoms(path, chains=None, resnames=None, exclude_resnames=(), keep_water=False, keep_ions=False,
               keep_hydrogens=False, model=None):
    """Heavy atoms of one side of an interface (examples/calculate_cms_sc.py:43-97).

    Returns xyz (N, 3) float64, radii (N,) float64 from get_radii_from_names,
    and info: one (chain, residue number + insertion code, residue name, atom
    name, element) tuple per atom.
    """
    from ._radii import get_radii_from_names
    if isinstance(chains, str):
        chains = [c for c in chains.split(",") if c] if "," in chains else list(chains)
    if model is None:
        model = parse_structure(path)
    xyz, res_names, atom_names, info = [], [], [], []
    for chain, residues in model.chains.items():
        if chains is not None and chain not in chains:
            continue
        for res in residues.values():
            if resnames is not None and res.resname not in resnames:
                continue
            if res.resname in exclude_resnames:
                continue
            if not keep_water and (res.het == "W" or res.resname in WATER_NAMES):
                continue
            if not keep_ions and res.het.startswith("H_") and len(res) == 1:
                continue
            for atom in res.atoms.values():
                if _is_hydrogen(atom.name, atom.element) and not keep_hydrogens:
                    continue
                xyz.append(atom.xyz)
                res_names.append(res.resname)
                atom_names.append(atom.name)
                info.append((chain, "%d%s" % (res.resseq, res.icode), res.resname, atom.name, atom.element))
    xyz = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
    radii = get_radii_from_names(res_names, atom_names)
    return xyz, radii, info


# ------------------------------------------------------------ float64 reader

def read_pdb_atoms(path, keep_hetero=True):
    """Heavy atoms of the first model as dicts (chain, resname, resseq,
    name, xyz) in file order, coordinates parsed as float64; alternate
    locations other than blank / A / 1, waters and hydrogens dropped."""
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
            element = _element(name, line[76:78] if len(line) >= 77 else "")
            if _is_hydrogen(name.strip(), element) or res in WATER_NAMES:
                continue
            if line[16] not in (" ", "A", "1"):
                continue
            atoms.append(dict(chain=line[21], resname=res, resseq=line[22:27].strip(), name=name.strip(),
                              xyz=(float(line[30:38]), float(line[38:46]), float(line[46:54]))))
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
