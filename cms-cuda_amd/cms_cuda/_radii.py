"""Atomic radius lookup for CMS / SC (host side).

The radius table is the Lawrence & Colman SC radius library used by the
reference (``cms-cuda/_core.py:172-269``); its rows, their order and the
wildcard rules are part of the contract, because the first matching row wins.

``get_radii_from_names`` follows ``_core.py:87-112``: for every (residue, atom)
name pair, walk the table in order and return the radius of the first row
whose residue pattern and atom pattern both match; 0.0 when nothing matches.
Pattern matching follows ``wildcard_match`` (``_core.py:719-758``): characters
must match one by one, ``*`` in the pattern matches the rest of the query,
a space in the query matches the end of the pattern, and at most ``len+1``
characters are compared.  Results are memoised per name pair, since real
inputs repeat a few hundred distinct pairs many thousands of times.
"""

import functools

import numpy as np

# (residue pattern, atom pattern, radius in Angstrom), in table order.
_TABLE_TEXT = """
ALA CB 1.95
ARG NH* 1.70
ARG CZ 1.80
ARG NE 1.65
ARG CD 1.90
ARG CG 1.90
ASN ND2 1.70
ASN OD1 1.60
ASN CG 1.80
ASP OD* 1.60
ASP CG 1.80
GLN NE2 1.70
GLN OE1 1.60
GLN CD 1.80
GLN CG 1.90
GLU OE* 1.60
GLU CD 1.80
GLU CG 1.90
GLY CA 1.90
HIS CD2 1.90
HIS NE2 1.65
HIS CE1 1.90
HIS ND1 1.65
HIS CG 1.80
HOH O* 1.70
ILE CD1 1.95
ILE CG1 1.90
ILE CB 1.85
ILE CG2 1.95
LEU CD* 1.95
LEU CG 1.85
LYS NZ 1.75
LYS CE 1.90
LYS CD 1.90
LYS CG 1.90
MET CE 1.95
MET CG 1.90
PHE CD* 1.90
PHE CE* 1.90
PHE CZ 1.90
PHE CG 1.80
PRO CD 1.90
PRO CG 1.90
SER OG 1.70
SUL S 1.90
SUL O* 1.65
THR CG2 1.95
THR OG1 1.70
THR CB 1.85
TRP CE2 1.80
TRP CE3 1.90
TRP CD1 1.90
TRP CD2 1.80
TRP CZ* 1.90
TRP CH2 1.90
TRP NE1 1.65
TRP CG 1.80
TYR OH 1.70
TYR CD* 1.90
TYR CE* 1.90
TYR CZ 1.80
TYR CG 1.80
VAL CG* 1.95
VAL CB 1.85
WAT O 1.70
WAT O* 1.70
LG* HARO 1.00
LG* AROC 2.00
LG* ARO 2.00
LG* NHIS 1.75
LG* NHI 1.75
LG* COO 1.55
LG* OCBB 1.55
LG* OCB 1.55
LG* OOC 1.55
LG* NLYS 1.75
* H 0.50
* H* 0.50
* 1H* 0.50
* 2H* 0.50
* 3H* 0.50
* C 1.80
* CA 1.85
* CB 1.90
* C* 1.85
* N 1.65
* N* 1.65
* O 1.60
* O* 1.60
* OXT 1.60
* OT* 1.60
* S* 1.90
* P 2.15
* P* 2.15
* F* 1.50
"""


def _parse_table(text):
    rows = []
    for line in text.splitlines():
        parts = line.split()
        if len(parts) != 3:
            continue
        rows.append((parts[0], parts[1], float(parts[2])))
    return tuple(rows)


RADIUS_TABLE = _parse_table(_TABLE_TEXT)


def name_matches(query, pattern):
    """True when ``query`` matches ``pattern`` under the SC radius-library rules.

    Up to ``len(query) + 1`` character positions are compared.  A position
    matches when the characters are equal, the pattern holds ``*``, or the
    query holds a space where the pattern has ended.  A ``*`` that is the
    last pattern character accepts whatever follows.  Exhausted strings read
    as NUL.  Mirrors ``_core.py:719-758`` including its trailing-``*`` fix
    (``_core.py:738-742``).
    """
    qn, pn = len(query), len(pattern)
    qi = pi = 0
    for _ in range(qn + 1):
        q = query[qi] if qi < qn else "\0"
        p = pattern[pi] if pi < pn else "\0"
        if not (q == p or p == "*" or (q == " " and p == "\0")):
            return False
        if p == "*" and pi + 1 >= pn:
            return True
        if q != "\0":
            qi += 1
        if p != "\0":
            pi += 1
    return True


@functools.lru_cache(maxsize=65536)
def radius_for(residue, atom):
    """Radius of the first table row matching (residue, atom); 0.0 if none."""
    for res_pat, atom_pat, radius in RADIUS_TABLE:
        if name_matches(residue, res_pat) and name_matches(atom, atom_pat):
            return radius
    return 0.0


def get_radii_from_names(res_names, atom_names):
    """Radii for parallel lists of stripped residue and atom names.

    Same contract as ``_core.get_radii_from_names`` (``_core.py:87-112``):
    returns a float64 array of length N, 0.0 for names with no table row.
    """
    res_names = list(res_names)
    atom_names = list(atom_names)
    out = np.zeros(len(res_names), dtype=np.float64)
    for i, (res, atom) in enumerate(zip(res_names, atom_names)):
        out[i] = radius_for(str(res), str(atom))
    return out
