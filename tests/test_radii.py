"""get_radii_from_names: table rows and wildcard rules (_core.py:87-112,
172-269, 719-758)."""

import numpy as np

from cms_cuda import get_radii_from_names
from cms_cuda._radii import name_matches, RADIUS_TABLE


def test_table_size_and_order():
    assert len(RADIUS_TABLE) == 95
    assert RADIUS_TABLE[0] == ("ALA", "CB", 1.95)
    assert RADIUS_TABLE[-1] == ("*", "F*", 1.50)


def test_known_values():
    res = ["ALA", "GLY", "LYS", "LYS", "ARG", "ARG", "TRP", "SER", "XYZ", "LIG", "LIG", "LIG", "LIG", "HOH", "MET"]
    atm = ["CB", "CA", "NZ", "CA", "NH1", "NE", "CZ3", "OG", "N", "C", "N", "O", "S", "O", "SD"]
    expect = [1.95, 1.90, 1.75, 1.85, 1.70, 1.65, 1.90, 1.70, 1.65, 1.80, 1.65, 1.60, 1.90, 1.70, 1.90]
    assert np.array_equal(get_radii_from_names(res, atm), np.array(expect))


def test_first_match_wins_and_unmatched_zero():
    # LEU CD1 matches "LEU CD*" before the generic "* C*"
    assert get_radii_from_names(["LEU"], ["CD1"])[0] == 1.95
    # generic carbon via "* C*" is 1.85, plain "* C" 1.80
    assert get_radii_from_names(["ALA"], ["C"])[0] == 1.80
    assert get_radii_from_names(["ALA"], ["CX"])[0] == 1.85
    # nothing matches an element-less name
    assert get_radii_from_names(["ALA"], ["ZZ"])[0] == 0.0
    # hydrogens
    assert get_radii_from_names(["ALA", "ALA"], ["H", "1HB"]).tolist() == [0.5, 0.5]


def test_wildcard_rules():
    assert name_matches("P", "P*")          # trailing * needs no extra character
    assert name_matches("NH2", "NH*")
    assert not name_matches("N", "NH*")
    assert name_matches("LG1", "LG*")
    assert not name_matches("LIG", "LG*")
    assert name_matches("CA ", "CA")        # query space matches end of pattern
    assert not name_matches("CAX", "CA")
