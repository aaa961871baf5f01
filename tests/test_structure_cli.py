"""The PDB / mmCIF loader's selection rules (those of the reference's
examples/calculate_cms_sc.py:43-97 through Bio.PDB) and the command line's
helpers, on the CPU."""

import os

import numpy as np
import pytest

from cms_cuda import cli
from cms_cuda.structure import load_atoms

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


def pdb_line(rec, serial, name, alt, res, chain, seq, x, y, z, occ, element):
    nm = name if len(name) == 4 else " " + name.ljust(3)
    return "%-6s%5d %4s%1s%3s %1s%4d    %8.3f%8.3f%8.3f%6.2f%6.2f          %2s\n" % (
        rec, serial, nm, alt, res, chain, seq, x, y, z, occ, 10.0, element)


def test_altlocs_waters_ions_hydrogens(tmp_path):
    lines = [
        pdb_line("ATOM", 1, "N", " ", "SER", "A", 1, 1.0, 2.0, 3.0, 1.0, "N"),
        pdb_line("ATOM", 2, "CA", "A", "SER", "A", 1, 2.0, 2.0, 3.0, 0.40, "C"),
        pdb_line("ATOM", 3, "CA", "B", "SER", "A", 1, 2.5, 2.0, 3.0, 0.60, "C"),   # higher occupancy wins
        pdb_line("ATOM", 4, "CB", "A", "SER", "A", 1, 3.0, 2.0, 3.0, 0.50, "C"),
        pdb_line("ATOM", 5, "CB", "B", "SER", "A", 1, 3.5, 2.0, 3.0, 0.50, "C"),   # a tie keeps the first
        pdb_line("ATOM", 6, "H", " ", "SER", "A", 1, 1.0, 1.0, 3.0, 1.0, "H"),
        pdb_line("ATOM", 7, "1HB", " ", "SER", "A", 1, 1.0, 1.0, 1.0, 1.0, ""),
        pdb_line("HETATM", 8, "O", " ", "HOH", "A", 101, 9.0, 9.0, 9.0, 1.0, "O"),
        pdb_line("HETATM", 9, "ZN", " ", "ZN", "A", 102, 8.0, 8.0, 8.0, 1.0, "ZN"),
        pdb_line("ATOM", 10, "N", " ", "GLY", "B", 1, 5.0, 5.0, 5.0, 1.0, "N"),
        "ENDMDL\n",
        pdb_line("ATOM", 11, "N", " ", "GLY", "C", 1, 7.0, 7.0, 7.0, 1.0, "N"),      # second model: dropped
    ]
    p = tmp_path / "t.pdb"
    p.write_text("MODEL        1\n" + "".join(lines))
    xyz, r, info = load_atoms(str(p), chains="A")
    assert [i[3] for i in info] == ["N", "CA", "CB"]
    assert np.array_equal(xyz[:, 0], np.float32([1.0, 2.5, 3.0]).astype(np.float64))
    xyz, r, info = load_atoms(str(p), chains="A", keep_water=True, keep_ions=True, keep_hydrogens=True)
    assert [i[3] for i in info] == ["N", "CA", "CB", "H", "1HB", "O", "ZN"]
    xyz, r, info = load_atoms(str(p))
    assert [i[0] for i in info] == ["A", "A", "A", "B"]
    assert load_atoms(str(p), resnames={"GLY"})[2][0][:3] == ("B", "1", "GLY")


def write_cif(pdb_path, cif_path):
    """A minimal mmCIF _atom_site loop with the ATOM / HETATM records of a PDB file."""
    out = ["data_test", "loop_"]
    cols = ["group_PDB", "id", "type_symbol", "label_atom_id", "label_alt_id", "label_comp_id", "label_asym_id",
            "label_seq_id", "pdbx_PDB_ins_code", "Cartn_x", "Cartn_y", "Cartn_z", "occupancy", "auth_seq_id",
            "auth_asym_id", "pdbx_PDB_model_num"]
    out += ["_atom_site." + c for c in cols]
    with open(pdb_path) as fh:
        for ln in fh:
            if not ln.startswith(("ATOM", "HETATM")):
                continue
            alt = ln[16].strip() or "."
            ic = ln[26].strip() or "?"
            out.append(" ".join([ln[:6].strip(), ln[6:11].strip(), ln[76:78].strip() or "?", '"%s"' % ln[12:16].strip(),
                                 alt, ln[17:20].strip(), ln[21], ln[22:26].strip(), ic, ln[30:38].strip(),
                                 ln[38:46].strip(), ln[46:54].strip(), ln[54:60].strip(), ln[22:26].strip(),
                                 ln[21], "1"]))
    out.append("#")
    with open(cif_path, "w") as fh:
        fh.write("\n".join(out) + "\n")


@pytest.mark.parametrize("name,bsel,tsel", [
    ("Nanobody", dict(chains="B"), dict(chains="A")),
    ("7jkv-smallmol", dict(chains="A", exclude_resnames={"V7G"}), dict(resnames={"V7G"})),
])
def test_mmcif_reads_like_pdb(tmp_path, name, bsel, tsel):
    pdb = os.path.join(GOLD, "example_%s.pdb" % name)
    cif = str(tmp_path / (name + ".cif"))
    write_cif(pdb, cif)
    for sel in (bsel, tsel):
        a = load_atoms(pdb, **sel)
        b = load_atoms(cif, **sel)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        assert [i[:4] for i in a[2]] == [i[:4] for i in b[2]]


def test_cli_helpers():
    info = [("A", "1", "GLY", "N", "N"), ("A", "1", "GLY", "CA", "C"), ("A", "2", "SER", "N", "N"),
            ("B", "1", "GLY", "N", "N")]
    assert cli.per_residue(info, [1.0, 2.0, 3.0, 4.0]) == [("A", "1", "GLY", 3.0), ("A", "2", "SER", 3.0),
                                                           ("B", "1", "GLY", 4.0)]
    assert cli.unmatched_atoms(info, np.array([1.6, 0.0, 1.7, 0.0])) == [("GLY", "CA"), ("GLY", "N")]
    args = cli.parser().parse_args(["x.pdb", "--binder-chains", "H,L", "--target-resnames", "BTN", "--max-cms"])
    s = cli.selection(args, "binder")
    assert s["chains"] == ["H", "L"] and s["resnames"] is None and s["exclude_resnames"] == set()
    t = cli.selection(args, "target")
    assert t["chains"] is None and t["resnames"] == {"BTN"}
    with pytest.raises(SystemExit):
        cli.main(["x.pdb", "--binder-chains", "B"])
