"""The command line (python -m cms_cuda.cli) on the reference's example
structures: printed CMS / SC values equal the reference calculator's
vectors (tests/golden/ref_examples.npz) rounded as printed, the max-CMS of
the 4,167-atom target equals its vector, the CSV and per-residue tables are
written, and a structure with an empty selection gets an error row without
disturbing the others."""

import csv
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ref():
    with np.load(os.path.join(GOLD, "ref_examples.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_cli_on_examples(ref, tmp_path):
    from cms_cuda import cli
    names = ["linearpeptide", "Nanobody", "binder", "scFv"]
    paths = [os.path.join(GOLD, "example_%s.pdb" % n) for n in names]
    paths.append(os.path.join(GOLD, "sc_example.pdb"))           # chains A / B: a normal row
    empty = tmp_path / "empty.pdb"
    empty.write_text("END\n")
    paths.append(str(empty))
    out_csv = tmp_path / "rows.csv"
    per = tmp_path / "per"
    say("cli: %d structures, --max-cms" % len(paths))
    rows = cli.main(paths + ["--binder-chains", "B", "--target-chains", "A", "--max-cms", "--csv", str(out_csv),
                             "--per-residue", str(per)])
    assert len(rows) == len(paths)
    for n, row in zip(names, rows):
        g = lambda k: ref["%s_%s" % (n, k)]  # noqa: E731
        assert row["binder_atoms"] == len(g("binder_radii")) and row["target_atoms"] == 4167
        assert row["cms"] == round(float(g("cms_target")), 4)
        assert row["cms_binder"] == round(float(g("cms_binder")), 4)
        sc = g("sc")
        assert (row["sc"], row["sc_area"], row["sc_median_dist"]) == tuple(round(float(v), 4) for v in sc)
        # the fixture's max-CMS is that of the scFv file's chain A (the files
        # hold different poses of the 4,167-atom target)
        if n == "scFv":
            assert row["max_cms_target"] == round(float(ref["maxcms_total"]), 4)
            assert row["cms_fraction_target"] == round(float(g("cms_target")) / float(ref["maxcms_total"]), 4)
        assert row["max_cms_target"] > 15000 and row["max_cms_binder"] > row["cms_binder"]
        # per-residue tables sum to the per-atom values
        with open(per / ("example_%s_target_cms.csv" % n)) as fh:
            tab = list(csv.reader(fh))
        assert tab[0] == ["chain", "resnum", "resname", "cms"]
        assert abs(sum(float(r[3]) for r in tab[1:]) - float(g("cms_target"))) < 1e-4 * len(tab)
    assert "error" not in rows[4] and rows[4]["binder_atoms"] == 242
    assert "empty selection" in rows[5]["error"]
    with open(out_csv) as fh:
        got = list(csv.DictReader(fh))
    assert len(got) == len(paths) and float(got[0]["cms"]) == rows[0]["cms"]


def test_cli_small_molecule(ref):
    from cms_cuda import cli
    say("cli: protein-ligand example")
    rows = cli.main([os.path.join(GOLD, "example_7jkv-smallmol.pdb"), "--binder-chains", "A", "--binder-exclude",
                     "V7G", "--target-resnames", "V7G"])
    k = "7jkv_smallmol_"
    assert rows[0]["binder_atoms"] == 2329 and rows[0]["target_atoms"] == 41
    assert rows[0]["cms"] == round(float(ref[k + "cms_target"]), 4)
    assert rows[0]["sc"] == round(float(ref[k + "sc"][0]), 4)
