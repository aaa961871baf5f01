import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "cms-cuda_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: CPU test of several seconds")


@pytest.fixture(scope="session")
def golden_complex():
    """The reference's known-answer complex (test/sc_example.pdb): chain A is
    the binder (molecule 0), chain B the target (molecule 1)."""
    from cms_cuda.structure import read_pdb_atoms, split_chains
    from cms_cuda import get_radii_from_names
    atoms = read_pdb_atoms(os.path.join(ROOT, "tests", "golden", "sc_example.pdb"))
    (bx, brn, ban), (tx, trn, tan) = split_chains(atoms, "A")
    return dict(binder_xyz=bx, binder_radii=get_radii_from_names(brn, ban),
                target_xyz=tx, target_radii=get_radii_from_names(trn, tan))


@pytest.fixture(scope="session")
def golden_oracle(golden_complex):
    from oracle.cms_oracle import OracleCalculator
    c = golden_complex
    o = OracleCalculator()
    o.add_binder_and_target(c["binder_xyz"], c["binder_radii"], c["target_xyz"], c["target_radii"])
    cms = o.contact_ms(True)
    cms_b = o.contact_ms(False)
    sc = o.shape_complementarity()
    return dict(calc=o, cms=cms, cms_binder=cms_b, sc=sc)
