"""Pins the oracle (oracle/cms_oracle.py) before it is trusted as the checker.

1. The reference's known-answer complex (test/sc_example.pdb) against the
   reference's golden run: CMS / SC / area / distance scalars and every
   recorded count.
2. Closed forms that the golden complex does not exercise: the grouped-data
   median, NumPy pairwise summation, a two-atom molecule's exact discrete
   surface area, far-apart molecules, an isolated atom.
"""

import math

import numpy as np
import pytest

from oracle import cms_oracle as O

# Scalars and counts of the reference's golden run on sc_example.pdb
# (test/golden.pkl, test/sc_golden.pkl; written by the reference's
# test/cms_test.py:294-298 and test/sc_test.py:302-307).
GOLDEN = dict(
    cms=275.5999647907913,
    sc=0.6297053428112819,
    area=529.8866053867879,
    distance=0.601465993751502,
    convex=11806, toroidal=13304, concave=10054,
    dots=(15158, 20006), buried=(6218, 6697),
    types=((5566, 5825, 3767), (6240, 7479, 6287)),
    probes=490, trimmed=(3866, 4118),
    atten5=281, atten1=111, access=258, radmax=1.95,
)


def rel(a, b):
    return abs(a - b) / max(1.0, abs(a), abs(b))


def test_golden_scalars(golden_oracle):
    cms, _ = golden_oracle["cms"]
    sc, area, dist = golden_oracle["sc"]
    assert rel(cms, GOLDEN["cms"]) <= 1e-12
    assert rel(sc, GOLDEN["sc"]) <= 1e-12
    assert rel(area, GOLDEN["area"]) <= 1e-12
    assert rel(dist, GOLDEN["distance"]) <= 1e-12


def test_golden_counts(golden_oracle):
    o = golden_oracle["calc"]
    assert o.counts == dict(convex=GOLDEN["convex"], toroidal=GOLDEN["toroidal"], concave=GOLDEN["concave"])
    assert tuple(len(d) for d in o.dots) == GOLDEN["dots"]
    assert tuple(int(d.buried.sum()) for d in o.dots) == GOLDEN["buried"]
    for m in (0, 1):
        assert tuple(np.bincount(o.dots[m].type, minlength=4)[1:]) == GOLDEN["types"][m]
    assert len(o.pr_h) == GOLDEN["probes"]
    assert tuple(golden_oracle["calc"].sc_results["n_trimmed"]) == GOLDEN["trimmed"]
    assert int((o.atten == 5).sum()) == GOLDEN["atten5"]
    assert int((o.atten == 1).sum()) == GOLDEN["atten1"]
    assert int(o.access.sum()) == GOLDEN["access"]
    assert o.radmax == GOLDEN["radmax"]


def test_golden_per_atom_sums(golden_oracle):
    cms, per_atom = golden_oracle["cms"]
    assert per_atom.shape == (242,)
    assert rel(per_atom.sum(), cms) <= 1e-12
    cmsb, per_b = golden_oracle["cms_binder"]
    assert per_b.shape == (150,)
    assert cmsb > 0


def test_binned_median_closed_forms():
    # three values in three bins: the middle bin crosses 50 %, interpolation
    # 1*0.02 + (50 - 33.33..)*0.02/33.33.. = 0.03
    assert abs(O.binned_median(np.array([0.01, 0.03, 0.05]), 0.02) - 0.03) < 1e-15
    # all values in one bin b: median = b*bw + 50*bw/100
    v = np.full(7, 0.4501)
    assert abs(O.binned_median(v, 0.02) - (22 * 0.02 + 0.01)) < 1e-15
    # negative values floor towards -inf
    v = np.array([-0.011, -0.011, 0.5])
    # bins -1 (2 values), 25 (1 value); first bin reaches 66.7 %
    expect = -1 * 0.02 + (50.0 - 0.0) * 0.02 / (2 * 100.0 / 3)
    assert abs(O.binned_median(v, 0.02) - expect) < 1e-15
    assert O.binned_median(np.zeros(0), 0.02) == 0.0


def pairwise_ref(a):
    """Recursive statement of NumPy's pairwise summation."""
    n = len(a)
    if n < 8:
        s = 0.0
        for x in a:
            s += x
        return s
    if n <= 128:
        r = list(a[:8])
        i = 8
        while i < n - (n % 8):
            for j in range(8):
                r[j] += a[i + j]
            i += 8
        s = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]))
        while i < n:
            s += a[i]
            i += 1
        return s
    n2 = n // 2
    n2 -= n2 % 8
    return pairwise_ref(a[:n2]) + pairwise_ref(a[n2:])


def test_numpy_pairwise_sum_order():
    rng = np.random.default_rng(5)
    for n in (0, 1, 7, 8, 9, 127, 128, 129, 1000, 4099):
        v = rng.standard_normal(n) * 10.0 ** rng.integers(-6, 6, size=n)
        assert O.pairwise_sum(v) == pairwise_ref(list(v))


def two_atom_area(d, r0, r1, rp=1.7, density=15.0):
    """Exact discrete max-CMS of a two-atom molecule.  Each atom's convex cap
    runs from the pole facing away from the other atom to the probe-contact
    direction at angle phi = pi - alpha, cos(alpha) = (er_i^2 + d^2 -
    er_j^2) / (2 er_i d).  Latitudes sit at a_t = delta (t + 1/2) <= phi,
    delta = 1/(sqrt(density) r); every circle's points sum to 2 pi rad_t *
    (r phi / L), rad_t = r sin(a_t)."""
    total = 0.0
    for ri, rj in ((r0, r1), (r1, r0)):
        eri, erj = ri + rp, rj + rp
        alpha = math.acos((eri * eri + d * d - erj * erj) / (2 * eri * d))
        phi = math.pi - alpha
        delta = 1.0 / (math.sqrt(density) * ri)
        lat = [delta * (t + 0.5) for t in range(100) if delta * (t + 0.5) <= phi]
        L = len(lat)
        cs = ri * phi / L
        for a in lat:
            rad = ri * math.sin(a)
            cd = 1.0 / (math.sqrt(density) * rad)
            P = len([k for k in range(100) if cd * (k + 0.5) <= 2 * math.pi])
            if P:
                total += 2 * math.pi * rad * cs
    return total


@pytest.mark.parametrize("d,r0,r1", [(3.0, 1.8, 1.6), (2.2, 1.9, 1.9), (4.5, 1.65, 1.95)])
def test_two_atom_molecule_area(d, r0, r1):
    xyz = np.array([[0.0, 0.0, 0.0], [d, 0.0, 0.0]])
    rot = np.linalg.qr(np.random.default_rng(int(d * 10)).standard_normal((3, 3)))[0]
    xyz = xyz @ rot.T + np.array([1.0, -2.0, 0.5])
    cms, per_atom, calc = O.calculate_maximum_possible_contact_ms(xyz, np.array([r0, r1]))
    expect = two_atom_area(d, r0, r1)
    assert abs(cms - expect) <= 1e-9 * expect
    assert calc.counts["toroidal"] == 0 and calc.counts["concave"] == 0
    assert len(calc.pr_h) == 0


def test_far_apart_molecules_give_zero():
    rng = np.random.default_rng(2)
    a = rng.uniform(-3, 3, size=(12, 3))
    a = a[np.argsort(a[:, 0])]
    b = a + np.array([40.0, 0, 0])
    r = np.full(12, 1.8)
    o = O.OracleCalculator()
    o.add_binder_and_target(a, r, b, r)
    cms, per_atom = o.contact_ms(True)
    assert cms == 0.0 and not per_atom.any()
    assert sum(len(d) for d in o.dots) == 0
    sc, area, dist = o.shape_complementarity()
    assert (sc, area, dist) == (0.0, 0.0, 0.0)


def test_isolated_atom_has_no_dots():
    cms, per_atom, calc = O.calculate_maximum_possible_contact_ms(np.zeros((1, 3)), np.array([1.8]))
    assert cms == 0.0 and per_atom.shape == (1,)
    assert calc.access[0] == 1


def test_coincident_atoms_raise():
    xyz = np.array([[0.0, 0, 0], [1.5, 0, 0], [1.5, 0.005, 0]])
    with pytest.raises(RuntimeError, match="Coincident atoms: 1:: == 2::"):
        O.calculate_maximum_possible_contact_ms(xyz, np.full(3, 1.8))
