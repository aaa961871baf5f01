"""Seeded synthetic protein-like complexes for tests and benchmarks.

Each molecule is a compact globule: heavy atoms placed by random sequential
addition (Poisson disk, minimum spacing 2.2 A) inside an ellipsoid whose
volume holds N atoms at the requested number density.  Atom names come from
standard amino-acid heavy-atom templates (residue types drawn uniformly from
the 20 standard amino acids, the last residue truncated at N atoms), or from
C/N/O/S for small-molecule-like ligands (residue name LIG), so radii come from
``get_radii_from_names``.

A complex is a target globule at the origin and a binder globule that is
randomly rotated (uniform on SO(3)) and moved along a random unit direction
until the minimum distance between the two ellipsoid surfaces equals ``gap``
(bisection on the offset; the surface distance is the minimum over a dense
sample of binder-surface points of the exact point-to-ellipsoid distance).
"""

import numpy as np
from scipy.spatial import cKDTree

from ._radii import get_radii_from_names

AA_TEMPLATES = {
    "ALA": ["N", "CA", "C", "O", "CB"],
    "ARG": ["N", "CA", "C", "O", "CB", "CG", "CD", "NE", "CZ", "NH1", "NH2"],
    "ASN": ["N", "CA", "C", "O", "CB", "CG", "OD1", "ND2"],
    "ASP": ["N", "CA", "C", "O", "CB", "CG", "OD1", "OD2"],
    "CYS": ["N", "CA", "C", "O", "CB", "SG"],
    "GLN": ["N", "CA", "C", "O", "CB", "CG", "CD", "OE1", "NE2"],
    "GLU": ["N", "CA", "C", "O", "CB", "CG", "CD", "OE1", "OE2"],
    "GLY": ["N", "CA", "C", "O"],
    "HIS": ["N", "CA", "C", "O", "CB", "CG", "ND1", "CD2", "CE1", "NE2"],
    "ILE": ["N", "CA", "C", "O", "CB", "CG1", "CG2", "CD1"],
    "LEU": ["N", "CA", "C", "O", "CB", "CG", "CD1", "CD2"],
    "LYS": ["N", "CA", "C", "O", "CB", "CG", "CD", "CE", "NZ"],
    "MET": ["N", "CA", "C", "O", "CB", "CG", "SD", "CE"],
    "PHE": ["N", "CA", "C", "O", "CB", "CG", "CD1", "CD2", "CE1", "CE2", "CZ"],
    "PRO": ["N", "CA", "C", "O", "CB", "CG", "CD"],
    "SER": ["N", "CA", "C", "O", "CB", "OG"],
    "THR": ["N", "CA", "C", "O", "CB", "OG1", "CG2"],
    "TRP": ["N", "CA", "C", "O", "CB", "CG", "CD1", "CD2", "NE1", "CE2", "CE3",
            "CZ2", "CZ3", "CH2"],
    "TYR": ["N", "CA", "C", "O", "CB", "CG", "CD1", "CD2", "CE1", "CE2", "CZ", "OH"],
    "VAL": ["N", "CA", "C", "O", "CB", "CG1", "CG2"],
}
AA_NAMES = sorted(AA_TEMPLATES)
LIGAND_ELEMENTS = ("C", "N", "O", "S")
LIGAND_WEIGHTS = (0.70, 0.13, 0.15, 0.02)
MIN_SPACING = 2.2


def residue_names(n, rng, ligand=False):
    """(residue names, atom names) for n atoms."""
    if ligand:
        el = rng.choice(len(LIGAND_ELEMENTS), size=n, p=LIGAND_WEIGHTS)
        return ["LIG"] * n, [LIGAND_ELEMENTS[e] for e in el]
    res, atoms = [], []
    while len(atoms) < n:
        aa = AA_NAMES[int(rng.integers(len(AA_NAMES)))]
        for name in AA_TEMPLATES[aa]:
            if len(atoms) == n:
                break
            res.append(aa)
            atoms.append(name)
    return res, atoms


def ellipsoid_axes(n, density, ratio):
    """Semi-axes (a, a, c = ratio*a) of the ellipsoid of volume n/density."""
    vol = n / density
    a = (3.0 * vol / (4.0 * np.pi * ratio)) ** (1.0 / 3.0)
    return np.array([a, a, ratio * a])


def poisson_disk(n, axes, rng, spacing=MIN_SPACING, max_rounds=5000):
    """n points inside the axis-aligned ellipsoid with pairwise distance >=
    spacing.  Random sequential addition in rounds: each round proposes a
    batch of uniform points, drops those within spacing of an accepted point,
    and within the batch keeps a proposal only if no earlier proposal of the
    batch is within spacing (so every round stays vectorised)."""
    pts = np.zeros((0, 3))
    tree = None
    for _ in range(max_rounds):
        need = n - len(pts)
        if need <= 0:
            break
        m = max(2048, 2 * need)
        cand = rng.uniform(-1.0, 1.0, size=(m, 3))
        cand = cand[np.sum(cand * cand, axis=1) <= 1.0] * axes
        if tree is not None and len(cand):
            d, _ = tree.query(cand, k=1)
            cand = cand[d >= spacing]
        if len(cand) == 0:
            continue
        pairs = cKDTree(cand).query_pairs(spacing, output_type="ndarray")
        keep = np.ones(len(cand), dtype=bool)
        if len(pairs):
            keep[np.maximum(pairs[:, 0], pairs[:, 1])] = False
        acc = cand[keep][:need]
        pts = np.concatenate([pts, acc])
        tree = cKDTree(pts)
    if len(pts) < n:
        raise RuntimeError("Poisson-disk sampling could not place %d points" % n)
    return pts


def random_rotation(rng):
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def point_ellipsoid_distance(p, axes, iters=40):
    """Exact distance from points outside an axis-aligned ellipsoid to its
    surface.  Newton on the Lagrange multiplier t of the closest point
    a_i^2 p_i / (t + a_i^2): f(t) = sum (a_i p_i / (t + a_i^2))^2 - 1 is
    convex and decreasing for t >= 0 and f(0) > 0 outside, so Newton from
    t = 0 climbs monotonically to the root."""
    a2 = axes ** 2
    t = np.zeros(len(p))
    for _ in range(iters):
        q = axes * p / (t[:, None] + a2)
        qq = q * q
        f = qq.sum(axis=1) - 1.0
        df = -2.0 * (qq / (t[:, None] + a2)).sum(axis=1)
        step = f / df
        t = t - step
        if np.all(np.abs(step) <= 1e-13 * (1.0 + np.abs(t))):
            break
    close = a2 * p / (t[:, None] + a2)
    return np.linalg.norm(p - close, axis=1)


def _surface_points(axes, n=3000):
    k = np.arange(n) + 0.5
    phi = np.arccos(1 - 2 * k / n)
    th = np.pi * (1 + 5 ** 0.5) * k
    s = np.stack([np.cos(th) * np.sin(phi), np.sin(th) * np.sin(phi), np.cos(phi)], axis=1)
    return s * axes


def place_binder(t_axes, b_axes, rot, direction, gap):
    """Offset along direction at which the binder ellipsoid (rotated by rot)
    is gap away from the target ellipsoid: Brent's method (to 1e-9 A) on the
    minimum over the binder-surface sample facing the target of the exact
    point-to-ellipsoid distance."""
    surf = _surface_points(b_axes) @ rot.T
    surf = surf[surf @ direction < 0.25 * np.linalg.norm(surf, axis=1)]

    def surface_gap(off):
        p = surf + off * direction
        inside = np.sum((p / t_axes) ** 2, axis=1) < 1.0
        if inside.any():
            return -1.0
        return float(point_ellipsoid_distance(p, t_axes).min())

    from scipy.optimize import brentq
    lo, hi = 0.0, float(np.max(t_axes) + np.max(b_axes) + gap + 1.0)
    while surface_gap(hi) < gap:
        hi *= 2.0
    return brentq(lambda off: surface_gap(off) - gap, lo, hi, xtol=1e-9, rtol=1e-15, maxiter=200)


def make_molecule(n, rng, density=0.05, ratio=1.3, ligand=False):
    axes = ellipsoid_axes(n, density, ratio)
    xyz = poisson_disk(n, axes, rng)
    res, names = residue_names(n, rng, ligand=ligand)
    return xyz, res, names, axes


def make_complex(n_target, n_binder, seed, gap=4.0, ratio_target=1.3, ratio_binder=1.3,
                 density_target=0.05, density_binder=None, ligand=None):
    """One seeded complex.  Binders of at most 60 atoms are ligand-like
    (C/N/O/S names, 0.06 atoms/A^3) unless told otherwise."""
    rng = np.random.Generator(np.random.PCG64(seed))
    if ligand is None:
        ligand = n_binder <= 60
    if density_binder is None:
        density_binder = 0.06 if ligand else 0.05
    tx, tres, tnames, t_axes = make_molecule(n_target, rng, density_target, ratio_target)
    bx, bres, bnames, b_axes = make_molecule(n_binder, rng, density_binder, ratio_binder, ligand)
    rot = random_rotation(rng)
    u = rng.standard_normal(3)
    u /= np.linalg.norm(u)
    off = place_binder(t_axes, b_axes, rot, u, gap)
    bx = bx @ rot.T + off * u
    return dict(
        binder_xyz=bx, binder_radii=get_radii_from_names(bres, bnames),
        target_xyz=tx, target_radii=get_radii_from_names(tres, tnames),
        binder_names=(bres, bnames), target_names=(tres, tnames), seed=seed, gap=gap,
    )


# -------------------------------------------------------------- workloads

def config1(seed=1):
    """400-atom binder + 1,000-atom target, axis ratio 1:1:1.3, gap 4.0 A."""
    return make_complex(1000, 400, seed, gap=4.0)


def config2(seed=2):
    """1,800-atom scFv-like binder (1:1:1.6) + 4,000-atom target, gap 4.0 A."""
    return make_complex(4000, 1800, seed, gap=4.0, ratio_binder=1.6)


def batch_sizes(n, seed, binder_lo=20, binder_hi=1800, target_lo=800, target_hi=6000,
                gap_lo=3.5, gap_hi=5.0):
    """Per-complex (n_target, n_binder, gap) for a mixed batch: binder sizes
    log-uniform, target sizes uniform, gaps uniform."""
    rng = np.random.Generator(np.random.PCG64(seed))
    nb = np.exp(rng.uniform(np.log(binder_lo), np.log(binder_hi), size=n)).round().astype(int)
    nt = rng.integers(target_lo, target_hi + 1, size=n)
    gaps = rng.uniform(gap_lo, gap_hi, size=n)
    return nt, nb, gaps


def _make_one(args):
    nt, nb, seed, gap = args
    return make_complex(nt, nb, seed, gap=gap)


def generate(jobs, workers=1):
    """Complexes for jobs (n_target, n_binder, seed, gap); workers > 1 uses a
    process pool (spawned, so it is safe after GPU initialisation)."""
    jobs = list(jobs)
    if workers <= 1 or len(jobs) < 4:
        return [_make_one(j) for j in jobs]
    import concurrent.futures
    import multiprocessing
    ctx = multiprocessing.get_context("spawn")
    with concurrent.futures.ProcessPoolExecutor(max_workers=workers, mp_context=ctx) as ex:
        return list(ex.map(_make_one, jobs, chunksize=max(1, len(jobs) // (4 * workers))))


def mixed_jobs(n, seed=3, **size_kw):
    nt, nb, gaps = batch_sizes(n, seed, **size_kw)
    return [(int(nt[i]), int(nb[i]), seed * 100003 + i, float(gaps[i])) for i in range(n)]


def mixed_batch(n, seed=3, workers=1, **size_kw):
    """n complexes with batch_sizes()."""
    return generate(mixed_jobs(n, seed, **size_kw), workers)


def c3_batch(seed=3, workers=1, n=2048):
    """2,048 complexes: binders log-uniform 20-1,800 atoms (ligand-like up to
    60), targets uniform 800-6,000, gaps uniform 3.5-5.0 A."""
    return mixed_batch(n, seed, workers)


def c4_batch(seed=3, workers=1, n=2048, n_small=64):
    """The c3 batch plus 64 complexes of the 400 + 1,000-atom size (gap 4 A)."""
    jobs = mixed_jobs(n, seed) + [(1000, 400, seed * 100003 + n + i, 4.0) for i in range(n_small)]
    return generate(jobs, workers)


def c5_batch(seed=5, workers=1, n=256):
    """256 complexes: 1,500-atom targets, binders log-uniform 80-600 atoms,
    gaps uniform 3.5-5.0 A (the differentiation workload)."""
    return mixed_batch(n, seed, workers, binder_lo=80, binder_hi=600, target_lo=1500, target_hi=1500)
