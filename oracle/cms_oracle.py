"""Oracle: NumPy FP64 restatement of the contact-molecular-surface (CMS) and
shape-complementarity (SC) pipeline.

TEST INFRASTRUCTURE ONLY.  This module is the checker for the HIP
implementation: only ``tests/``, ``__graft_entry__.smoke()`` and the CPU
baseline leg of ``bench.py`` may import it, and never as the thing measured
or shipped.  The product path (``cms_cuda``) never imports it.

What it restates: the reference calculator ``cms-cuda/_core.py`` stage by
stage (line numbers below are that file's).  Data containers differ (compact
neighbour rows with explicit masks instead of NaN-padded tables, stages
chunked to bound memory), but every floating-point quantity is formed with
the same operations in the same order, so results agree with the reference
to the last bit wherever NumPy's own elementwise functions agree:

* 3-component sums: ``np.sum``/``linalg.norm``/``cdist`` add (x + y) + z
  ("seq"); ``np.einsum('ij,ij->i')`` adds (x + z) + y ("alt").  Both were
  checked against numpy 2.2.6 on this image.
* ``a ** 2`` on float64 arrays is ``a * a``.
* no fused multiply-add anywhere.

How it is pinned (see tests/test_oracle.py): the reference's own known-answer
complex ``test/sc_example.pdb`` with the golden CMS / SC / area / distance
scalars and dot, probe and trimmed-dot counts recorded from the reference's
golden run; the vectors written by ``tests/golden/make_golden.py``; and closed
forms (isolated atom, two-atom molecule cap areas, far-apart molecules, all
radii zero, binned median on hand-made histograms, bincount order).
"""

import math

import numpy as np
from scipy.spatial.distance import cdist

ATTEN_BLOCKER = 1
ATTEN_2 = 2
ATTEN_BURIED = 5
MAX_SUBDIV = 100
TWO_PI = 2 * np.pi

# --------------------------------------------------------------------------
# 3-vector helpers (order of additions matters; see module docstring)
# --------------------------------------------------------------------------


def dot_seq(a, b):
    """(a0*b0 + a1*b1) + a2*b2, like np.sum(a*b, axis=-1)."""
    p = a * b
    return (p[..., 0] + p[..., 1]) + p[..., 2]


def dot_alt(a, b):
    """(a0*b0 + a2*b2) + a1*b1, like np.einsum('...i,...i')."""
    p = a * b
    return (p[..., 0] + p[..., 2]) + p[..., 1]


def norm_seq(a):
    return np.sqrt(dot_seq(a, a))


def cross(a, b):
    """a x b with NumPy's component formulas (a1*b2 - a2*b1, ...)."""
    return np.cross(a, b)


def unit_rows(a):
    return a / norm_seq(a)[..., None]


def pairwise_sum(v):
    """np.sum of a 1-D float64 array (pairwise summation from 0.0)."""
    return float(np.sum(np.asarray(v, dtype=np.float64)))


# --------------------------------------------------------------------------
# Arc / circle subdivision   (_core.py:2407-2527)
# --------------------------------------------------------------------------


def arc_angle(x, v, axis):
    """Angle from x to v about axis, in [0, 2pi); NaN when v has no component
    in the plane (both projections |.| <= 1e-8).  Returns (angle, y) with
    y = axis x x.  _core.py:2407-2427."""
    y = cross(axis, x)
    dt1 = dot_seq(v, x)
    dt2 = dot_seq(v, y)
    angle = np.arctan2(dt2, dt1)
    angle = np.where(angle < 0.0, angle + TWO_PI, angle)
    degenerate = (np.abs(dt1) <= 1e-8) & (np.abs(dt2) <= 1e-8)
    angle = np.where(degenerate, np.nan, angle)
    return angle, y


def subdivide(cen, rad, x, y, angle, density):
    """Points cen + x*rad*cos(a_i) + y*rad*sin(a_i) for a_i = delta*(i+1/2)
    <= angle, delta = 1/(sqrt(density)*rad), i < MAX_SUBDIV.

    Returns (ps, pts, mask): ps = rad*angle/count (0 without points), pts
    (E, K, 3) and mask (E, K) of the points that exist.  _core.py:2429-2486.
    """
    rad = np.asarray(rad, dtype=np.float64)
    angle = np.asarray(angle, dtype=np.float64)
    density = np.asarray(density, dtype=np.float64)
    E = rad.shape[0]
    with np.errstate(divide="ignore", invalid="ignore"):
        delta = 1.0 / (np.sqrt(density) * rad)
        raw = angle / delta
    ok = np.isfinite(raw) & (raw > 0)
    if not ok.any():
        return np.zeros(E), np.zeros((E, 0, 3)), np.zeros((E, 0), dtype=bool)
    K = int(min(MAX_SUBDIV, math.ceil(float(raw[ok].max()))))
    if K == 0:
        return np.zeros(E), np.zeros((E, 0, 3)), np.zeros((E, 0), dtype=bool)
    with np.errstate(invalid="ignore"):
        a = delta[:, None] * (np.arange(K) + 0.5)
        mask = a <= angle[:, None]
        c = rad[:, None] * np.cos(a)
        s = rad[:, None] * np.sin(a)
        pts = cen[:, None, :] + x[:, None, :] * c[..., None] + y[:, None, :] * s[..., None]
    counts = mask.sum(axis=-1)
    with np.errstate(invalid="ignore"):
        ps = np.where(counts > 0, (rad * angle) / np.maximum(counts, 1), 0.0)
    return ps, pts, mask


def circle_frame(axis):
    """In-plane basis (x, y) of the circle about axis.  _core.py:2488-2523."""
    axis = unit_rows(axis)
    a0, a1, a2 = axis[:, 0], axis[:, 1], axis[:, 2]
    v1 = np.stack([a1 * a1 + a2 * a2, a0 * a0 + a2 * a2, a0 * a0 + a1 * a1], axis=-1)
    v1 = unit_rows(v1)
    dt = dot_seq(v1, axis)
    v1 = np.where((np.abs(dt) > 0.99)[:, None], np.array([1.0, 0.0, 0.0]), v1)
    v2 = unit_rows(cross(axis, v1))
    x = unit_rows(cross(axis, v2))
    y = cross(axis, x)
    return x, y


def dist_to_line(cen, axis, pnt):
    """Distance from pnt to the line through cen along unit axis, 0 when the
    rounded square goes negative.  _core.py:2380-2386."""
    vec = pnt - cen
    dt = dot_seq(vec, axis)
    d2 = dot_seq(vec, vec) - dt * dt
    with np.errstate(invalid="ignore"):
        return np.where(d2 < 0.0, 0.0, np.sqrt(np.maximum(d2, 0.0)))


def binned_median(values, binwidth):
    """Grouped-data median: bin by floor(v/binwidth), walk bins upward until
    the cumulative percentage reaches 50, interpolate inside that bin.
    _core.py:1080-1104."""
    n = len(values)
    if n == 0:
        return 0.0
    ibin = np.floor(values / binwidth).astype(np.int64)
    bins, counts = np.unique(ibin, return_counts=True)
    pct = 100.0 / n
    cum = np.cumsum(counts) * pct
    before = cum - counts * pct
    j = int(np.searchsorted(cum, 50.0, side="left"))
    j = min(j, len(bins) - 1)
    return float(bins[j] * binwidth + (50.0 - before[j]) * binwidth / (counts[j] * pct))


# --------------------------------------------------------------------------
# Containers
# --------------------------------------------------------------------------


class Dots:
    """Dots of one molecule in emission order."""

    FIELDS = ("coor", "nml", "area", "buried", "type", "atom")

    def __init__(self):
        self._parts = []

    def extend(self, coor, nml, area, buried, type_, atom):
        if len(coor):
            self._parts.append((coor, nml, area, buried, type_, atom))

    def finalize(self):
        if self._parts:
            cols = list(zip(*self._parts))
            self.coor = np.concatenate(cols[0]).reshape(-1, 3)
            self.nml = np.concatenate(cols[1]).reshape(-1, 3)
            self.area = np.concatenate(cols[2]).astype(np.float64)
            self.buried = np.concatenate(cols[3]).astype(np.int8)
            self.type = np.concatenate(cols[4]).astype(np.int8)
            self.atom = np.concatenate(cols[5]).astype(np.int32)
        else:
            self.coor = np.zeros((0, 3))
            self.nml = np.zeros((0, 3))
            self.area = np.zeros(0)
            self.buried = np.zeros(0, dtype=np.int8)
            self.type = np.zeros(0, dtype=np.int8)
            self.atom = np.zeros(0, dtype=np.int32)
        del self._parts
        return self

    def select(self, mask):
        out = Dots()
        out._parts = []
        out.finalize()
        for f in self.FIELDS:
            setattr(out, f, getattr(self, f)[mask])
        return out

    def __len__(self):
        return len(self.area)


# --------------------------------------------------------------------------
# Calculator
# --------------------------------------------------------------------------


class OracleCalculator:
    """FP64 CPU restatement of MolecularSurfaceCalculator (_core.py:762-2527).

    Usage mirrors the reference: add_binder_and_target / add_molecule, then
    contact_ms(target_side), max_contact_ms(), shape_complementarity().
    """

    def __init__(self, rp=1.7, density=15.0, band=1.5, sep=8.0, weight=0.5,
                 binwidth_dist=0.02, binwidth_norm=0.02, chunk=4096):
        self.rp = float(rp)
        self.density = float(density)
        self.band = float(band)
        self.sep = float(sep)
        self.weight = float(weight)
        self.binwidth_dist = float(binwidth_dist)
        self.binwidth_norm = float(binwidth_norm)
        self.chunk = int(chunk)
        self.reset()

    # ---------------------------------------------------------- loading
    def reset(self):
        self.xyz = np.zeros((0, 3))
        self.radius = np.zeros(0)
        self.dens = np.zeros(0)
        self.mol = np.zeros(0, dtype=np.int8)
        self.generated = None
        self.counts = dict(convex=0, toroidal=0, concave=0)
        self.sc_results = None

    def add_molecule(self, molecule, xyz, radii):
        """_core.py:923-943: append atoms; natom = position in the array."""
        xyz = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
        radii = np.asarray(radii, dtype=np.float64).reshape(-1)
        m = 1 if molecule == 1 else 0
        self.xyz = np.concatenate([self.xyz, xyz])
        self.radius = np.concatenate([self.radius, radii])
        self.dens = np.concatenate([self.dens, np.full(len(radii), self.density)])
        self.mol = np.concatenate([self.mol, np.full(len(radii), m, dtype=np.int8)])

    def add_binder_and_target(self, bx, br, tx, tr):
        self.reset()
        self.add_molecule(0, bx, br)
        self.add_molecule(1, tx, tr)

    # ---------------------------------------------------------- surface
    def generate(self, all_atoms=False):
        """Attention, neighbours, pairs, probes, convex, toroidal, concave.
        _core.py:865-909, 1197-1231."""
        if self.generated is not None:
            if self.generated != all_atoms:
                raise RuntimeError("surface already generated with all_atoms=%r" % self.generated)
            return
        N = len(self.radius)
        assert N > 0
        self.access = np.zeros(N, dtype=np.int8)
        self._attention(all_atoms)
        self.radmax = float(self.radius.max())
        good = self._neighbors()
        self._pairs(np.nonzero(good)[0])
        self._probes()
        self.dots = [Dots(), Dots()]
        convex_queue = np.nonzero(good & (self.access != 0) & (self.atten > ATTEN_BLOCKER))[0]
        self.convex_queue = convex_queue
        self._convex(convex_queue)
        self._toroidal()
        if self.rp > 0:
            self._concave()
        for d in self.dots:
            d.finalize()
        self.generated = all_atoms

    def _attention(self, all_atoms):
        """_core.py:1150-1195."""
        N = len(self.radius)
        self.atten = np.zeros(N, dtype=np.int8)
        if all_atoms:
            self.atten[:] = ATTEN_BURIED
            return
        i0 = np.nonzero(self.mol == 0)[0]
        i1 = np.nonzero(self.mol == 1)[0]
        if len(i0) and len(i1):
            d = cdist(self.xyz[i0], self.xyz[i1])
            min0, min1 = d.min(axis=1), d.min(axis=0)
        else:
            min0 = np.full(len(i0), 99999.0)
            min1 = np.full(len(i1), 99999.0)
        self.atten[i0] = np.where(min0 >= self.sep, ATTEN_BLOCKER, ATTEN_BURIED)
        self.atten[i1] = np.where(min1 >= self.sep, ATTEN_BLOCKER, ATTEN_BURIED)

    def _neighbors(self):
        """Same-molecule neighbour rows sorted by (d^2, index); other-molecule
        buried rows ascending; coincident-atom check.  _core.py:1251-1349."""
        N = len(self.radius)
        rp = self.rp
        xyz, r, mol, atten = self.xyz, self.radius, self.mol, self.atten
        d2 = cdist(xyz, xyz, metric="sqeuclidean")
        bridge2 = ((r[:, None] + r[None, :]) + 2.0 * rp) ** 2
        active = atten > 0
        same = mol[:, None] == mol[None, :]
        same_off = same.copy()
        np.fill_diagonal(same_off, False)
        both = active[:, None] & active[None, :]
        coincident = same_off & both & (d2 <= 0.0001)
        if coincident.any():
            i0, j0 = map(int, np.argwhere(coincident)[0])
            raise RuntimeError("Coincident atoms: %d:: == %d::" % (i0, j0))
        nbr = both & same_off & (d2 < bridge2)
        bur = active[:, None] & (atten >= ATTEN_BURIED)[None, :] & ~same & (d2 < bridge2)
        nn = nbr.sum(axis=1)
        nb = bur.sum(axis=1)
        self.access |= (active & (nn == 0)).astype(np.int8)
        K = max(int(nn.max()) if N else 0, 1)
        Kb = max(int(nb.max()) if N else 0, 1)
        self.nbr = np.full((N, K), -1, dtype=np.int32)
        self.bur = np.full((N, Kb), -1, dtype=np.int32)
        for i in np.nonzero(active)[0]:
            j = np.nonzero(nbr[i])[0]
            if len(j):
                j = j[np.argsort(d2[i, j], kind="stable")]
                self.nbr[i, :len(j)] = j
            b = np.nonzero(bur[i])[0]
            self.bur[i, :len(b)] = b
        self.nn = nn.astype(np.int32)
        self.nb = nb.astype(np.int32)
        return active & (nn > 0)

    def _pairs(self, good):
        """Forward neighbour pairs, torus geometry, single-neighbour shortcut
        and the toroid queue.  _core.py:1351-1442."""
        rp = self.rp
        xyz, r = self.xyz, self.radius
        rows = self.nbr[good]
        fwd = rows > good[:, None]
        gi, kk = np.nonzero(fwd)
        i = good[gi]
        j = rows[gi, kk]
        diff = xyz[j] - xyz[i]
        dij_sq = dot_alt(diff, diff)
        dij = np.sqrt(dij_sq)
        eri = r[i] + rp
        erj = r[j] + rp
        asymm = (eri * eri - erj * erj) / dij
        between = np.abs(asymm) < dij
        tij = (xyz[i] + xyz[j]) * 0.5 + (diff / dij[:, None]) * (asymm * 0.5)[:, None]
        far_sq = (eri + erj) ** 2 - dij_sq
        cont_sq = dij_sq - (r[i] - r[j]) ** 2
        geom_ok = (far_sq > 0.0) & (cont_sq > 0.0)
        # single-neighbour shortcut
        any_ok = np.zeros(len(good), dtype=bool)
        any_ok[gi[geom_ok]] = True
        shortcut = (self.nn[good] <= 1) & any_ok
        if shortcut.any():
            self.access[good[shortcut]] = 1
            for g in np.nonzero(shortcut)[0]:
                sel = np.nonzero((gi == g) & geom_ok)[0]
                self.access[j[sel[0]]] = 1
        keep = geom_ok & ~shortcut[gi]
        self.p_i = i[keep].astype(np.int32)
        self.p_j = j[keep].astype(np.int32)
        d = dij[keep]
        self.p_u = diff[keep] / d[:, None]
        self.p_t = tij[keep]
        self.p_r = (0.5 * np.sqrt(far_sq[keep])) * np.sqrt(cont_sq[keep]) / d
        self.p_between = between[keep]
        at = self.atten
        self.p_tor = (at[self.p_i] > ATTEN_BLOCKER) | ((at[self.p_j] > ATTEN_BLOCKER) & (rp > 0.0))

    def _probes(self):
        """Probe placement on atom triples with pair-kill on degenerate
        triples and neighbour collision.  _core.py:1445-1643."""
        rp = self.rp
        xyz, r, at = self.xyz, self.radius, self.atten
        plus, minus = [], []
        M = len(self.p_i)
        for s in range(0, M, self.chunk):
            e = min(M, s + self.chunk)
            pi, pj = self.p_i[s:e], self.p_j[s:e]
            uij, tij, rij = self.p_u[s:e], self.p_t[s:e], self.p_r[s:e]
            rows = self.nbr[pi]
            valid = (rows > pj[:, None]) & (rows >= 0)
            safe = np.where(rows >= 0, rows, 0)
            xk = xyz[safe]
            erk = r[safe] + rp
            eri = r[pi] + rp
            erj = r[pj] + rp
            xi, xj = xyz[pi], xyz[pj]
            djk = norm_seq(xk - xj[:, None, :])
            valid &= djk < (erj[:, None] + erk)
            dik = norm_seq(xk - xi[:, None, :])
            valid &= dik < (eri[:, None] + erk)
            blocked = (at[pi] <= ATTEN_BLOCKER)[:, None] & (at[pj] <= ATTEN_BLOCKER)[:, None] & (at[safe] <= ATTEN_BLOCKER)
            valid &= ~blocked
            pq, kq = np.nonzero(valid)
            if len(pq) == 0:
                continue
            xk_q = xk[pq, kq]
            erk_q = erk[pq, kq]
            dik_q = dik[pq, kq]
            xi_q = xi[pq]
            eri_q = eri[pq]
            uij_q, tij_q, rij_q = uij[pq], tij[pq], rij[pq]
            uik = (xk_q - xi_q) / dik_q[:, None]
            dt = dot_alt(uij_q, uik)
            w = np.arccos(np.clip(dt, -1.0, 1.0))
            sw = np.sin(w)
            degenerate = (dt >= 1.0) | (dt <= -1.0) | (w <= 0.0) | (sw <= 0.0)
            killed = np.zeros(e - s, dtype=bool)
            if degenerate.any():
                dk = dot_seq(tij_q[degenerate] - xk_q[degenerate], tij_q[degenerate] - xk_q[degenerate])
                kills = dk < erk_q[degenerate] ** 2 - rij_q[degenerate] ** 2
                killed[pq[degenerate][kills]] = True
            keep = ~degenerate & ~killed[pq]
            if not keep.any():
                continue
            q = np.nonzero(keep)[0]
            uik, sw = uik[q], sw[q]
            uij_q, tij_q = uij_q[q], tij_q[q]
            xi_q, xk_q = xi_q[q], xk_q[q]
            eri_q, erk_q, dik_q = eri_q[q], erk_q[q], dik_q[q]
            uijk = cross(uij_q, uik) / sw[:, None]
            utb = cross(uijk, uij_q)
            asymm = (eri_q ** 2 - erk_q ** 2) / dik_q
            tik = (xi_q + xk_q) * 0.5 + uik * (asymm * 0.5)[:, None]
            dtb = dot_alt(uik, tik - tij_q)
            bijk = tij_q + utb * (dtb / sw)[:, None]
            hsq = eri_q ** 2 - dot_alt(bijk - xi_q, bijk - xi_q)
            vh = hsq > 0.0
            if not vh.any():
                continue
            q2 = q[vh]
            bijk, uijk, h = bijk[vh], uijk[vh], np.sqrt(hsq[vh])
            a_i = pi[pq[q2]]
            a_j = pj[pq[q2]]
            a_k = rows[pq[q2], kq[q2]]
            for sign, out in ((1.0, plus), (-1.0, minus)):
                p = bijk + uijk * (h * sign)[:, None]
                alt = uijk * sign
                hit = self._probe_collides(p, a_i, a_j, a_k)
                ok = ~hit
                if sign > 0:
                    a0, a1 = a_i, a_j
                else:
                    a0, a1 = a_j, a_i
                out.append((a0[ok], a1[ok], a_k[ok], h[ok], p[ok], alt[ok]))
        parts = plus + minus
        if parts:
            cols = list(zip(*parts))
            self.pr_a = np.stack([np.concatenate(cols[0]), np.concatenate(cols[1]),
                                  np.concatenate(cols[2])], axis=1).astype(np.int32)
            self.pr_h = np.concatenate(cols[3])
            self.pr_p = np.concatenate(cols[4]).reshape(-1, 3)
            self.pr_alt = np.concatenate(cols[5]).reshape(-1, 3)
        else:
            self.pr_a = np.zeros((0, 3), dtype=np.int32)
            self.pr_h = np.zeros(0)
            self.pr_p = np.zeros((0, 3))
            self.pr_alt = np.zeros((0, 3))
        if len(self.pr_h):
            self.access[np.unique(self.pr_a)] = 1

    def _probe_collides(self, p, a_i, a_j, a_k):
        """Probe centre within (r + rp) of any neighbour of a_i other than
        a_j and a_k (<=, einsum order).  _core.py:1600-1623."""
        rows = self.nbr[a_i]
        ok = (rows >= 0) & (rows != a_j[:, None]) & (rows != a_k[:, None])
        rr, cc = np.nonzero(ok)
        n = rows[rr, cc]
        dv = p[rr] - self.xyz[n]
        d2 = dot_alt(dv, dv)
        within = d2 <= (self.radius[n] + self.rp) ** 2
        hit = np.zeros(len(p), dtype=bool)
        hit[rr[within]] = True
        return hit

    def _add_dots(self, molecule, type_, coor, area, pcen, atom):
        """Outward normal and burial test against the owning atom's buried
        row (<=, seq order), then append per molecule.  _core.py:2337-2377."""
        n = len(coor)
        if n == 0:
            return
        rp = self.rp
        if rp <= 0:
            nml = coor - self.xyz[atom]
        else:
            nml = (pcen - coor) / rp
        rows = self.bur[atom]
        ok = rows >= 0
        rr, cc = np.nonzero(ok)
        b = rows[rr, cc]
        dv = self.xyz[b] - pcen[rr]
        d2 = dot_seq(dv, dv)
        within = d2 <= (self.radius[b] + rp) ** 2
        buried = np.zeros(n, dtype=np.int8)
        buried[rr[within]] = 1
        tarr = np.full(n, type_, dtype=np.int8)
        for m in (0, 1):
            sel = molecule == m
            if sel.any():
                self.dots[m].extend(coor[sel], nml[sel], area[sel], buried[sel], tarr[sel],
                                    atom[sel].astype(np.int32))

    def _convex(self, queue):
        """Convex (contact) dots on each accessible non-blocker atom.
        _core.py:1646-1840."""
        rp = self.rp
        for s in range(0, len(queue), self.chunk):
            a = queue[s:s + self.chunk]
            A = len(a)
            xa = self.xyz[a]
            ri = self.radius[a]
            eri = ri + rp
            dens = self.dens[a]
            n0 = self.nbr[a, 0]
            x0 = self.xyz[n0]
            r0 = self.radius[n0]
            north = unit_rows(xa - x0)
            vt = np.stack([north[:, 1] ** 2 + north[:, 2] ** 2,
                           north[:, 0] ** 2 + north[:, 2] ** 2,
                           north[:, 0] ** 2 + north[:, 1] ** 2], axis=-1)
            vt = unit_rows(vt)
            dt = dot_seq(vt, north)
            vt[np.abs(dt) > 0.99] = np.array([1.0, 0.0, 0.0])
            eq = unit_rows(cross(north, vt))
            vql = cross(eq, north)
            erj = r0 + rp
            dij = norm_seq(x0 - xa)
            uij = (x0 - xa) / dij[:, None]
            asymm = (eri * eri - erj * erj) / dij
            tij = ((xa + x0) * 0.5) + uij * (asymm[:, None] * 0.5)
            far_sq = (eri + erj) ** 2 - dij * dij
            cont_sq = dij * dij - (ri - r0) ** 2
            gok = (far_sq > 0.0) & (cont_sq > 0.0)
            far = np.sqrt(np.where(gok, far_sq, 1.0))
            cont = np.sqrt(np.where(gok, cont_sq, 1.0))
            rij = 0.5 * far * cont / dij
            pij = tij + vql * rij[:, None]
            south = np.where(gok[:, None], (pij - xa) / eri[:, None], np.array([0.0, 0.0, -1.0]))
            # latitudes
            angle, y = arc_angle(north, south, eq)
            cs, lats, lmask = subdivide(np.zeros((A, 3)), ri, north, y, angle, dens)
            L = lats.shape[1]
            if L == 0:
                continue
            with np.errstate(invalid="ignore"):
                dtl = dot_seq(lats, north[:, None, :])
                cen = xa[:, None, :] + dtl[..., None] * north[:, None, :]
                rsq = ri[:, None] ** 2 - dtl * dtl
            rsq = np.where(lmask, rsq, np.nan)
            live = rsq > 0
            rad = np.where(live, np.sqrt(np.where(live, rsq, 1.0)), 0.0)
            ai, li = np.nonzero(live)
            if len(ai) == 0:
                continue
            cx, cy = circle_frame(north[ai])
            ps, pts, pmask = subdivide(cen[ai, li], rad[ai, li], cx, cy,
                                       np.full(len(ai), TWO_PI), dens[ai])
            area_row = ps * cs[ai]
            rr, kk = np.nonzero(pmask)          # row-major over (atom, lat, pt)
            own = ai[rr]
            pt = pts[rr, kk]
            xo = self.xyz[a[own]]
            pc = xo + (pt - xo) * (eri[own] / ri[own])[:, None]
            hit = self._convex_collides(pc, a[own])
            k = ~hit
            if not k.any():
                continue
            self.counts["convex"] += int(k.sum())
            at = a[own[k]]
            self._add_dots(self.mol[at], 1, pt[k], area_row[rr[k]], pc[k], at)

    def _convex_collides(self, pc, atom):
        """pcen within (r + rp) of any neighbour except the nearest one
        (slot 0 skipped), <=, seq order.  _core.py:1834-1840."""
        rows = self.nbr[atom]
        ok = rows >= 0
        ok[:, 0] = False
        rr, cc = np.nonzero(ok)
        n = rows[rr, cc]
        dv = self.xyz[n] - pc[rr]
        d2 = dot_seq(dv, dv)
        within = d2 <= (self.radius[n] + self.rp) ** 2
        hit = np.zeros(len(pc), dtype=bool)
        hit[rr[within]] = True
        return hit

    def _toroidal(self):
        """Toroidal (reentrant saddle) dots for every queued pair, atom-1 arcs
        before atom-2 arcs.  _core.py:1234-1248, 1847-2067."""
        rp = self.rp
        q = np.nonzero(self.p_tor)[0]
        self.tq = q
        arc1, arc2 = [], []
        has1 = np.zeros(len(q), dtype=bool)
        has2 = np.zeros(len(q), dtype=bool)
        for s in range(0, len(q), self.chunk):
            e_idx = q[s:s + self.chunk]
            i, j = self.p_i[e_idx], self.p_j[e_idx]
            uij, tij, rij = self.p_u[e_idx], self.p_t[e_idx], self.p_r[e_idx]
            between = self.p_between[e_idx]
            xi, xj = self.xyz[i], self.xyz[j]
            ri, rj = self.radius[i], self.radius[j]
            dens = (self.dens[i] + self.dens[j]) * 0.5
            eri, erj = ri + rp, rj + rp
            rci = rij * ri / eri
            rcj = rij * rj / erj
            rb = np.maximum(rij - rp, 0.0)
            rs = (rci + 2 * rb + rcj) * 0.25
            e = rs / rij
            edens = e * e * dens
            cx, cy = circle_frame(uij)
            ts, subs, smask = subdivide(tij, rij, cx, cy, np.full(len(e_idx), TWO_PI), edens)
            ei, si = np.nonzero(smask)
            if len(ei) == 0:
                continue
            p = subs[ei, si]
            # collision with atom1's neighbours other than atom2 (strict <, einsum order)
            rows = self.nbr[i[ei]]
            ok = (rows >= 0) & (rows != j[ei][:, None])
            rr, cc = np.nonzero(ok)
            n = rows[rr, cc]
            dv = p[rr] - self.xyz[n]
            d2 = dot_alt(dv, dv)
            within = d2 < (self.radius[n] + rp) ** 2
            hit = np.zeros(len(p), dtype=bool)
            hit[rr[within]] = True
            live = ~hit
            if not live.any():
                continue
            ei, p = ei[live], p[live]
            has1[s + ei] = True
            has2[s + ei] = True
            pi_ = (xi[ei] - p) / eri[ei][:, None]
            pj_ = (xj[ei] - p) / erj[ei][:, None]
            axis = unit_rows(cross(pi_, pj_))
            dtq = rp ** 2 - rij[ei] ** 2
            cusp = (dtq > 0) & between[ei]
            pqi = np.empty_like(pi_)
            pqj = np.empty_like(pi_)
            if cusp.any():
                dq = np.sqrt(dtq[cusp])
                qij = tij[ei[cusp]] - uij[ei[cusp]] * dq[:, None]
                pqi[cusp] = (qij - p[cusp]) / rp
                pqj[cusp] = 0.0
            if (~cusp).any():
                tmp = unit_rows(pi_[~cusp] + pj_[~cusp])
                pqi[~cusp] = tmp
                pqj[~cusp] = tmp
            dt1 = dot_seq(pqi, pi_)
            dt2 = dot_seq(pqj, pj_)
            va = (np.abs(dt1) < 1.0) & (np.abs(dt2) < 1.0)
            ei, p, axis, pi_, pj_, pqi, pqj = ei[va], p[va], axis[va], pi_[va], pj_[va], pqi[va], pqj[va]
            if len(ei) == 0:
                continue
            for which, out in ((1, arc1), (2, arc2)):
                owner = i[ei] if which == 1 else j[ei]
                m = self.atten[owner] >= ATTEN_2
                if not m.any():
                    continue
                if which == 1:
                    ang, y = arc_angle(pi_[m], pqi[m], axis[m])
                    x = pi_[m]
                else:
                    ang, y = arc_angle(pqj[m], pj_[m], axis[m])
                    x = pqj[m]
                ps, pts, pm = subdivide(p[m], np.full(int(m.sum()), rp), x, y, ang, dens[ei[m]])
                if pts.shape[1] == 0:
                    continue
                with np.errstate(invalid="ignore"):
                    dist = dist_to_line(tij[ei[m]][:, None, :], uij[ei[m]][:, None, :], pts)
                    areas = ps[:, None] * ts[ei[m], None] * dist / rij[ei[m], None]
                rr, kk = np.nonzero(pm)
                pc = np.broadcast_to(p[m][:, None, :], pts.shape)[rr, kk]
                out.append((self.mol[i[ei[m]]][rr], pts[rr, kk], areas[rr, kk], pc, owner[m][rr]))
        for out in (arc1, arc2):
            for mol, pts, areas, pc, own in out:
                self.counts["toroidal"] += len(pts)
                self._add_dots(mol, 2, pts, areas, pc, own)
        # access flags: fancy-index |= keeps the last write per atom
        a1 = self.p_i[q]
        a2 = self.p_j[q]
        self.access[a1] = self.access[a1] | has1.astype(np.int8)
        self.access[a2] = self.access[a2] | has2.astype(np.int8)

    def _concave(self):
        """Concave (reentrant probe) dots.  _core.py:2069-2333."""
        P = len(self.pr_h)
        if P == 0:
            return
        rp = self.rp
        rp2 = rp * rp
        A = self.pr_a
        xyzA = self.xyz[A]                 # (P, 3, 3)
        rA = self.radius[A]
        dens = ((self.dens[A[:, 0]] + self.dens[A[:, 1]]) + self.dens[A[:, 2]]) / 3
        pp, uu, hh = self.pr_p, self.pr_alt, self.pr_h
        low = np.nonzero(hh < rp)[0]
        lowp = pp[low]
        for s in range(0, P, self.chunk):
            e = min(P, s + self.chunk)
            idx = np.arange(s, e)
            p, u, xa = pp[s:e], uu[s:e], xyzA[s:e]
            vp = xa - p[:, None, :]
            vp = vp / norm_seq(vp)[..., None]
            vec = np.stack([cross(vp[:, 0], vp[:, 1]), cross(vp[:, 1], vp[:, 2]),
                            cross(vp[:, 2], vp[:, 0])], axis=1)
            vec = vec / norm_seq(vec)[..., None]
            dt = dot_seq(u[:, None, :], vp)
            mm = np.argmax(dt, axis=1)
            top = vp[np.arange(e - s), mm]
            south = -u
            axis = unit_rows(cross(top, south))
            ang, y = arc_angle(top, south, axis)
            cs, lats, lmask = subdivide(np.zeros((e - s, 3)), np.full(e - s, rp), top, y, ang, dens[s:e])
            if lats.shape[1] == 0:
                continue
            with np.errstate(invalid="ignore"):
                dtl = dot_seq(lats, south[:, None, :])
                cen = south[:, None, :] * dtl[..., None]
                rsq = rp2 - dtl ** 2
            rsq = np.where(lmask, rsq, np.nan)
            with np.errstate(invalid="ignore"):
                rad = np.sqrt(np.clip(rsq, 0, None))
                live = rad > 0
            pi_, li = np.nonzero(live)
            if len(pi_) == 0:
                continue
            cx, cy = circle_frame(south[pi_])
            ps, pts, pm = subdivide(cen[pi_, li], rad[pi_, li], cx, cy,
                                    np.full(len(pi_), TWO_PI), dens[s:e][pi_])
            area_row = ps * cs[pi_]
            rr, kk = np.nonzero(pm)
            own = pi_[rr]
            pt = pts[rr, kk]
            bail = np.any(dot_seq(pt[:, None, :], vec[own]) >= 0, axis=1)
            rr, own, pt = rr[~bail], own[~bail], pt[~bail]
            pt = pt + p[own]
            if len(low):
                lowself = hh[s:e] < rp
                cand = np.nonzero(lowself[own])[0]
                if len(cand):
                    reject = np.zeros(len(pt), dtype=bool)
                    ow = own[cand]
                    # near low probes of each candidate dot's probe
                    dlp = lowp[None, :, :] - p[np.unique(ow)][:, None, :]
                    uq, inv = np.unique(ow, return_inverse=True)
                    d2pl = dot_seq(dlp, dlp)
                    near = d2pl <= 4 * rp2
                    selfcol = (low[None, :] == (idx[uq])[:, None])
                    near &= ~selfcol
                    nr = near[inv]                       # (C, L)
                    rr2, ll = np.nonzero(nr)
                    dv = pt[cand][rr2] - lowp[ll]
                    d2 = dot_alt(dv, dv)
                    col = np.zeros(len(cand), dtype=bool)
                    col[rr2[d2 < rp2]] = True
                    reject[cand[col]] = True
                    rr, own, pt = rr[~reject], own[~reject], pt[~reject]
            if len(pt) == 0:
                continue
            dd = norm_seq(pt[:, None, :] - xa[own]) - rA[s:e][own]
            mc = np.argmin(dd, axis=1)
            at = A[s:e][own, mc]
            self.counts["concave"] += len(pt)
            self._add_dots(self.mol[at], 3, pt, area_row[rr], p[own], at)

    # ---------------------------------------------------------- readouts
    def _natoms(self, m):
        return int((self.mol == m).sum())

    def contact_ms(self, target_side=True):
        """Distance-weighted buried area of the query molecule on the CMS-mode
        surface.  _core.py:791-797, 946-1004."""
        self.generate(False)
        return self.contact_ms_of_surface(target_side)

    def contact_ms_of_surface(self, target_side=True):
        """The CMS readout on whatever surface was generated (no mode check,
        as the reference's calc_contact_molecular_surface).  _core.py:946-1004."""
        n1 = self._natoms(1)
        n0 = len(self.mol) - n1
        if target_side:
            q, rf, nq, off = self.dots[1], self.dots[0], n1, n0
        else:
            q, rf, nq, off = self.dots[0], self.dots[1], n0, 0
        zero = (0.0, np.zeros(nq))
        if len(rf) == 0:
            return zero
        bq = q.buried.astype(bool)
        br = rf.buried.astype(bool)
        if not bq.any() or not br.any():
            return zero
        d2 = cdist(rf.coor[br], q.coor[bq], metric="sqeuclidean").min(axis=0)
        w = q.area[bq] * np.exp(-d2 * self.weight)
        per_atom = np.bincount(q.atom[bq] - off, weights=w, minlength=nq).astype(np.float64)
        return float(per_atom.sum()), per_atom

    def max_contact_ms(self):
        """Total and per-atom area of molecule 0's whole surface.
        _core.py:799-805, 1007-1034."""
        self.generate(True)
        return self.max_contact_ms_of_surface()

    def max_contact_ms_of_surface(self):
        """Area readout of molecule 0's dots on whatever surface was
        generated.  _core.py:1007-1034."""
        n0 = self._natoms(0)
        d = self.dots[0]
        if len(d) == 0:
            return 0.0, np.zeros(n0)
        per_atom = np.bincount(d.atom, weights=d.area, minlength=n0).astype(np.float64)
        return float(per_atom.sum()), per_atom

    def _trim(self, d):
        """Keep buried dots farther than band from every accessible dot.
        _core.py:1037-1077."""
        buried = d.buried.astype(bool)
        if not buried.any():
            return 0.0, np.zeros(len(d), dtype=bool)
        acc = ~buried
        if not acc.any():
            keep = buried.copy()
        else:
            m = cdist(d.coor[buried], d.coor[acc], metric="sqeuclidean").min(axis=1)
            keep = np.zeros(len(d), dtype=bool)
            keep[np.nonzero(buried)[0][m > self.band ** 2]] = True
        return float(d.area[keep].sum()), keep

    def _side_stats(self, mine, theirs):
        """Nearest trimmed dot on the other side, distance and normal
        statistics.  _core.py:1107-1147."""
        if len(mine) == 0 or len(theirs) == 0:
            return None
        tb = theirs.buried.astype(bool)
        if not tb.any():
            return None
        d2 = cdist(theirs.coor[tb], mine.coor, metric="sqeuclidean")
        nearest = np.argmin(d2, axis=0)
        distmin = np.sqrt(d2[nearest, np.arange(len(mine))])
        r = dot_alt(mine.nml, theirs.nml[tb][nearest])
        r = r * np.exp(-np.square(distmin) * self.weight)
        r = np.clip(r, -0.999, 0.999)
        return dict(d_mean=float(distmin.mean()),
                    d_median=binned_median(distmin, self.binwidth_dist),
                    s_mean=float(-r.mean()),
                    s_median=float(-binned_median(r, self.binwidth_norm)),
                    nearest=nearest, distmin=distmin, r=r)

    def shape_complementarity(self):
        """_core.py:807-863.  Returns (sc, area, distance) and keeps per-side
        statistics in self.sc_results."""
        self.generate(False)
        trimmed, areas = [], []
        for m in (0, 1):
            area, keep = self._trim(self.dots[m])
            trimmed.append(self.dots[m].select(keep))
            areas.append(area)
        sides = [self._side_stats(trimmed[0], trimmed[1]), self._side_stats(trimmed[1], trimmed[0])]
        zero = dict(d_mean=0.0, d_median=0.0, s_mean=0.0, s_median=0.0)
        s = [x if x is not None else zero for x in sides]
        res = dict(
            sc=(s[0]["s_median"] + s[1]["s_median"]) / 2,
            area=areas[0] + areas[1],
            distance=(s[0]["d_median"] + s[1]["d_median"]) / 2,
            d_mean=(s[0]["d_mean"] + s[1]["d_mean"]) / 2,
            s_mean=(s[0]["s_mean"] + s[1]["s_mean"]) / 2,
            trimmed_area=areas, n_trimmed=[len(trimmed[0]), len(trimmed[1])],
            sides=sides, trimmed=trimmed)
        self.sc_results = res
        return res["sc"], res["area"], res["distance"]


# --------------------------------------------------------------------------
# Convenience entry points mirroring _core.calculate_*
# --------------------------------------------------------------------------


def calculate_contact_ms(bx, br, tx, tr, **settings):
    calc = OracleCalculator(**settings)
    calc.add_binder_and_target(bx, br, tx, tr)
    cms, per_atom = calc.contact_ms(True)
    return cms, per_atom, calc


def calculate_shape_complementarity(bx, br, tx, tr, **settings):
    calc = OracleCalculator(**settings)
    calc.add_binder_and_target(bx, br, tx, tr)
    sc, area, dist = calc.shape_complementarity()
    return sc, area, dist, calc


def calculate_maximum_possible_contact_ms(xyz, radii, **settings):
    calc = OracleCalculator(**settings)
    calc.add_molecule(0, xyz, radii)
    cms, per_atom = calc.max_contact_ms()
    return cms, per_atom, calc
