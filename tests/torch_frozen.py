"""FP64 torch re-implementation of the frozen-topology CMS (test reference).

With the surface topology frozen -- which dots exist, their burial, their
nearest partners and every subdivision index and count -- each dot's
position P and area A are smooth functions of the coordinates of the one to
three atoms that generate it, and the CMS of a side is

    sum over buried dots q of that side of  A_q * exp(-w |P_q - P_r(q)|^2)

with r(q) the nearest buried dot of the other molecule.  This module
recomputes P and A for every dot from the descriptors the HIP library
exports (cms_copy_dot_desc), with torch autograd supplying the derivatives,
so it checks the hand-written adjoint and dual-number kernels of the
backward pass independently.  The geometry follows the reference's surface
stages: convex dots (cms-cuda/_core.py:1646-1832, pole from the nearest
neighbour, latitude arcs, circles), toroidal arcs (_core.py:1847-2067, torus
circle subdivision, atom-1 / atom-2 arcs, cusp rows), concave dots
(_core.py:2069-2292, probe from the atom triple, spherical triangle) and the
subdivision helpers vec_sub_arc / vec_sub_div / vec_sub_cir
(_core.py:2407-2527).  Test infrastructure only.
"""

import torch

TWO_PI = 6.283185307179586
DK_CONVEX, DK_ARC1, DK_ARC2, DK_CONCAVE = 1, 2, 3, 4


def dseq(a, b):
    p = a * b
    return (p[..., 0] + p[..., 1]) + p[..., 2]


def dalt(a, b):
    p = a * b
    return (p[..., 0] + p[..., 2]) + p[..., 1]


def cross(a, b):
    return torch.stack([a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1],
                        a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
                        a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]], -1)


def norm(a):
    return torch.sqrt(dseq(a, a))


def unit(a):
    return a / norm(a)[..., None]


def col(v):
    return v[..., None]


def circle_frame(axis):
    axis = unit(axis)
    v1 = torch.stack([axis[..., 1] * axis[..., 1] + axis[..., 2] * axis[..., 2],
                      axis[..., 0] * axis[..., 0] + axis[..., 2] * axis[..., 2],
                      axis[..., 0] * axis[..., 0] + axis[..., 1] * axis[..., 1]], -1)
    v1 = unit(v1)
    far = dseq(v1, axis).abs() > 0.99
    ex = torch.zeros_like(v1)
    ex[..., 0] = 1.0
    v1 = torch.where(col(far), ex, v1)
    v2 = unit(cross(axis, v1))
    x = unit(cross(axis, v2))
    return x, cross(axis, x)


def arc_angle(x, v, axis):
    y = cross(axis, x)
    ang = torch.atan2(dseq(v, y), dseq(v, x))
    return torch.where(ang < 0.0, ang + TWO_PI, ang), y


def subdiv_point(cen, x, y, rad, delta, i):
    a = delta * (i.to(torch.float64) + 0.5)
    return (cen + x * col(rad * torch.cos(a))) + y * col(rad * torch.sin(a))


def safe_sqrt(v, ok):
    return torch.sqrt(torch.where(ok, v, torch.ones_like(v)))


def pair_geom(xi, xj, ri, rj, rp):
    eri, erj = ri + rp, rj + rp
    diff = xj - xi
    dsq = dalt(diff, diff)
    dij = torch.sqrt(dsq)
    asymm = (eri * eri - erj * erj) / dij
    between = asymm.abs() < dij
    u = diff / col(dij)
    t = (xi + xj) * 0.5 + u * col(asymm * 0.5)
    far_sq = (eri + erj) * (eri + erj) - dsq
    cont_sq = dsq - (ri - rj) * (ri - rj)
    rij = (0.5 * torch.sqrt(far_sq)) * torch.sqrt(cont_sq) / dij
    return u, t, rij, between


def convex(X, rad, dens, rp, d1, d2):
    g, n0 = d1[:, 1], d1[:, 2]
    t, k, L, Pc = d2[:, 0], d2[:, 1], d2[:, 2], d2[:, 3]
    ri = rad[g]
    eri = ri + rp
    r0 = rad[n0]
    erj = r0 + rp
    xi, x0 = X[g], X[n0]
    north = unit(xi - x0)
    vt = torch.stack([north[:, 1] * north[:, 1] + north[:, 2] * north[:, 2],
                      north[:, 0] * north[:, 0] + north[:, 2] * north[:, 2],
                      north[:, 0] * north[:, 0] + north[:, 1] * north[:, 1]], -1)
    vt = unit(vt)
    ex = torch.zeros_like(vt)
    ex[:, 0] = 1.0
    vt = torch.where(col(dseq(vt, north).abs() > 0.99), ex, vt)
    eq = unit(cross(north, vt))
    vql = cross(eq, north)
    dij = norm(x0 - xi)
    uij = (x0 - xi) / col(dij)
    asymm = (eri * eri - erj * erj) / dij
    tij = ((xi + x0) * 0.5) + uij * col(asymm * 0.5)
    far_sq = (eri + erj) * (eri + erj) - dij * dij
    cont_sq = dij * dij - (ri - r0) * (ri - r0)
    ok = (far_sq > 0.0) & (cont_sq > 0.0)
    rij = 0.5 * safe_sqrt(far_sq, ok) * safe_sqrt(cont_sq, ok) / dij
    south = torch.where(col(ok), ((tij + vql * col(rij)) - xi) / col(eri),
                        torch.tensor([0.0, 0.0, -1.0], dtype=X.dtype).expand_as(tij))
    ang, y = arc_angle(north, south, eq)
    delta = 1.0 / (torch.sqrt(dens[g]) * ri)
    cs = (ri * ang) / L.to(torch.float64)
    cx, cy = circle_frame(north)
    lat = subdiv_point(torch.zeros_like(xi), north, y, ri, delta, t)
    dt = dseq(lat, north)
    cen = xi + north * col(dt)
    r = torch.sqrt(ri * ri - dt * dt)
    cdelta = 1.0 / (torch.sqrt(dens[g]) * r)
    ps = (r * TWO_PI) / Pc.to(torch.float64)
    return subdiv_point(cen, cx, cy, r, cdelta, k), ps * cs


def toroidal(X, rad, dens, rp, d1, d2):
    i, j = d1[:, 1], d1[:, 2]
    arc1 = (d1[:, 0] & 15) == DK_ARC1
    m, k, M, n = d2[:, 0], d2[:, 1], d2[:, 2], d2[:, 3]
    ri, rj = rad[i], rad[j]
    eri, erj = ri + rp, rj + rp
    xi, xj = X[i], X[j]
    uij, tij, rij, between = pair_geom(xi, xj, ri, rj, rp)
    dn = (dens[i] + dens[j]) * 0.5
    rci = rij * ri / eri
    rcj = rij * rj / erj
    rb = torch.where(rij - rp > 0.0, rij - rp, torch.zeros_like(rij))
    rs = (rci + 2 * rb + rcj) * 0.25
    e = rs / rij
    edens = e * e * dn
    cx, cy = circle_frame(uij)
    tdelta = 1.0 / (torch.sqrt(edens) * rij)
    ts = (rij * TWO_PI) / M.to(torch.float64)
    pm = subdiv_point(tij, cx, cy, rij, tdelta, m)
    pi = (xi - pm) / col(eri)
    pj = (xj - pm) / col(erj)
    axis = unit(cross(pi, pj))
    dtq = rp * rp - rij * rij
    cusp = (dtq > 0.0) & between
    qij = tij - uij * col(safe_sqrt(dtq, cusp))
    tmp = unit(pi + pj)
    pqi = torch.where(col(cusp), (qij - pm) / rp, tmp)
    pqj = torch.where(col(cusp), torch.zeros_like(tmp), tmp)
    adelta = 1.0 / (torch.sqrt(dn) * rp)
    a1, y1 = arc_angle(pi, pqi, axis)
    a2, y2 = arc_angle(pqj, pj, axis)
    ang = torch.where(arc1, a1, a2)
    start = torch.where(col(arc1), pi, pqj)
    y = torch.where(col(arc1), y1, y2)
    pt = subdiv_point(pm, start, y, torch.full_like(rij, rp), adelta, k)
    ps = (rp * ang) / n.to(torch.float64)
    v = pt - tij
    dt = dseq(v, uij)
    d2_ = dseq(v, v) - dt * dt
    dist = torch.where(d2_ < 0.0, torch.zeros_like(d2_), safe_sqrt(d2_, d2_ >= 0.0))
    return pt, ps * ts * dist / rij


def concave(X, rad, dens, rp, d1, d2):
    mm = (d1[:, 0] >> 4) & 3
    a0, a1, a2 = d1[:, 1], d1[:, 2], d1[:, 3]
    t, k, L, Pc = d2[:, 0], d2[:, 1], d2[:, 2], d2[:, 3]
    plus = a0 < a1
    x0, x1, x2 = X[a0], X[a1], X[a2]
    xi = torch.where(col(plus), x0, x1)
    xj = torch.where(col(plus), x1, x0)
    xk = x2
    ri = rad[torch.where(plus, a0, a1)]
    rj = rad[torch.where(plus, a1, a0)]
    rk = rad[a2]
    eri, erk = ri + rp, rk + rp
    uij, tij, rij, _ = pair_geom(xi, xj, ri, rj, rp)
    dik = norm(xk - xi)
    uik = (xk - xi) / col(dik)
    dt = dalt(uij, uik)
    wk = torch.acos(torch.clamp(dt, -1.0, 1.0))
    sw = torch.sin(wk)
    uijk = cross(uij, uik) / col(sw)
    utb = cross(uijk, uij)
    asymm = (eri * eri - erk * erk) / dik
    tik = (xi + xk) * 0.5 + uik * col(asymm * 0.5)
    dtb = dalt(uik, tik - tij)
    bijk = tij + utb * col(dtb / sw)
    bi = bijk - xi
    h = torch.sqrt(eri * eri - dalt(bi, bi))
    p = torch.where(col(plus), bijk + uijk * col(h), bijk + uijk * col(-h))
    alt = torch.where(col(plus), uijk, -uijk)
    dn = ((dens[a0] + dens[a1]) + dens[a2]) / 3.0
    xa = torch.where(col(mm == 0), x0, torch.where(col(mm == 1), x1, x2))
    top = unit(xa - p)
    south = -alt
    axis = unit(cross(top, south))
    ang, y = arc_angle(top, south, axis)
    delta = 1.0 / (torch.sqrt(dn) * rp)
    cs = (rp * ang) / L.to(torch.float64)
    cx, cy = circle_frame(south)
    lat = subdiv_point(torch.zeros_like(p), top, y, torch.full_like(dn, rp), delta, t)
    dtl = dseq(lat, south)
    cen = south * col(dtl)
    rsq = rp * rp - dtl * dtl
    r = torch.where(rsq > 0.0, safe_sqrt(rsq, rsq > 0.0), torch.zeros_like(rsq))
    cdelta = 1.0 / (torch.sqrt(dn) * r)
    ps = (r * TWO_PI) / Pc.to(torch.float64)
    return subdiv_point(cen, cx, cy, r, cdelta, k) + p, ps * cs


def dots(X, rad, dens, rp, d1, d2):
    """Positions (n, 3) and areas (n,) of dots with descriptors d1, d2
    (int64 (n, 4) each) at atom coordinates X (N, 3)."""
    kind = d1[:, 0] & 15
    P = torch.zeros((len(d1), 3), dtype=X.dtype)
    A = torch.zeros(len(d1), dtype=X.dtype)
    for sel, fn in ((kind == DK_CONVEX, convex), ((kind == DK_ARC1) | (kind == DK_ARC2), toroidal),
                    (kind == DK_CONCAVE, concave)):
        idx = torch.nonzero(sel).flatten()
        if len(idx):
            p, a = fn(X, rad, dens, rp, d1[idx], d2[idx])
            P = P.index_put((idx,), p)
            A = A.index_put((idx,), a)
    return P, A


def frozen_cms(X, rad, dens, rp, weight, q_desc, r_desc, partner):
    """CMS of one side: query dots (buried, this side) with descriptors
    q_desc = (d1, d2), reference dots (buried, other side) r_desc, and the
    frozen nearest partner index of every query dot."""
    Pq, Aq = dots(X, rad, dens, rp, *q_desc)
    Pr, _ = dots(X, rad, dens, rp, *r_desc)
    d = Pq - Pr[partner]
    return torch.sum(Aq * torch.exp(-weight * dseq(d, d)))
