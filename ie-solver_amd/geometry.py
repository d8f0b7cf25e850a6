"""Boundary discretizations (SoA arrays) for the RS factor + solve path.

Each generator returns a dict with the arrays the path consumes, in the
layout of the reference's Boundary (ie_solver/boundaries/boundary.h:34-41):

    points      (N, 2)  node coordinates
    normals     (N, 2)  unit normals, pointing out of the fluid/potential domain
    curvatures  (N,)    signed curvature (1/R for a circle with outward normal)
    weights     (N,)    trapezoid weights |x'(t)| * 2*pi/N_curve
    holes       {"centers": (h,2), "radii": (h,)}  (multiply connected only)

plus boundary data helpers.  Geometry generation is outside the timed
factor+solve region (as in bie_time_trial, ie_solver/linear_solve.cpp:156-185).
"""
import math

import numpy as np


def _radial_curve(n, center, r, dr, ddr, t):
    """Closed curve x(t) = c + r(t) (cos t, sin t), counter-clockwise."""
    c, s = np.cos(t), np.sin(t)
    x = np.stack([center[0] + r * c, center[1] + r * s], axis=1)
    dx = np.stack([dr * c - r * s, dr * s + r * c], axis=1)
    ddx = np.stack([ddr * c - 2 * dr * s - r * c, ddr * s + 2 * dr * c - r * s], axis=1)
    speed = np.sqrt(dx[:, 0] ** 2 + dx[:, 1] ** 2)
    normals = np.stack([dx[:, 1] / speed, -dx[:, 0] / speed], axis=1)
    curv = (dx[:, 0] * ddx[:, 1] - dx[:, 1] * ddx[:, 0]) / speed ** 3
    weights = speed * (2 * math.pi / n)
    return x, normals, curv, weights


def circle(n, center=(0.5, 0.5), radius=0.5):
    """Circle with equispaced nodes theta_i = 2*pi*i/N (ie_solver/boundaries/circle.cpp:15-25)."""
    i = np.arange(n, dtype=np.float64)
    ang = i * 2.0 * math.pi / n
    pts = np.stack([center[0] + radius * np.cos(ang), center[1] + radius * np.sin(ang)], axis=1)
    nrm = np.stack([np.cos(ang), np.sin(ang)], axis=1)
    return dict(points=pts, normals=nrm, curvatures=np.full(n, 1.0 / radius),
                weights=np.full(n, 2.0 * math.pi * radius / n))


def reference_circle(n):
    """The reference's default Circle: centre (0.5,0.5), radius 0.25, w = pi/(2N)."""
    g = circle(n, (0.5, 0.5), 0.25)
    g["curvatures"] = np.full(n, 4.0)
    g["weights"] = np.full(n, math.pi / (n * 2))
    return g


def star(n, center=(0.5, 0.5), r0=0.3, amp=0.2, k=7):
    """r(t) = r0 (1 + amp cos(k t)), equispaced t."""
    t = np.arange(n, dtype=np.float64) * (2 * math.pi / n)
    r = r0 * (1 + amp * np.cos(k * t))
    dr = -r0 * amp * k * np.sin(k * t)
    ddr = -r0 * amp * k * k * np.cos(k * t)
    p, nm, cv, w = _radial_curve(n, center, r, dr, ddr, t)
    return dict(points=p, normals=nm, curvatures=cv, weights=w)


def bump_star(n, bump, t0, width=0.4, center=(0.5, 0.5), r0=0.3, amp=0.2, k=7):
    """star() with r(t) raised by r0 * bump * (1 - u^2)^3 on the arc
    |u| < 1, u = (t - t0) / width (t - t0 wrapped to [-pi, pi)): a local,
    twice continuously differentiable deformation.  Nodes off the arc are
    bit-identical to star()'s."""
    t = np.arange(n, dtype=np.float64) * (2 * math.pi / n)
    r = r0 * (1 + amp * np.cos(k * t))
    dr = -r0 * amp * k * np.sin(k * t)
    ddr = -r0 * amp * k * k * np.cos(k * t)
    _add_bump(t, r, dr, ddr, r0 * bump, t0, width)
    p, nm, cv, w = _radial_curve(n, center, r, dr, ddr, t)
    return dict(points=p, normals=nm, curvatures=cv, weights=w)


def _add_bump(t, r, dr, ddr, h, t0, width):
    """r += h (1 - u^2)^3 on |u| < 1, u = (t - t0) / width wrapped, with its
    first and second derivatives; r, dr, ddr are updated in place."""
    u = (np.mod(t - t0 + math.pi, 2 * math.pi) - math.pi) / width
    on = np.abs(u) < 1
    uo = u[on]
    g = 1 - uo * uo
    r[on] += h * g ** 3
    dr[on] += h * (-6 * uo * g * g / width)
    ddr[on] += h * (-6 * g * (1 - 5 * uo * uo) / (width * width))


def bump_wavy(n, bump, t0, width=0.2, seed=0, center=(0.5, 0.5), r0=0.3):
    """wavy() with the local bump of bump_star(); nodes off the arc are
    bit-identical to wavy()'s."""
    ks, a = wavy_coefficients(seed)
    t = np.arange(n, dtype=np.float64) * (2 * math.pi / n)
    kt = np.outer(t, ks)
    r = r0 * (1 + np.cos(kt) @ a)
    dr = -r0 * (np.sin(kt) @ (a * ks))
    ddr = -r0 * (np.cos(kt) @ (a * ks * ks))
    _add_bump(t, r, dr, ddr, r0 * bump, t0, width)
    p, nm, cv, w = _radial_curve(n, center, r, dr, ddr, t)
    return dict(points=p, normals=nm, curvatures=cv, weights=w)


def wavy_coefficients(seed, kmin=2, kmax=12):
    """a_k ~ N(0, 0.01/k) (variance 0.01/k), k = kmin..kmax."""
    rng = np.random.default_rng(seed)
    ks = np.arange(kmin, kmax + 1)
    return ks, rng.normal(0.0, np.sqrt(0.01 / ks))


def wavy(n, seed=0, center=(0.5, 0.5), r0=0.3):
    """r(t) = r0 (1 + sum_{k=2..12} a_k cos(k t)), a_k seeded."""
    ks, a = wavy_coefficients(seed)
    t = np.arange(n, dtype=np.float64) * (2 * math.pi / n)
    kt = np.outer(t, ks)
    r = r0 * (1 + np.cos(kt) @ a)
    dr = -r0 * (np.sin(kt) @ (a * ks))
    ddr = -r0 * (np.cos(kt) @ (a * ks * ks))
    p, nm, cv, w = _radial_curve(n, center, r, dr, ddr, t)
    return dict(points=p, normals=nm, curvatures=cv, weights=w)


def _hole_nodes(n, center, radius):
    i = np.arange(n, dtype=np.float64)
    ang = i * 2.0 * math.pi / n
    pts = np.stack([center[0] + radius * np.cos(ang), center[1] + radius * np.sin(ang)], axis=1)
    nrm = np.stack([-np.cos(ang), -np.sin(ang)], axis=1)
    return pts, nrm, np.full(n, -1.0 / radius), np.full(n, 2 * radius * math.pi / n)


def circle_with_holes(n_total, n_holes=8, seed=0, center=(0.5, 0.5), radius=0.5,
                      rmin=0.02, rmax=0.05, clearance=0.02):
    """Outer circle plus seeded circular holes (normals into each hole, as
    ie_solver/boundaries/annulus.cpp:40-55).  Radii U[rmin, rmax]; centres by
    rejection sampling, non-overlapping and clear of the outer wall.  The
    N nodes are split between curves in proportion to perimeter; the outer
    curve comes first."""
    rng = np.random.default_rng(seed)
    radii = rng.uniform(rmin, rmax, n_holes)
    centers = []
    for h in range(n_holes):
        for _ in range(100000):
            rho = (radius - radii[h] - clearance) * math.sqrt(rng.uniform())
            phi = rng.uniform(0, 2 * math.pi)
            c = (center[0] + rho * math.cos(phi), center[1] + rho * math.sin(phi))
            if all(math.hypot(c[0] - o[0], c[1] - o[1]) > radii[h] + radii[j] + clearance
                   for j, o in enumerate(centers)):
                centers.append(c)
                break
        else:
            raise RuntimeError("could not place hole")
    perim = np.concatenate([[radius], radii]) * 2 * math.pi
    counts = np.floor(n_total * perim / perim.sum()).astype(int)
    counts[0] += n_total - counts.sum()
    g = circle(int(counts[0]), center, radius)
    parts = [(g["points"], g["normals"], g["curvatures"], g["weights"])]
    for h in range(n_holes):
        parts.append(_hole_nodes(int(counts[h + 1]), centers[h], radii[h]))
    return dict(points=np.concatenate([p[0] for p in parts]),
                normals=np.concatenate([p[1] for p in parts]),
                curvatures=np.concatenate([p[2] for p in parts]),
                weights=np.concatenate([p[3] for p in parts]),
                holes=dict(centers=np.array(centers), radii=radii),
                counts=counts)


def annulus(n):
    """The reference's Annulus (ie_solver/boundaries/annulus.cpp:9-55): unit
    circle about (0.5,0.5) with holes r=0.2 at (0.9,0.5) and (0.1,0.5), N/5
    nodes each."""
    outer = dict(points=None)
    i = np.arange(n, dtype=np.float64)
    ang = i * 2.0 * math.pi / n
    pts = [np.stack([0.5 + np.cos(ang), 0.5 + np.sin(ang)], axis=1)]
    nrm = [np.stack([np.cos(ang), np.sin(ang)], axis=1)]
    cv = [np.ones(n)]
    w = [np.full(n, 2 * math.pi / n)]
    centers = np.array([[0.9, 0.5], [0.1, 0.5]])
    radii = np.array([0.2, 0.2])
    hn = n // 5
    for h in range(2):
        p, q, c, ww = _hole_nodes(hn, centers[h], radii[h])
        pts.append(p)
        nrm.append(q)
        cv.append(c)
        w.append(ww)
    del outer
    return dict(points=np.concatenate(pts), normals=np.concatenate(nrm),
                curvatures=np.concatenate(cv), weights=np.concatenate(w),
                holes=dict(centers=centers, radii=radii))


def donut(n):
    """The reference's Donut (ie_solver/boundaries/donut.cpp:9-52): circle of
    radius 1 about (0.5,0.5) (n nodes) around a hole of radius 0.5 at the
    same centre (n/5 nodes, inward normals, curvature -2)."""
    i = np.arange(n, dtype=np.float64)
    ang = i * 2.0 * math.pi / n
    hn = n // 5
    a2 = np.arange(hn, dtype=np.float64) * 2.0 * math.pi / hn
    pts = np.concatenate([np.stack([0.5 + np.cos(ang), 0.5 + np.sin(ang)], 1),
                          np.stack([0.5 + 0.5 * np.cos(a2), 0.5 + 0.5 * np.sin(a2)], 1)])
    nrm = np.concatenate([np.stack([np.cos(ang), np.sin(ang)], 1),
                          np.stack([-np.cos(a2), -np.sin(a2)], 1)])
    return dict(points=pts, normals=nrm,
                curvatures=np.concatenate([np.ones(n), np.full(hn, -2.0)]),
                weights=np.concatenate([np.full(n, 2 * math.pi / n), np.full(hn, math.pi / hn)]),
                holes=dict(centers=np.array([[0.5, 0.5]]), radii=np.array([0.5])),
                counts=np.array([n, hn]))


def star_batch(n_problems, seed=0, nmin=2048, nmax=32768):
    """Independent star boundaries for a batch: sizes log-uniform in
    [nmin, nmax], k uniform in {3..9}, amplitude U[0.05, 0.25], seeded.
    Returns the geometries and each one's (sizes, k, amp)."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(math.log(nmin), math.log(nmax), n_problems))
    sizes = np.clip(np.rint(sizes).astype(int), nmin, nmax)
    ks = rng.integers(3, 10, n_problems)
    amps = rng.uniform(0.05, 0.25, n_problems)
    geoms = [star(int(n), amp=float(a), k=int(k)) for n, k, a in zip(sizes, ks, amps)]
    return geoms, dict(sizes=sizes, k=ks, amp=amps)


def batch_data(geoms, seed=0):
    """Harmonic data per problem (4 seeded exterior sources each),
    concatenated."""
    return np.concatenate([harmonic_data(g["points"], *exterior_sources(seed + 1000 * i))
                           for i, g in enumerate(geoms)])


def exterior_sources(seed, n_src=4, center=(0.5, 0.5), rmin=0.7, rmax=1.0):
    """Seeded point sources outside the curve: radius U[rmin,rmax] from the
    centre, angle U[0,2pi), strengths N(0,1)."""
    rng = np.random.default_rng(seed)
    rho = rng.uniform(rmin, rmax, n_src)
    phi = rng.uniform(0, 2 * math.pi, n_src)
    ys = np.stack([center[0] + rho * np.cos(phi), center[1] + rho * np.sin(phi)], axis=1)
    q = rng.normal(0.0, 1.0, n_src)
    return ys, q


def harmonic_data(points, sources, strengths):
    """f(x) = sum_k q_k log|x - y_k| (harmonic inside the curve)."""
    f = np.zeros(points.shape[0])
    for y, q in zip(sources, strengths):
        f += q * np.log(np.hypot(points[:, 0] - y[0], points[:, 1] - y[1]))
    return f


def interior_points(seed, n=100, center=(0.5, 0.5), radius=0.4):
    """Area-uniform points in a disc: r = radius*sqrt(U), phi = 2*pi*V."""
    rng = np.random.default_rng(seed)
    u = rng.uniform(size=n)
    v = rng.uniform(size=n)
    r = radius * np.sqrt(u)
    phi = 2 * math.pi * v
    return np.stack([center[0] + r * np.cos(phi), center[1] + r * np.sin(phi)], axis=1)


def tangent_data(normals):
    """Stokes boundary velocity (-n_y, n_x) per node, interleaved by DOF
    (ie_solver/boundaries/boundary.cpp:70-74)."""
    f = np.zeros(2 * normals.shape[0])
    f[0::2] = -normals[:, 1]
    f[1::2] = normals[:, 0]
    return f
