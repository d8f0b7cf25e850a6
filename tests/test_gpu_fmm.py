"""The fast multipole form of the FP64 operator apply (rs_fmm.hip) against
the direct sum (rs_matvec.hip, itself checked against the oracle's dense
kernel matrix in test_gpu_operators.py and test_gpu_status.py).

* Laplace, Laplace Neumann and Stokes at up to 65,536 nodes: FMM (order 48)
  within 1e-13 relative of the direct apply; orders 32 and 64 within their
  truncation bounds.
* Coincident nodes (the diagonal rule), a boundary too small for more than
  one level, and a geometry update (the tree follows the new nodes).
* FP32 factorization + FP64 refinement with the FMM residual converges to
  the FP64 solution; a batch keeps the direct sum.
"""
import numpy as np
import pytest

import geometry as G
import iesolver

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _both(s, x, alpha=None, **fmm):
    s.set_matvec(iesolver.MATVEC_DIRECT)
    yd, zd = s.apply(x, alpha)
    assert s.matvec_info()[0] == iesolver.MATVEC_DIRECT
    s.set_matvec(iesolver.MATVEC_FMM, **fmm)
    yf, zf = s.apply(x, alpha)
    info = s.matvec_info()
    assert info[0] == iesolver.MATVEC_FMM and info[1] >= 2 and info[3] > 0
    return yd, yf, info


@pytest.mark.parametrize("pde,geom,n", [
    (iesolver.PDE_LAPLACE, "star", 65536),
    (iesolver.PDE_LAPLACE, "wavy", 65536),
    (iesolver.PDE_LAPLACE_NEUMANN, "star", 32768),
    (iesolver.PDE_STOKES, "star", 65536),
    (iesolver.PDE_STOKES, "wavy", 32768),
])
def test_fmm_matches_direct(pde, geom, n):
    g = G.star(n) if geom == "star" else G.wavy(n, seed=2)
    s = iesolver.RSSolver(0)
    s.set_geometry(g, pde)
    x = np.random.default_rng(n + pde).standard_normal(s.ndof)
    yd, yf, info = _both(s, x)
    err = _rel(yf, yd)
    print(f"pde {pde} {geom} N={n}: depth {info[1]}, {info[3]} leaves, rel {err:.2e}", flush=True)
    assert err <= 1e-13


@pytest.mark.parametrize("order,bound", [(32, 1e-9), (64, 1e-13)])
def test_fmm_orders(order, bound):
    g = G.star(32768)
    s = iesolver.RSSolver(0)
    s.set_geometry(g, iesolver.PDE_STOKES)
    x = np.random.default_rng(3).standard_normal(s.ndof)
    yd, yf, _ = _both(s, x, order=order, leaf_points=48)
    err = _rel(yf, yd)
    print(f"order {order}: rel {err:.2e}", flush=True)
    assert err <= bound


@pytest.mark.parametrize("pde", [iesolver.PDE_LAPLACE, iesolver.PDE_STOKES])
def test_fmm_coincident_and_tiny(pde):
    g = G.star(4096)
    g["points"][1] = g["points"][0]  # a coincident pair: the diagonal rule
    s = iesolver.RSSolver(0)
    s.set_geometry(g, pde)
    x = np.random.default_rng(4).standard_normal(s.ndof)
    yd, yf, _ = _both(s, x)
    assert _rel(yf, yd) <= 1e-13
    t = G.circle(40)  # one level of boxes only
    s.set_geometry(t, pde)
    x = np.random.default_rng(5).standard_normal(s.ndof)
    yd, yf, info = _both(s, x)
    assert info[1] == 2
    assert _rel(yf, yd) <= 1e-13


def test_fmm_follows_geometry_update():
    g = G.star(32768)
    s = iesolver.RSSolver(0)
    s.set_geometry(g)
    s.build_tree(64)
    x = np.random.default_rng(6).standard_normal(s.ndof)
    _both(s, x)
    g2 = G.bump_star(32768, 0.02, 1.0)
    s.update_geometry(g2)
    yd, yf, _ = _both(s, x)
    assert _rel(yf, yd) <= 1e-13
    s.set_geometry(g2)
    s.set_matvec(iesolver.MATVEC_DIRECT)
    y2, _ = s.apply(x)
    assert np.array_equal(y2, yd)


def test_refinement_with_fmm_residual():
    g = G.wavy(65536, seed=0)
    ys, q = G.exterior_sources(0)
    f = G.harmonic_data(g["points"], ys, q)
    s = iesolver.RSSolver(0)
    s.set_geometry(g)
    s.build_tree(64)
    s.factor(1e-10)
    mu64, _ = s.solve(f)
    s.factor(1e-6, iesolver.FP32)
    s.set_matvec(iesolver.MATVEC_AUTO)
    mu, _, it, rr = s.solve_refined(f, rtol=1e-12, max_iter=20)
    assert s.matvec_info()[0] == iesolver.MATVEC_FMM
    print(f"FP32 + FMM refinement: {it} corrections, relres {rr:.2e}, "
          f"vs FP64 solve {_rel(mu, mu64):.2e}", flush=True)
    assert rr <= 1e-12
    assert _rel(mu, mu64) <= 1e-9


def test_batch_keeps_direct_sum():
    geoms, _ = G.star_batch(4, seed=1, nmin=2048, nmax=4096)
    s = iesolver.RSSolver(0)
    s.set_geometry_batch(geoms)
    s.set_matvec(iesolver.MATVEC_FMM)
    x = np.random.default_rng(7).standard_normal(s.ndof)
    s.apply(x)
    assert s.matvec_info()[0] == iesolver.MATVEC_DIRECT
