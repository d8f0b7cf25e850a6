"""Laplace Neumann on the GPU (ie_solver/kernel/kernel.cpp:80-102, 168-209,
594-646; initialization.cpp:66-97): the operator is the adjoint double
layer (target normal) plus the rank-one term w_j, the proxy block follows
the same formula, and the interior field is the single layer.  Parity with
the CPU oracle at 10 x tol, the reference's backward-error and donut tests,
and the randomized and FP32 forms on the same kernel."""
import math

import numpy as np
import pytest

import geometry as G
import iesolver
import oracle

pytestmark = pytest.mark.gpu

NEU = iesolver.PDE_LAPLACE_NEUMANN


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _gpu(g, tol, precision=iesolver.FP64, method=None):
    s = iesolver.RSSolver(0)
    s.set_geometry(g, NEU)
    s.build_tree(64)
    if "holes" in g:
        s.set_holes(g["holes"]["centers"], g["holes"]["radii"])
        assert s.hole_dim == 0  # no border for Laplace Neumann
    if method is not None:
        s.set_id_method(*method)
    s.factor(tol, precision)
    return s


def test_neumann_vs_oracle():
    tol = 1e-10
    g = G.star(4096, amp=0.2, k=7)
    f = G.harmonic_data(g["points"], *G.exterior_sources(5))
    s = _gpu(g, tol)
    mu, _ = s.solve(f)
    o = oracle.OracleProblem(g, pde=oracle.PDE_LAPLACE_NEUMANN)
    o.factor(tol, 4)
    mo, _ = o.solve(f, 4)
    err = _rel(mu, mo)
    print(f"neumann star rel err vs oracle {err:.2e}", flush=True)
    assert err <= 10 * tol
    gi, oi = s.box_info(), o.box_info()
    comp = (gi["rank"] > 0) | (oi["rank"] > 0)
    assert np.abs(gi["rank"][comp] - oi["rank"][comp]).max() <= 2


@pytest.mark.parametrize("shape", ["circle", "star", "annulus"])
def test_reference_neumann_backward_error(shape):
    # ie_solver_test.cpp:124-135, 165-176, 209-220 (check_solve_err): all
    # ones, N = 512, tol 1e-6, |K mu - f| <= 50 tol.
    tol = 1e-6
    g = {"circle": G.circle(512), "star": G.star(512), "annulus": G.annulus(512)}[shape]
    f = np.ones(g["points"].shape[0])
    s = _gpu(g, tol)
    mu, alpha = s.solve(f)
    assert alpha is None
    K = oracle.dense_matrix(g, oracle.PDE_LAPLACE_NEUMANN)
    assert np.abs(K @ mu - f).max() <= 50 * tol


def test_neumann_apply_and_evaluate_match_oracle():
    g = G.star(3000, amp=0.15, k=4)
    x = np.random.default_rng(4).standard_normal(3000)
    s = iesolver.RSSolver(0)
    s.set_geometry(g, NEU)
    y, _ = s.apply(x)
    K = oracle.dense_matrix(g, oracle.PDE_LAPLACE_NEUMANN)
    assert _rel(y, K @ x) <= 1e-13
    tg = G.interior_points(3, 50, radius=0.15)
    u = s.evaluate(tg, x)
    assert _rel(u, oracle.evaluate(g, tg, x, oracle.PDE_LAPLACE_NEUMANN)) <= 1e-13


def test_reference_neumann_donut_analytic():
    # ie_solver_test.cpp:433-471, 510-525: the field minus log r is constant.
    tol = 1e-6
    g = G.donut(512)
    n, hn = g["counts"]
    f = np.concatenate([np.ones(n), np.full(hn, -2.0)])
    s = _gpu(g, tol)
    mu, _ = s.solve(f)
    rng = np.random.default_rng(2)
    rr = rng.uniform(0.6, 0.9, 100)
    th = rng.uniform(0, 2 * math.pi, 100)
    tg = np.stack([0.5 + rr * np.cos(th), 0.5 + rr * np.sin(th)], 1)
    d = s.evaluate(tg, mu) - np.log(rr)
    print(f"donut deviation {np.abs(d - d.mean()).max():.2e}", flush=True)
    assert np.abs(d - d.mean()).max() <= 10 * tol


def test_neumann_randomized_and_fp32_forms():
    tol = 1e-10
    g = G.wavy(16384, seed=6)
    f = G.harmonic_data(g["points"], *G.exterior_sources(6))
    ref = _gpu(g, tol)
    mu, _ = ref.solve(f)
    r = _gpu(g, tol, method=(iesolver.ID_RANDOMIZED, 32))
    mr, _ = r.solve(f)
    y, _ = r.apply(mr)
    assert np.abs(y - f).max() <= 50 * tol * np.abs(f).max()
    assert _rel(mr, mu) <= 10 * tol
    m = _gpu(g, 1e-6, precision=iesolver.FP32)
    x, _, it, rel = m.solve_refined(f, rtol=1e-12, max_iter=10)
    print(f"neumann fp32: {it} corrections, rel residual {rel:.2e}", flush=True)
    assert rel <= 1e-12 and _rel(x, mu) <= 10 * tol
