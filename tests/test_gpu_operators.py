"""GPU operator apply, interior evaluation and FP32-preconditioned FP64
iterative refinement, against the oracle and dense LAPACK solves."""
import numpy as np
import pytest

import geometry as G
import iesolver
import oracle

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _solver(g, pde=iesolver.PDE_LAPLACE, tol=1e-10, precision=iesolver.FP64, leaf=64):
    s = iesolver.RSSolver(0)
    s.set_geometry(g, pde)
    s.build_tree(leaf)
    if g.get("holes") is not None:
        s.set_holes(g["holes"]["centers"], g["holes"]["radii"])
    s.factor(tol, precision)
    return s


@pytest.mark.parametrize("pde", [iesolver.PDE_LAPLACE, iesolver.PDE_STOKES])
def test_apply_matches_dense(pde):
    g = G.star(3000)
    s = _solver(g, pde)
    K = oracle.dense_matrix(g, pde)
    x = np.random.default_rng(0).standard_normal(K.shape[0])
    y, _ = s.apply(x)
    err = np.abs(y - K @ x).max() / np.abs(K @ x).max()
    print("apply rel err", err, flush=True)
    assert err < 1e-13


def test_apply_bordered_matches_dense():
    g = G.circle_with_holes(1600, 8, seed=1)
    s = _solver(g, iesolver.PDE_STOKES)
    o = oracle.OracleProblem(g, pde=oracle.PDE_STOKES)
    U, Psi = o.U_Psi()
    K = oracle.dense_matrix(g, oracle.PDE_STOKES)
    rng = np.random.default_rng(1)
    x = rng.standard_normal(K.shape[0])
    a = rng.standard_normal(U.shape[1])
    y, z = s.apply(x, a)
    assert np.abs(y - (K @ x + U @ a)).max() < 1e-12 * np.abs(K @ x).max()
    assert np.abs(z - (Psi @ x - a)).max() < 1e-12 * np.abs(a).max()


@pytest.mark.parametrize("pde", [iesolver.PDE_LAPLACE, iesolver.PDE_STOKES])
def test_evaluate_matches_oracle(pde):
    g = G.star(4096)
    s = _solver(g, pde)
    mu = np.random.default_rng(2).standard_normal(s.ndof)
    tg = G.interior_points(3, 300, radius=0.2)
    u = s.evaluate(tg, mu)
    uo = oracle.evaluate(g, tg, mu, pde)
    assert _rel(u, uo) < 1e-13


def test_evaluate_holes_and_mask():
    g = G.annulus(1000)
    s = _solver(g)
    mu = np.random.default_rng(4).standard_normal(s.ndof)
    alpha = np.array([0.3, -0.7])
    tg = np.array([[0.5, 0.9], [0.5, 0.2], [0.9, 0.5]])  # last one is a hole centre
    mask = np.array([1, 1, 0])
    u = s.evaluate(tg, mu, alpha, mask)
    c = g["holes"]["centers"]
    fwd = np.log(np.hypot(tg[:, None, 0] - c[None, :, 0], tg[:, None, 1] - c[None, :, 1])) @ alpha
    uo = oracle.evaluate(g, tg, mu) + fwd
    assert np.allclose(u[:2], uo[:2], rtol=1e-13, atol=0)
    assert u[2] == 0.0


def test_fp32_preconditioned_refinement_reaches_1e10():
    g = G.wavy(16384, seed=3)
    ys, q = G.exterior_sources(3)
    f = G.harmonic_data(g["points"], ys, q)
    s = _solver(g, tol=1e-6, precision=iesolver.FP32)
    mu, _, iters, rr = s.solve_refined(f, rtol=1e-12, max_iter=8)
    o = oracle.OracleProblem(g)
    o.factor(1e-12, 4)
    mo, _ = o.solve(f, 4)
    err = _rel(mu, mo)
    print(f"refined fp32: {iters} corrections, rel residual {rr:.2e}, rel err vs oracle {err:.2e}",
          flush=True)
    assert rr <= 1e-12
    assert err <= 1e-10


def test_refinement_stokes_holes():
    g = G.circle_with_holes(2048, 8, seed=0)
    f = G.tangent_data(g["normals"])
    f[2 * g["counts"][0]:] = 0
    s = _solver(g, iesolver.PDE_STOKES, tol=1e-6, precision=iesolver.FP32)
    mu, alpha, iters, rr = s.solve_refined(f, rtol=1e-12, max_iter=10)
    o = oracle.OracleProblem(g, pde=oracle.PDE_STOKES)
    U, Psi = o.U_Psi()
    K = oracle.dense_matrix(g, oracle.PDE_STOKES)
    h = U.shape[1]
    D = np.block([[K, U], [Psi, -np.eye(h)]])
    xd = np.linalg.solve(D, np.concatenate([f, np.zeros(h)]))
    err = _rel(np.concatenate([mu, alpha]), xd)
    print(f"stokes holes refined: {iters} corrections, res {rr:.2e}, err {err:.2e}", flush=True)
    assert rr <= 1e-12
    assert err <= 1e-9


def test_fp64_refinement_is_noop_when_converged():
    g = G.star(4096)
    ys, q = G.exterior_sources(1)
    f = G.harmonic_data(g["points"], ys, q)
    s = _solver(g)
    mu0, _ = s.solve(f)
    mu, _, iters, rr = s.solve_refined(f, rtol=1e-9, max_iter=5)
    assert iters == 0 and rr <= 1e-9
    assert np.array_equal(mu, mu0)


def test_stokes_holes_full_size_fp32_refined():
    # The Stokes workload at its full size: outer circle r = 0.5 plus 8
    # seeded holes (radii U[0.02, 0.05]), N = 131072 points (262,144 + 24
    # unknowns).  The FP32 factorization preconditions FP64 refinement on
    # the bordered operator down to a relative residual of 1e-10; its
    # interior field is checked against the exact discrete solution (the
    # plain FP64 factorization refined to a relative residual of 1e-15).
    g = G.circle_with_holes(131072, 8, seed=0)
    f = G.tangent_data(g["normals"])
    f[2 * g["counts"][0]:] = 0
    m = _solver(g, iesolver.PDE_STOKES, tol=1e-6, precision=iesolver.FP32)
    mu, alpha, iters, rr = m.solve_refined(f, rtol=1e-10, max_iter=10)
    print(f"stokes N=131072 holes: {iters} corrections, rel residual {rr:.2e}", flush=True)
    assert rr <= 1e-10 and iters <= 6
    p = _solver(g, iesolver.PDE_STOKES, tol=1e-10)
    mp, ap = p.solve(f)
    mx, ax, _, rx = p.solve_refined(f, rtol=1e-15, max_iter=6)
    assert rx <= 1e-12  # the FP64 operator's floor on this bordered system is ~5e-13
    rng = np.random.default_rng(3)
    tg = 0.5 + np.stack([rng.uniform(-0.4, 0.4, 400), rng.uniform(-0.4, 0.4, 400)], 1)
    c, r = g["holes"]["centers"], g["holes"]["radii"]
    inside = np.hypot(tg[:, 0] - 0.5, tg[:, 1] - 0.5) < 0.45
    for k in range(len(r)):
        inside &= np.hypot(tg[:, 0] - c[k, 0], tg[:, 1] - c[k, 1]) > r[k] + 0.01
    tg = tg[inside]
    ux = p.evaluate(tg, mx, ax)
    err = _rel(p.evaluate(tg, mu, alpha), ux)
    err_plain = _rel(p.evaluate(tg, mp, ap), ux)
    print(f"interior field vs exact over {len(tg)} points: FP32-refined {err:.2e}, "
          f"plain FP64 {err_plain:.2e}", flush=True)
    assert err <= 1e-9
