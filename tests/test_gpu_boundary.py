"""The drop-in boundary's multi-column solve, caller-supplied border blocks
and per-box factor read-back, on the GPU.

* A multi-column b is solved column by column as the reference's sweeps do
  (apply_sweep_matrix takes every column, skel_factorization.cpp:388-406):
  each column equals that column solved alone, bit for bit.
* skeletonize reads the caller's U and Psi (linear_solve.cpp:256-257): the
  solution satisfies the caller's bordered system (dense check).
* A box's T is the interpolative decomposition of its ID matrix: the
  oracle's ID matrix of the same box (the trees are identical) satisfies
  A(:, redundant) = A(:, skeleton) T to the ID tolerance.
"""
import numpy as np
import pytest

import geometry as G
import iesolver
import oracle

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _solver(g, pde=iesolver.PDE_LAPLACE, holes=True, tol=TOL, precision=iesolver.FP64):
    s = iesolver.RSSolver(0)
    s.set_geometry(g, pde)
    s.build_tree(64)
    if holes and "holes" in g:
        s.set_holes(g["holes"]["centers"], g["holes"]["radii"])
    s.factor(tol, precision)
    return s


@pytest.mark.parametrize("case", ["star", "star_fp32", "annulus", "stokes_holes"])
def test_multi_column_solve_equals_columns(case):
    if case == "stokes_holes":
        g, pde = G.circle_with_holes(4096, 4, 1), iesolver.PDE_STOKES
    elif case == "annulus":
        g, pde = G.annulus(4000), iesolver.PDE_LAPLACE
    else:
        g, pde = G.star(8192), iesolver.PDE_LAPLACE
    prec = iesolver.FP32 if case.endswith("fp32") else iesolver.FP64
    s = _solver(g, pde, tol=1e-6 if prec == iesolver.FP32 else TOL, precision=prec)
    B = np.random.default_rng(5).standard_normal((s.ndof, 3))
    MU, AL = s.solve_many(B)
    for j in range(3):
        mu, al = s.solve(B[:, j])
        assert np.array_equal(MU[:, j], mu), j
        if AL is not None:
            assert np.array_equal(AL[:, j], al), j


def test_callers_border_blocks_are_used():
    g = G.annulus(2000)
    s = iesolver.RSSolver(0)
    s.set_geometry(g)
    s.build_tree(64)
    s.set_holes(g["holes"]["centers"], g["holes"]["radii"])
    o = oracle.OracleProblem(g)
    U, Psi = o.U_Psi()
    U2 = 2 * U + 0.25  # not the formula's blocks
    s.set_border(U2, Psi)
    s.factor(TOL)
    f = G.harmonic_data(g["points"], *G.exterior_sources(3))
    mu, alpha = s.solve(f)
    K = oracle.dense_matrix(g)
    r1 = K @ mu + U2 @ alpha - f
    r2 = Psi @ mu - alpha
    err = max(np.abs(r1).max(), np.abs(r2).max())
    print(f"caller's border: residual {err:.2e}", flush=True)
    assert err <= 50 * TOL
    with pytest.raises(iesolver.IeGpuError):  # the border's width must match
        s.set_border(U2[:, :1], Psi[:1])


def test_box_factors_t_is_the_interpolative_decomposition():
    g = G.star(8192)
    s = _solver(g)
    o = oracle.OracleProblem(g, record_id_mats=True)
    o.factor(TOL, 4)
    info = s.box_info()
    nl = info["level"].max()
    boxes = [b for b in np.flatnonzero((info["level"] == nl) & (info["rank"] > 0))][:12]
    assert boxes
    for b in boxes:
        fac = s.box_factors(int(b))
        k, r = fac["k"], fac["r"]
        assert k == info["rank"][b] and k + r == info["n_active"][b]
        sk, rd = s.box_list(int(b), 3), s.box_list(int(b), 4)
        A = o.id_mat(int(b))
        pos = {d: i for i, d in enumerate(o.box_list(int(b), 2))}
        As, Ar = A[:, [pos[d] for d in sk]], A[:, [pos[d] for d in rd]]
        err = np.linalg.norm(Ar - As @ fac["T"]) / np.linalg.norm(A)
        assert err <= 20 * TOL, (b, err)
        # X_rr^{-1} U = X_rr^{-1} K_rs and L = K_sr X_rr^{-1}: both factor
        # through the same X_rr^{-1}, so L X_rr^{-1}^{-1} U = K_sr X^{-1} K_rs,
        # the Schur part of the carried block; shapes are the reference's
        assert fac["L"].shape == (k, r) and fac["U"].shape == (r, k)
        assert fac["Xinv"].shape == (r, r) and fac["S"].shape == (k, k)
        assert np.all(np.isfinite(fac["Xinv"]))
