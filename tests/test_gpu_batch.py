"""A batch of independent problems factored and solved in one context.

The batch's quadtree is the problems' trees merged level by level, so each
stage of a level covers every problem in one launch.  Parity: each
problem's slice of the batch solution against the same problem factored
alone (same kernels, same per-box inputs: ranks equal, solutions to
rounding) and against the CPU oracle (relative L2 <= 10 x ID tolerance).
"""
import numpy as np
import pytest

import geometry as G
import iesolver
import oracle

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _batch(geoms, tol=TOL, precision=iesolver.FP64, method=None):
    s = iesolver.RSSolver(0)
    s.set_geometry_batch(geoms)
    s.build_tree(64)
    if method is not None:
        s.set_id_method(*method)
    s.factor(tol, precision)
    return s


def _single(g, tol=TOL):
    s = iesolver.RSSolver(0)
    s.set_geometry(g)
    s.build_tree(64)
    s.factor(tol)
    return s


def test_batch_matches_problems_solved_alone():
    geoms, info = G.star_batch(6, seed=1, nmin=512, nmax=4096)
    print("sizes", info["sizes"].tolist(), "k", info["k"].tolist(), flush=True)
    f = G.batch_data(geoms, seed=1)
    s = _batch(geoms)
    mu, _ = s.solve(f)
    off = s.dof_offsets
    bi = s.box_info()
    id0 = 0
    for p, g in enumerate(geoms):
        t = _single(g)
        sl = slice(off[p], off[p + 1])
        mp, _ = t.solve(f[sl])
        ti = t.box_info()
        nb = len(ti["rank"])
        assert np.array_equal(bi["rank"][id0:id0 + nb], ti["rank"]), p
        assert np.array_equal(bi["n_active"][id0:id0 + nb], ti["n_active"]), p
        err = _rel(mu[sl], mp)
        print(f"problem {p} n={len(mp)} batch vs alone {err:.2e}", flush=True)
        assert err <= 1e-12
        id0 += nb


def test_batch_with_large_roots_matches_problems_alone():
    # Two of the problems have root skeletons above 192 DOFs (blocked
    # Gauss-Jordan root inverse), the others below (register inverse).
    geoms = [G.star(4096, amp=0.25, k=9), G.star(1500, amp=0.1, k=4),
             G.star(8192, amp=0.233, k=8), G.star(700, amp=0.05, k=3)]
    f = G.batch_data(geoms, seed=7)
    s = _batch(geoms)
    mu, _ = s.solve(f)
    off = s.dof_offsets
    roots = s.box_info()["n_active"][s.box_list(0, 5)]
    print("root sizes", roots.tolist(), flush=True)
    assert (roots > 192).sum() == 2
    for p, g in enumerate(geoms):
        t = _single(g)
        sl = slice(off[p], off[p + 1])
        mp, _ = t.solve(f[sl])
        err = _rel(mu[sl], mp)
        print(f"problem {p} batch vs alone {err:.2e}", flush=True)
        assert err <= 1e-12


def test_batch_vs_oracle():
    geoms, _ = G.star_batch(3, seed=2, nmin=1024, nmax=3072)
    f = G.batch_data(geoms, seed=2)
    s = _batch(geoms)
    mu, _ = s.solve(f)
    off = s.dof_offsets
    for p, g in enumerate(geoms):
        sl = slice(off[p], off[p + 1])
        o = oracle.OracleProblem(g)
        o.factor(TOL, 4)
        mo, _ = o.solve(f[sl], 4)
        err = _rel(mu[sl], mo)
        print(f"problem {p} rel err vs oracle {err:.2e}", flush=True)
        assert err <= 10 * TOL


def test_batch_operator_is_block_diagonal():
    geoms, _ = G.star_batch(4, seed=3, nmin=300, nmax=2000)
    s = iesolver.RSSolver(0)
    s.set_geometry_batch(geoms)
    x = np.random.default_rng(3).standard_normal(s.ndof)
    y, _ = s.apply(x)
    off = s.dof_offsets
    for p, g in enumerate(geoms):
        t = iesolver.RSSolver(0)
        t.set_geometry(g)
        sl = slice(off[p], off[p + 1])
        yp, _ = t.apply(x[sl])
        # same sources in the same order per target: identical sums
        assert np.array_equal(y[sl], yp), p


def test_batch_fp32_refined_and_randomized():
    geoms, _ = G.star_batch(5, seed=4, nmin=1024, nmax=8192)
    f = G.batch_data(geoms, seed=4)
    ref = _batch(geoms)
    mu, _ = ref.solve(f)
    # randomized ID (32-row sketches) in FP64
    r = _batch(geoms, method=(iesolver.ID_RANDOMIZED, 32))
    mr, _ = r.solve(f)
    assert _rel(mr, mu) <= 10 * TOL
    # FP32 factorization preconditioning FP64 refinement on the batch's
    # block-diagonal operator
    m = _batch(geoms, tol=1e-6, precision=iesolver.FP32)
    x, _, it, rel = m.solve_refined(f, rtol=1e-12, max_iter=10)
    print(f"fp32 batch: {it} corrections, rel residual {rel:.2e}, "
          f"vs fp64 {_rel(x, mu):.2e}", flush=True)
    assert rel <= 1e-12 and it <= 6
    assert _rel(x, mu) <= 10 * TOL


def test_batch_graph_replays_are_exact():
    geoms, _ = G.star_batch(4, seed=5, nmin=512, nmax=4096)
    f = G.batch_data(geoms, seed=5)
    s = iesolver.RSSolver(0)
    s.set_geometry_batch(geoms)
    s.build_tree(64)
    s.factor(TOL)
    m0, _ = s.solve(f)
    s.set_graphs(True)
    for _ in range(3):
        s.factor(TOL)
        m1, _ = s.solve(f)
        assert np.array_equal(m0, m1)


def test_autotune_and_launch_caps_keep_results():
    # The caps change launch widths only: the solution is bit-identical for
    # any caps, including the autotuner's choice and one workgroup each.
    geoms, _ = G.star_batch(6, seed=6, nmin=1024, nmax=8192)
    f = G.batch_data(geoms, seed=6)
    s = _batch(geoms)
    m0, _ = s.solve(f)
    caps, ms = s.autotune(TOL, reps=1)
    print("autotuned caps", caps.tolist(), f"{ms:.3f} ms", flush=True)
    assert set(caps[:6].tolist()) <= {64, 256, 1024, 4096} and ms > 0
    assert caps[6] in (1, 2048, 8192, 1 << 20)
    assert caps[7] in (1, 2048, 4096, 1 << 20)
    assert np.array_equal(s.launch_caps(), caps)
    m1, _ = s.solve(f)
    # The decouple form (cap 7) moves a level between two eliminations with
    # the same factors up to rounding; the other caps are launch widths.
    assert np.linalg.norm(m1 - m0) <= 1e-11 * np.linalg.norm(m0)
    s.set_launch_caps([1] * 7 + [4096])
    s.factor(TOL)
    m2, _ = s.solve(f)
    assert np.array_equal(m0, m2)


@pytest.mark.parametrize("pde", [iesolver.PDE_STOKES, iesolver.PDE_LAPLACE_NEUMANN])
def test_batch_operator_with_rank_one_terms(pde):
    # Stokes (its n n^T part) and Laplace Neumann (its w_j part) sum a
    # global term per problem: the batch's operator equals each problem's
    # own, and FP32 + refinement converges on the batch.
    geoms, _ = G.star_batch(3, seed=10, nmin=400, nmax=1500)
    s = iesolver.RSSolver(0)
    s.set_geometry_batch(geoms, pde)
    x = np.random.default_rng(10).standard_normal(s.ndof)
    y, _ = s.apply(x)
    off = s.dof_offsets
    for p, g in enumerate(geoms):
        t = iesolver.RSSolver(0)
        t.set_geometry(g, pde)
        sl = slice(off[p], off[p + 1])
        assert np.array_equal(y[sl], t.apply(x[sl])[0]), p
    s.build_tree(64)
    s.factor(1e-6, iesolver.FP32)
    f = s.apply(np.ones(s.ndof))[0]
    xr, _, it, rel = s.solve_refined(f, rtol=1e-12, max_iter=10)
    assert rel <= 1e-12 and _rel(xr, np.ones(s.ndof)) <= 1e-9
