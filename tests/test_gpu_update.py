"""Incremental re-factorization after a geometry update on the GPU
(QuadTree::perturb, quadtree.cpp:520-801, then skeletonize skipping the
boxes still compressed, skel_factorization.cpp:157).  With every box
marked, or none, the result is bit-identical to a full factorization of
the new nodes; after a local deformation the interior field stays within
the reference's bound of the field from a fresh tree
(ie_solver_test.cpp:577-729: 100 x tol, Stokes 300 x tol)."""
import numpy as np
import pytest
# torch's HIP runtime has to initialize before this library loads the
# system one, so it is imported when the module is.
import torch

import geometry as G
import iesolver
import oracle

pytestmark = pytest.mark.gpu

FP64 = iesolver.FP64


def _solver(g, pde=iesolver.PDE_LAPLACE, method=None):
    s = iesolver.RSSolver(0)
    if isinstance(g, list):
        s.set_geometry_batch(g, pde)
    else:
        s.set_geometry(g, pde)
        if "holes" in g:
            s.set_holes(g["holes"]["centers"], g["holes"]["radii"])
    s.build_tree(64)
    if method is not None:
        s.set_id_method(*method)
    return s


@pytest.mark.parametrize("method", [None, (iesolver.ID_RANDOMIZED, 32)])
def test_update_marking_every_box_is_bit_exact(method):
    tol = 1e-10
    g0 = G.wavy(32768, seed=3)
    g1 = dict(g0, weights=g0["weights"] * (1 + 1e-6))  # every node changes
    f = G.harmonic_data(g0["points"], *G.exterior_sources(3))
    s = _solver(g0, method=method)
    s.factor(tol, FP64)
    assert s.update_geometry(g1)
    n_dirty, _, marks = s.update_info()
    assert marks.all() and n_dirty == marks.size
    s.factor(tol, FP64)
    assert s.update_info()[1]  # the incremental path ran
    mu, _ = s.solve(f)
    fresh = _solver(g1, method=method)
    fresh.factor(tol, FP64)
    mf, _ = fresh.solve(f)
    assert np.array_equal(mu, mf, equal_nan=True)


def test_update_without_changes_is_bit_exact():
    tol = 1e-10
    g = G.wavy(32768, seed=4)
    f = G.harmonic_data(g["points"], *G.exterior_sources(4))
    s = _solver(g)
    s.factor(tol, FP64)
    m0, _ = s.solve(f)
    assert s.update_geometry(g)
    s.factor(tol, FP64)
    n_dirty, inc, _ = s.update_info()
    assert inc and n_dirty == 1
    m1, _ = s.solve(f)
    assert np.array_equal(m0, m1)


CASES = [
    ("laplace", iesolver.PDE_LAPLACE, 1e-10, 100),
    ("laplace", iesolver.PDE_LAPLACE, 1e-6, 100),
    ("stokes", iesolver.PDE_STOKES, 1e-6, 300),
    ("neumann", iesolver.PDE_LAPLACE_NEUMANN, 1e-6, 100),
]


@pytest.mark.parametrize("name,pde,tol,factor", CASES)
def test_update_local_bump_matches_fresh_tree(name, pde, tol, factor):
    n = 8192
    g0 = G.star(n)
    tg = G.interior_points(5, 100, radius=0.2)
    s = _solver(g0, pde)
    s.factor(tol, FP64)
    for frame, bump in enumerate([0.01, 0.03, -0.02]):
        g1 = G.bump_star(n, bump, 1.0 + 0.7 * frame)
        if pde == iesolver.PDE_STOKES:
            f = G.tangent_data(g1["normals"]).reshape(-1)
        elif pde == iesolver.PDE_LAPLACE_NEUMANN:
            f = np.ones(n)
        else:
            f = G.harmonic_data(g1["points"], *G.exterior_sources(5))
        assert s.update_geometry(g1)
        s.factor(tol, FP64)
        n_dirty, inc, marks = s.update_info()
        assert inc and n_dirty < marks.size
        mu, _ = s.solve(f)
        u = s.evaluate(tg, mu)
        # every box whose ID or decouple inputs changed was re-factored, so
        # a full factorization on the same tree gives the same bits
        s.discard_update()  # the next factorization covers every box
        s.factor(tol, FP64)
        assert not s.update_info()[1]
        m_full, _ = s.solve(f)
        assert np.array_equal(mu, m_full)
        fresh = _solver(g1, pde)
        fresh.factor(tol, FP64)
        mf, _ = fresh.solve(f)
        uf = fresh.evaluate(tg, mf)
        d = np.abs(u - uf).max()
        print(f"{name} tol {tol:g} frame {frame}: {n_dirty} of {marks.size} boxes "
              f"re-factored, max field difference {d:.2e}", flush=True)
        assert d <= factor * tol


def test_update_one_problem_of_a_batch():
    tol = 1e-10
    geoms = [G.star(4096), G.wavy(8192, seed=1), G.star(2048, amp=0.1, k=5), G.wavy(4096, seed=2)]
    f = G.batch_data(geoms, seed=7)
    s = _solver(geoms)
    s.factor(tol, FP64)
    m0, _ = s.solve(f)
    g1 = list(geoms)
    g1[0] = G.bump_star(4096, 0.03, 1.0)
    assert s.update_geometry(g1)
    s.factor(tol, FP64)
    assert s.update_info()[1]
    m1, _ = s.solve(f)
    off = s.dof_offsets
    # the other problems' boxes and roots are re-used or re-formed from
    # unchanged data: their solutions do not move
    assert np.array_equal(m1[off[1]:], m0[off[1]:])
    one = _solver(g1[0])
    one.factor(tol, FP64)
    m_one, _ = one.solve(f[:off[1]])
    tg = G.interior_points(5, 100, radius=0.2)
    d = np.abs(one.evaluate(tg, m1[:off[1]]) - one.evaluate(tg, m_one)).max()
    assert d <= 100 * tol


def test_update_with_holes_rebuilds_border():
    tol = 1e-8
    g0 = G.circle_with_holes(8192, n_holes=4, seed=2)
    sel = np.zeros(g0["points"].shape[0], bool)
    sel[100:300] = True  # part of the outer circle
    g1 = dict(g0, weights=np.where(sel, g0["weights"] * 1.01, g0["weights"]))
    f = G.tangent_data(g0["normals"]).reshape(-1)
    s = _solver(g0, iesolver.PDE_STOKES)
    s.factor(tol, FP64)
    assert s.update_geometry(g1)
    # setting the holes again (as after moving them) keeps the update
    # incremental: they enter only the border blocks
    s.set_holes(g0["holes"]["centers"], g0["holes"]["radii"])
    s.factor(tol, FP64)
    n_dirty, inc, marks = s.update_info()
    assert inc and n_dirty < marks.size
    mu, alpha = s.solve(f)
    fresh = _solver(g1, iesolver.PDE_STOKES)
    fresh.factor(tol, FP64)
    mf, af = fresh.solve(f)
    assert np.abs(mu - mf).max() <= 300 * tol * np.abs(mf).max()
    assert np.abs(alpha - af).max() <= 300 * tol * max(np.abs(af).max(), 1.0)


def test_update_that_rebuilds_the_tree_factors_every_box():
    tol = 1e-10
    g0 = G.star(4096)
    g1 = {k: v.copy() for k, v in g0.items()}
    g1["points"][7] = [1.5, 1.5]  # outside the root cell: the tree is rebuilt
    f = np.ones(4096)
    s = _solver(g0)
    s.factor(tol, FP64)
    assert not s.update_geometry(g1)
    s.factor(tol, FP64)
    assert not s.update_info()[1]
    mu, _ = s.solve(f)
    fresh = _solver(g1)
    fresh.factor(tol, FP64)
    mf, _ = fresh.solve(f)
    assert np.array_equal(mu, mf, equal_nan=True)


@pytest.mark.parametrize("method", [None, (iesolver.ID_RANDOMIZED, 32)])
def test_update_split_and_merge_equal_full_refactor(method):
    # Nodes crowded into a small square subdivide a leaf several levels
    # deep and empty the leaves they left (merged into their parents); then
    # they go back.  The kept boxes' factors move to their new slots, the
    # new and changed boxes are re-factored, and each incremental
    # factorization equals a full one on the same tree bit for bit.
    tol = 1e-10
    n = 16384
    g0 = G.wavy(n, seed=6)
    g1 = {k: v.copy() for k, v in g0.items()}
    c = g0["points"][4000]
    g1["points"][3900:4100] = c + 2e-4 * np.stack([np.arange(200) % 15, np.arange(200) // 15], 1)
    f0 = G.harmonic_data(g0["points"], *G.exterior_sources(6))
    s = _solver(g0, method=method)
    s.factor(tol, FP64)
    for frame, g in enumerate([g1, g0]):
        levels0, boxes0 = s.tree_info()[:2]
        assert s.update_geometry(g)
        n_split, n_merged, old_id = s.update_tree_info()
        levels1, boxes1 = s.tree_info()[:2]
        s.factor(tol, FP64)
        n_dirty, inc, marks = s.update_info()
        print(f"frame {frame}: {n_split} split, {n_merged} merged, boxes {boxes0} -> {boxes1}, "
              f"levels {levels0} -> {levels1}, {n_dirty} re-factored, "
              f"{int((old_id < 0).sum())} new", flush=True)
        assert n_split + n_merged > 0 and boxes1 != boxes0
        assert inc and n_dirty < marks.size and np.all(marks[old_id < 0] == 1)
        mu, _ = s.solve(f0)
        s.discard_update()
        s.factor(tol, FP64)
        m_full, _ = s.solve(f0)
        assert np.array_equal(mu, m_full), frame
    # back on the first nodes: within the reference's bound of a fresh tree
    fresh = _solver(g0, method=method)
    fresh.factor(tol, FP64)
    mf, _ = fresh.solve(f0)
    tg = G.interior_points(6, 100, radius=0.2)
    d = np.abs(s.evaluate(tg, mu) - fresh.evaluate(tg, mf)).max()
    assert d <= 100 * tol


def test_update_then_fused_factor_solve():
    # iegpu_factor_solve_device after an update re-factors the marked boxes
    # and gives the same x as iegpu_factor + iegpu_solve_device after the
    # same update.
    tol = 1e-10
    n = 8192
    g0, g1 = G.star(n), G.bump_star(n, 0.03, 1.0)
    f = G.harmonic_data(g1["points"], *G.exterior_sources(5))
    dev = torch.device("cuda:0")
    fd = torch.from_numpy(f).to(dev)
    xa, xb = torch.zeros_like(fd), torch.zeros_like(fd)
    torch.cuda.synchronize()  # the library works on its own streams
    a, b = _solver(g0), _solver(g0)
    a.factor(tol, FP64)
    b.factor(tol, FP64)
    assert a.update_geometry(g1) and b.update_geometry(g1)
    a.factor_solve_device(tol, FP64, fd.data_ptr(), xa.data_ptr())
    b.factor(tol, FP64)
    b.solve_device(fd.data_ptr(), xb.data_ptr())
    torch.cuda.synchronize()
    assert a.update_info()[1] and b.update_info()[1]
    assert np.array_equal(xa.cpu().numpy(), xb.cpu().numpy())


@pytest.mark.parametrize("learned", [False, True])
def test_update_randomized_forms_equal_full_refactor(learned):
    # The marks cover every box whose ID inputs changed, so in the
    # randomized forms too an incremental factorization equals a full one
    # on the same tree bit for bit.
    tol = 1e-10
    n = 65536
    s = _solver(G.wavy(n, seed=0))
    if learned:
        import os
        model = np.load(os.path.join(os.path.dirname(iesolver.__file__),
                                     "rank_model_default.npy")).astype(np.float32)
        s.set_rank_model(model, margin=3)
        s.set_id_method(iesolver.ID_RANDOMIZED, 0)
    else:
        s.set_id_method(iesolver.ID_RANDOMIZED, 32)
    s.factor(tol, FP64)
    for frame in range(2):
        g1 = G.bump_wavy(n, 3e-4 * (1 - 2 * frame), 1.0 + 0.5 * frame, seed=0)
        f = G.harmonic_data(g1["points"], *G.exterior_sources(0))
        assert s.update_geometry(g1)
        s.factor(tol, FP64)
        n_dirty, inc, marks = s.update_info()
        assert inc and n_dirty < marks.size
        mu, _ = s.solve(f)
        s.discard_update()
        s.factor(tol, FP64)
        m_full, _ = s.solve(f)
        assert np.array_equal(mu, m_full), (learned, frame)


def test_many_incremental_updates_keep_the_pool_bounded():
    # 24 consecutive incremental factorizations with no full one between
    # them: re-factored boxes reuse their kept slots, and the pool top stays
    # within twice a full factorization's (past it, the next factorization
    # is a full one); the last solution equals a full factorization's.
    tol = 1e-10
    n = 16384
    s = _solver(G.star(n))
    s.factor(tol, FP64)
    full_bytes = s.timings()["factor_bytes"]
    incs = 0
    for frame in range(24):
        g1 = G.bump_star(n, 0.004 * (1 + frame % 3), 0.3 + 0.25 * frame)
        assert s.update_geometry(g1)
        s.factor(tol, FP64)
        incs += s.update_info()[1]
        fb = s.timings()["factor_bytes"]
        assert fb <= 2 * full_bytes * 1.25, (frame, fb, full_bytes)
    assert incs >= 20
    f = G.harmonic_data(g1["points"], *G.exterior_sources(5))
    mu, _ = s.solve(f)
    s.discard_update()
    s.factor(tol, FP64)
    m_full, _ = s.solve(f)
    assert np.array_equal(mu, m_full)


def test_update_marks_cover_changed_box_inputs():
    # Full factorizations of the tree before and after a bump: every box
    # whose active list or skeleton differs between them (its ID or decouple
    # inputs changed), and every new box, carries a mark, and the marks stay
    # a small part of the tree.  Boxes are matched through the ids the
    # update reports.
    tol = 1e-10
    n = 8192
    g0 = G.star(n)
    s = _solver(g0)
    s.factor(tol, FP64)
    nb0 = s.tree_info()[1]
    before = [(s.box_list(b, 2).tolist(), s.box_list(b, 3).tolist()) for b in range(nb0)]
    g1 = G.bump_star(n, 0.02, 1.3)
    assert s.update_geometry(g1)
    _, _, marks = s.update_info()
    old_id = s.update_tree_info()[2]
    s.discard_update()
    s.factor(tol, FP64)
    nb = s.tree_info()[1]
    after = [(s.box_list(b, 2).tolist(), s.box_list(b, 3).tolist()) for b in range(nb)]
    changed = np.array([old_id[b] < 0 or before[old_id[b]] != after[b] for b in range(nb)])
    print(f"{changed.sum()} boxes changed ({int((old_id < 0).sum())} new), "
          f"{marks.sum()} marked of {nb}", flush=True)
    assert changed.any()
    assert np.all(marks[changed] == 1)
    assert marks.sum() <= max(4 * changed.sum(), nb // 4)


@pytest.mark.parametrize("pde", [iesolver.PDE_LAPLACE, iesolver.PDE_STOKES])
def test_update_incremental_matches_oracle_incremental(pde):
    # The oracle restates QuadTree::perturb (its tree equals the library's)
    # and skeletonize's skip of compressed boxes; given the library's marks
    # it re-factors the same boxes.  The two incremental solutions agree to
    # the ID tolerance (leaf DOF order differs: the reference appends moved
    # DOFs at the end of a leaf's list).
    tol = 1e-10
    n = 8192
    g0 = G.star(n)
    g1 = G.bump_star(n, 0.03, 1.0 if pde == iesolver.PDE_LAPLACE else 2.5)
    if pde == iesolver.PDE_STOKES:
        f = G.tangent_data(g1["normals"]).reshape(-1)
    else:
        f = G.harmonic_data(g1["points"], *G.exterior_sources(5))
    s = _solver(g0, pde)
    s.factor(tol, FP64)
    assert s.update_geometry(g1)
    n_split, n_merged, _ = s.update_tree_info()
    marks = s.update_info()[2]
    s.factor(tol, FP64)
    assert s.update_info()[1]
    mu, _ = s.solve(f)
    o = oracle.OracleProblem(g0, pde=pde)
    o.factor(tol)
    o.perturb(g1)
    o.mark(marks)
    o.factor(tol)
    mo, _ = o.solve(f)
    d = np.linalg.norm(mu - mo) / np.linalg.norm(mo)
    print(f"pde {pde}: {n_split} split, {n_merged} merged, {int(marks.sum())} of {marks.size} "
          f"re-factored; incremental vs oracle incremental {d:.2e}", flush=True)
    assert n_split + n_merged > 0
    assert d <= 10 * tol
