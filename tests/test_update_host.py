"""Geometry updates on the host (QuadTree::perturb, quadtree.cpp:520-801,
and mark_neighbors_and_parents, quadtree.cpp:435-461): the oracle's
marking rule on a hand-built tree, then iegpu_update_geometry's tree and
marks in a host-only context against that rule."""
import numpy as np
import pytest

import geometry as G
import iesolver
import oracle


def test_oracle_update_marks_hand_built_tree():
    # root 0; level 1: 1..4 (a 2 x 2 grid, all adjacent); box 1 has children
    # 5..8, box 4 has children 9..12; 6 and 9 touch across the 1 | 4 edge.
    parent = [-1, 0, 0, 0, 0, 1, 1, 1, 1, 4, 4, 4, 4]
    nbrs = {0: [], 1: [2, 3, 4], 2: [1, 3, 4], 3: [1, 2, 4], 4: [1, 2, 3],
            5: [6, 7, 8], 6: [5, 7, 8, 9], 7: [5, 6, 8], 8: [5, 6, 7],
            9: [10, 11, 12, 6], 10: [9, 11, 12], 11: [9, 10, 12], 12: [9, 10, 11]}
    nb = [nbrs[b] for b in range(13)]
    # a change in leaf 6: 6 and its neighbours (9 included), then box 1 and
    # its neighbours, then the root; 9's siblings stay
    m = oracle.update_marks(parent, nb, [6])
    assert m.tolist() == [1] * 10 + [0, 0, 0]
    # a change in leaf 3 (level 1): 3, its neighbours, the root
    m = oracle.update_marks(parent, nb, [3])
    assert m.tolist() == [1, 1, 1, 1, 1] + [0] * 8
    # a point moving from leaf 11 to leaf 7: both walks
    m = oracle.update_marks(parent, nb, [11, 7])
    assert m.tolist() == [1] * 13
    assert oracle.update_marks(parent, nb, []).sum() == 0


def _host(g, pde=iesolver.PDE_LAPLACE):
    s = iesolver.RSSolver(device=-1)
    s.set_geometry(g, pde)
    s.build_tree(64)
    return s


def _tree(s):
    bi = s.box_info()
    nb = len(bi["level"])
    nbrs = [s.box_list(b, 1).tolist() for b in range(nb)]
    leaf_dofs = {b: s.box_list(b, 0) for b in range(nb) if bi["is_leaf"][b]}
    return bi, nbrs, leaf_dofs


def _leaf_of(leaf_dofs, n_points, sd):
    out = np.full(n_points, -1)
    for b, d in leaf_dofs.items():
        out[d[::sd] // sd] = b
    return out


def test_update_identical_geometry_keeps_everything():
    g = G.star(8192)
    s = _host(g)
    before = [s.box_list(b, 0).tolist() for b in range(s.tree_info()[1])]
    assert s.update_geometry(g)
    n, _, marks = s.update_info()
    assert n == 1 and marks.sum() == 1 and marks[0] == 1  # the root only
    after = [s.box_list(b, 0).tolist() for b in range(s.tree_info()[1])]
    assert before == after


@pytest.mark.parametrize("pde,bump", [(iesolver.PDE_LAPLACE, 0.002), (iesolver.PDE_LAPLACE, 0.03),
                                      (iesolver.PDE_STOKES, 0.03)])
def test_update_marks_follow_reference_rule(pde, bump):
    sd = 2 if pde == iesolver.PDE_STOKES else 1
    n = 8192
    g0 = G.star(n)
    g1 = G.bump_star(n, bump, 1.0)
    s = _host(g0, pde)
    bi0, nbrs, leaf0 = _tree(s)
    assert s.update_geometry(g1)
    bi1, nbrs1, leaf1 = _tree(s)
    n_split, n_merged, old_id = s.update_tree_info()
    nb1 = len(bi1["level"])
    survivors = np.flatnonzero(old_id >= 0)
    if n_split == 0 and n_merged == 0:
        # the cells, levels and neighbours are kept
        assert np.array_equal(old_id, np.arange(nb1))
        assert nbrs1 == nbrs
        for k in ("level", "parent", "is_leaf", "children"):
            assert np.array_equal(bi0[k], bi1[k])
    else:
        assert survivors.size < nb1 or nb1 < len(bi0["level"])
        assert np.array_equal(bi0["level"][old_id[survivors]], bi1["level"][survivors])
    # every point lies in its new leaf's cell (the oracle's cells of the
    # old tree: surviving boxes keep theirs)
    o = oracle.OracleProblem(g0, pde=pde)
    bl = o.box_info()["bl_side"].reshape(-1, 3)
    lo = _leaf_of(leaf1, n, sd)
    assert np.all(lo >= 0)
    p = g1["points"]
    kept_leaf = old_id[lo] >= 0
    x0, y0, side = (bl[old_id[lo][kept_leaf], k] for k in range(3))
    pk = p[kept_leaf]
    assert np.all((pk[:, 0] >= x0) & (pk[:, 0] <= x0 + side) & (pk[:, 1] >= y0) & (pk[:, 1] <= y0 + side))
    # marks cover the reference's walk from the old and new leaf of every
    # changed node (boxes of the old tree, by their ids now), and every new box
    changed = np.flatnonzero(np.any(g0["points"] != g1["points"], 1) |
                             np.any(g0["normals"] != g1["normals"], 1) |
                             (g0["weights"] != g1["weights"]) | (g0["curvatures"] != g1["curvatures"]))
    la = _leaf_of(leaf0, n, sd)
    moved = int(np.sum(old_id[lo[changed]] != la[changed]))
    new_of_old = np.full(len(bi0["level"]), -1)
    new_of_old[old_id[survivors]] = survivors
    leaves = [int(v) for v in la[changed] if new_of_old[v] >= 0]
    leaves += [int(old_id[v]) for v in lo[changed] if old_id[v] >= 0]
    want = oracle.update_marks(bi0["parent"].tolist(), nbrs, leaves)
    nd, _, marks = s.update_info()
    print(f"bump {bump}: {changed.size} changed nodes, {moved} change leaf, {n_split} leaves "
          f"split, {n_merged} merged, {nd} of {marks.size} boxes marked "
          f"({want.sum()} by the reference's walk)", flush=True)
    want_now = want[old_id[survivors]] == 1
    assert np.all(marks[survivors[want_now]] == 1)
    assert np.all(marks[old_id < 0] == 1)
    assert nd == marks.sum() < marks.size
    if bump > 0.01:
        assert moved > 0


def test_update_splits_a_leaf_that_overflows():
    n = 4096
    g0 = G.star(n)
    g1 = {k: v.copy() for k, v in g0.items()}
    # crowd 200 nodes of the arc into a tiny square: the leaf there is
    # subdivided until its boxes hold at most leaf_cap DOFs, and the leaves
    # they left merge where they fall under it (quadtree.cpp:726-741)
    g1["points"][:200] = 0.8 + 1e-4 * np.stack([np.arange(200) % 14, np.arange(200) // 14], 1)
    s = _host(g0)
    levels0, boxes0 = s.tree_info()[:2]
    assert s.update_geometry(g1)
    n_split, n_merged, old_id = s.update_tree_info()
    levels1, boxes1 = s.tree_info()[:2]
    print(f"{n_split} leaves split, {n_merged} merged: {boxes0} -> {boxes1} boxes, "
          f"{levels0} -> {levels1} levels", flush=True)
    assert n_split >= 1 and levels1 > levels0
    bi = s.box_info()
    leaves = np.flatnonzero(bi["is_leaf"])
    sizes = np.array([s.box_list(b, 0).size for b in leaves])
    assert sizes.max() <= 64 and sizes.sum() == n
    # the leaves that hold the crowded nodes are new boxes, and marked
    _, _, marks = s.update_info()
    assert np.all(marks[old_id < 0] == 1)
    crowded = [int(b) for b in leaves if np.any(s.box_list(b, 0) < 200)]
    assert all(old_id[b] < 0 for b in crowded)


def test_update_rebuilds_tree_when_a_node_leaves_the_root():
    n = 4096
    g0 = G.star(n)
    g1 = {k: v.copy() for k, v in g0.items()}
    g1["points"][7] = [3.0, 3.0]  # outside the root cell
    s = _host(g0)
    assert not s.update_geometry(g1)
    n_marked, _, marks = s.update_info()
    assert n_marked == marks.size and marks.all()
    assert np.all(s.update_tree_info()[2] == -1)
    fresh = _host(g1)
    assert s.tree_info() == fresh.tree_info()
    for b in range(s.tree_info()[1]):
        assert np.array_equal(s.box_list(b, 0), fresh.box_list(b, 0))


def test_update_argument_errors():
    g = G.star(1024)
    s = iesolver.RSSolver(device=-1)
    s.set_geometry(g)
    with pytest.raises(iesolver.IeGpuError):
        s.update_geometry(g)  # no tree yet
    s.build_tree(64)
    with pytest.raises(iesolver.IeGpuError):
        s.update_geometry(G.star(1000))  # point count changes
    bad = {k: v.copy() for k, v in g.items()}
    bad["weights"][3] = np.nan
    with pytest.raises(iesolver.IeGpuError):
        s.update_geometry(bad)
