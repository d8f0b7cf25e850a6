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
    # the cells, levels and neighbours are kept
    assert nbrs1 == nbrs
    for k in ("level", "parent", "is_leaf", "children"):
        assert np.array_equal(bi0[k], bi1[k])
    # every point lies in its new leaf's cell
    o = oracle.OracleProblem(g0, pde=pde)
    bl = o.box_info()["bl_side"].reshape(-1, 3)
    lo = _leaf_of(leaf1, n, sd)
    p = g1["points"]
    x0, y0, side = bl[lo, 0], bl[lo, 1], bl[lo, 2]
    assert np.all((p[:, 0] >= x0) & (p[:, 0] <= x0 + side) & (p[:, 1] >= y0) & (p[:, 1] <= y0 + side))
    # marks = the reference's walk from the old and new leaf of every
    # changed node
    changed = np.flatnonzero(np.any(g0["points"] != g1["points"], 1) |
                             np.any(g0["normals"] != g1["normals"], 1) |
                             (g0["weights"] != g1["weights"]) | (g0["curvatures"] != g1["curvatures"]))
    la = _leaf_of(leaf0, n, sd)
    moved = int(np.sum(la[changed] != lo[changed]))
    leaves = np.concatenate([la[changed], lo[changed]]).tolist()
    want = oracle.update_marks(bi0["parent"].tolist(), nbrs, leaves)
    nd, _, marks = s.update_info()
    print(f"bump {bump}: {changed.size} changed nodes, {moved} change leaf, "
          f"{nd} of {marks.size} boxes marked ({want.sum()} by the reference's walk)", flush=True)
    # every box the reference marks is marked; the library adds boxes whose
    # ID row sources (not always neighbours) hold changed nodes
    assert np.all(marks >= want)
    assert nd == marks.sum() < marks.size
    if bump > 0.01:
        assert moved > 0


def test_update_rebuilds_tree_when_a_leaf_overflows():
    n = 4096
    g0 = G.star(n)
    g1 = {k: v.copy() for k, v in g0.items()}
    # crowd 200 nodes of the arc into a tiny square: no leaf of the kept
    # tree can take them (a kept leaf holds at most 2 leaf_cap DOFs)
    g1["points"][:200] = 0.8 + 1e-4 * np.stack([np.arange(200) % 14, np.arange(200) // 14], 1)
    s = _host(g0)
    assert not s.update_geometry(g1)
    n_marked, _, marks = s.update_info()
    assert n_marked == marks.size and marks.all()
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
