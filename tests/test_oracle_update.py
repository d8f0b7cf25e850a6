"""Geometry updates against the oracle's restatement of QuadTree::perturb
(quadtree.cpp:520-801: nodes re-inserted, leaves over the capacity
subdivided, boxes under it merged, neighbour lists formed again) and of
skeletonize's skip of compressed boxes (skel_factorization.cpp:157).

* The library's updated tree (host-only context) equals the oracle's
  perturbed tree: ids, levels, parents, children, leaf flags, level orders
  and neighbour lists in order; leaf DOF lists as sets (the reference
  appends moved DOFs at the end of a leaf's list, the library keeps DOF
  order).
* With the library's marks, the oracle's incremental factorization equals
  its full factorization of the same tree bit for bit: the marks cover every
  box whose ID or decouple inputs changed.  With the reference's own marks
  it does not always (boxes whose ID rows hold moved nodes keep stale IDs).
"""
import numpy as np
import pytest

import geometry as G
import iesolver
import oracle


def _crowd(g, start, count, spacing):
    g1 = {k: v.copy() for k, v in g.items()}
    c = g["points"][start + count // 2]
    k = np.arange(count)
    g1["points"][start:start + count] = c + spacing * np.stack([k % 15, k // 15], 1)
    return g1


CASES = [
    ("bump 0.002", iesolver.PDE_LAPLACE, lambda n: G.bump_star(n, 0.002, 1.0)),
    ("bump 0.03", iesolver.PDE_LAPLACE, lambda n: G.bump_star(n, 0.03, 1.0)),
    ("bump -0.02", iesolver.PDE_LAPLACE, lambda n: G.bump_star(n, -0.02, 2.0)),
    ("crowd", iesolver.PDE_LAPLACE, lambda n: _crowd(G.star(n), 3000, 200, 2e-4)),
    ("stokes bump", iesolver.PDE_STOKES, lambda n: G.bump_star(n, 0.03, 2.5)),
]


def _pair(g0, g1, pde):
    s = iesolver.RSSolver(device=-1)
    s.set_geometry(g0, pde)
    s.build_tree(64)
    o = oracle.OracleProblem(g0, pde=pde)
    return s, o


@pytest.mark.parametrize("name,pde,make", CASES, ids=[c[0] for c in CASES])
def test_updated_tree_matches_oracle_perturb(name, pde, make):
    n = 8192
    g0 = G.star(n)
    g1 = make(n)
    s, o = _pair(g0, g1, pde)
    assert s.update_geometry(g1)
    o.perturb(g1)
    n_split, n_merged, _ = s.update_tree_info()
    bi, oi = s.box_info(), o.box_info()
    nb = len(bi["level"])
    print(f"{name}: {n_split} split, {n_merged} merged, {nb} boxes", flush=True)
    assert nb == o.num_boxes() and s.tree_info()[0] == o.num_levels()
    for k in ("level", "parent", "is_leaf"):
        assert np.array_equal(bi[k], oi[k]), k
    assert np.array_equal(bi["children"].reshape(-1, 4), oi["children"])
    for lv in range(o.num_levels()):
        assert np.array_equal(s.box_list(lv, 5), o.level_order(lv)), lv
    for b in range(nb):
        assert np.array_equal(s.box_list(b, 1), o.box_list(b, 1)), b
        if bi["is_leaf"][b]:
            assert np.array_equal(s.box_list(b, 0), np.sort(o.box_list(b, 0))), b
    if name == "crowd":
        assert n_split > 0 and n_merged > 0


@pytest.mark.parametrize("name,pde,make", CASES[1:4], ids=[c[0] for c in CASES[1:4]])
def test_oracle_incremental_factor_with_library_marks_is_full(name, pde, make):
    n = 8192
    tol = 1e-10
    g0 = G.star(n)
    g1 = make(n)
    f = G.harmonic_data(g1["points"], *G.exterior_sources(5))
    s = iesolver.RSSolver(device=-1)
    s.set_geometry(g0, pde)
    s.build_tree(64)
    assert s.update_geometry(g1)
    marks = s.update_info()[2]
    out = {}
    for mode in ("reference", "library", "full"):
        o = oracle.OracleProblem(g0, pde=pde)
        o.factor(tol)
        o.perturb(g1)
        if mode == "library":
            ref_kept = o.compressed()
            o.mark(marks)
            # the library marks every box the reference marks
            assert np.all(o.compressed() <= ref_kept)
        elif mode == "full":
            o.mark(None)
        kept = int(o.compressed().sum())
        o.factor(tol)
        out[mode] = (o.solve(f)[0], kept)
    d_ref = np.abs(out["reference"][0] - out["full"][0]).max() / np.abs(out["full"][0]).max()
    print(f"{name}: boxes kept {out['library'][1]} (library marks), {out['reference'][1]} "
          f"(reference marks); reference-marks result differs from full by {d_ref:.1e}",
          flush=True)
    assert out["library"][1] > 0
    assert np.array_equal(out["library"][0], out["full"][0])
