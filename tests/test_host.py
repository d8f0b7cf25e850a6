"""Host-side checks of the product library (CPU only, no GPU calls)."""
import ctypes
import math
import os
import re
import subprocess

import numpy as np
import pytest

import geometry as G
import iesolver
import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "iegpu.h")).read()
    declared = set(re.findall(r"^(?:int|const char\*)\s+(iegpu_\w+)\s*\(", header, re.M))
    out = subprocess.run(["nm", "-D", "--defined-only", iesolver.LIB_PATH],
                         stdout=subprocess.PIPE, text=True, check=True).stdout
    exported = set(re.findall(r" T (iegpu_\w+)", out))
    assert declared, "no declarations parsed"
    assert declared <= exported, declared - exported
    assert set(iesolver.EXPORTS) <= exported
    iesolver.load()


@pytest.mark.parametrize("case", ["circle", "star", "wavy", "holes_stokes", "annulus_stokes"])
def test_host_tree_matches_oracle(case):
    if case == "circle":
        g, pde, cap = G.circle(4096), 0, 64
    elif case == "star":
        g, pde, cap = G.star(20000), 0, 64
    elif case == "wavy":
        g, pde, cap = G.wavy(30000, seed=3), 0, 64
    elif case == "holes_stokes":
        g, pde, cap = G.circle_with_holes(8000, 8, 0), 2, 64
    else:
        g, pde, cap = G.annulus(2000), 2, 32
    s = iesolver.RSSolver(device=-1)
    s.set_geometry(g, pde)
    s.build_tree(cap)
    o = oracle.OracleProblem(g, pde=pde, leaf_cap=cap)
    nl, nb, nd, _ = s.tree_info()
    assert (nl, nb) == (o.num_levels(), o.num_boxes())
    bi, oi = s.box_info(), o.box_info()
    for k in ("level", "parent", "is_leaf", "children"):
        assert np.array_equal(bi[k], oi[k]), k
    for b in range(nb):
        assert np.array_equal(s.box_list(b, 0), o.box_list(b, 0))
        assert np.array_equal(s.box_list(b, 1), o.box_list(b, 1))
    for lv in range(nl):
        assert np.array_equal(s.box_list(lv, 5), o.level_order(lv))


def test_error_codes():
    s = iesolver.RSSolver(device=-1)
    g = G.circle(64)
    with pytest.raises(iesolver.IeGpuError) as e:
        s.set_geometry(g, pde=3)  # 0 Laplace, 1 Laplace Neumann, 2 Stokes
    assert e.value.code == 1
    iesolver.RSSolver(device=-1).set_geometry(g, pde=iesolver.PDE_LAPLACE_NEUMANN)
    bad = dict(g)
    bad["points"] = g["points"].copy()
    bad["points"][3, 0] = np.nan
    with pytest.raises(iesolver.IeGpuError) as e:
        s.set_geometry(bad)
    assert e.value.code == 1
    with pytest.raises(iesolver.IeGpuError) as e:
        s.build_tree(64)  # no geometry yet
    assert e.value.code == 1
    s.set_geometry(g)
    with pytest.raises(iesolver.IeGpuError) as e:
        s.build_tree(0)
    assert e.value.code == 1
    s.build_tree(16)
    with pytest.raises(iesolver.IeGpuError) as e:
        s.factor(1e-10)
    assert e.value.code == 7  # host-only context
    with pytest.raises(iesolver.IeGpuError) as e:
        s.factor(-1.0)
    assert e.value.code in (1, 7)
    L = iesolver.load()
    assert L.iegpu_factor(None, 1e-10, 0) == 1
    assert L.iegpu_destroy(None) == 1


def test_id_method_arguments():
    s = iesolver.RSSolver(device=-1)
    s.set_id_method(iesolver.ID_RANDOMIZED, 32)
    s.set_id_method(iesolver.ID_RANDOMIZED, 16)
    s.set_id_method(iesolver.ID_QRCP)
    for method, rows in [(7, 32), (iesolver.ID_RANDOMIZED, 24), (iesolver.ID_RANDOMIZED, 48),
                         (iesolver.ID_RANDOMIZED, -16), (iesolver.ID_RANDOMIZED, 0)]:
        with pytest.raises(iesolver.IeGpuError) as e:
            s.set_id_method(method, rows)
        assert e.value.code == 1, (method, rows)
    with pytest.raises(iesolver.IeGpuError) as e:
        s.id_stats()  # nothing factored
    assert e.value.code == 1
    L = iesolver.load()
    assert L.iegpu_set_id_method(None, 0, 0) == 1
    assert L.iegpu_id_stats(None, None, None, None) == 1


def test_geometry_circle_closed_form():
    g = G.circle(1000, (0.5, 0.5), 0.5)
    assert np.allclose(np.hypot(*(g["points"] - 0.5).T), 0.5, atol=1e-15)
    assert np.allclose(g["normals"], (g["points"] - 0.5) / 0.5, atol=1e-14)
    assert abs(g["weights"].sum() - math.pi) < 1e-12


@pytest.mark.parametrize("gen", [lambda n: G.star(n), lambda n: G.wavy(n, seed=7)])
def test_geometry_curvature_and_weights(gen):
    n = 4000
    g = gen(n)
    p = g["points"]
    # perimeter vs polygon length
    poly = np.hypot(*(np.roll(p, -1, 0) - p).T).sum()
    assert abs(g["weights"].sum() - poly) < 1e-5
    # outward unit normals: orthogonal to the chord, pointing away from the centre
    chord = np.roll(p, -1, 0) - np.roll(p, 1, 0)
    assert np.abs(np.einsum("ij,ij->i", chord, g["normals"])).max() < 1e-3 * np.abs(chord).max()
    assert (np.einsum("ij,ij->i", p - 0.5, g["normals"]) > 0).all()
    # curvature: turning of the tangent per unit length
    t = chord / np.hypot(*chord.T)[:, None]
    ta, tb = np.roll(t, 1, 0), np.roll(t, -1, 0)
    turn = np.arctan2(ta[:, 0] * tb[:, 1] - ta[:, 1] * tb[:, 0], np.einsum("ij,ij->i", ta, tb))
    ds = np.hypot(*(np.roll(p, -1, 0) - np.roll(p, 1, 0)).T)
    assert np.abs(turn / ds - g["curvatures"]).max() < 1e-2 * np.abs(g["curvatures"]).max()


def test_holes_are_inside_and_disjoint():
    g = G.circle_with_holes(20000, 8, seed=0)
    c, r = g["holes"]["centers"], g["holes"]["radii"]
    assert ((0.02 <= r) & (r <= 0.05)).all()
    for i in range(8):
        assert math.hypot(*(c[i] - 0.5)) + r[i] < 0.5
        for j in range(i):
            assert math.hypot(*(c[i] - c[j])) > r[i] + r[j]
    assert g["points"].shape[0] == 20000


def test_forest_tree_is_the_problems_trees_merged():
    # A batch's quadtree is each problem's tree, merged level by level, with
    # global DOF numbers and creation ids numbered problem after problem.
    geoms, _ = G.star_batch(5, seed=3, nmin=300, nmax=3000)
    s = iesolver.RSSolver(device=-1)
    s.set_geometry_batch(geoms)
    s.build_tree(64)
    nl, nb, nd, _ = s.tree_info()
    bi = s.box_info()
    id0, dof0, levels = 0, 0, []
    for g in geoms:
        t = iesolver.RSSolver(device=-1)
        t.set_geometry(g)
        t.build_tree(64)
        l1, b1, d1, _ = t.tree_info()
        levels.append(l1)
        ti = t.box_info()
        sl = slice(id0, id0 + b1)
        assert np.array_equal(bi["level"][sl], ti["level"])
        assert np.array_equal(bi["is_leaf"][sl], ti["is_leaf"])
        par = np.where(ti["parent"] >= 0, ti["parent"] + id0, -1)
        assert np.array_equal(bi["parent"][sl], par)
        ch = np.where(ti["children"] >= 0, ti["children"] + id0, -1)
        assert np.array_equal(bi["children"][sl], ch)
        for b in range(b1):
            assert np.array_equal(s.box_list(id0 + b, 0), t.box_list(b, 0) + dof0)
            assert np.array_equal(s.box_list(id0 + b, 1), t.box_list(b, 1) + id0)
        id0 += b1
        dof0 += d1
    assert (nl, nb, nd) == (max(levels), id0, dof0)
    # level 0 holds the roots, in problem order
    roots = s.box_list(0, 5)
    assert len(roots) == len(geoms) and (bi["level"][roots] == 0).all()


def test_batch_arguments():
    s = iesolver.RSSolver(device=-1)
    g = G.circle(64)
    L = s.L
    off = np.array([0, 64], dtype=np.int32)
    bad = np.array([0, 0, 64], dtype=np.int32)
    a = [np.ascontiguousarray(g["points"]).reshape(-1), np.ascontiguousarray(g["normals"]).reshape(-1),
         g["curvatures"], g["weights"]]
    ptr = [x.ctypes.data_as(iesolver._dp) for x in a]
    assert L.iegpu_set_geometry_batch(s.h, 0, 0, off.ctypes.data_as(iesolver._ip), *ptr) != 0
    assert L.iegpu_set_geometry_batch(s.h, 0, 2, bad.ctypes.data_as(iesolver._ip), *ptr) != 0
    assert L.iegpu_set_geometry_batch(s.h, 0, 1, off.ctypes.data_as(iesolver._ip), *ptr) == 0


def test_launch_caps_arguments():
    s = iesolver.RSSolver(device=-1)
    assert list(s.launch_caps()) == [1024] * 6 + [8192, 4096]
    s.set_launch_caps([64, 256, 1024, 4096, 1, 2, 3, 2])
    assert list(s.launch_caps()) == [64, 256, 1024, 4096, 1, 2, 3, 2]
    with pytest.raises(iesolver.IeGpuError):
        s.set_launch_caps([0] * iesolver.NUM_LAUNCH_CAPS)
    with pytest.raises(iesolver.IeGpuError):
        s.set_launch_caps([64] * 7 + [(1 << 20) + 1])
    assert list(s.launch_caps()) == [64, 256, 1024, 4096, 1, 2, 3, 2]
    with pytest.raises(iesolver.IeGpuError):
        s.set_launch_caps([64] * 3)


def test_batch_update_keeps_each_problem_size():
    # A batch update whose total matches but whose problems trade points is
    # rejected (the kept tree's point offsets would assign nodes to the
    # wrong problem).
    geoms = [G.star(300), G.star(500, amp=0.1, k=5)]
    s = iesolver.RSSolver(device=-1)
    s.set_geometry_batch(geoms)
    s.build_tree(64)
    shifted = [G.star(301), G.star(499, amp=0.1, k=5)]
    with pytest.raises(iesolver.IeGpuError, match="number of points"):
        s.update_geometry(shifted)
    with pytest.raises(iesolver.IeGpuError):
        s.update_geometry(geoms[:1])
    s.discard_update()


def test_matvec_arguments():
    s = iesolver.RSSolver(device=-1)
    assert s.matvec_info() == (0, 0, 0, 0)  # nothing applied, no FMM tree
    s.set_matvec(iesolver.MATVEC_FMM, 64, 16)
    s.set_matvec(iesolver.MATVEC_AUTO)
    s.set_matvec(iesolver.MATVEC_DIRECT, 32, 4096)
    for args in [(3, 0, 0), (-1, 0, 0), (2, 40, 0), (2, 16, 0), (2, 48, -1), (2, 48, 5000)]:
        with pytest.raises(iesolver.IeGpuError) as e:
            s.set_matvec(*args)
        assert e.value.code == 1, args
    g = G.star(256)
    s.set_geometry(g)
    with pytest.raises(iesolver.IeGpuError) as e:
        s.apply(np.ones(256))
    assert e.value.code == 7  # host-only context
    L = iesolver.load()
    assert L.iegpu_set_matvec(None, 0, 0, 0) == 1
    assert L.iegpu_matvec_info(None, None, 0) == 1


def test_fmm_tree_and_translations_cpu():
    """tests/cpp/fmm_tree_test.cpp: the FMM tree's lists count every pair of
    leaves once, and its translations (in the kernels' order) reproduce the
    direct sum of the three kernels to 1e-13."""
    exe = os.path.join(ROOT, "tests", "cpp", "fmm_tree_test.bin")
    if not os.path.exists(exe):
        import build
        build.build_cpp_tests()
    p = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120)
    print(p.stdout)
    assert p.returncode == 0, p.stdout


def test_tree_local_updates_cpu():
    """tests/cpp/tree_update_test.cpp: after local deformations that split
    and merge leaves, the lists formed only where they can change equal the
    tree-wide pass, the cells follow the capacity rule (a fresh build's
    cells), every point lies in its leaf and surviving boxes keep their
    cells and report their old ids."""
    exe = os.path.join(ROOT, "tests", "cpp", "tree_update_test.bin")
    if not os.path.exists(exe):
        import build
        build.build_cpp_tests()
    p = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120)
    print(p.stdout)
    assert p.returncode == 0, p.stdout
