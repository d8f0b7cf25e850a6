"""Status codes of the failure cases the reference handles by exit or
assert, and the kernel's coincident-node rule, on the GPU.

* No compression possible with holes (skel_factorization.cpp:209-212, the
  reference prints "No compression possible" and exits):
  IEGPU_ERR_NO_COMPRESSION.
* An exactly singular root block (LAPACK dgetrf info > 0, asserted by the
  reference at ie_mat.cpp:449-455): IEGPU_ERR_SINGULAR.
* Coincident nodes take the diagonal value of kernel.cpp:149-157 for the
  source's weight and curvature.
* Boxes under MIN_DOFS_TO_COMPRESS (skel_factorization.h:11,
  skel_factorization.cpp:157-163) are not compressed.

Every case returns an ordinary status; no kernel faults or hangs.
"""
import math

import numpy as np
import pytest

import geometry as G
import iesolver
import oracle

pytestmark = pytest.mark.gpu


def test_no_compression_with_holes_is_reported():
    # 48 nodes in one leaf: the root is the only box, every DOF a skeleton
    # DOF, and the bordered system has no redundant part.
    g = G.annulus(40)
    s = iesolver.RSSolver(0)
    s.set_geometry(g)
    s.build_tree(64)
    assert s.tree_info()[0] == 1
    s.set_holes(g["holes"]["centers"], g["holes"]["radii"])
    with pytest.raises(iesolver.IeGpuError) as e:
        s.factor(1e-10)
    assert e.value.code == 2 and "No compression possible" in str(e.value)
    # the context stays usable: without holes the same tree factors
    s.set_holes(np.zeros((0, 2)), np.zeros(0))
    s.factor(1e-10)


def _pair_geometry(n):
    """A circle whose normals are zero (every off-diagonal entry 0) with
    node 1 on node 0: rows 0 and 1 are both [1/2, 1/2, 0, ...], an exactly
    singular matrix whose elimination gives an exact zero pivot."""
    g = G.circle(n)
    g["normals"][:] = 0.0
    g["curvatures"][:] = 0.0
    g["points"][1] = g["points"][0]
    return g


@pytest.mark.parametrize("precision", [iesolver.FP64, iesolver.FP32])
def test_singular_root_block_is_reported(precision):
    g = _pair_geometry(48)
    s = iesolver.RSSolver(0)
    s.set_geometry(g)
    s.build_tree(64)
    with pytest.raises(iesolver.IeGpuError) as e:
        s.factor(1e-10, precision)
    assert e.value.code == 3
    with pytest.raises(iesolver.IeGpuError):  # no factorization is stored
        s.solve(np.ones(48))
    # a regular boundary in the same context factors again
    g2 = G.circle(48)
    s.set_geometry(g2)
    s.build_tree(64)
    s.factor(1e-10, precision)
    mu, _ = s.solve(np.ones(48))
    assert np.all(np.isfinite(mu))


def test_coincident_nodes_take_the_diagonal_rule():
    g = G.star(512)
    g["points"][1] = g["points"][0]
    s = iesolver.RSSolver(0)
    s.set_geometry(g)
    x = np.zeros(512)
    x[1] = 1.0
    col, _ = s.apply(x)
    w, k = g["weights"][1], g["curvatures"][1]
    diag = 0.5 + 0.5 * k * w / (2 * math.pi)
    assert abs(col[0] - diag) <= 1e-14 and abs(col[1] - diag) <= 1e-14
    # the whole operator against the oracle's dense kernel, same rule
    K = oracle.dense_matrix(g)
    y = np.random.default_rng(1).standard_normal(512)
    gy, _ = s.apply(y)
    assert np.linalg.norm(gy - K @ y) <= 1e-13 * np.linalg.norm(K @ y)


def test_small_boxes_are_not_compressed():
    g = G.star(16384)
    s = iesolver.RSSolver(0)
    s.set_geometry(g)
    s.build_tree(64)
    s.factor(1e-10)
    info = s.box_info()
    below = (info["n_active"] < 16) & (info["level"] > 0)
    assert below.any()
    assert np.all(info["rank"][below] == 0)
    assert np.all(info["m_rows"][below] == 0)
