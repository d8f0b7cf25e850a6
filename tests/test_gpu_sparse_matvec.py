"""Compressed forward apply on the GPU (SkelFactorization::sparse_matvec,
skel_factorization.cpp:429-504): parity with the oracle's restatement, the
inverse of the solve to rounding, the dense operator to the ID error."""
import numpy as np
import pytest

import geometry as G
import iesolver
import oracle

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _gpu(g, pde, tol, precision=iesolver.FP64, method=None, batch=False):
    s = iesolver.RSSolver(0)
    if batch:
        s.set_geometry_batch(g, pde)
    else:
        s.set_geometry(g, pde)
    s.build_tree(64)
    if method is not None:
        s.set_id_method(*method)
    s.set_keep_operator(True)
    s.factor(tol, precision)
    return s


@pytest.mark.parametrize("pde", [iesolver.PDE_LAPLACE, iesolver.PDE_LAPLACE_NEUMANN,
                                 iesolver.PDE_STOKES])
def test_sparse_matvec_vs_oracle_and_dense(pde):
    tol = 1e-10
    g = G.star(2048) if pde != iesolver.PDE_STOKES else G.circle(1024)
    s = _gpu(g, pde, tol)
    x = np.random.default_rng(7).standard_normal(s.ndof)
    b = s.sparse_matvec(x)
    o = oracle.OracleProblem(g, pde=pde)
    o.factor(tol, 4)
    bo = o.sparse_matvec(x)
    K = oracle.dense_matrix(g, pde)
    e_or, e_k = _rel(b, bo), _rel(b, K @ x)
    print(f"pde {pde}: vs oracle {e_or:.2e}, vs dense {e_k:.2e}", flush=True)
    assert e_or <= 50 * tol and e_k <= 50 * tol
    y, _ = s.solve(b)
    assert _rel(y, x) <= 1e-12


@pytest.mark.parametrize("form", ["randomized", "fp32"])
def test_sparse_matvec_inverts_solve(form):
    g = G.wavy(32768, seed=8)
    if form == "fp32":
        s = _gpu(g, iesolver.PDE_LAPLACE, 1e-6, precision=iesolver.FP32)
        bound = 1e-5
    else:
        s = _gpu(g, iesolver.PDE_LAPLACE, 1e-10, method=(iesolver.ID_RANDOMIZED, 32))
        bound = 1e-12
    x = np.random.default_rng(8).standard_normal(s.ndof)
    y, _ = s.solve(s.sparse_matvec(x))
    assert _rel(y, x) <= bound
    # the compressed operator against the FP64 matrix-free one
    assert _rel(s.sparse_matvec(x), s.apply(x)[0]) <= (1e-5 if form == "fp32" else 1e-8)


def test_sparse_matvec_batch_and_errors():
    geoms, _ = G.star_batch(4, seed=9, nmin=512, nmax=4096)
    s = _gpu(geoms, iesolver.PDE_LAPLACE, 1e-10, batch=True)
    x = np.random.default_rng(9).standard_normal(s.ndof)
    b = s.sparse_matvec(x)
    off = s.dof_offsets
    for p, g in enumerate(geoms):
        t = _gpu(g, iesolver.PDE_LAPLACE, 1e-10)
        sl = slice(off[p], off[p + 1])
        assert _rel(b[sl], t.sparse_matvec(x[sl])) <= 1e-12
    u = iesolver.RSSolver(0)
    u.set_geometry(geoms[0])
    u.build_tree(64)
    u.factor(1e-10)  # without the kept operator
    with pytest.raises(iesolver.IeGpuError):
        u.sparse_matvec(np.zeros(u.ndof))
