"""The benchmark configurations at their full sizes on the GPU, each against
the CPU oracle (oracle/, the restatement of the reference) on the same
seeded inputs.

* Star N = 65,536, plain FP64 (the plain-GPU configuration): solution
  relative L2 <= 10 x ID tolerance, ranks within +-2 per box.
* Wavy N = 262,144, plain FP64 and learned-rank randomized ID (the headline
  configuration): each within 10 x tol of the oracle's solution.
* 256 star problems of 2,048 to 32,768 nodes in one batch: every
  problem's backward error through the block-diagonal FP64 operator
  <= 50 x tol (the reference's bound, ie_solver_test.cpp:84-95), and a
  seeded sample of 8 problems within 10 x tol of the oracle.

The circle (N = 4,096, closed form) and Stokes with 8 holes (N = 131,072,
FP32 + refinement) run in test_gpu_parity.py and test_gpu_operators.py.
"""
import os

import numpy as np
import pytest

import geometry as G
import iesolver
import oracle

pytestmark = pytest.mark.gpu

TOL = 1e-10
THREADS = min(16, os.cpu_count() or 1)
MODEL = os.path.join(os.path.dirname(iesolver.__file__), "rank_model_default.npy")


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _oracle(g, f):
    o = oracle.OracleProblem(g)
    o.factor(TOL, THREADS)
    mo, _ = o.solve(f, THREADS)
    return o, mo


def test_star65k_plain_vs_oracle():
    g = G.star(65536)
    ys, q = G.exterior_sources(0)
    f = G.harmonic_data(g["points"], ys, q)
    s = iesolver.RSSolver(0)
    s.set_geometry(g)
    s.build_tree(64)
    s.factor(TOL)
    mu, _ = s.solve(f)
    o, mo = _oracle(g, f)
    err = _rel(mu, mo)
    print(f"star65k plain: rel err vs oracle {err:.3e}", flush=True)
    assert err <= 10 * TOL
    gi, oi = s.box_info(), o.box_info()
    assert np.array_equal(gi["level"], oi["level"])
    comp = (gi["rank"] > 0) | (oi["rank"] > 0)
    assert np.abs(gi["rank"][comp] - oi["rank"][comp]).max() <= 2


@pytest.fixture(scope="module")
def wavy262k():
    g = G.wavy(262144, seed=0)
    ys, q = G.exterior_sources(0)
    f = G.harmonic_data(g["points"], ys, q)
    _, mo = _oracle(g, f)
    return g, f, mo


@pytest.mark.parametrize("form", ["plain", "learned"])
def test_wavy262k_vs_oracle(wavy262k, form):
    g, f, mo = wavy262k
    s = iesolver.RSSolver(0)
    s.set_geometry(g)
    s.build_tree(64)
    if form == "learned":
        s.set_rank_model(np.load(MODEL), margin=3)
        s.set_id_method(iesolver.ID_RANDOMIZED, 0)
    s.factor(TOL)
    mu, _ = s.solve(f)
    err = _rel(mu, mo)
    print(f"wavy262k {form}: rel err vs oracle {err:.3e}", flush=True)
    assert err <= 10 * TOL
    # backward error through the FP64 operator (ie_solver_test.cpp:84-95)
    y, _ = s.apply(mu)
    assert _rel(y, f) <= 50 * TOL


def test_batch256_backward_error_and_oracle_sample():
    geoms, meta = G.star_batch(256, seed=0, nmin=2048, nmax=32768)
    f = G.batch_data(geoms, seed=0)
    s = iesolver.RSSolver(0)
    s.set_geometry_batch(geoms)
    s.build_tree(64)
    s.factor(TOL)
    mu, _ = s.solve(f)
    y, _ = s.apply(mu)
    off = s.dof_offsets
    worst = 0.0
    for p in range(256):
        sl = slice(off[p], off[p + 1])
        worst = max(worst, _rel(y[sl], f[sl]))
    print(f"batch256: {int(off[-1])} unknowns, worst backward error {worst:.3e}", flush=True)
    assert worst <= 50 * TOL
    sample = np.random.default_rng(0).choice(256, 8, replace=False)
    for p in sample:
        sl = slice(off[p], off[p + 1])
        _, mo = _oracle(geoms[p], f[sl])
        err = _rel(mu[sl], mo)
        print(f"  problem {p} (N={meta['sizes'][p]}): rel err vs oracle {err:.3e}", flush=True)
        assert err <= 10 * TOL
