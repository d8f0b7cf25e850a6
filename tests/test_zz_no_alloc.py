"""The allocation-free interior evaluation (iegpu_evaluate_noalloc) against
iegpu_evaluate, through the C ABI, on two calls in a row (the second with
fewer targets): both run the same kernels on the same partial-sum layout,
so the fields agree bit for bit.  The partial sums live in the library's
static device scratch; a call that would not fit it is rejected."""
import ctypes

import numpy as np
import pytest
# torch's HIP runtime initializes before the library loads (device buffers)
import torch

import geometry as G
import iesolver

pytestmark = pytest.mark.gpu


def _noalloc(s, tg, mu, alpha=None, mask=None):
    L = s.L
    L.iegpu_evaluate_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_void_p]
    dev = torch.device("cuda:0")
    dt = torch.from_numpy(np.ascontiguousarray(tg, dtype=np.float64)).to(dev)
    dm = torch.from_numpy(np.ascontiguousarray(mu, dtype=np.float64)).to(dev)
    da = None if alpha is None else torch.from_numpy(np.asarray(alpha, np.float64)).to(dev)
    dk = None if mask is None else torch.from_numpy(np.asarray(mask, np.int32)).to(dev)
    sd = 2 if s.pde == iesolver.PDE_STOKES else 1
    out = torch.full((tg.shape[0] * sd,), np.nan, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    st = L.iegpu_evaluate_noalloc(s.h, tg.shape[0], dt.data_ptr(),
                                  None if dk is None else dk.data_ptr(), dm.data_ptr(),
                                  None if da is None else da.data_ptr(), out.data_ptr(), None)
    torch.cuda.synchronize()
    return st, out.cpu().numpy()


@pytest.mark.parametrize("pde", [iesolver.PDE_LAPLACE, iesolver.PDE_STOKES])
def test_noalloc_evaluate_matches_evaluate_twice(pde):
    g = G.star(20000)
    s = iesolver.RSSolver(0)
    s.set_geometry(g, pde)
    rng = np.random.default_rng(7)
    mu = rng.standard_normal(s.ndof)
    for nt in (5000, 700):  # the second call smaller than the first
        tg = G.interior_points(nt, nt, radius=0.2)
        st, u = _noalloc(s, tg, mu)
        assert st == 0, s.L.iegpu_last_error(s.h)
        ref = s.evaluate(tg, mu)
        assert np.array_equal(u, ref), nt


def test_noalloc_evaluate_holes_mask_and_capacity():
    g = G.annulus(1000)
    s = iesolver.RSSolver(0)
    s.set_geometry(g)
    s.build_tree(64)
    s.set_holes(g["holes"]["centers"], g["holes"]["radii"])
    mu = np.random.default_rng(4).standard_normal(s.ndof)
    alpha = np.array([0.3, -0.7])
    for tg, mask in ((G.interior_points(9, 300, radius=0.3), None),
                     (np.array([[0.5, 0.9], [0.5, 0.2], [0.9, 0.5]]), np.array([1, 1, 0]))):
        st, u = _noalloc(s, tg, mu, alpha, mask)
        assert st == 0
        assert np.array_equal(u, s.evaluate(tg, mu, alpha, mask))
    # more partial sums than the static scratch holds: a status, nothing run
    big = G.interior_points(1, (1 << 19) + 1, radius=0.3)
    st, u = _noalloc(s, big, mu, alpha)
    assert st == 6 and "static scratch" in s.L.iegpu_last_error(s.h).decode()
    assert np.isnan(u).all()
