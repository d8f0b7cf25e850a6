"""Factorization replayed from a captured HIP graph (iegpu_set_graphs): every
replay must give the eager factorization's solution bit for bit, with the
program's own eager work (solves, operator applies, synchronizations)
between the replays, for the plain and the randomized forms."""
import numpy as np
import pytest
# torch's HIP runtime has to initialize before this library loads the
# system one (as in bench.py), so it is imported when the module is.
import torch

import geometry as G
import iesolver

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("form", ["plain", "randomized"])
def test_graph_replays_match_eager(form):
    g = G.wavy(16384, seed=1)
    f = np.random.default_rng(2).standard_normal(16384)
    s = iesolver.RSSolver(0)
    s.set_geometry(g, iesolver.PDE_LAPLACE)
    s.build_tree(64)
    if form == "randomized":
        s.set_id_method(iesolver.ID_RANDOMIZED, 32)
    s.factor(1e-10)
    mu0, _ = s.solve(f)
    s.set_graphs(True)
    for rep in range(4):
        s.factor(1e-10)  # first call captures, later calls replay
        mu, _ = s.solve(f)
        print(f"{form} replay {rep}: max |diff| {np.abs(mu - mu0).max():.1e}", flush=True)
        assert np.array_equal(mu, mu0)
        y, _ = s.apply(mu)  # eager work between replays
        assert np.linalg.norm(y - f) <= 1e-9 * np.linalg.norm(f)
    # a tolerance change recaptures; going back replays the first state again
    s.factor(1e-8)
    mu8, _ = s.solve(f)
    assert not np.array_equal(mu8, mu0)
    s.factor(1e-10)
    assert np.array_equal(s.solve(f)[0], mu0)
    s.set_graphs(False)
    s.factor(1e-10)
    assert np.array_equal(s.solve(f)[0], mu0)


@pytest.mark.parametrize("case", ["plain", "randomized", "stokes_holes", "fp32"])
def test_overlap_matches_single_stream(case):
    # Overlapping each level's ID with the decouple of the level below (a
    # second stream) changes only the order in which independent kernels
    # run: the solution is bit-identical to the single-stream run, eager and
    # replayed from a graph.
    pde, prec, tol = iesolver.PDE_LAPLACE, iesolver.FP64, 1e-10
    if case == "stokes_holes":
        g = G.circle_with_holes(8192, 8, seed=2)
        f = G.tangent_data(g["normals"])
        f[2 * g["counts"][0]:] = 0
        pde = iesolver.PDE_STOKES
    else:
        g = G.wavy(32768, seed=4)
        f = G.harmonic_data(g["points"], *G.exterior_sources(4))
    if case == "fp32":
        prec, tol = iesolver.FP32, 1e-6
    s = iesolver.RSSolver(0)
    s.set_geometry(g, pde)
    s.build_tree(64)
    if "holes" in g:
        s.set_holes(g["holes"]["centers"], g["holes"]["radii"])
    if case == "randomized":
        s.set_id_method(iesolver.ID_RANDOMIZED, 32)
    s.set_overlap(False)
    s.factor(tol, prec)
    m0, a0 = s.solve(f)
    s.set_overlap(True)
    for graphs in (False, True, True, True):
        s.set_graphs(graphs)
        s.factor(tol, prec)
        m1, a1 = s.solve(f)
        assert np.array_equal(m0, m1), (case, graphs)
        if a0 is not None:
            assert np.array_equal(a0, a1)


@pytest.mark.parametrize("case", ["plain", "randomized", "fp32", "batch"])
def test_fused_factor_solve_matches_separate(case):
    # iegpu_factor_solve_device runs each level's upward sweep beside the
    # factorization: the same kernels on the same data, so x is
    # bit-identical to iegpu_factor + iegpu_solve_device, eager and replayed.
    prec, tol = iesolver.FP64, 1e-10
    s = iesolver.RSSolver(0)
    if case == "batch":
        geoms, _ = G.star_batch(4, seed=11, nmin=1024, nmax=8192)
        s.set_geometry_batch(geoms)
        f = G.batch_data(geoms, seed=11)
    else:
        g = G.wavy(32768, seed=12)
        s.set_geometry(g)
        f = G.harmonic_data(g["points"], *G.exterior_sources(12))
    s.build_tree(64)
    if case == "randomized":
        s.set_id_method(iesolver.ID_RANDOMIZED, 32)
    if case == "fp32":
        prec, tol = iesolver.FP32, 1e-6
    dev = torch.device("cuda:0")
    fd = torch.from_numpy(f).to(dev)
    x0 = torch.empty_like(fd)
    # the library works on its own non-blocking stream: torch's fills must
    # be finished before a buffer is handed over
    torch.cuda.synchronize()
    s.factor(tol, prec)
    s.solve_device(fd.data_ptr(), x0.data_ptr())
    torch.cuda.synchronize()
    ref = x0.cpu().numpy()
    for graphs in (False, True, True):
        s.set_graphs(graphs)
        x1 = torch.zeros_like(fd)
        torch.cuda.synchronize()
        s.factor_solve_device(tol, prec, fd.data_ptr(), x1.data_ptr())
        torch.cuda.synchronize()
        assert np.array_equal(x1.cpu().numpy(), ref), (case, graphs)
        # the stored factorization is the usual one
        x2 = torch.zeros_like(fd)
        torch.cuda.synchronize()
        s.solve_device(fd.data_ptr(), x2.data_ptr())
        torch.cuda.synchronize()
        assert np.array_equal(x2.cpu().numpy(), ref)
