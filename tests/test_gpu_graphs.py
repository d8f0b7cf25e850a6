"""Factorization replayed from a captured HIP graph (iegpu_set_graphs): every
replay must give the eager factorization's solution bit for bit, with the
program's own eager work (solves, operator applies, synchronizations)
between the replays, for the plain and the randomized forms."""
import numpy as np
import pytest

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
