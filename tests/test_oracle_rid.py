"""Pins of the oracle's randomized-ID restatement (CPU only).

The randomized ID has no reference counterpart; its checker
(oracle.randomized_id, oracle.gaussian_table) is pinned here by published
known answers (SplitMix64), distribution statistics, exact-rank recovery,
the fallback rule, and agreement with LAPACK's pivoted QR ID on the
reference-defined ID matrices of real boxes.
"""
import numpy as np
import pytest

import geometry as G
import oracle


def test_splitmix64_known_answers():
    # SplitMix64 (Steele, Lea & Flood 2014) from state 0: the first outputs
    # are mix(g), mix(2g), mix(3g) with g = 0x9E3779B97F4A7C15.
    g = np.uint64(0x9E3779B97F4A7C15)
    with np.errstate(over="ignore"):
        outs = [int(oracle._splitmix64(np.uint64(i) * g)) for i in range(3)]
    assert outs == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_gaussian_table_is_standard_normal():
    t = oracle.gaussian_table()
    assert t.shape == (oracle.RID_TABLE_ROWS, oracle.RID_ROWS)
    assert np.array_equal(t, oracle.gaussian_table())  # deterministic
    x = t.ravel()
    assert abs(x.mean()) < 0.03 and abs(x.std() - 1) < 0.02
    assert abs(np.mean(x ** 4) - 3) < 0.15  # Gaussian kurtosis
    # columns are not copies of one another
    c = np.corrcoef(t.T)
    assert np.abs(c - np.eye(c.shape[0])).max() < 0.15


def _lowrank(m, n, rank, decay, seed):
    rng = np.random.default_rng(seed)
    U = np.linalg.qr(rng.standard_normal((m, rank)))[0]
    V = np.linalg.qr(rng.standard_normal((n, rank)))[0]
    return U @ np.diag(np.logspace(0, decay, rank)) @ V.T


def test_randomized_id_exact_rank():
    A = _lowrank(400, 40, 12, -3, 0)
    k, perm, T = oracle.randomized_id(A, 1e-10, 32)
    assert k == 12
    assert sorted(perm) == list(range(40))
    err = np.linalg.norm(A[:, perm[k:]] - A[:, perm[:k]] @ T) / np.linalg.norm(A)
    assert err < 1e-12


def test_randomized_id_falls_back_when_the_sketch_is_too_narrow():
    A = _lowrank(400, 60, 40, -6, 1)
    assert oracle.randomized_id(A, 1e-10, 32) is None
    assert oracle.randomized_id(A, 1e-10, 16) is None
    k, _, _ = oracle.randomized_id(A[:, :20], 1e-10, 32)  # n <= rows - 6: every column
    assert k == 20


@pytest.mark.parametrize("pde,geom", [(0, "star"), (2, "circle")])
def test_randomized_id_tracks_lapack_qrcp_on_box_id_matrices(pde, geom):
    g = G.star(4096) if geom == "star" else G.circle(2048)
    tol = 1e-10
    o = oracle.OracleProblem(g, pde=pde, record_id_mats=True)
    o.factor(tol, 4)
    bi = o.box_info()
    table = oracle.gaussian_table()
    seen = 0
    for b in range(o.num_boxes()):
        if bi["rank"][b] <= 0:
            continue
        A = o.id_mat(b)
        m, n = A.shape
        if n > 64 or m > oracle.RID_TABLE_ROWS:
            continue
        r = oracle.randomized_id(A, tol, 32, table)
        if r is None:
            continue
        k, perm, T = r
        ka = oracle.interp_decomp(A, tol)[0]
        assert -1 <= k - ka <= 3, (b, k, ka)
        if k < n:
            err = np.linalg.norm(A[:, perm[k:]] - A[:, perm[:k]] @ T, 2) / np.linalg.norm(A, 2)
            assert err <= 2 * tol, (b, err)
        seen += 1
    assert seen >= 10
