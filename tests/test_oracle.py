"""Pins of the CPU oracle (oracle/rs_oracle.cpp) — CPU only.

The reference stores no golden vectors (its .gitignore drops *.txt); its
known-answer tests are analytic agreements and a backward-error bound in
ie_solver/ie_solver_test.cpp.  Those are restated here on the oracle,
together with closed forms, dense LAPACK solves, brute-force tree checks
and the LAPACK pivoted QR, each chosen so that a dropped term, a wrong
sign or index, or a transposed block fails.
"""
import math

import numpy as np
import pytest
import scipy.linalg as sla

import geometry as G
import oracle

TOL = 1e-10


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


# --------------------------------------------------------------------------
# Kernel entries (kernel.cpp:123-165, 212-290)
# --------------------------------------------------------------------------

def test_laplace_circle_closed_form():
    # Circle radius R, N nodes: off-diagonal w/(4 pi R) = 1/(2N), diagonal
    # 1/2 + kappa w/(4 pi) = 1/2 + 1/(2N): K = I/2 + 11^T/(2N).
    for n, rad in ((256, 0.5), (512, 0.25)):
        g = G.circle(n, (0.5, 0.5), rad)
        K = oracle.dense_matrix(g)
        assert np.abs(K - (0.5 * np.eye(n) + 1.0 / (2 * n))).max() < 1e-13


def test_laplace_gauss_row_sums():
    # Double layer of a constant density: every row sums to 1 on a smooth
    # closed curve (sign of the normal term and the curvature diagonal).
    K = oracle.dense_matrix(G.star(2048))
    assert np.abs(K.sum(axis=1) - 1).max() < 1e-11


def test_stokes_translation_identity():
    # Stokes double layer (+ w n n^T) of a rigid translation e equals -e.
    g = G.star(1024)
    K = oracle.dense_matrix(g, oracle.PDE_STOKES)
    for comp in (0, 1):
        e = np.zeros(2048)
        e[comp::2] = 1
        assert np.abs(K @ e + e).max() < 1e-11


def test_stokes_entry_structure():
    # Off-diagonal 2x2 block: (w/pi) (r.n)/|r|^4 r r^T + w n_t n_s^T.
    g = G.star(64)
    p, nm, w = g["points"], g["normals"], g["weights"]
    i, j = 3, 17
    K = oracle.kernel_block(g, [2 * i, 2 * i + 1], [2 * j, 2 * j + 1], oracle.PDE_STOKES)
    r = p[i] - p[j]
    expect = w[j] / math.pi * (r @ nm[j]) / (r @ r) ** 2 * np.outer(r, r) + w[j] * np.outer(nm[i], nm[j])
    assert np.abs(K - expect).max() < 1e-14


def test_proxy_block_formula():
    # Rows [0,128): proxy <- box; rows [128,256): (box <- proxy)^T with
    # 128 points on radius r, normal radial, weight 2 pi r / 128.
    g = G.star(512)
    box = np.arange(40, 70)
    cx, cy, r = 0.6, 0.5, 0.07
    M = oracle.proxy_block(g, cx, cy, r, box)
    ang = 2 * math.pi * np.arange(128) / 128
    pp = np.stack([cx + r * np.cos(ang), cy + r * np.sin(ang)], 1)
    pn = np.stack([np.cos(ang), np.sin(ang)], 1)
    pw = 2 * math.pi * r / 128
    p, nm, w = g["points"], g["normals"], g["weights"]
    d = pp[:, None, :] - p[None, box, :]
    top = -w[box] / (2 * math.pi) * np.einsum("pbk,bk->pb", d, nm[box]) / np.einsum("pbk,pbk->pb", d, d)
    e = p[None, box, :] - pp[:, None, :]
    bot = -pw / (2 * math.pi) * np.einsum("pbk,pk->pb", e, pn) / np.einsum("pbk,pbk->pb", e, e)
    assert np.abs(M[:128] - top).max() < 1e-12
    assert np.abs(M[128:] - bot).max() < 1e-12


# --------------------------------------------------------------------------
# Interpolative decomposition (ie_mat.cpp:548-588)
# --------------------------------------------------------------------------

@pytest.mark.parametrize("seed", [0, 1, 2])
def test_id_reconstruction_and_rank(seed):
    rng = np.random.default_rng(seed)
    m, n, rk = 300, 48, 17
    A = rng.standard_normal((m, rk)) @ np.diag(10.0 ** -np.linspace(0, 14, rk)) @ rng.standard_normal((rk, n))
    k, perm, T = oracle.interp_decomp(A, TOL)
    # Rank cut rule: first i >= 1 with |R_ii| < tol |R_00| of LAPACK dgeqp3.
    R = sla.qr(A, pivoting=True, mode="r")[0]
    d = np.abs(np.diag(R))
    expect = next(i for i in range(1, n) if d[i] < TOL * d[0])
    assert k == expect
    sk, rd = perm[:k], perm[k:]
    err = np.linalg.norm(A[:, rd] - A[:, sk] @ T, 2)
    assert err <= 50 * TOL * np.linalg.norm(A, 2)


def test_id_full_rank_returns_zero():
    A = np.random.default_rng(3).standard_normal((80, 20))
    k, _, _ = oracle.interp_decomp(A, TOL)
    assert k == 0


# --------------------------------------------------------------------------
# Quadtree (quadtree.cpp:39-432) against brute force
# --------------------------------------------------------------------------

def test_tree_partition_and_neighbours():
    g = G.wavy(6000, seed=2)
    o = oracle.OracleProblem(g, leaf_cap=64)
    info = o.box_info()
    nb = o.num_boxes()
    leaves = np.where(info["is_leaf"] == 1)[0]
    seen = np.concatenate([o.box_list(b, 0) for b in leaves])
    assert np.array_equal(np.sort(seen), np.arange(6000))
    for b in range(nb):
        cnt = len(o.box_list(b, 0))
        if info["is_leaf"][b]:
            assert cnt <= 64
        else:
            assert cnt > 64
    # Same-level neighbours = boxes at Chebyshev centre distance <= side.
    bl = info["bl_side"]
    for lv in range(o.num_levels()):
        ids = o.level_order(lv)
        for a in ids:
            got = {x for x in o.box_list(a, 1) if info["level"][x] == lv}
            want = {c for c in ids if c != a and
                    max(abs(bl[a, 0] - bl[c, 0]), abs(bl[a, 1] - bl[c, 1])) <= bl[a, 2] * 1.0000001}
            assert got == want, (a, got ^ want)


# --------------------------------------------------------------------------
# Factor + solve (skel_factorization.cpp) — closed forms, dense LAPACK,
# and the reference's own analytic / backward-error tests
# --------------------------------------------------------------------------

def test_circle_closed_form_solution():
    # mu = 2 f - mean(f) for K = I/2 + 11^T/(2N).
    g = G.circle(4096)
    f = np.log(np.hypot(g["points"][:, 0] - 2, g["points"][:, 1] - 2))
    o = oracle.OracleProblem(g)
    assert o.factor(TOL, 4) == 0
    mu, _ = o.solve(f, 4)
    assert rel(mu, 2 * f - f.mean()) <= 10 * TOL
    tg = G.interior_points(0, 100, radius=0.4)
    u = oracle.evaluate(g, tg, mu)
    assert rel(u, np.log(np.hypot(tg[:, 0] - 2, tg[:, 1] - 2))) <= 10 * TOL


@pytest.mark.parametrize("shape", ["star", "wavy"])
def test_dense_solve_agreement(shape):
    g = G.star(3000) if shape == "star" else G.wavy(3000, seed=4)
    ys, q = G.exterior_sources(5)
    f = G.harmonic_data(g["points"], ys, q)
    o = oracle.OracleProblem(g)
    o.factor(TOL, 4)
    mu, _ = o.solve(f, 4)
    md = np.linalg.solve(oracle.dense_matrix(g), f)
    assert rel(mu, md) <= 10 * TOL


def test_reference_laplace_circle_analytic():
    # ie_solver_test.cpp:404-416: SINGLE_ELECTRON data on the default
    # circle, interior u = log|x-(-3,-2)|/(2 pi); here at tol 1e-10.
    g = G.reference_circle(512)
    src = lambda x: np.log(np.hypot(x[:, 0] + 3, x[:, 1] + 2)) / (2 * math.pi)
    o = oracle.OracleProblem(g, leaf_cap=32)
    o.factor(TOL, 4)
    mu, _ = o.solve(src(g["points"]), 4)
    tg = G.interior_points(1, 200, radius=0.2)
    assert rel(oracle.evaluate(g, tg, mu), src(tg)) <= 10 * TOL


def test_reference_backward_error_bound():
    # ie_solver_test.cpp:84-95: ||K mu - f||_inf <= 50 tol.
    g = G.star(2048)
    ys, q = G.exterior_sources(6)
    f = G.harmonic_data(g["points"], ys, q)
    o = oracle.OracleProblem(g)
    o.factor(TOL, 4)
    mu, _ = o.solve(f, 4)
    assert np.abs(oracle.dense_matrix(g) @ mu - f).max() <= 50 * TOL


def _u_forward_laplace(targets, centers):
    return np.log(np.hypot(targets[:, None, 0] - centers[None, :, 0],
                           targets[:, None, 1] - centers[None, :, 1]))


def test_reference_laplace_annulus_analytic():
    # ie_solver_test.cpp:419-431 (annulus with two holes, SINGLE_ELECTRON),
    # evaluated as K_domain mu + U_forward alpha (linear_solve.cpp:148-152).
    g = G.annulus(1000)
    src = lambda x: np.log(np.hypot(x[:, 0] + 3, x[:, 1] + 2)) / (2 * math.pi)
    o = oracle.OracleProblem(g)
    o.factor(TOL, 4)
    mu, alpha = o.solve(src(g["points"]), 4)
    rng = np.random.default_rng(0)
    tg = []
    while len(tg) < 200:
        x = rng.uniform(-0.4, 1.4, 2)
        if (np.hypot(*(x - 0.5)) < 0.9 and all(np.hypot(*(x - c)) > 0.3 for c in g["holes"]["centers"])):
            tg.append(x)
    tg = np.array(tg)
    u = oracle.evaluate(g, tg, mu) + _u_forward_laplace(tg, g["holes"]["centers"]) @ alpha
    assert rel(u, src(tg)) <= 10 * TOL


def test_reference_stokes_circle_rotation():
    # ie_solver_test.cpp:528-543: tangential data on a circle gives rigid
    # rotation inside, |u| = |r| / R.
    g = G.circle(1024, (0.5, 0.5), 0.5)
    f = G.tangent_data(g["normals"])
    o = oracle.OracleProblem(g, pde=oracle.PDE_STOKES)
    o.factor(TOL, 4)
    mu, _ = o.solve(f, 4)
    tg = G.interior_points(2, 100, radius=0.35)
    u = oracle.evaluate(g, tg, mu, oracle.PDE_STOKES).reshape(-1, 2)
    r = tg - 0.5
    truth = np.stack([-r[:, 1], r[:, 0]], 1) / 0.5
    assert np.abs(u - truth).max() <= 10 * TOL * 10


def _u_forward_stokes(targets, centers):
    # initialize_U_mat on targets (linear_solve.cpp:37-60).
    nt, h = targets.shape[0], centers.shape[0]
    U = np.zeros((2 * nt, 3 * h))
    s = 1 / (4 * math.pi)
    for k in range(h):
        r = targets - centers[k]
        rn = np.hypot(r[:, 0], r[:, 1])
        U[0::2, 3 * k] = s * (np.log(1 / rn) + r[:, 0] ** 2 / rn ** 2)
        U[1::2, 3 * k] = s * (r[:, 1] * r[:, 0] / rn ** 2)
        U[0::2, 3 * k + 1] = s * (r[:, 0] * r[:, 1] / rn ** 2)
        U[1::2, 3 * k + 1] = s * (np.log(1 / rn) + r[:, 1] ** 2 / rn ** 2)
        U[0::2, 3 * k + 2] = r[:, 1] * s / rn ** 2
        U[1::2, 3 * k + 2] = -r[:, 0] * s / rn ** 2
    return U


def test_reference_stokes_donut_couette():
    # ie_solver_test.cpp:545-562: donut (outer r=1, hole r=0.5 about
    # (0.5,0.5)), tangential data; Couette flow u_theta = 2r - 1/r.
    n = 2048
    i = np.arange(n)
    ang = i * 2.0 * math.pi / n
    hn = n // 5
    a2 = np.arange(hn) * 2.0 * math.pi / hn
    pts = np.concatenate([np.stack([0.5 + np.cos(ang), 0.5 + np.sin(ang)], 1),
                          np.stack([0.5 + 0.5 * np.cos(a2), 0.5 + 0.5 * np.sin(a2)], 1)])
    nrm = np.concatenate([np.stack([np.cos(ang), np.sin(ang)], 1),
                          np.stack([-np.cos(a2), -np.sin(a2)], 1)])
    g = dict(points=pts, normals=nrm,
             curvatures=np.concatenate([np.ones(n), np.full(hn, -2.0)]),
             weights=np.concatenate([np.full(n, 2 * math.pi / n), np.full(hn, math.pi / hn)]),
             holes=dict(centers=np.array([[0.5, 0.5]]), radii=np.array([0.5])))
    f = G.tangent_data(nrm)
    o = oracle.OracleProblem(g, pde=oracle.PDE_STOKES)
    o.factor(TOL, 4)
    mu, alpha = o.solve(f, 4)
    rng = np.random.default_rng(1)
    rr = rng.uniform(0.6, 0.9, 100)
    th = rng.uniform(0, 2 * math.pi, 100)
    tg = np.stack([0.5 + rr * np.cos(th), 0.5 + rr * np.sin(th)], 1)
    u = (oracle.evaluate(g, tg, mu, oracle.PDE_STOKES) +
         _u_forward_stokes(tg, g["holes"]["centers"]) @ alpha).reshape(-1, 2)
    p = -1 / rr + 2 * rr
    truth = np.stack([-np.sin(th), np.cos(th)], 1) * p[:, None]
    assert np.abs(u - truth).max() <= 1e-7


def test_stokes_holes_dense_bordered():
    # Bordered system [[K, U], [Psi, -I]] (ie_solver_test.cpp:43-82).
    g = G.circle_with_holes(1600, 8, seed=1)
    f = G.tangent_data(g["normals"])
    f[2 * g["counts"][0]:] = 0
    o = oracle.OracleProblem(g, pde=oracle.PDE_STOKES)
    o.factor(TOL, 4)
    mu, alpha = o.solve(f, 4)
    U, Psi = o.U_Psi()
    K = oracle.dense_matrix(g, oracle.PDE_STOKES)
    h = U.shape[1]
    D = np.block([[K, U], [Psi, -np.eye(h)]])
    x = np.linalg.solve(D, np.concatenate([f, np.zeros(h)]))
    assert rel(np.concatenate([mu, alpha]), x) <= 100 * TOL


def test_id_matrix_stacking():
    # ID rows of a level >= 3 box = [K(ic, box); K(box, ic)^T; proxy]
    # (kernel.cpp:384-418), rebuilt here from the dense K.
    g = G.star(4096)
    o = oracle.OracleProblem(g, record_id_mats=True)
    o.factor(TOL, 2)
    info = o.box_info()
    K = oracle.dense_matrix(g)
    # a box of the deepest level with a recorded matrix
    lv = info["level"].max()
    b = next(x for x in o.level_order(lv) if info["m_rows"][x] > 0)
    M = o.id_mat(b)
    act = o.box_list(b, 2)
    near = o.box_list(b, 4)
    x0, y0, side = info["bl_side"][b]
    c = np.array([x0 + side / 2, y0 + side / 2])
    p = g["points"]
    ic = [d for d in near if math.hypot(*(c - p[d])) < 1.5 * side]
    nr = len(ic)
    assert M.shape == (2 * nr + 256, len(act))
    assert np.array_equal(M[:nr], K[np.ix_(ic, act)])
    assert np.array_equal(M[nr:2 * nr], K[np.ix_(act, ic)].T)
    assert np.allclose(M[2 * nr:], oracle.proxy_block(g, c[0], c[1], 1.5 * side, act), rtol=0, atol=1e-14)


# --------------------------------------------------------------------------
# Laplace Neumann (kernel.cpp:80-102, 168-209, 594-646)
# --------------------------------------------------------------------------

NEU = oracle.PDE_LAPLACE_NEUMANN


def test_neumann_circle_closed_form():
    # On the circle r = 0.5, (t - s) . n_t / |t - s|^2 = 1/(2R) for any two
    # points, so K = (w + 1/(2N)) 11^T - I/2 with w = pi/N.
    n = 256
    g = G.circle(n)
    K = oracle.dense_matrix(g, NEU)
    want = np.full((n, n), math.pi / n + 1 / (2 * n)) - 0.5 * np.eye(n)
    assert np.abs(K - want).max() <= 1e-12


def test_neumann_is_adjoint_double_layer_plus_weights():
    # Off the diagonal K_N(i,j) - w_j = (w_j / w_i) K_L(j,i): the target's
    # normal replaces the source's.  On it K_N(i,i) = K_L(i,i) + w_i - 1.
    g = G.star(300, amp=0.2, k=5)
    KN = oracle.dense_matrix(g, NEU)
    KL = oracle.dense_matrix(g, oracle.PDE_LAPLACE)
    w = g["weights"]
    off = ~np.eye(300, dtype=bool)
    lhs = (KN - w[None, :])[off]
    rhs = ((w[None, :] / w[:, None]) * KL.T)[off]
    assert np.abs(lhs - rhs).max() <= 1e-12 * np.abs(rhs).max()
    assert np.abs(np.diag(KN) - (np.diag(KL) + w - 1)).max() <= 1e-14


def test_neumann_proxy_block_formula():
    g = G.star(512)
    box = np.arange(40, 70)
    cx, cy, r = 0.6, 0.5, 0.07
    M = oracle.proxy_block(g, cx, cy, r, box, NEU)
    ang = 2 * math.pi * np.arange(128) / 128
    pp = np.stack([cx + r * np.cos(ang), cy + r * np.sin(ang)], 1)
    pn = np.stack([np.cos(ang), np.sin(ang)], 1)
    pw = 2 * math.pi * r / 128
    p, nm, w = g["points"], g["normals"], g["weights"]
    d = pp[:, None, :] - p[None, box, :]  # proxy target <- box source
    top = w[box] + w[box] / (2 * math.pi) * np.einsum("pbk,pk->pb", d, pn) / np.einsum("pbk,pbk->pb", d, d)
    e = p[None, box, :] - pp[:, None, :]  # box target <- proxy source
    bot = pw + pw / (2 * math.pi) * np.einsum("pbk,bk->pb", e, nm[box]) / np.einsum("pbk,pbk->pb", e, e)
    assert np.abs(M[:128] - top).max() < 1e-12
    assert np.abs(M[128:] - bot).max() < 1e-12


@pytest.mark.parametrize("shape", ["circle", "star", "annulus"])
def test_reference_neumann_backward_error(shape):
    # ie_solver_test.cpp:124-135, 165-176, 209-220 through check_solve_err
    # (:29-96): all-ones data, N = 512, tol 1e-6, |K mu - f| <= 50 tol; the
    # annulus's holes take no border for Laplace Neumann.
    tol = 1e-6
    g = {"circle": G.circle(512), "star": G.star(512), "annulus": G.annulus(512)}[shape]
    n = g["points"].shape[0]
    f = np.ones(n)
    o = oracle.OracleProblem(g, pde=NEU)
    assert o.hole_dim == 0
    o.factor(tol, 4)
    mu, alpha = o.solve(f, 4)
    assert alpha is None or len(alpha) == 0
    K = oracle.dense_matrix(g, NEU)
    assert np.abs(K @ mu - f).max() <= 50 * tol


def test_reference_neumann_donut_analytic():
    # ie_solver_test.cpp:433-471, 510-525: donut, data 1 on the outer
    # circle and -1/r_hole on the hole; the single-layer field is log r plus
    # a constant.
    tol = 1e-6
    g = G.donut(512)
    n, hn = g["counts"]
    f = np.concatenate([np.ones(n), np.full(hn, -2.0)])
    o = oracle.OracleProblem(g, pde=NEU)
    o.factor(tol, 4)
    mu, _ = o.solve(f, 4)
    rng = np.random.default_rng(2)
    rr = rng.uniform(0.6, 0.9, 100)
    th = rng.uniform(0, 2 * math.pi, 100)
    tg = np.stack([0.5 + rr * np.cos(th), 0.5 + rr * np.sin(th)], 1)
    d = oracle.evaluate(g, tg, mu, NEU) - np.log(rr)
    assert np.abs(d - d.mean()).max() <= 10 * tol


# --------------------------------------------------------------------------
# Compressed forward apply (skel_factorization.cpp:429-504)
# --------------------------------------------------------------------------

@pytest.mark.parametrize("pde", [oracle.PDE_LAPLACE, oracle.PDE_LAPLACE_NEUMANN, oracle.PDE_STOKES])
def test_sparse_matvec_inverts_solve_and_approximates_K(pde):
    # The forward apply undoes the solve step by step, so solve(A x) = x to
    # rounding; and A equals the dense operator up to the ID error.
    tol = 1e-10
    g = G.star(2048) if pde != oracle.PDE_STOKES else G.circle(1024)
    o = oracle.OracleProblem(g, pde=pde)
    o.factor(tol, 4)
    x = np.random.default_rng(7).standard_normal(o.ndof)
    b = o.sparse_matvec(x)
    y, _ = o.solve(b, 4)
    assert rel(y, x) <= 1e-12
    K = oracle.dense_matrix(g, pde)
    assert rel(b, K @ x) <= 50 * tol
