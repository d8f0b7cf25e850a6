"""ctypes wrapper of the CPU oracle (liboracle_rs.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.  See
oracle/rs_oracle.cpp for the reference functions each entry restates.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_rs.so")
_lib = None

PDE_LAPLACE = 0
PDE_LAPLACE_NEUMANN = 1
PDE_STOKES = 2

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build(force=False):
    """Compile the oracle with its Makefile (CPU only)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"],
                       check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.rso_create.restype = ctypes.c_void_p
        L.rso_create.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp,
                                 ctypes.c_int]
        L.rso_destroy.argtypes = [ctypes.c_void_p]
        L.rso_perturb.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp]
        L.rso_mark.argtypes = [ctypes.c_void_p, _ip]
        L.rso_compressed.argtypes = [ctypes.c_void_p, _ip]
        L.rso_record_id_mats.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.rso_set_holes.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
        L.rso_hole_dim.argtypes = [ctypes.c_void_p]
        L.rso_get_U_Psi.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.rso_factor.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int]
        L.rso_solve.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, ctypes.c_int]
        L.rso_sparse_matvec.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.rso_num_levels.argtypes = [ctypes.c_void_p]
        L.rso_num_boxes.argtypes = [ctypes.c_void_p]
        L.rso_box_info.argtypes = [ctypes.c_void_p, _ip, _ip, _ip, _ip, _dp,
                                   _ip, _ip, _ip]
        L.rso_box_list.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                   _ip, ctypes.c_int]
        L.rso_allskel.argtypes = [ctypes.c_void_p, _ip, ctypes.c_int]
        L.rso_get_id_mat.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp,
                                     ctypes.c_long, _ip, _ip]
        L.rso_kernel_block.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp,
                                       _dp, ctypes.c_int, _ip, ctypes.c_int, _ip,
                                       _dp]
        L.rso_proxy_block.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp,
                                      _dp, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_int, _ip, _dp]
        L.rso_id.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_double,
                             _ip, _dp]
        L.rso_eval.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp,
                               ctypes.c_int, _dp, _dp, _dp]
        _lib = L
    return _lib


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _pd(a):
    return a.ctypes.data_as(_dp)


def _pi(a):
    return a.ctypes.data_as(_ip)


class OracleProblem:
    """One boundary, its quadtree and (after factor()) its RS factorization."""

    def __init__(self, geom, pde=PDE_LAPLACE, leaf_cap=64, record_id_mats=False):
        self.geom = geom
        self.pde = pde
        self.sol_dim = 2 if pde == PDE_STOKES else 1
        self.n_points = geom["points"].shape[0]
        self._keep = [_d(geom["points"]).reshape(-1), _d(geom["normals"]).reshape(-1),
                      _d(geom["curvatures"]), _d(geom["weights"])]
        self.h = lib().rso_create(pde, self.n_points, *[_pd(a) for a in self._keep],
                                  leaf_cap)
        if not self.h:
            raise ValueError("rso_create failed")
        if record_id_mats:
            lib().rso_record_id_mats(self.h, 1)
        self.hole_dim = 0
        holes = geom.get("holes")
        if holes is not None and len(holes["radii"]) > 0:
            c = _d(holes["centers"]).reshape(-1)
            r = _d(holes["radii"])
            lib().rso_set_holes(self.h, len(r), _pd(c), _pd(r))
            self.hole_dim = lib().rso_hole_dim(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().rso_destroy(self.h)
            self.h = None

    @property
    def ndof(self):
        return self.n_points * self.sol_dim

    def perturb(self, geom):
        """New data for the same nodes (QuadTree::perturb): the tree is
        updated and the boxes the reference marks lose their factors."""
        self.geom = geom
        self._keep = [_d(geom["points"]).reshape(-1), _d(geom["normals"]).reshape(-1),
                      _d(geom["curvatures"]), _d(geom["weights"])]
        lib().rso_perturb(self.h, *[_pd(a) for a in self._keep])

    def mark(self, marks=None):
        """Boxes (by id) that lose their factors; None: every box."""
        if marks is None:
            lib().rso_mark(self.h, None)
        else:
            m = np.ascontiguousarray(marks, dtype=np.int32)
            assert m.size == self.num_boxes()
            lib().rso_mark(self.h, _pi(m))

    def compressed(self):
        out = np.zeros(self.num_boxes(), np.int32)
        lib().rso_compressed(self.h, _pi(out))
        return out

    def factor(self, id_tol, num_threads=1):
        return lib().rso_factor(self.h, id_tol, num_threads)

    def solve(self, f, num_threads=1):
        f = _d(f)
        mu = np.zeros(self.ndof)
        if self.hole_dim:
            alpha = np.zeros(self.hole_dim)
            st = lib().rso_solve(self.h, _pd(f), _pd(mu), _pd(alpha), num_threads)
            if st:
                raise RuntimeError("oracle solve failed")
            return mu, alpha
        st = lib().rso_solve(self.h, _pd(f), _pd(mu), None, num_threads)
        if st:
            raise RuntimeError("oracle solve failed")
        return mu, None

    def U_Psi(self):
        if not self.hole_dim:
            return None, None
        U = np.zeros((self.hole_dim, self.ndof)).T.copy(order="F")
        Psi = np.zeros((self.hole_dim, self.ndof), order="F")
        Ub = np.zeros(self.ndof * self.hole_dim)
        Pb = np.zeros(self.ndof * self.hole_dim)
        lib().rso_get_U_Psi(self.h, _pd(Ub), _pd(Pb))
        U = Ub.reshape(self.hole_dim, self.ndof).T
        Psi = Pb.reshape(self.ndof, self.hole_dim).T
        return U, Psi

    def sparse_matvec(self, x):
        """b = A x with the compressed operator (skel_factorization.cpp:429-504)."""
        x = _d(x)
        b = np.zeros(self.ndof)
        if lib().rso_sparse_matvec(self.h, _pd(x), _pd(b)) != 0:
            raise RuntimeError("factor first")
        return b

    def num_levels(self):
        return lib().rso_num_levels(self.h)

    def num_boxes(self):
        return lib().rso_num_boxes(self.h)

    def box_info(self):
        nb = self.num_boxes()
        out = dict(level=np.zeros(nb, np.int32), parent=np.zeros(nb, np.int32),
                   is_leaf=np.zeros(nb, np.int32), children=np.zeros(4 * nb, np.int32),
                   bl_side=np.zeros(3 * nb), n_active=np.zeros(nb, np.int32),
                   rank=np.zeros(nb, np.int32), m_rows=np.zeros(nb, np.int32))
        lib().rso_box_info(self.h, _pi(out["level"]), _pi(out["parent"]),
                           _pi(out["is_leaf"]), _pi(out["children"]),
                           _pd(out["bl_side"]), _pi(out["n_active"]),
                           _pi(out["rank"]), _pi(out["m_rows"]))
        out["children"] = out["children"].reshape(nb, 4)
        out["bl_side"] = out["bl_side"].reshape(nb, 3)
        return out

    def box_list(self, box, which):
        n = lib().rso_box_list(self.h, box, which, None, 0)
        if n < 0:
            raise IndexError(box)
        buf = np.zeros(max(n, 1), np.int32)
        lib().rso_box_list(self.h, box, which, _pi(buf), n)
        return buf[:n].copy()

    def level_order(self, level):
        return self.box_list(level, 5)

    def allskel(self):
        n = lib().rso_allskel(self.h, None, 0)
        buf = np.zeros(max(n, 1), np.int32)
        lib().rso_allskel(self.h, _pi(buf), n)
        return buf[:n].copy()

    def id_mat(self, box):
        m = ctypes.c_int(0)
        n = ctypes.c_int(0)
        lib().rso_get_id_mat(self.h, box, None, 0, ctypes.byref(m), ctypes.byref(n))
        buf = np.zeros(max(m.value * n.value, 1))
        lib().rso_get_id_mat(self.h, box, _pd(buf), buf.size, ctypes.byref(m),
                             ctypes.byref(n))
        return buf[: m.value * n.value].reshape(n.value, m.value).T.copy()


def kernel_block(geom, I, J, pde=PDE_LAPLACE):
    pts = [_d(geom["points"]).reshape(-1), _d(geom["normals"]).reshape(-1),
           _d(geom["curvatures"]), _d(geom["weights"])]
    I = _i(I)
    J = _i(J)
    out = np.zeros(len(I) * len(J))
    lib().rso_kernel_block(pde, geom["points"].shape[0], *[_pd(a) for a in pts],
                           len(I), _pi(I), len(J), _pi(J), _pd(out))
    return out.reshape(len(J), len(I)).T.copy()


def proxy_block(geom, cx, cy, r, box, pde=PDE_LAPLACE):
    pts = [_d(geom["points"]).reshape(-1), _d(geom["normals"]).reshape(-1),
           _d(geom["curvatures"]), _d(geom["weights"])]
    box = _i(box)
    rows = 512 if pde == PDE_STOKES else 256
    out = np.zeros(rows * len(box))
    lib().rso_proxy_block(pde, geom["points"].shape[0], *[_pd(a) for a in pts],
                          cx, cy, r, len(box), _pi(box), _pd(out))
    return out.reshape(len(box), rows).T.copy()


def interp_decomp(A, tol):
    A = np.asfortranarray(A, dtype=np.float64)
    m, n = A.shape
    flat = np.ascontiguousarray(A.T.reshape(-1))
    perm = np.zeros(n, np.int32)
    T = np.zeros(n * n)
    k = lib().rso_id(m, n, _pd(flat), tol, _pi(perm), _pd(T))
    if k == 0:
        return 0, None, None
    return k, perm.copy(), T[: k * (n - k)].reshape(n - k, k).T.copy()


def evaluate(geom, targets, mu, pde=PDE_LAPLACE):
    pts = _d(geom["points"]).reshape(-1)
    nrm = _d(geom["normals"]).reshape(-1)
    w = _d(geom["weights"])
    tg = _d(targets).reshape(-1)
    nt = tg.size // 2
    mu = _d(mu)
    u = np.zeros(nt * (2 if pde == PDE_STOKES else 1))
    lib().rso_eval(pde, geom["points"].shape[0], _pd(pts), _pd(nrm), _pd(w), nt,
                   _pd(tg), _pd(mu), _pd(u))
    return u


def dense_matrix(geom, pde=PDE_LAPLACE):
    n = geom["points"].shape[0] * (2 if pde == PDE_STOKES else 1)
    idx = np.arange(n, dtype=np.int32)
    return kernel_block(geom, idx, idx, pde)


# ---------------------------------------------------------------------------
# Randomized interpolative decomposition.  No reference counterpart: the
# reference always factors the full ID matrix (ie_mat.cpp:548-588).  This is
# a numpy restatement of the product's IEGPU_ID_RANDOMIZED form (Gaussian
# sketch, pivoted QR of the sketch with the reference's rank rule), written
# as the checker of that form.
# ---------------------------------------------------------------------------
RID_ROWS = 32          # sketch rows held per box
RID_OVER = 6           # sketch rows kept beyond the last accepted pivot
RID_TABLE_ROWS = 1024  # Omega rows (ID rows of an eligible box)
RID_SEED = 0x5EED


def _splitmix64(x):
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return x ^ (x >> np.uint64(31))


def gaussian_table(rows=RID_TABLE_ROWS, cols=RID_ROWS, seed=RID_SEED):
    """Omega, row-major (rows x cols): entry i = r * cols + c is
    sqrt(-2 ln u1) cos(2 pi u2) (Box-Muller) with u1 = (h1 >> 11 + 1) 2^-53,
    u2 = (h2 >> 11) 2^-53 from SplitMix64 hashes of seed ^ 2i, seed ^ (2i+1)."""
    with np.errstate(over="ignore"):
        i = np.arange(rows * cols, dtype=np.uint64)
        s = np.uint64(seed)
        h1 = _splitmix64(s ^ (np.uint64(2) * i))
        h2 = _splitmix64(s ^ (np.uint64(2) * i + np.uint64(1)))
    u1 = ((h1 >> np.uint64(11)) + np.uint64(1)).astype(np.float64) * 2.0 ** -53
    u2 = (h2 >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    return (np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)).reshape(rows, cols)


def randomized_id(A, tol, rows, table=None, over=RID_OVER):
    """ID of A (m x n) from its sketch Y = Omega[:m, :rows]^T A.

    Column-pivoted Householder QR of Y: the pivot is the largest remaining
    column norm (first index on ties); stop at the first j >= 1 with
    2 |R_jj| sqrt(rows / (rows - j)) < tol |R_00| (the sketch keeps rows - j
    dimensions at step j); at most min(n, rows - over) pivots.  Returns
    (k, perm, T) with perm = skeleton columns in pivot order, then the
    redundant columns in increasing order, and T = R11^{-1} R12 (k x n-k);
    (n, perm, empty) when every column is taken; None when the sketch runs
    out before a cut (the full QRCP then takes the box)."""
    if table is None:
        table = gaussian_table()
    A = np.asarray(A, dtype=np.float64)
    m, n = A.shape
    Y = table[:m, :rows].T @ A
    avail = np.ones(n, bool)
    piv = []
    thresh = 0.0
    k = None
    jlim = min(n, rows - over)
    for j in range(jlim):
        nrm = np.where(avail, np.sum(Y[j:, :] ** 2, axis=0), -1.0)
        p = int(np.argmax(nrm))
        x = Y[j:, p].copy()
        alpha, xt = x[0], float(np.sum(x[1:] ** 2))
        beta = -np.copysign(np.sqrt(alpha * alpha + xt), alpha) if xt > 0 else alpha
        if j == 0:
            thresh = abs(tol * beta)
        elif 2 * abs(beta) * np.sqrt(rows / (rows - j)) < thresh:
            k = j
            break
        if xt > 0:
            tau = (beta - alpha) / beta
            v = x / (alpha - beta)
            v[0] = 1.0
            Y[j:, :] -= tau * np.outer(v, v @ Y[j:, :])
        Y[j, p] = beta
        Y[j + 1:, p] = 0.0
        piv.append(p)
        avail[p] = False
    if k is None:
        if jlim < n:
            return None
        k = n
    red = [c for c in range(n) if avail[c]]
    perm = np.array(piv[:k] + red, dtype=np.int64)
    if k == n:
        return k, perm, np.zeros((k, 0))
    R11 = Y[:k, :][:, perm[:k]]
    R12 = Y[:k, :][:, red]
    T = np.linalg.solve(np.triu(R11), R12)
    return k, perm, T


# ---------------------------------------------------------------------------
# Rank model (no reference counterpart): numpy restatement of the product's
# MLP (9 features -> 32 tanh -> 32 tanh -> 1) and of its training loss, the
# checker of the GPU forward / backward kernels.  Blob layout as iegpu.h.
# ---------------------------------------------------------------------------
RANK_FEATURES = 9
RANK_HID = 32
_F, _H = RANK_FEATURES, RANK_HID
_MU, _SD = 0, _F
_W1 = 2 * _F
_B1 = _W1 + _F * _H
_W2 = _B1 + _H
_B2 = _W2 + _H * _H
_W3 = _B2 + _H
_B3 = _W3 + _H
_YMU = _B3 + 1
_YSD = _YMU + 1
RANK_MODEL_LEN = _YSD + 1
RANK_PARAMS = _YMU - _W1


def _mlp_parts(model):
    m = np.asarray(model, dtype=np.float64)
    return dict(mu=m[_MU:_SD], sd=m[_SD:_W1], W1=m[_W1:_B1].reshape(_F, _H), b1=m[_B1:_W2],
                W2=m[_W2:_B2].reshape(_H, _H), b2=m[_B2:_W3], w3=m[_W3:_B3], b3=m[_B3],
                ymu=m[_YMU], ysd=m[_YSD])


def _mlp_forward(p, X):
    x = (np.asarray(X, np.float64) - p["mu"]) / p["sd"]
    h1 = np.tanh(x @ p["W1"] + p["b1"])
    h2 = np.tanh(h1 @ p["W2"] + p["b2"])
    return x, h1, h2, h2 @ p["w3"] + p["b3"]


def mlp_predict(model, X):
    """Predicted ranks of feature rows X."""
    p = _mlp_parts(model)
    return _mlp_forward(p, X)[3] * p["ysd"] + p["ymu"]


def mlp_loss_grad(model, X, y):
    """Mean squared error of the standardized rank and its gradient with
    respect to the trained parameters (blob order W1, b1, W2, b2, w3, b3)."""
    p = _mlp_parts(model)
    x, h1, h2, out = _mlp_forward(p, X)
    e = out - (np.asarray(y, np.float64) - p["ymu"]) / p["ysd"]
    n = len(e)
    dout = 2 * e / n
    d2 = np.outer(dout, p["w3"]) * (1 - h2 ** 2)
    d1 = (d2 @ p["W2"].T) * (1 - h1 ** 2)
    g = np.concatenate([(x.T @ d1).ravel(), d1.sum(0), (h1.T @ d2).ravel(), d2.sum(0),
                        h2.T @ dout, [dout.sum()]])
    return float(np.mean(e ** 2)), g


def update_marks(parent, neighbors, leaves):
    """Boxes marked for re-factorization after a geometry update, given the
    leaves that gain, lose or hold a changed point (QuadTree::perturb,
    quadtree.cpp:680-697 for additions, the deletion walk after it), each
    passed to mark_neighbors_and_parents (quadtree.cpp:435-461): the box
    and its neighbours are marked, then the same for its parent, up to the
    root.  parent[b] is -1 at a root; neighbors[b] lists box ids.  Returns
    0/1 per box id."""
    marks = np.zeros(len(parent), np.int32)
    for leaf in leaves:
        node = leaf
        while node >= 0:
            marks[node] = 1
            for nb in neighbors[node]:
                marks[nb] = 1
            node = parent[node]
    return marks

