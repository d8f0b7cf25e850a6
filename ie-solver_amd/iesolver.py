"""Python host binding of the C ABI in include/iegpu.h (ctypes).

Mirrors the reference's factorize-then-solve call sequence
(ie_solver/linear_solve.cpp:156-185, bie_time_trial):

    s = RSSolver()                      # iegpu_create
    s.set_geometry(geom, pde)           # Kernel + Boundary arrays
    s.build_tree(leaf_cap)              # QuadTree::initialize_tree
    s.set_holes(centers, radii)         # quadtree->U / Psi (optional)
    s.factor(id_tol)                    # SkelFactorization::skeletonize
    mu, alpha = s.solve(f)              # linear_solve

The shared library is built in-tree (ie-solver_amd/build.py).  There is no
CPU fallback: loading fails loudly when the library is missing, and every
non-zero status raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libiesolver_amd.so")

PDE_LAPLACE = 0
PDE_LAPLACE_NEUMANN = 1
PDE_STOKES = 2
FP64 = 0
FP32 = 1
ID_QRCP = 0
ID_RANDOMIZED = 1
MATVEC_AUTO = 0
MATVEC_DIRECT = 1
MATVEC_FMM = 2
FMM_AUTO_POINTS = 16384

STATUS = {
    0: "OK", 1: "INVALID", 2: "NO_COMPRESSION", 3: "SINGULAR", 4: "OOM",
    5: "HIP", 6: "CAPACITY", 7: "NO_DEVICE",
}

EXPORTS = [
    "iegpu_create", "iegpu_destroy", "iegpu_last_error", "iegpu_set_geometry",
    "iegpu_build_tree", "iegpu_set_holes", "iegpu_factor", "iegpu_solve",
    "iegpu_solve_device", "iegpu_tree_info", "iegpu_box_info", "iegpu_box_list",
    "iegpu_update_tree_info",
    "iegpu_last_timings", "iegpu_set_profiling", "iegpu_profile", "iegpu_evaluate",
    "iegpu_solve_refined", "iegpu_apply", "iegpu_set_id_method", "iegpu_id_stats",
    "iegpu_set_graphs", "iegpu_record_features", "iegpu_box_features", "iegpu_rank_model_init",
    "iegpu_rank_model_train", "iegpu_rank_model_grad", "iegpu_set_rank_model",
    "iegpu_set_geometry_batch", "iegpu_set_launch_caps", "iegpu_get_launch_caps",
    "iegpu_autotune", "iegpu_set_overlap", "iegpu_set_keep_operator", "iegpu_sparse_matvec",
    "iegpu_factor_solve_device", "iegpu_update_geometry", "iegpu_update_info",
    "iegpu_update_geometry_batch", "iegpu_discard_update", "iegpu_evaluate_noalloc",
    "iegpu_solve_nrhs", "iegpu_set_border", "iegpu_box_factors", "iegpu_set_matvec",
    "iegpu_matvec_info", "iegpu_set_launch_plan", "iegpu_launch_plan_info",
]
NUM_LAUNCH_CAPS = 8
RANK_FEATURES = 9
RANK_MODEL_LEN = 1429
RANK_PARAMS = 1409

STAGES = ["active", "rows", "scan", "assemble", "rid", "qrcp", "decouple", "root",
          "sweep_up", "root_solve", "sweep_down", "border"]

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lib = None


class IeGpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with ie-solver_amd/build.py")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    L.iegpu_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int]
    L.iegpu_destroy.argtypes = [vp]
    L.iegpu_last_error.argtypes = [vp]
    L.iegpu_last_error.restype = ctypes.c_char_p
    L.iegpu_set_geometry.argtypes = [vp, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp]
    L.iegpu_set_geometry_batch.argtypes = [vp, ctypes.c_int, ctypes.c_int, _ip, _dp, _dp,
                                           _dp, _dp]
    L.iegpu_build_tree.argtypes = [vp, ctypes.c_int]
    L.iegpu_set_holes.argtypes = [vp, ctypes.c_int, _dp, _dp]
    L.iegpu_factor.argtypes = [vp, ctypes.c_double, ctypes.c_int]
    L.iegpu_solve.argtypes = [vp, _dp, _dp, _dp]
    L.iegpu_solve_device.argtypes = [vp, vp, vp, vp, vp]
    L.iegpu_factor_solve_device.argtypes = [vp, ctypes.c_double, ctypes.c_int, vp, vp]
    L.iegpu_tree_info.argtypes = [vp, _ip, _ip, _ip, _ip]
    L.iegpu_box_info.argtypes = [vp, _ip, _ip, _ip, _ip, _ip, _ip, _ip]
    L.iegpu_box_list.argtypes = [vp, ctypes.c_int, ctypes.c_int, _ip, ctypes.c_int]
    L.iegpu_last_timings.argtypes = [vp, _dp, ctypes.c_int]
    L.iegpu_evaluate.argtypes = [vp, ctypes.c_int, _dp, _ip, _dp, _dp, _dp]
    L.iegpu_solve_refined.argtypes = [vp, _dp, _dp, _dp, ctypes.c_double, ctypes.c_int,
                                      _ip, _dp]
    L.iegpu_apply.argtypes = [vp, _dp, _dp, _dp, _dp]
    L.iegpu_set_matvec.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.iegpu_matvec_info.argtypes = [vp, _ip, ctypes.c_int]
    L.iegpu_set_launch_plan.argtypes = [vp, ctypes.c_int]
    L.iegpu_launch_plan_info.argtypes = [vp, _ip, ctypes.c_int]
    L.iegpu_set_profiling.argtypes = [vp, ctypes.c_int]
    L.iegpu_profile.argtypes = [vp, ctypes.c_int, ctypes.c_int, _dp, _ip]
    L.iegpu_set_id_method.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.iegpu_id_stats.argtypes = [vp, _ip, _ip, _dp]
    fp = ctypes.POINTER(ctypes.c_float)
    L.iegpu_record_features.argtypes = [vp, ctypes.c_int]
    L.iegpu_set_graphs.argtypes = [vp, ctypes.c_int]
    L.iegpu_box_features.argtypes = [vp, fp, ctypes.c_int]
    L.iegpu_rank_model_init.argtypes = [vp, fp, fp, ctypes.c_int, ctypes.c_uint, fp]
    L.iegpu_rank_model_train.argtypes = [vp, fp, fp, ctypes.c_int, fp, ctypes.c_int,
                                         ctypes.c_float, ctypes.c_uint, fp]
    L.iegpu_rank_model_grad.argtypes = [vp, fp, fp, fp, ctypes.c_int, fp, fp]
    L.iegpu_set_rank_model.argtypes = [vp, fp, ctypes.c_int]
    L.iegpu_set_overlap.argtypes = [vp, ctypes.c_int]
    L.iegpu_set_keep_operator.argtypes = [vp, ctypes.c_int]
    L.iegpu_sparse_matvec.argtypes = [vp, _dp, _dp]
    L.iegpu_set_launch_caps.argtypes = [vp, _ip, ctypes.c_int]
    L.iegpu_get_launch_caps.argtypes = [vp, _ip, ctypes.c_int]
    L.iegpu_autotune.argtypes = [vp, ctypes.c_double, ctypes.c_int, ctypes.c_int, _ip, _dp]
    L.iegpu_update_geometry.argtypes = [vp, ctypes.c_int, _dp, _dp, _dp, _dp, _ip]
    L.iegpu_update_geometry_batch.argtypes = [vp, ctypes.c_int, _ip, _dp, _dp, _dp, _dp, _ip]
    L.iegpu_update_info.argtypes = [vp, _ip, _ip, _ip]
    L.iegpu_update_tree_info.argtypes = [vp, _ip, _ip, _ip]
    L.iegpu_discard_update.argtypes = [vp]
    L.iegpu_evaluate_noalloc.argtypes = [vp, ctypes.c_int, vp, vp, vp, vp, vp, vp]
    L.iegpu_solve_nrhs.argtypes = [vp, ctypes.c_int, _dp, _dp, _dp]
    L.iegpu_set_border.argtypes = [vp, ctypes.c_int, _dp, _dp]
    L.iegpu_box_factors.argtypes = [vp, ctypes.c_int, _ip, _ip, _dp, _dp, _dp, _dp, _dp]
    _lib = L
    return L


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _pd(a):
    return a.ctypes.data_as(_dp)


def _pi(a):
    return a.ctypes.data_as(_ip)


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _pf(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class RSSolver:
    """One boundary, its quadtree and its factorization on one GPU."""

    def __init__(self, device=0):
        self.L = load()
        h = ctypes.c_void_p()
        st = self.L.iegpu_create(ctypes.byref(h), device)
        if st != 0:
            raise IeGpuError(st, "iegpu_create failed")
        self.h = h
        self.hole_dim = 0
        self.ndof = 0

    def _check(self, st):
        if st != 0:
            raise IeGpuError(st, self.L.iegpu_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.L.iegpu_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_geometry(self, geom, pde=PDE_LAPLACE):
        self._arrays = [_d(geom["points"]).reshape(-1), _d(geom["normals"]).reshape(-1),
                        _d(geom["curvatures"]), _d(geom["weights"])]
        n = geom["points"].shape[0]
        self._check(self.L.iegpu_set_geometry(self.h, pde, n, *[_pd(a) for a in self._arrays]))
        self.pde = pde
        self.ndof = n * (2 if pde == PDE_STOKES else 1)
        self.hole_dim = 0
        self.dof_offsets = np.array([0, self.ndof], dtype=np.int64)

    def set_geometry_batch(self, geoms, pde=PDE_LAPLACE):
        """A batch of independent boundaries factored and solved together.
        Right-hand sides and solutions are the problems' vectors concatenated;
        `dof_offsets[p]:dof_offsets[p+1]` is problem p's slice."""
        sd = 2 if pde == PDE_STOKES else 1
        npts = [g["points"].shape[0] for g in geoms]
        off = np.zeros(len(geoms) + 1, dtype=np.int32)
        off[1:] = np.cumsum(npts)
        self._arrays = [
            _d(np.concatenate([g["points"] for g in geoms])).reshape(-1),
            _d(np.concatenate([g["normals"] for g in geoms])).reshape(-1),
            _d(np.concatenate([g["curvatures"] for g in geoms])),
            _d(np.concatenate([g["weights"] for g in geoms]))]
        self._poff = np.ascontiguousarray(off)
        self._check(self.L.iegpu_set_geometry_batch(self.h, pde, len(geoms), _pi(self._poff),
                                                    *[_pd(a) for a in self._arrays]))
        self.pde = pde
        self.ndof = int(off[-1]) * sd
        self.hole_dim = 0
        self.dof_offsets = off.astype(np.int64) * sd

    def update_geometry(self, g):
        """New node data for the same points (QuadTree::perturb): a geometry
        dict, or a list of them for a batch.  Returns True when the tree is
        unchanged; the next factor then re-factors only the marked boxes
        (update_info)."""
        batch = isinstance(g, (list, tuple))
        geoms = g if batch else [g]
        keys = ("points", "normals", "curvatures", "weights")
        if batch:
            arrs = [_d(np.concatenate([q[k] for q in geoms])).reshape(-1) for k in keys]
        else:
            arrs = [_d(g[k]).reshape(-1) for k in keys]
        kept = ctypes.c_int(0)
        if batch:
            # every problem keeps its number of points (checked by the library)
            off = np.zeros(len(geoms) + 1, dtype=np.int32)
            off[1:] = np.cumsum([q["points"].shape[0] for q in geoms])
            off = np.ascontiguousarray(off)
            self._check(self.L.iegpu_update_geometry_batch(
                self.h, len(geoms), _pi(off), *[_pd(a) for a in arrs], ctypes.byref(kept)))
        else:
            self._check(self.L.iegpu_update_geometry(self.h, arrs[2].shape[0],
                                                     *[_pd(a) for a in arrs],
                                                     ctypes.byref(kept)))
        self._arrays = arrs
        return bool(kept.value)

    def discard_update(self):
        """Drop a pending update's marks: the next factor covers every box."""
        self._check(self.L.iegpu_discard_update(self.h))

    def update_info(self):
        """(boxes marked by the last update, whether the last factorization
        re-factored only them, per-box-id marks)."""
        nb = self.tree_info()[1]
        n, inc = ctypes.c_int(0), ctypes.c_int(0)
        marks = np.zeros(nb, np.int32)
        self._check(self.L.iegpu_update_info(self.h, ctypes.byref(n), ctypes.byref(inc),
                                             _pi(marks)))
        return n.value, bool(inc.value), marks

    def update_tree_info(self):
        """(leaves the last update subdivided, boxes it merged into leaves,
        per box id the id before the update, -1 for new boxes)."""
        nb = self.tree_info()[1]
        ns, nm = ctypes.c_int(0), ctypes.c_int(0)
        old = np.zeros(nb, np.int32)
        self._check(self.L.iegpu_update_tree_info(self.h, ctypes.byref(ns), ctypes.byref(nm),
                                                  _pi(old)))
        return ns.value, nm.value, old

    def build_tree(self, leaf_cap=64):
        self._check(self.L.iegpu_build_tree(self.h, leaf_cap))

    def set_holes(self, centers, radii):
        c = _d(centers).reshape(-1)
        r = _d(radii)
        self._check(self.L.iegpu_set_holes(self.h, len(r), _pd(c), _pd(r)))
        self.hole_dim = self.tree_info()[3]

    def factor(self, id_tol, precision=FP64):
        self._check(self.L.iegpu_factor(self.h, id_tol, precision))

    def set_id_method(self, method, sketch_rows=32):
        """ID_QRCP (the reference's full pivoted QR) or ID_RANDOMIZED (Gaussian
        sketch of sketch_rows rows, full-QRCP fallback)."""
        self._check(self.L.iegpu_set_id_method(self.h, method, sketch_rows))

    def set_keep_operator(self, on=True):
        """Keep X_rr and the root blocks at the next factor (for sparse_matvec)."""
        self._check(self.L.iegpu_set_keep_operator(self.h, 1 if on else 0))

    def sparse_matvec(self, x):
        """b = A x with the compressed operator (SkelFactorization::sparse_matvec)."""
        x = _d(x)
        b = np.zeros(self.ndof)
        self._check(self.L.iegpu_sparse_matvec(self.h, _pd(x), _pd(b)))
        return b

    def set_overlap(self, on=True):
        """Overlap each level's ID with the decouple of the level below
        (second stream); results are identical either way."""
        self._check(self.L.iegpu_set_overlap(self.h, 1 if on else 0))

    def set_launch_caps(self, caps):
        c = np.ascontiguousarray(caps, dtype=np.int32)
        self._check(self.L.iegpu_set_launch_caps(self.h, _pi(c), len(c)))

    def launch_caps(self):
        c = np.zeros(NUM_LAUNCH_CAPS, dtype=np.int32)
        self._check(self.L.iegpu_get_launch_caps(self.h, _pi(c), len(c)))
        return c

    def autotune(self, id_tol, precision=FP64, reps=2):
        """Pick the launch caps for this workload; returns (caps, best ms)."""
        c = np.zeros(NUM_LAUNCH_CAPS, dtype=np.int32)
        ms = ctypes.c_double(0)
        self._check(self.L.iegpu_autotune(self.h, id_tol, precision, reps, _pi(c),
                                          ctypes.byref(ms)))
        return c, ms.value

    def set_graphs(self, on=True):
        """Replay the factorization as a captured HIP graph (iegpu_set_graphs)."""
        self._check(self.L.iegpu_set_graphs(self.h, 1 if on else 0))

    def record_features(self, on=True):
        self._check(self.L.iegpu_record_features(self.h, 1 if on else 0))

    def box_features(self):
        """(n_boxes, RANK_FEATURES) float32 features of the last factorization."""
        nb = self.tree_info()[1]
        out = np.zeros(nb * RANK_FEATURES, np.float32)
        self._check(self.L.iegpu_box_features(self.h, _pf(out), out.size))
        return out.reshape(nb, RANK_FEATURES)

    def rank_model_init(self, X, y, seed=1):
        X, y = _f(X), _f(y)
        m = np.zeros(RANK_MODEL_LEN, np.float32)
        self._check(self.L.iegpu_rank_model_init(self.h, _pf(X), _pf(y), len(y), seed, _pf(m)))
        return m

    def rank_model_train(self, model, X, y, steps=4000, lr=3e-3, seed=1):
        """Trains a copy of model on the GPU; returns (model, final loss)."""
        X, y, m = _f(X), _f(y), _f(model).copy()
        loss = ctypes.c_float(0)
        self._check(self.L.iegpu_rank_model_train(self.h, _pf(X), _pf(y), len(y), _pf(m), steps,
                                                  lr, seed, ctypes.byref(loss)))
        return m, loss.value

    def rank_model_grad(self, model, X, y):
        """(loss, gradient) of the mean squared error, from the GPU kernels."""
        X, y, m = _f(X), _f(y), _f(model)
        g = np.zeros(RANK_PARAMS, np.float32)
        loss = ctypes.c_float(0)
        self._check(self.L.iegpu_rank_model_grad(self.h, _pf(m), _pf(X), _pf(y), len(y), _pf(g),
                                                 ctypes.byref(loss)))
        return loss.value, g

    def set_rank_model(self, model, margin=2):
        m = _f(model)
        self._check(self.L.iegpu_set_rank_model(self.h, _pf(m), margin))

    def id_stats(self, with_flops=False):
        """(boxes whose ID came from the sketch, boxes that fell back[, sketch
        matrix-core flops])."""
        a = ctypes.c_int(0)
        b = ctypes.c_int(0)
        f = ctypes.c_double(0)
        self._check(self.L.iegpu_id_stats(self.h, ctypes.byref(a), ctypes.byref(b),
                                          ctypes.byref(f)))
        return (a.value, b.value, f.value) if with_flops else (a.value, b.value)

    def solve(self, b):
        b = _d(b)
        mu = np.zeros(self.ndof)
        if self.hole_dim:
            alpha = np.zeros(self.hole_dim)
            self._check(self.L.iegpu_solve(self.h, _pd(b), _pd(mu), _pd(alpha)))
            return mu, alpha
        self._check(self.L.iegpu_solve(self.h, _pd(b), _pd(mu), None))
        return mu, None

    def solve_many(self, B):
        """Solve for the columns of B (ndof x nrhs); returns (MU, ALPHA)."""
        B = np.asfortranarray(B, dtype=np.float64)
        nrhs = B.shape[1]
        MU = np.zeros((self.ndof, nrhs), order="F")
        AL = np.zeros((self.hole_dim, nrhs), order="F") if self.hole_dim else None
        self._check(self.L.iegpu_solve_nrhs(self.h, nrhs, B.ctypes.data_as(_dp),
                                            MU.ctypes.data_as(_dp),
                                            None if AL is None else AL.ctypes.data_as(_dp)))
        return MU, AL

    def set_border(self, U, Psi):
        """Caller-supplied border blocks U (ndof x h) and Psi (h x ndof)."""
        U = np.asfortranarray(U, dtype=np.float64)
        P = np.asfortranarray(Psi, dtype=np.float64)
        self._check(self.L.iegpu_set_border(self.h, U.shape[1], U.ctypes.data_as(_dp),
                                            P.ctypes.data_as(_dp)))

    def box_factors(self, box):
        """dict(k, r, T, L, U, Xinv, S) of one box (column-major blocks)."""
        k, r = ctypes.c_int(0), ctypes.c_int(0)
        self._check(self.L.iegpu_box_factors(self.h, box, ctypes.byref(k), ctypes.byref(r),
                                             None, None, None, None, None))
        k, r = k.value, r.value
        out = dict(k=k, r=r)
        if r == 0:
            return out
        T, Lm, U = (np.zeros((k, r), order="F"), np.zeros((k, r), order="F"),
                    np.zeros((r, k), order="F"))
        X, S = np.zeros((r, r), order="F"), np.zeros((k, k), order="F")
        self._check(self.L.iegpu_box_factors(self.h, box, None, None, *[a.ctypes.data_as(_dp)
                                                                        for a in (T, Lm, U, X, S)]))
        out.update(T=T, L=Lm, U=U, Xinv=X, S=S)
        return out

    def solve_refined(self, b, rtol=1e-10, max_iter=10):
        """FP64 iterative refinement preconditioned by the stored factor.
        Returns (mu, alpha, corrections, relative residual)."""
        b = _d(b)
        mu = np.zeros(self.ndof)
        alpha = np.zeros(self.hole_dim) if self.hole_dim else None
        it = ctypes.c_int(0)
        rr = ctypes.c_double(0)
        self._check(self.L.iegpu_solve_refined(self.h, _pd(b), _pd(mu),
                                               None if alpha is None else _pd(alpha),
                                               rtol, max_iter, ctypes.byref(it), ctypes.byref(rr)))
        return mu, alpha, it.value, rr.value

    def set_matvec(self, method=MATVEC_AUTO, order=0, leaf_points=0):
        """Operator form of apply and solve_refined: MATVEC_DIRECT (O(N^2)
        sum), MATVEC_FMM (fast multipole, `order` 32/48/64 coefficients,
        about `leaf_points` nodes per leaf; 0 = defaults 48 and 64) or
        MATVEC_AUTO (FMM from FMM_AUTO_POINTS nodes of one boundary)."""
        self._check(self.L.iegpu_set_matvec(self.h, method, order, leaf_points))

    def set_launch_plan(self, mode=1):
        """1: skip the ID-stage kernels whose lists were empty in the last
        full factorization of the same inputs (checked on the device); 0:
        launch every kernel; 2: take every list as empty once (the check
        repeats the factorization in full)."""
        self._check(self.L.iegpu_set_launch_plan(self.h, mode))

    def launch_plan_info(self):
        """(last factorization used the plan, repeats after a miss,
        skipped kernel groups in the plan)."""
        v = np.zeros(3, dtype=np.int32)
        self._check(self.L.iegpu_launch_plan_info(self.h, _pi(v), 3))
        return tuple(int(t) for t in v)

    def matvec_info(self):
        """(form used by the last apply / solve_refined, FMM leaf level,
        FMM boxes, FMM leaves)."""
        v = np.zeros(4, dtype=np.int32)
        self._check(self.L.iegpu_matvec_info(self.h, _pi(v), 4))
        return tuple(int(t) for t in v)

    def apply(self, x, alpha=None):
        """FP64 operator: K x (+ U alpha), and Psi x - alpha with holes."""
        x = _d(x)
        y = np.zeros(self.ndof)
        if self.hole_dim and alpha is not None:
            a = _d(alpha)
            z = np.zeros(self.hole_dim)
            self._check(self.L.iegpu_apply(self.h, _pd(x), _pd(a), _pd(y), _pd(z)))
            return y, z
        self._check(self.L.iegpu_apply(self.h, _pd(x), None, _pd(y), None))
        return y, None

    def evaluate(self, targets, mu, alpha=None, in_domain=None):
        """Interior field K_domain mu + U_forward alpha at targets (n, 2)."""
        tg = _d(targets).reshape(-1)
        nt = tg.size // 2
        mu = _d(mu)
        sd = 2 if self.pde == PDE_STOKES else 1
        out = np.zeros(nt * sd)
        mask = None if in_domain is None else np.ascontiguousarray(in_domain, dtype=np.int32)
        al = None if alpha is None else _d(alpha)
        self._check(self.L.iegpu_evaluate(self.h, nt, _pd(tg), None if mask is None else _pi(mask),
                                          _pd(mu), None if al is None else _pd(al), _pd(out)))
        return out

    def factor_solve_device(self, id_tol, precision, b_ptr, x_ptr):
        """Factor and solve one right-hand side (device pointers, distinct);
        each level's upward sweep runs beside the factorization."""
        self._check(self.L.iegpu_factor_solve_device(self.h, id_tol, precision, b_ptr, x_ptr))

    def solve_device(self, b_ptr, mu_ptr, alpha_ptr=None, stream=None):
        self._check(self.L.iegpu_solve_device(self.h, b_ptr, mu_ptr, alpha_ptr, stream))

    def tree_info(self):
        v = [ctypes.c_int(0) for _ in range(4)]
        self._check(self.L.iegpu_tree_info(self.h, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def box_info(self):
        nl, nb, _, _ = self.tree_info()
        out = dict(level=np.zeros(nb, np.int32), parent=np.zeros(nb, np.int32),
                   is_leaf=np.zeros(nb, np.int32), children=np.zeros(4 * nb, np.int32),
                   n_active=np.zeros(nb, np.int32), rank=np.zeros(nb, np.int32),
                   m_rows=np.zeros(nb, np.int32))
        self._check(self.L.iegpu_box_info(self.h, *[_pi(out[k]) for k in
                                                    ("level", "parent", "is_leaf", "children",
                                                     "n_active", "rank", "m_rows")]))
        out["children"] = out["children"].reshape(nb, 4)
        return out

    def box_list(self, box, which):
        n = self.L.iegpu_box_list(self.h, box, which, None, 0)
        if n < 0:
            raise IeGpuError(-n, self.L.iegpu_last_error(self.h).decode())
        buf = np.zeros(max(n, 1), np.int32)
        self.L.iegpu_box_list(self.h, box, which, _pi(buf), n)
        return buf[:n].copy()

    def timings(self):
        t = np.zeros(4)
        self._check(self.L.iegpu_last_timings(self.h, _pd(t), 4))
        return dict(factor_ms=t[0], solve_ms=t[1], root_ms=t[2], factor_bytes=t[3])

    def set_profiling(self, on=True):
        self._check(self.L.iegpu_set_profiling(self.h, 1 if on else 0))

    def profile(self, stage, level=-1):
        """(milliseconds, launches) of one stage (name or index)."""
        if isinstance(stage, str):
            stage = STAGES.index(stage)
        ms = ctypes.c_double(0)
        n = ctypes.c_int(0)
        self._check(self.L.iegpu_profile(self.h, stage, level, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def profile_table(self):
        return {name: self.profile(i) for i, name in enumerate(STAGES)}
