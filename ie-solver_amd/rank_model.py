"""Rank model for the randomized ID: training data from plain (full QRCP)
factorizations and training on the GPU, both through the C ABI.

The features are computed by the factorization on the GPU (iegpu.h,
iegpu_record_features); the targets are the ranks the full QRCP finds.  The
trained blob goes to iegpu_set_rank_model, after which
set_id_method(ID_RANDOMIZED, 0) sizes each box's sketch from the predicted
rank (or sends the box straight to the full QRCP).
"""
import numpy as np

import geometry as G
import iesolver


def _factor_plain(geom, pde, tol, leaf=64, device=0):
    s = iesolver.RSSolver(device)
    s.set_geometry(geom, pde)
    s.build_tree(leaf)
    holes = geom.get("holes")
    if holes is not None:
        s.set_holes(holes["centers"], holes["radii"])
    s.record_features(True)
    s.factor(tol)
    return s


def collect(cases, leaf=64, device=0):
    """Features (float32, n x RANK_FEATURES) and full-QRCP ranks of the
    compressed boxes of each (geometry, pde, tol) case."""
    Xs, ys = [], []
    for geom, pde, tol in cases:
        s = _factor_plain(geom, pde, tol, leaf, device)
        bi = s.box_info()
        F = s.box_features()
        sel = (bi["level"] >= 1) & (bi["m_rows"] > 0) & (bi["rank"] > 0) & np.isfinite(F).all(1)
        Xs.append(F[sel])
        ys.append(bi["rank"][sel].astype(np.float32))
        s.close()
    return np.concatenate(Xs), np.concatenate(ys)


def training_cases(tols=(1e-6, 1e-8, 1e-10)):
    """Boundaries the default model is trained on (none of them is the
    benchmark's wavy boundary of seed 0)."""
    geoms = [(G.star(16384, k=k, amp=a), iesolver.PDE_LAPLACE)
             for k, a in [(3, 0.1), (5, 0.2), (9, 0.15)]]
    geoms += [(G.wavy(65536, seed=s), iesolver.PDE_LAPLACE) for s in (1, 2)]
    geoms += [(G.circle_with_holes(16384, 8, 1), iesolver.PDE_STOKES),
              (G.star(8192, k=7, amp=0.2), iesolver.PDE_STOKES)]
    return [(g, pde, tol) for g, pde in geoms for tol in tols]


def train(X, y, steps=6000, lr=3e-3, seed=1, device=0):
    """(model blob, final training loss) trained on the GPU."""
    s = iesolver.RSSolver(device)
    m0 = s.rank_model_init(X, y, seed)
    m, loss = s.rank_model_train(m0, X, y, steps, lr, seed)
    s.close()
    return m, loss


def main(argv=None):
    """python ie-solver_amd/rank_model.py --out FILE: collect the training
    boxes, train on GPU 0, save the model blob (float32 .npy)."""
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--steps", type=int, default=6000)
    args = ap.parse_args(argv)
    X, y = collect(training_cases())
    model, loss = train(X, y, steps=args.steps)
    np.save(args.out, model.astype(np.float32))
    print(f"trained on {len(y)} boxes, final loss {loss:.4f} -> {args.out}")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
