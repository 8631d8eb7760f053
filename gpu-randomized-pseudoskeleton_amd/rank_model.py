"""Training data and training of the block rank predictor.

Labels come from the plain batched CUR itself: every admissible block of a
seeded random ellipsoid (semi-axes U[0.5, 3] lambda, icosphere level 4 or 5,
+-1% jitter, 0.5 lambda leaves) is compressed to tol with the reference's
adaptive loop, and the eps-rank of its final intersection matrix
(singular values >= eps * sigma_max) is the label; the features are the
library's eight block features (farcur_batch_features). Training runs on the
GPU through the library's own forward/backward/Adam kernels.

    python rank_model.py --blocks 200000 --out models/rank_mlp.json
"""
from __future__ import annotations

import argparse
import pathlib
import sys
import time

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
DEFAULT_MODEL = HERE / "models" / "rank_mlp.json"


def _fc():
    import importlib.util

    if "farcur" in sys.modules:
        return sys.modules["farcur"]
    spec = importlib.util.spec_from_file_location("farcur", HERE / "farcur.py")
    mod = importlib.util.module_from_spec(spec)
    sys.modules["farcur"] = mod
    spec.loader.exec_module(mod)
    return mod


def ellipsoid_blocks(fc, ctx, axes, level, seed, tol=1e-4, leaf=0.5, eta_adm=2.0):
    """Features and eps-rank labels of every far block of one ellipsoid."""
    mesh = fc.Mesh.icosphere(*axes, level, 0.01, seed).build_rwg()
    part = fc.Partition(mesh, leaf, eta_adm)
    if len(part.far_pairs) == 0:
        return np.zeros((0, 8), np.float32), np.zeros(0, np.float32)
    ctx.set_mesh(mesh)
    batch = fc.Batch(ctx, *part.far_blocks())
    batch.compress(fc.MODE_BATCHED, tol=tol, seed=seed, precision=fc.F32, pinv_rtol=1e-6, label_eps=tol)
    rank = batch.info_arrays()["rank_eps"].astype(np.float64)
    X = batch.features()
    batch.close()
    return X, np.log(np.maximum(rank, 1.0)).astype(np.float32)


def training_set(fc, ctx, n_blocks=200_000, seed=1, levels=(4, 5), tol=1e-4, log=print):
    rng = np.random.default_rng(seed)
    Xs, ys, total, k = [], [], 0, 0
    while total < n_blocks:
        axes = tuple(float(a) for a in rng.uniform(0.5, 3.0, 3))
        level = int(rng.choice(levels))
        t0 = time.perf_counter()
        X, y = ellipsoid_blocks(fc, ctx, axes, level, seed * 1000 + k, tol)
        k += 1
        Xs.append(X)
        ys.append(y)
        total += len(y)
        log(f"ellipsoid {k}: axes {axes[0]:.2f} {axes[1]:.2f} {axes[2]:.2f} level {level}: {len(y)} blocks "
            f"({total} total, {time.perf_counter() - t0:.1f} s)")
    return np.concatenate(Xs)[:n_blocks], np.concatenate(ys)[:n_blocks]


def train(fc, ctx, X, y, epochs=40, batch=4096, lr=2e-3, seed=1, log=print):
    model = fc.Model(ctx, seed)
    losses = model.train(X, y, epochs=epochs, batch=batch, lr=lr, seed=seed)
    log(f"training loss (normalised MSE) first {losses[0]:.4f} last {losses[-1]:.4f}")
    return model, losses


def load(fc, ctx, path=DEFAULT_MODEL):
    return fc.Model(ctx).load(path)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--blocks", type=int, default=200_000)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--out", default=str(DEFAULT_MODEL))
    ap.add_argument("--holdout", type=float, default=0.1, help="fraction of blocks kept out of training")
    args = ap.parse_args()
    fc = _fc()
    ctx = fc.Context(0)
    X, y = training_set(fc, ctx, args.blocks, args.seed, log=lambda s: print(s, flush=True))
    rng = np.random.default_rng(args.seed)
    perm = rng.permutation(len(y))
    nte = int(args.holdout * len(y))
    te, tr = perm[:nte], perm[nte:]
    model, losses = train(fc, ctx, X[tr], y[tr], epochs=args.epochs, seed=args.seed, log=lambda s: print(s, flush=True))
    pred = model.predict(X[te])
    rk, rp = np.exp(y[te]), np.exp(pred)
    print(f"held-out blocks {nte}: median |log r_pred - log r| {np.median(np.abs(pred - y[te])):.3f}, "
          f"rank within +-10: {np.mean(np.abs(rp - rk) <= 10):.3f}, under-predicted by >10: {np.mean(rk - rp > 10):.3f}",
          flush=True)
    pathlib.Path(args.out).parent.mkdir(parents=True, exist_ok=True)
    model.save(args.out)
    print(f"saved {args.out}", flush=True)


if __name__ == "__main__":
    main()
