"""Learned warm start: data collection, normalisation and GPU training.

Samples are (features, converged impulses / force scale) pairs recorded by the
batched step in warm mode 3 while the plain solver (warm-started from the
matched previous-step impulses) iterates to a tight LCP residual, i.e. the
targets are impulses the plain solver itself converged to.  Training runs on
the GPU in odx_mlp.hip (bf16 MFMA forward/backward, Adam on fp32 weights).
"""
import numpy as np

from . import odx


def collect(batch, steps, tol=1e-5, max_iters=3000, check_every=5, order=1, every=1, max_samples=None):
    """Steps `batch` (odx.Batch) and returns stacked features X[N,16] and
    targets T[N,3]."""
    batch.set_mlp(None)
    batch.set_solver(order=order, max_iters=max_iters, tol=tol, check_every=check_every, warm=3)
    xs, ts, n = [], [], 0
    for s in range(steps):
        batch.step(1)
        if s % every:
            continue
        X, T = batch.samples()
        xs.append(X)
        ts.append(T)
        n += len(X)
        if max_samples and n >= max_samples:
            break
    X = np.concatenate(xs) if xs else np.zeros((0, 16), np.float32)
    T = np.concatenate(ts) if ts else np.zeros((0, 3), np.float32)
    if max_samples:
        X, T = X[:max_samples], T[:max_samples]
    return X, T


def normalization(X, T):
    """Per-feature mean/scale of inputs and targets (norm38 layout)."""
    mu = X.mean(0)
    sig = X.std(0)
    sig = np.where(sig > 1e-6, sig, 1.0)
    tmu = T.mean(0)
    tsig = T.std(0)
    tsig = np.where(tsig > 1e-6, tsig, 1.0)
    return np.concatenate([mu, sig, tmu, tsig]).astype(np.float32)


def residual_targets(X, T):
    """The network learns the change from the matched previous impulses
    (features 7-9, zero for an unmatched contact): an untrained network that
    outputs the mean change reproduces the plain warm start."""
    return (np.asarray(T, np.float32) - np.asarray(X, np.float32)[:, 7:10]).astype(np.float32)


def clip_targets(R, k=4.0):
    """Clips each target column to its median +- k robust standard deviations
    ((p95 - p5) / 3.29): rare impact impulses then do not set the output scale,
    whose bf16-limited resolution would otherwise add noise to every contact."""
    R = np.array(R, np.float32, copy=True)
    for o in range(R.shape[1]):
        med = np.median(R[:, o])
        s = max((np.percentile(R[:, o], 95) - np.percentile(R[:, o], 5)) / 3.29, 1e-6)
        np.clip(R[:, o], med - k * s, med + k * s, out=R[:, o])
    return R


def train(X, T, steps=3000, batch=4096, lr=2e-3, seed=1, lr_final=1e-4, clip=None):
    """Trains a fresh network on (X, T); returns (mlp, losses).  clip: robust
    clipping of the residual targets (clip_targets) at that many sigmas."""
    mlp = odx.Mlp(seed)
    R = residual_targets(X, T)
    if clip:
        R = clip_targets(R, clip)
    mlp.set_normalization(normalization(X, R))
    mlp.set_data(X, R)
    losses = []
    # step-decayed learning rate in four stages
    stages = 4
    for k in range(stages):
        f = k / max(1, stages - 1)
        lr_k = lr * (lr_final / lr) ** f
        losses.append(mlp.train(steps // stages, batch=batch, lr=lr_k, seed=seed + 101 * k))
    return mlp, np.concatenate(losses)
