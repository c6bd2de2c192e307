"""Learned PTT proposal: training pairs, training on the GPU and export.

No counterpart in the reference.  At every non-initial sub-step the
reference's PTT getter measures the probe support of 24 disc vertices (8 ODF
taps each) to build its face-CDF proposal (cuslines/cuda_c/ptt.cu:232-300).
The learned proposal predicts those 24 supports with a BF16 MLP
(40 -> 128 -> 128 -> 24, ReLU) from two taps per vertex; the reference's face
CDF, candidate draw and acceptance rule (support >= 0.4, first lane wins)
then run unchanged on the prediction, and a round without a winner falls back
to the plain getter.  Inputs and parameter layout: ``include/cuslines_hip.h``.

The plain getter records (inputs, measured supports) pairs while it tracks
(``cuslines_set_recording``); the network trains on them with the library's
own kernels (``cuslines_mlp_loss_grad_async``: squared error,
``cuslines_mlp_adam_async``).
"""

from __future__ import annotations

import numpy as np

from . import _lib

N_IN, N_H, N_OUT = 40, 128, 24
W1, B1 = 0, N_H * N_IN
W2, B2 = B1 + N_H, B1 + N_H + N_H * N_H
W3, B3 = B2 + N_H, B2 + N_H + N_OUT * N_H
NPARAM = B3 + N_OUT
assert NPARAM == _lib.MLP_NPARAM
PAIR = _lib.PAIR_FLOATS


def init_params(seed: int = 0) -> np.ndarray:
    """He-initialised layers, zero biases."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, dtype=np.float32)
    p[W1:B1] = rng.normal(0, np.sqrt(2.0 / N_IN), N_H * N_IN)
    p[W2:B2] = rng.normal(0, np.sqrt(2.0 / N_H), N_H * N_H)
    p[W3:B3] = rng.normal(0, np.sqrt(1.0 / N_H), N_OUT * N_H)
    return p


def to_bf16(p: np.ndarray) -> np.ndarray:
    """float32 -> bf16 bit patterns (uint16), round to nearest, ties to even."""
    b = np.ascontiguousarray(p, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = (b + 0x7FFF + ((b >> 16) & 1)) >> 16
    finite = (b & 0x7F800000) != 0x7F800000
    r = np.where(finite, r, b >> 16)
    return r.astype(np.uint16)


def from_bf16(w: np.ndarray) -> np.ndarray:
    return (np.asarray(w, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


def record_pairs(dt, seeds: np.ndarray, capacity: int, rate: float, rng_chunk: int = 25000):
    """Track ``seeds`` with the plain getter, keeping each non-initial
    sub-step with probability ``rate``.  Returns (device buffer of pairs,
    count)."""
    pairs = dt.alloc(4 * PAIR * capacity)
    cnt = dt.alloc(4)
    try:
        cnt.upload(np.zeros(1, np.int32))
        dt.set_proposal(None)
        dt.set_recording(pairs.ptr, capacity, cnt.ptr, rate)
        try:
            dt.track(seeds, rng_chunk=rng_chunk)
        finally:
            dt.set_recording(None)
        n = int(cnt.download(np.zeros(1, np.int32))[0])
    except BaseException:
        pairs.free()
        raise
    finally:
        cnt.free()
    return pairs, min(n, capacity)


class Trainer:
    """Adam on float32 master parameters on the tracker's device."""

    def __init__(self, dt, params: np.ndarray | None = None, batch: int = 4096):
        self.dt = dt
        self.batch = int(batch)
        p0 = init_params() if params is None else np.ascontiguousarray(params, dtype=np.float32)
        self.p = dt.alloc(4 * NPARAM)
        self.g = dt.alloc(4 * NPARAM)
        self.m = dt.alloc(4 * NPARAM)
        self.v = dt.alloc(4 * NPARAM)
        self.loss = dt.alloc(4)
        self.work = dt.alloc(dt.mlp_workspace_bytes(self.batch))
        self.p.upload(p0)
        z = np.zeros(NPARAM, np.float32)
        self.m.upload(z)
        self.v.upload(z)
        self.step = 0

    def loss_grad(self, pairs_ptr: int, n: int):
        """(mean loss, gradient) of n <= batch pairs at the current parameters."""
        self.dt.mlp_loss_grad_async(self.p.ptr, pairs_ptr, n, self.loss.ptr, self.g.ptr, self.work.ptr)
        self.dt.synchronize()
        return float(self.loss.download(np.zeros(1, np.float32))[0]), self.g.download(np.zeros(NPARAM, np.float32))

    def fit(self, pairs, n_pairs: int, epochs: int = 3, lr: float = 1e-3, lr_final: float = 1e-4,
            holdout: int = 0):
        """Mini-batch Adam over the pairs in recording order (the hash already
        interleaves seeds), learning rate decayed geometrically from ``lr`` to
        ``lr_final``; the last ``holdout`` pairs are not trained on.  Returns
        the mean loss of each epoch's last 20 batches."""
        n_train = n_pairs - holdout
        if n_train <= 0:
            raise ValueError(f"no pairs left to train on ({n_pairs} pairs, holdout {holdout})")
        nb = max(1, n_train // self.batch)
        total = epochs * nb
        hist = []
        stride = 4 * PAIR
        for ep in range(epochs):
            tail = []
            for b in range(nb):
                n = min(self.batch, n_train - b * self.batch)
                rate = lr * (lr_final / lr) ** (self.step / max(1, total - 1))
                self.dt.mlp_loss_grad_async(self.p.ptr, pairs.ptr + b * self.batch * stride, n, self.loss.ptr,
                                            self.g.ptr, self.work.ptr)
                self.step += 1
                self.dt.mlp_adam_async(self.p.ptr, self.g.ptr, self.m.ptr, self.v.ptr, self.step, rate)
                if b >= nb - 20:
                    self.dt.synchronize()
                    tail.append(float(self.loss.download(np.zeros(1, np.float32))[0]))
            hist.append(float(np.mean(tail)))
        self.dt.synchronize()
        return hist

    def params(self) -> np.ndarray:
        return self.p.download(np.zeros(NPARAM, np.float32))

    def free(self):
        for b in (self.p, self.g, self.m, self.v, self.loss, self.work):
            b.free()


def train(dt, seeds: np.ndarray, rate: float = 0.06, capacity: int = 20_000_000, epochs: int = 4,
          batch: int = 4096, rng_chunk: int = 25000, verbose: bool = False):
    """Record pairs with the plain getter on ``seeds`` and train a network on
    the GPU.  Returns (float32 parameters, info dict)."""
    pairs, n = record_pairs(dt, seeds, capacity, rate, rng_chunk)
    try:
        if n == 0:
            raise RuntimeError("no training pairs were recorded (no seed took a non-initial sub-step)")
        tr = Trainer(dt, batch=batch)
        try:
            loss0 = tr.loss_grad(pairs.ptr, min(batch, n))[0]
            hist = tr.fit(pairs, n, epochs=epochs)
            params = tr.params()
        finally:
            tr.free()
    finally:
        pairs.free()
    if verbose:
        print(f"[learned] {n} pairs, loss {loss0:.4f} -> {hist}", flush=True)
    return params, dict(pairs=n, loss=[loss0] + hist)


# ---------------------------------------------------------------------------
# float64 restatement of the training objective (the checker of the GPU
# gradient in the tests; finite differences pin it)
# ---------------------------------------------------------------------------
def forward_np(params: np.ndarray, X: np.ndarray):
    p = np.asarray(params, dtype=np.float64)
    w1 = p[W1:B1].reshape(N_H, N_IN); b1 = p[B1:W2]
    w2 = p[W2:B2].reshape(N_H, N_H); b2 = p[B2:W3]
    w3 = p[W3:B3].reshape(N_OUT, N_H); b3 = p[B3:]
    h1 = np.maximum(X @ w1.T + b1, 0)
    h2 = np.maximum(h1 @ w2.T + b2, 0)
    return h1, h2, h2 @ w3.T + b3


def loss_np(params: np.ndarray, X: np.ndarray, T: np.ndarray) -> float:
    """Mean over pairs of the squared error averaged over the 24 outputs."""
    return float(np.mean((forward_np(params, X)[2] - T) ** 2))


def grad_np(params: np.ndarray, X: np.ndarray, T: np.ndarray) -> np.ndarray:
    """Analytic gradient of loss_np (back-propagation in float64)."""
    p = np.asarray(params, dtype=np.float64)
    n = len(X)
    h1, h2, O = forward_np(p, X)
    dO = 2.0 * (O - T) / (N_OUT * n)
    w2 = p[W2:B2].reshape(N_H, N_H)
    w3 = p[W3:B3].reshape(N_OUT, N_H)
    g = np.zeros(NPARAM)
    g[W3:B3] = (dO.T @ h2).ravel(); g[B3:] = dO.sum(0)
    dh2 = (dO @ w3) * (h2 > 0)
    g[W2:B2] = (dh2.T @ h1).ravel(); g[B2:W3] = dh2.sum(0)
    dh1 = (dh2 @ w2) * (h1 > 0)
    g[W1:B1] = (dh1.T @ X).ravel(); g[B1:W2] = dh1.sum(0)
    return g
