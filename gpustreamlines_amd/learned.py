"""Learned PTT proposal: features, training pairs, training and export.

No counterpart in the reference.  A BF16 MLP (40 -> 128 -> 128 -> 40, ReLU)
maps the 28 order-6 SH coefficients of the ODF at the point, the incoming
direction and the 9-float PTT frame to an 8-component von Mises-Fisher
mixture on the hemisphere around the frame tangent; its orthographic
projection onto the (k1, k2) curvature disc proposes PTT candidates, and the
reference's acceptance rule (probe support >= 0.4, first candidate wins)
still decides each one.  Parameter layout: ``include/cuslines_hip.h``.

Training runs on the GPU through the library's own kernels
(``cuslines_mlp_loss_grad_async``, ``cuslines_mlp_adam_async``) on pairs the
plain getter records while it tracks (``cuslines_set_recording``).
"""

from __future__ import annotations

import numpy as np

from . import _lib, shm

N_IN, N_H, N_OUT, N_K = 40, 128, 40, 8
W1, B1 = 0, N_H * N_IN
W2, B2 = B1 + N_H, B1 + N_H + N_H * N_H
W3, B3 = B2 + N_H, B2 + N_H + N_OUT * N_H
NPARAM = B3 + N_OUT
assert NPARAM == _lib.MLP_NPARAM


def sh_coefficients(odf: np.ndarray, vertices: np.ndarray, order: int = 6, chunk: int = 65536) -> np.ndarray:
    """Least-squares real SH coefficients (order 6: 28) of the ODF sampled on
    the sphere's vertices: [X, Y, Z, 28] float32."""
    _, th, ph = shm.cart2sphere(*np.asarray(vertices, dtype=np.float64).T)
    B, _, _ = shm.real_sh_descoteaux(order, th, ph)
    P = np.linalg.pinv(B).T  # [V, 28]
    flat = odf.reshape(-1, odf.shape[-1])
    out = np.empty((flat.shape[0], P.shape[1]), dtype=np.float32)
    for i in range(0, flat.shape[0], chunk):
        out[i:i + chunk] = flat[i:i + chunk].astype(np.float64) @ P
    return out.reshape(odf.shape[:-1] + (P.shape[1],))


def init_params(seed: int = 0) -> np.ndarray:
    """He-initialised hidden layers; the output layer starts near a single
    broad lobe around the tangent (mean bias +1 on the tangent axis, log
    concentration bias 2)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, dtype=np.float32)
    p[W1:B1] = rng.normal(0, np.sqrt(2.0 / N_IN), N_H * N_IN)
    p[W2:B2] = rng.normal(0, np.sqrt(2.0 / N_H), N_H * N_H)
    p[W3:B3] = rng.normal(0, 0.1 * np.sqrt(1.0 / N_H), N_OUT * N_H)
    b3 = p[B3:B3 + N_OUT].reshape(N_K, 5)
    b3[:, 0] = 1.0
    b3[:, 1:3] = rng.normal(0, 0.3, (N_K, 2))
    b3[:, 3] = 2.0
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
    """Track ``seeds`` with the plain getter and keep each accepted sub-step
    with probability ``rate``.  Returns (device buffer of pairs, count)."""
    pairs = dt.alloc(4 * _lib.PAIR_FLOATS * capacity)
    cnt = dt.alloc(4)
    cnt.upload(np.zeros(1, np.int32))
    dt.set_proposal(None)
    dt.set_recording(pairs.ptr, capacity, cnt.ptr, rate)
    try:
        dt.track(seeds, rng_chunk=rng_chunk)
    finally:
        dt.set_recording(None)
    n = int(cnt.download(np.zeros(1, np.int32))[0])
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

    def fit(self, pairs, n_pairs: int, epochs: int = 3, lr: float = 2e-3, log_every: int = 0):
        """Mini-batch Adam over the pairs in order (the recording order is
        already shuffled across seeds by the hash).  Returns the mean loss of
        each epoch's last 50 batches."""
        nb = max(1, n_pairs // self.batch)
        hist = []
        stride = 4 * _lib.PAIR_FLOATS
        for ep in range(epochs):
            tail = []
            for b in range(nb):
                n = min(self.batch, n_pairs - b * self.batch)
                self.dt.mlp_loss_grad_async(self.p.ptr, pairs.ptr + b * self.batch * stride, n, self.loss.ptr,
                                            self.g.ptr, self.work.ptr)
                self.step += 1
                self.dt.mlp_adam_async(self.p.ptr, self.g.ptr, self.m.ptr, self.v.ptr, self.step, lr)
                if b >= nb - 50 or (log_every and b % log_every == 0):
                    self.dt.synchronize()
                    val = float(self.loss.download(np.zeros(1, np.float32))[0])
                    if b >= nb - 50:
                        tail.append(val)
                    if log_every and b % log_every == 0:
                        print(f"[train] epoch {ep} batch {b}/{nb} loss {val:.4f}", flush=True)
            hist.append(float(np.mean(tail)))
        self.dt.synchronize()
        return hist

    def params(self) -> np.ndarray:
        return self.p.download(np.zeros(NPARAM, np.float32))

    def free(self):
        for b in (self.p, self.g, self.m, self.v, self.loss, self.work):
            b.free()


# ---------------------------------------------------------------------------
# float64 restatement of the training objective (used by the tests as the
# checker of the GPU gradient; finite differences pin it)
# ---------------------------------------------------------------------------
def forward_np(params: np.ndarray, X: np.ndarray):
    p = np.asarray(params, dtype=np.float64)
    w1 = p[W1:B1].reshape(N_H, N_IN); b1 = p[B1:W2]
    w2 = p[W2:B2].reshape(N_H, N_H); b2 = p[B2:W3]
    w3 = p[W3:B3].reshape(N_OUT, N_H); b3 = p[B3:]
    h1 = np.maximum(X @ w1.T + b1, 0)
    h2 = np.maximum(h1 @ w2.T + b2, 0)
    return h1, h2, h2 @ w3.T + b3


def loss_np(params: np.ndarray, X: np.ndarray, U: np.ndarray):
    """Mean negative log-likelihood of unit targets U under the mixture."""
    _, _, O = forward_np(params, X)
    return float(np.mean(_nll_rows(O, U)[0]))


def _nll_rows(O, U):
    O = O.reshape(len(O), N_K, 5)
    m = O[:, :, :3]
    nrm = np.linalg.norm(m, axis=2)
    mh = m / nrm[:, :, None]
    z = O[:, :, 3]
    kap = np.exp(np.clip(z, -4, 6))
    d = np.einsum("nkc,nc->nk", mh, U)
    a = O[:, :, 4]
    lpi = a - _lse(a)[:, None]
    logc = np.log(kap) - np.log(2 * np.pi) - kap - np.log1p(-np.exp(-2 * kap))
    ell = lpi + logc + kap * d
    L = _lse(ell)
    return -L, dict(mh=mh, nrm=nrm, z=z, kap=kap, d=d, lpi=lpi, ell=ell, L=L)


def _lse(a):
    mx = a.max(axis=1)
    return mx + np.log(np.exp(a - mx[:, None]).sum(axis=1))


def grad_np(params: np.ndarray, X: np.ndarray, U: np.ndarray) -> np.ndarray:
    """Analytic gradient of loss_np (back-propagation in float64)."""
    p = np.asarray(params, dtype=np.float64)
    n = len(X)
    h1, h2, O = forward_np(p, X)
    _, c = _nll_rows(O, U)
    r = np.exp(c["ell"] - c["L"][:, None])
    pi = np.exp(c["lpi"])
    kap, d, e2 = c["kap"], c["d"], np.exp(-2 * c["kap"])
    dO = np.zeros((n, N_K, 5))
    dO[:, :, 4] = pi - r
    dl = 1 - 2 * kap * e2 / (1 - e2) - kap + kap * d
    inside = (c["z"] >= -4) & (c["z"] <= 6)
    dO[:, :, 3] = np.where(inside, -r * dl, 0.0)
    dO[:, :, :3] = (-r * kap / c["nrm"])[:, :, None] * (U[:, None, :] - d[:, :, None] * c["mh"])
    dO = dO.reshape(n, N_OUT) / n
    w2 = p[W2:B2].reshape(N_H, N_H)
    w3 = p[W3:B3].reshape(N_OUT, N_H)
    g = np.zeros(NPARAM)
    g[W3:B3] = (dO.T @ h2).ravel(); g[B3:] = dO.sum(0)
    dh2 = (dO @ w3) * (h2 > 0)
    g[W2:B2] = (dh2.T @ h1).ravel(); g[B2:W3] = dh2.sum(0)
    dh1 = (dh2 @ w2) * (h1 > 0)
    g[W1:B1] = (dh1.T @ X).ravel(); g[B1:W2] = dh1.sum(0)
    return g
