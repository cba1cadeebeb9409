"""CPU oracle of the density-compensation part (test infrastructure only).

Restates in numpy what nfft_amd/csrc/density.hip computes, so the GPU
library can be checked against it:

* ``features``: radius 2|x|, log(1 + c / cbar) of the node counts in the
  5 x 5 coarse cells around each node at grids of 32, 64, 128 cells, and the
  kernel density estimate log(v / mean v), v = Re(B B^T 1) with the oracle's
  B and B^T of the plan (N, n, m).
* ``pipe_menon``: w <- w / Re(B B^T w) with the oracle's own B and B^T
  (oracle.stage_B / stage_BT, which use the reference's unscaled window;
  the weights agree with the library's after normalisation to mean 1).
* ``mlp_forward`` / ``mlp_loss_grad``: the 80 -> 64 -> 64 -> 1 ReLU network
  in FP64, optionally rounding to BF16 exactly where the GPU path does
  (inputs and both hidden activations, the layer-2 input gradient and the
  two MFMA layer weights).

The reference has no such component (its MRI applications read weights
from files made by applications/mri/mri2d/precompute_weights.m), so these
functions are pinned by mathematics instead: finite differences for the
gradient, brute-force counting for the features and a dense B built from the
window function for the density estimate, constant weights for a Cartesian
node set for Pipe-Menon.
"""
import numpy as np

import oracle as O

NFEAT = 80
NRAW = 77
H = 64
GRIDS = (32, 64, 128)
P_W1 = 0
P_B1 = P_W1 + NFEAT * H
P_W2 = P_B1 + H
P_B2 = P_W2 + H * H
P_W3 = P_B2 + H
P_B3 = P_W3 + H
P_COUNT = P_B3 + 1


def bf16_round(a):
    """Round float32 values to the nearest bfloat16 (ties to even), returned
    as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def _coarse_cell(x, G):
    c = np.floor((x.astype(np.float32) + np.float32(0.5)) * np.float32(G)).astype(np.int64)
    return np.clip(c, 0, G - 1)


def kde(N, n, m, x):
    """log(v / mean v), v = Re(B B^T 1): the first Pipe-Menon denominator."""
    x = np.asarray(x, dtype=np.float64)
    v = O.stage_B(N, n, m, x, O.stage_BT(N, n, m, x, np.ones(x.shape[0], dtype=complex))).real
    return np.log(v / v.mean())


def kde_bruteforce(N, n, m, x, _dm=0):
    """The same from a dense B assembled tap by tap from the window function
    (small M and n only).  _dm evaluates the window with cut-off m + _dm on
    the m taps (a planted mistake for the tests)."""
    x = np.asarray(x, dtype=np.float64)
    M, d = x.shape
    b = [O.kb_b(n[t] / N[t]) for t in range(d)]
    B = np.zeros((M, int(np.prod(n))))
    for j in range(M):
        per = []
        for t in range(d):
            u = int(np.floor(n[t] * x[j, t])) - m
            per.append([((l % n[t]), O.phi(n[t], x[j, t] - l / n[t], m + _dm, b[t]))
                        for l in range(u, u + 2 * m + 2)])
        for (i0, w0) in per[0]:
            for (i1, w1) in per[1]:
                B[j, i0 * n[1] + i1] += w0 * w1
    v = B @ (B.T @ np.ones(M))
    return np.log(v / v.mean())


def features(x, N, n, m):
    """x: (M, 2) nodes in [-1/2, 1/2) of a plan (N, n, m).  Returns (M, 80)
    float64."""
    x = np.asarray(x, dtype=np.float32)
    M = x.shape[0]
    out = np.zeros((M, NFEAT))
    out[:, 0] = 2.0 * np.sqrt(x[:, 0].astype(np.float64) ** 2 + x[:, 1].astype(np.float64) ** 2)
    q = 1
    for G in GRIDS:
        cx = _coarse_cell(x[:, 0], G)
        cy = _coarse_cell(x[:, 1], G)
        counts = np.zeros((G, G), dtype=np.int64)
        np.add.at(counts, (cx, cy), 1)
        cbar = M / (G * G)
        for a in range(-2, 3):
            for b in range(-2, 3):
                out[:, q] = np.log1p(counts[(cx + a) % G, (cy + b) % G] / cbar)
                q += 1
    out[:, q] = kde(N, n, m, x)
    return out


def features_bruteforce(x, N, n, m):
    """Same features by direct counting per node and a dense B (for small M)."""
    x = np.asarray(x, dtype=np.float32)
    M = x.shape[0]
    out = np.zeros((M, NFEAT))
    for j in range(M):
        out[j, 0] = 2.0 * np.hypot(float(x[j, 0]), float(x[j, 1]))
        q = 1
        for G in GRIDS:
            cj = _coarse_cell(x[j], G)
            call = _coarse_cell(x, G)
            for a in range(-2, 3):
                for b in range(-2, 3):
                    tx, ty = (cj[0] + a) % G, (cj[1] + b) % G
                    c = np.count_nonzero((call[:, 0] == tx) & (call[:, 1] == ty))
                    out[j, q] = np.log1p(c * G * G / M)
                    q += 1
    out[:, q] = kde_bruteforce(N, n, m, x)
    return out


def pipe_menon(N, n, m, x, iters, w0=None):
    """Pipe-Menon fixed point with the oracle's B, B^T; weights normalised to
    mean 1."""
    x = np.asarray(x, dtype=np.float64)
    w = np.ones(x.shape[0]) if w0 is None else np.array(w0, dtype=np.float64)
    for _ in range(iters):
        g = O.stage_BT(N, n, m, x, w.astype(complex))
        v = O.stage_B(N, n, m, x, g).real
        w = np.where(v > 0, w / np.where(v > 0, v, 1.0), w)
    return w / w.mean()


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[P_W1:P_B1].reshape(NFEAT, H), p[P_B1:P_W2], p[P_W2:P_B2].reshape(H, H),
            p[P_B2:P_W3], p[P_W3:P_B3], p[P_B3])


def _r(a, bf16):
    return bf16_round(a).astype(np.float64) if bf16 else np.asarray(a, dtype=np.float64)


def mlp_forward(p, X, bf16=False):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    Xr = _r(X, bf16)
    H1 = _r(np.maximum(Xr @ _r(W1, bf16) + b1, 0.0), bf16)
    H2 = _r(np.maximum(H1 @ _r(W2, bf16) + b2, 0.0), bf16)
    return H2 @ w3 + b3, (Xr, H1, H2)


def mlp_loss_grad(p, X, t, bf16=False):
    """Mean squared error and its gradient (flat, parameter layout)."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    out, (Xr, H1, H2) = mlp_forward(p, X, bf16)
    B = X.shape[0]
    err = out - np.asarray(t, dtype=np.float64)
    loss = float(np.mean(err ** 2))
    dout = 2.0 * err / B
    dH2 = _r(np.where(H2 > 0, dout[:, None] * w3[None, :], 0.0), bf16)
    dH1 = _r(np.where(H1 > 0, dH2 @ _r(W2, bf16).T, 0.0), bf16)
    g = np.zeros(P_COUNT)
    g[P_W1:P_B1] = (Xr.T @ dH1).ravel()
    g[P_B1:P_W2] = dH1.sum(0)
    g[P_W2:P_B2] = (H1.T @ dH2).ravel()
    g[P_B2:P_W3] = dH2.sum(0)
    g[P_W3:P_B3] = H2.T @ dout
    g[P_B3] = dout.sum()
    return loss, g


def init_params(seed=0):
    """He-initialised parameters (any deterministic init serves the oracle)."""
    r = np.random.default_rng(seed)
    p = np.zeros(P_COUNT)
    p[P_W1:P_W1 + NRAW * H] = r.standard_normal(NRAW * H) * np.sqrt(2.0 / NRAW)
    p[P_W2:P_B2] = r.standard_normal(H * H) * np.sqrt(2.0 / H)
    p[P_W3:P_B3] = r.standard_normal(H) * np.sqrt(1.0 / H)
    p[P_B1:P_W2] = 0.01 * r.standard_normal(H)
    p[P_B2:P_W3] = 0.01 * r.standard_normal(H)
    return p
