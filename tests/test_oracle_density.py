"""Pins of the density-compensation oracle (oracle/density.py) on the CPU.

The reference has no density-weight computation on this path, so the
oracle is pinned by mathematics: central finite differences for the MLP
gradient (FP64, no BF16 rounding), brute-force counting for the features,
torch's own bfloat16 conversion for the rounding helper, and the closed form
of Pipe-Menon on a Cartesian node set (all weights equal).
"""
import numpy as np
import pytest

from oracle import density as D
from nfft_amd import workloads as W


def test_bf16_round_matches_torch():
    torch = pytest.importorskip("torch")
    r = np.random.default_rng(1)
    a = np.concatenate([r.standard_normal(10000).astype(np.float32) * 10.0 ** r.integers(-8, 8, 10000),
                        np.float32([0.0, -0.0, 1.0, 1.00390625, 1.0078125, -2.5])]).astype(np.float32)
    ref = torch.from_numpy(a).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(D.bf16_round(a), ref)


def test_features_match_bruteforce():
    x = W.clustered_2d(400, seed=3)
    fa = D.features(x, [16, 16], [32, 32], 4)
    fb = D.features_bruteforce(x, [16, 16], [32, 32], 4)
    assert fa.shape == (400, D.NFEAT)
    assert np.allclose(fa, fb, rtol=0, atol=1e-12)
    assert np.all(fa[:, D.NRAW:] == 0)
    # the density estimate follows the clustering: nodes near the centre of
    # the cluster see more neighbours
    r = np.hypot(x[:, 0], x[:, 1])
    assert fa[r < 0.05, D.NRAW - 1].mean() > fa[r > 0.3, D.NRAW - 1].mean() + 1.0


def test_kde_planted_window_mistake_is_caught():
    """A plausible mistake in assembling B (the window evaluated with cut-off
    m - 1) moves the density estimate far beyond the 1e-12 agreement asserted
    above."""
    x = W.clustered_2d(60, seed=4)
    good = D.kde([16, 16], [32, 32], 4, x)
    assert np.allclose(good, D.kde_bruteforce([16, 16], [32, 32], 4, x), atol=1e-12)
    bad = D.kde_bruteforce([16, 16], [32, 32], 4, x, _dm=-1)
    assert np.max(np.abs(good - bad)) > 1e-3


def test_features_uniform_density_near_log2():
    # a Cartesian node set at the cell centres of the oversampled grid: every
    # coarse cell holds the mean count, so every count feature is log(1 + 1),
    # and every node has the same density estimate, log(v / mean v) = 0
    g = (np.arange(128) + 0.5) / 128 - 0.5
    x = np.stack(np.meshgrid(g, g, indexing="ij"), -1).reshape(-1, 2)
    f = D.features(x, [64, 64], [128, 128], 4)
    assert np.allclose(f[:, 1:D.NRAW - 1], np.log(2.0), atol=1e-12)
    assert np.allclose(f[:, D.NRAW - 1], 0.0, atol=1e-12)


@pytest.mark.parametrize("bf16", [False])
def test_mlp_gradient_finite_differences(bf16):
    r = np.random.default_rng(7)
    p = D.init_params(3)
    X = r.standard_normal((37, D.NFEAT))
    X[:, D.NRAW:] = 0
    t = r.standard_normal(37)
    loss, g = D.mlp_loss_grad(p, X, t, bf16=bf16)
    idx = np.concatenate([r.choice(np.arange(D.P_W1, D.P_W1 + D.NRAW * D.H), 40, replace=False),
                          r.choice(np.arange(D.P_B1, D.P_COUNT), 40, replace=False)])
    h = 1e-6
    for i in idx:
        pp = p.copy()
        pp[i] += h
        lp, _ = D.mlp_loss_grad(pp, X, t)
        pp[i] -= 2 * h
        lm, _ = D.mlp_loss_grad(pp, X, t)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_mlp_planted_gradient_mistake_is_caught():
    """A plausible mistake (ReLU mask of layer 1 taken from layer 2's input
    instead of its output) fails the finite-difference check."""
    r = np.random.default_rng(8)
    p = D.init_params(4)
    X = r.standard_normal((20, D.NFEAT))
    t = r.standard_normal(20)
    _, g = D.mlp_loss_grad(p, X, t)
    W1, b1, W2, b2, w3, b3 = D.unpack(p)
    out, (Xr, H1, H2) = D.mlp_forward(p, X)
    dout = 2 * (out - t) / 20
    dH2 = np.where(H2 > 0, dout[:, None] * w3[None, :], 0.0)
    dH1_wrong = dH2 @ W2.T  # mask dropped
    gw = (Xr.T @ dH1_wrong).ravel()
    assert np.linalg.norm(gw - g[D.P_W1:D.P_B1]) > 1e-3 * np.linalg.norm(g[D.P_W1:D.P_B1])


def test_pipe_menon_cartesian_is_constant():
    # nodes on the oversampled-grid-compatible Cartesian lattice of N x N:
    # by symmetry B B^T 1 is the same at every node, so Pipe-Menon weights
    # are constant
    N = [16, 16]
    g = np.arange(16) / 16 - 0.5
    x = np.stack(np.meshgrid(g, g, indexing="ij"), -1).reshape(-1, 2)
    w = D.pipe_menon(N, [32, 32], 4, x, 5)
    assert np.allclose(w, 1.0, atol=1e-10)


def test_pipe_menon_radial_downweights_centre():
    N = [16, 16]
    x = W.golden_radial(24, 32)
    w = D.pipe_menon(N, [32, 32], 4, x, 10)
    r = np.hypot(x[:, 0], x[:, 1])
    assert w[r < 0.05].mean() < 0.2 * w[r > 0.3].mean()
