"""Seeded synthetic node sets and coefficient vectors of the benchmark and
parity workloads.  Everything is generated in FP64 on the host with numpy's
PCG64 so that the GPU library and the CPU oracle see identical inputs.

Node sets (x in [-0.5, 0.5)^d, row j = node j):
  uniform(M, d)              iid uniform
  clustered_2d(M)            half uniform, half N(0, 0.05^2) at the origin,
                             wrapped by x - round(x)
  kooshball(spokes, samples) 3-D radial spokes: directions uniform on the
                             sphere, radii -1/2 + k/samples, wrapped
  golden_radial(spokes, samples)  2-D golden-angle radial trajectory
  spiral(N, arms, alpha)     2-D variable-density spiral: arm b, sample i of
                             M/arms, t = (i / (M/arms))^(1/alpha),
                             x = A t (cos, sin)(2 pi (w t + b/arms)), A = 1/2,
                             w = 50 N / 64, M = N^2 (alpha = 2 is the spiral
                             of the reference's applications/mri/mri2d/
                             construct_knots_spiral.m)
  phantom(N, n_ellipses)     N x N image of random ellipses, intensities
                             U[0,1], row-major (the f_hat layout)
"""
import numpy as np


def rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def _wrap(x):
    x = x - np.round(x)
    # x - round(x) lies in [-0.5, 0.5]; fold the +0.5 edge
    x[x >= 0.5] -= 1.0
    return x


def uniform(M, d, seed=1):
    return rng(seed).uniform(-0.5, 0.5, size=(M, d))


def clustered_2d(M, seed=2, sd=0.05):
    r = rng(seed)
    h = M // 2
    a = r.uniform(-0.5, 0.5, size=(h, 2))
    b = _wrap(sd * r.standard_normal(size=(M - h, 2)))
    return np.concatenate([a, b], axis=0)


def kooshball(spokes=16384, samples=1024, seed=3):
    r = rng(seed)
    z = r.uniform(-1.0, 1.0, size=spokes)
    phi = r.uniform(0.0, 2 * np.pi, size=spokes)
    s = np.sqrt(1.0 - z * z)
    dirs = np.stack([s * np.cos(phi), s * np.sin(phi), z], axis=1)
    rad = -0.5 + np.arange(samples) / samples
    x = dirs[:, None, :] * rad[None, :, None]
    return _wrap(x.reshape(-1, 3))


def golden_radial(spokes=403, samples=512):
    golden = (1 + 5 ** 0.5) / 2
    theta = np.arange(spokes) * np.pi / golden
    rad = -0.5 + np.arange(samples) / samples
    dirs = np.stack([np.cos(theta), np.sin(theta)], axis=1)
    x = dirs[:, None, :] * rad[None, :, None]
    return _wrap(x.reshape(-1, 2))


def spiral(N, arms, alpha=2.0, M=None):
    M = N * N if M is None else M
    per = M // arms
    t = (np.arange(per) / per) ** (1.0 / alpha)
    w = N / 64 * 50
    ang = 2 * np.pi * (w * t[None, :] + np.arange(arms)[:, None] / arms)
    x = np.stack([0.5 * t[None, :] * np.cos(ang), 0.5 * t[None, :] * np.sin(ang)], -1)
    return _wrap(x.reshape(-1, 2))


def phantom(N, n_ellipses=20, seed=5):
    r = rng(seed)
    u = (np.arange(N) - N // 2) / N
    X, Y = np.meshgrid(u, u, indexing="ij")
    img = np.zeros((N, N))
    for _ in range(n_ellipses):
        cx, cy = r.uniform(-0.3, 0.3, 2)
        a, b = r.uniform(0.03, 0.2, 2)
        th = r.uniform(0, np.pi)
        c, s = np.cos(th), np.sin(th)
        xr = (X - cx) * c + (Y - cy) * s
        yr = -(X - cx) * s + (Y - cy) * c
        img[(xr / a) ** 2 + (yr / b) ** 2 <= 1.0] += r.uniform(0, 1)
    return img.ravel()


def complex_normal(n, seed=4):
    r = rng(seed)
    return r.standard_normal(n) + 1j * r.standard_normal(n)
