"""Training and evaluation of the learned density-compensation weights.

    python -m nfft_amd.dcf_train [--n-traj 2000] [--steps 3000] [--out PATH]

Training data: seeded 2-D trajectories for an N = 256 x 256 plan (sigma 2,
m = 4, float family): variable-density spirals with 8-48 interleaves and
density exponent alpha in [1, 3], and golden-angle radial sets of 100-400
spokes x 512 samples (for other N: spokes scaled by N / 256, 2N samples).
For each, the library computes Pipe-Menon weights (w <- w / (B B^T w),
20 iterations) and the node-density features on the GPU;
a random subset of nodes per trajectory enters the training set with target
log(w / mean w).  The MLP is trained with the library's BF16 MFMA forward and
backward kernels and Adam.  The parameters go to ``nfft_amd/data/dcf_model.npy``.

``cgnr_time_to`` runs the library's device CGNR (weighted when w is given)
until the normal-equation residual sqrt(z^H z / z0^H z0) falls below a
tolerance (z = A^H W r, the quantity CGNR drives to zero,
reference kernel/solver/solver.c:279-287) and returns iterations and
seconds.
"""
import argparse
import os
import time

import numpy as np

import nfft_amd as nf
from nfft_amd import density as D
from nfft_amd import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_MODEL = os.path.join(HERE, "data", "dcf_model.npy")
FLAGS = (nf.PRE_PHI_HUT | nf.PRE_PSI | nf.MALLOC_X | nf.MALLOC_F_HAT | nf.MALLOC_F
         | nf.FFTW_INIT | nf.FFT_OUT_OF_PLACE)


def trajectory(seed, N=256):
    """Seeded trajectory: even seeds spirals, odd seeds golden-angle radial."""
    r = np.random.default_rng(1000003 * seed + 17)
    if seed % 2 == 0:
        arms = int(r.integers(8, 49))
        alpha = float(r.uniform(1.0, 3.0))
        return W.spiral(N, arms, alpha, M=(N * N // arms) * arms), ("spiral", arms, alpha)
    spokes = int(r.integers(100, 401)) * N // 256
    x = W.golden_radial(spokes, 2 * N)
    # rotate by a seeded angle so that radial sets differ beyond their count
    th = r.uniform(0, np.pi)
    c, s = np.cos(th), np.sin(th)
    x = np.stack([c * x[:, 0] - s * x[:, 1], s * x[:, 0] + c * x[:, 1]], -1)
    x = x - np.round(x)
    x[x >= 0.5] -= 1.0
    return x, ("radial", spokes)


def make_plan(x, N=256, m=4):
    p = nf.Plan([N, N], x.shape[0], n=[2 * N, 2 * N], m=m, flags=FLAGS, precision="single")
    p.x[:] = x
    p.precompute()
    return p


def training_set(n_traj, N=256, per_traj=2048, pm_iters=20, seed0=0, log=print):
    import torch
    feats, targets = [], []
    g = torch.Generator(device="cpu").manual_seed(seed0)
    t0 = time.time()
    for k in range(n_traj):
        x, _ = trajectory(seed0 + k, N)
        p = make_plan(x, N)
        try:
            w = D.pipe_menon(p, pm_iters)
            f = D.features(p)
        finally:
            p.finalize()
        sel = torch.randperm(f.shape[0], generator=g)[:per_traj].cuda()
        feats.append(f[sel])
        t = np.log(w / w.mean()).astype(np.float32)
        targets.append(torch.from_numpy(t).cuda()[sel])
        if (k + 1) % 200 == 0:
            log("[dcf] %d / %d trajectories (%.1f s)" % (k + 1, n_traj, time.time() - t0))
    return torch.cat(feats).contiguous(), torch.cat(targets).contiguous()


def train(X, T, steps=3000, batch=8192, lr=2e-3, seed=0, log=print):
    import torch
    model = D.DcfModel(seed=seed)
    g = torch.Generator(device="cpu").manual_seed(seed)
    t0 = time.time()
    for it in range(steps):
        idx = torch.randint(0, X.shape[0], (batch,), generator=g).cuda()
        a = it / max(1, steps - 1)
        step_lr = lr * (0.05 + 0.95 * 0.5 * (1 + np.cos(np.pi * a)))  # cosine decay
        loss = model.train_step(X[idx].contiguous(), T[idx].contiguous(), step_lr)
        if (it + 1) % 500 == 0:
            log("[dcf] step %d loss %.4f (%.1f s)" % (it + 1, loss, time.time() - t0))
    return model


def cgnr_time_to(plan, y, w=None, tol=1e-6, max_iter=3000):
    """Device CGNR from f = 0 until sqrt(dot_z_hat / dot_z_hat_0) < tol.
    Returns (iterations, seconds, final relative residual, f_hat)."""
    flags = nf.CGNR | (nf.PRECOMPUTE_WEIGHT if w is not None else 0)
    s = nf.Solver(plan, flags)
    try:
        s.set_sync(0)
        s.y[:] = y
        s.f_hat_iter[:] = 0
        if w is not None:
            s.w[:] = w
        t0 = time.perf_counter()
        s.before_loop()
        z0 = float(s.s.dot_z_hat_iter)
        it, res = 0, 1.0
        while it < max_iter:
            s.loop_one_step()
            it += 1
            res = float(np.sqrt(max(float(s.s.dot_z_hat_iter), 0.0) / z0)) if z0 > 0 else 0.0
            if res < tol:
                break
        plan.synchronize()
        dt = time.perf_counter() - t0
        s.set_sync(1)
        s.sync()
        fh = s.f_hat_iter.astype(np.complex128).copy()
    finally:
        s.finalize()
    return it, dt, res, fh


def image_error_trace(plan, y, f_true, w=None, iters=300):
    """Relative l2 error of Re(f_hat_iter) against f_true after each CGNR step."""
    flags = nf.CGNR | (nf.PRECOMPUTE_WEIGHT if w is not None else 0)
    s = nf.Solver(plan, flags)
    try:
        s.set_sync(1)
        s.y[:] = y
        s.f_hat_iter[:] = 0
        if w is not None:
            s.w[:] = w
        s.before_loop()
        nt = np.linalg.norm(f_true)
        errs = []
        for _ in range(iters):
            s.loop_one_step()
            errs.append(float(np.linalg.norm(s.f_hat_iter.real - f_true) / nt))
        return errs
    finally:
        s.finalize()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n-traj", type=int, default=2000)
    ap.add_argument("--per-traj", type=int, default=2048)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--out", default=DEFAULT_MODEL)
    a = ap.parse_args()
    t0 = time.time()
    X, T = training_set(a.n_traj, per_traj=a.per_traj)
    t1 = time.time()
    print("[dcf] training set %d rows in %.1f s" % (X.shape[0], t1 - t0), flush=True)
    model = train(X, T, steps=a.steps)
    print("[dcf] trained in %.1f s, final full-set loss %.4f"
          % (time.time() - t1, model.loss_grad(X[:200000].contiguous(), T[:200000].contiguous())[0]),
          flush=True)
    os.makedirs(os.path.dirname(a.out), exist_ok=True)
    model.save(a.out)
    print("[dcf] saved %s" % a.out, flush=True)


if __name__ == "__main__":
    main()
