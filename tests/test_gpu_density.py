"""Density-compensation weights on the GPU (nfft_amd/csrc/density.hip through
the C ABI) against the CPU oracle (oracle/density.py):

  features         exact counts, log1p in FP32: |diff| <= 1e-6; the density
                   estimate log(v / mean v), v = B B^T 1 in FP32: <= 2e-5
  Pipe-Menon       mean-normalised weights, FP32 spreading/interpolation vs
                   FP64 oracle: rel l2 <= 1e-4
  MLP forward      BF16 MFMA with FP32 accumulation vs FP64 with the same BF16
                   roundings: rel l2 <= 1e-3
  MLP gradient     same, per parameter block: rel l2 <= 2e-2 (a BF16 rounding
                   of an intermediate may land on the other side of a tie)
  training         Adam steps reduce the loss on Pipe-Menon targets
  learned weights  CGNR weighted by the trained model reaches a better image in
                   its first iterations than unweighted CGNR
"""
import numpy as np
import pytest

import nfft_amd as nf
import oracle as O
from oracle import density as OD
from nfft_amd import density as D
from nfft_amd import workloads as W

torch = pytest.importorskip("torch")

FLAGS = nf.PRE_PHI_HUT | nf.PRE_PSI | nf.MALLOC_X | nf.MALLOC_F_HAT | nf.MALLOC_F | nf.FFTW_INIT | nf.FFT_OUT_OF_PLACE


def make_plan(N, x, m=4):
    p = nf.Plan(N, x.shape[0], n=[2 * t for t in N], m=m, flags=FLAGS, precision="single")
    p.x[:] = x
    p.precompute()
    return p


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.mark.gpu
@pytest.mark.parametrize("gen", ["clustered", "spiral", "radial"])
def test_features_vs_oracle(gen):
    x = {"clustered": lambda: W.clustered_2d(5000, seed=1),
         "spiral": lambda: W.spiral(64, 8, 2.5),
         "radial": lambda: W.golden_radial(60, 128)}[gen]()
    p = make_plan([64, 64], x)
    try:
        f = D.features(p).cpu().numpy()
        ref = OD.features(p.x.astype(np.float64), [64, 64], [128, 128], 4)
    finally:
        p.finalize()
    k = OD.NRAW - 1
    assert np.max(np.abs(f[:, :k] - ref[:, :k])) <= 1e-6
    # density estimate: FP32 spreading and interpolation vs the FP64 oracle
    print("kde max |diff| %.2e" % np.max(np.abs(f[:, k] - ref[:, k])))
    assert np.max(np.abs(f[:, k] - ref[:, k])) <= 2e-5
    assert np.all(f[:, OD.NRAW:] == 0)


@pytest.mark.gpu
def test_pipe_menon_vs_oracle():
    N = [32, 32]
    x = W.golden_radial(40, 64)
    p = make_plan(N, x)
    try:
        w = D.pipe_menon(p, 8)
        xr = p.x.astype(np.float64)
    finally:
        p.finalize()
    ref = OD.pipe_menon(N, [64, 64], 4, xr, 8)
    w = w / w.mean()
    print("pipe-menon rel l2 %.2e" % rel(w, ref))
    assert rel(w, ref) <= 1e-4


def _batch(seed, rows=1000):
    r = np.random.default_rng(seed)
    X = np.abs(r.standard_normal((rows, D.num_features()))).astype(np.float32)
    X[:, OD.NRAW:] = 0
    t = r.standard_normal(rows).astype(np.float32)
    return X, t


@pytest.mark.gpu
@pytest.mark.parametrize("rows", [16, 1000, 4099])
def test_mlp_forward_vs_oracle(rows):
    X, _ = _batch(11, rows)
    p0 = OD.init_params(5).astype(np.float32)
    model = D.DcfModel(params=p0)
    try:
        out = model.predict(torch.from_numpy(X).cuda()).cpu().numpy()
    finally:
        model.close()
    ref, _ = OD.mlp_forward(p0.astype(np.float64), X, bf16=True)
    print("forward rel l2 %.2e" % rel(out, ref))
    assert rel(out, ref) <= 1e-3


@pytest.mark.gpu
@pytest.mark.parametrize("rows", [64, 1000, 2500])
def test_mlp_gradient_vs_oracle(rows):
    X, t = _batch(12, rows)
    p0 = OD.init_params(6).astype(np.float32)
    model = D.DcfModel(params=p0)
    try:
        loss, g = model.loss_grad(torch.from_numpy(X).cuda(), torch.from_numpy(t).cuda())
        assert np.array_equal(model.params, p0)  # loss_grad leaves parameters alone
    finally:
        model.close()
    ref_loss, ref = OD.mlp_loss_grad(p0.astype(np.float64), X, t, bf16=True)
    assert abs(loss - ref_loss) <= 1e-3 * abs(ref_loss)
    blocks = {"W1": (OD.P_W1, OD.P_B1), "b1": (OD.P_B1, OD.P_W2), "W2": (OD.P_W2, OD.P_B2),
              "b2": (OD.P_B2, OD.P_W3), "w3": (OD.P_W3, OD.P_B3), "b3": (OD.P_B3, OD.P_COUNT)}
    for name, (a, b) in blocks.items():
        e = rel(g[a:b], ref[a:b])
        print("grad %s rel l2 %.2e" % (name, e))
        assert e <= 2e-2, name


@pytest.mark.gpu
def test_training_reduces_loss_on_pipe_menon_targets():
    feats, targets = [], []
    for s in range(4):
        x = W.spiral(64, 8 + 4 * s, 1.5 + 0.4 * s)
        p = make_plan([64, 64], x)
        try:
            w = D.pipe_menon(p, 10)
            feats.append(D.features(p))
            targets.append(torch.from_numpy(np.log(w / w.mean()).astype(np.float32)).cuda())
        finally:
            p.finalize()
    X = torch.cat(feats)
    T = torch.cat(targets)
    model = D.DcfModel(seed=1)
    try:
        first = model.loss_grad(X, T)[0]
        g = torch.Generator(device="cpu").manual_seed(0)
        for it in range(300):
            idx = torch.randint(0, X.shape[0], (4096,), generator=g).cuda()
            model.train_step(X[idx].contiguous(), T[idx].contiguous(), 3e-3)
        last = model.loss_grad(X, T)[0]
    finally:
        model.close()
    print("loss %.4f -> %.4f" % (first, last))
    assert last < 0.2 * first


def _cgnr_image_errors(p, y, w, f_true, iters):
    s = nf.Solver(p, nf.CGNR | (nf.PRECOMPUTE_WEIGHT if w is not None else 0))
    try:
        s.set_sync(1)
        s.y[:] = y
        s.f_hat_iter[:] = 0
        if w is not None:
            s.w[:] = w
        s.before_loop()
        errs = []
        for _ in range(iters):
            s.loop_one_step()
            fh = s.f_hat_iter.astype(np.complex128)
            errs.append(float(np.linalg.norm(fh.real - f_true) / np.linalg.norm(f_true)))
        return errs
    finally:
        s.finalize()


@pytest.mark.gpu
def test_learned_weights_speed_up_cgnr():
    """What the trained network gains: on held-out trajectories, CGNR weighted
    by the model's prediction reaches a better image in the first iterations
    than unweighted CGNR, close to what Pipe-Menon weights give (N = 64 model
    trained here on 40 trajectories; phantom data with 1% noise)."""
    from nfft_amd import dcf_train as T
    N = 64
    X, Tg = T.training_set(40, N=N, per_traj=1024, pm_iters=15, seed0=100, log=lambda *a: None)
    model = T.train(X, Tg, steps=600, batch=4096, lr=3e-3, seed=2, log=lambda *a: None)
    try:
        for seed in (1001, 1002):
            x, _ = T.trajectory(seed, N)
            p = T.make_plan(x, N)
            try:
                f_true = W.phantom(N, 12, seed=seed)
                y = O.nfft_trafo([N, N], [2 * N, 2 * N], 4, p.x.astype(np.float64), f_true)
                y = y + 0.01 * np.sqrt(np.mean(np.abs(y) ** 2)) * W.complex_normal(p.M, seed=seed)
                e_plain = _cgnr_image_errors(p, y, None, f_true, 5)
                e_pm = _cgnr_image_errors(p, y, D.pipe_menon(p, 15), f_true, 5)
                e_learned = _cgnr_image_errors(p, y, model.weights(p).cpu().numpy(), f_true, 5)
            finally:
                p.finalize()
            print("seed %d image error after 5 iterations: plain %.4f, pipe-menon %.4f, learned %.4f"
                  % (seed, e_plain[-1], e_pm[-1], e_learned[-1]))
            assert e_learned[-1] < 0.92 * e_plain[-1]
            assert e_learned[-1] < 1.1 * e_pm[-1]
    finally:
        model.close()
