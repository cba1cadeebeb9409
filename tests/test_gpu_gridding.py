"""Gridding reconstruction f_hat = A^H (w .* f) through the C ABI
(nfft[f]_gridding_device, nfft[f]_gridding), the weighted adjoint of the
reference's MRI applications (applications/mri/mri2d/
reconstruct_data_gridding.c:54-68).

Parity: against the oracle's adjoint of w .* f (FP64 <= 1e-12 relative l2,
FP32 <= 1e-5), host and device forms, w = NULL equal to the plain adjoint,
the caller's f left unchanged.  Use: with density-compensation weights
(Pipe-Menon, and the shipped learned model on a held-out N = 256 trajectory)
the gridding image is closer to the phantom than the unweighted adjoint.
"""
import os

import numpy as np
import pytest

import nfft_amd as nf
import oracle as O
from nfft_amd import workloads as W

torch = pytest.importorskip("torch")

FLAGS = nf.PRE_PHI_HUT | nf.PRE_PSI | nf.MALLOC_X | nf.MALLOC_F_HAT | nf.MALLOC_F | nf.FFTW_INIT | nf.FFT_OUT_OF_PLACE


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def scaled_err(img, ref):
    """Relative error of img after the least-squares complex scale."""
    img = np.asarray(img).ravel()
    ref = np.asarray(ref).ravel()
    a = np.vdot(img, ref) / np.vdot(img, img)
    return float(np.linalg.norm(a * img - ref) / np.linalg.norm(ref))


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("N,gen", [([32, 24], lambda: W.clustered_2d(3000, seed=41)),
                                   ([12, 16, 10], lambda: W.uniform(2000, 3, seed=42))],
                         ids=["2d", "3d"])
def test_gridding_vs_oracle(precision, N, gen):
    x = gen()
    n = [2 * t for t in N]
    m = 4
    p = nf.Plan(N, x.shape[0], n=n, m=m, flags=FLAGS, precision=precision)
    tol = 1e-12 if precision == "double" else 1e-5
    try:
        p.x[:] = x
        p.precompute()
        xr = p.x.astype(np.float64)
        f = W.complex_normal(p.M, seed=43).astype(p.cdtype)
        w = (0.1 + np.random.default_rng(44).random(p.M)).astype(p.rdtype)
        ref = O.nfft_adjoint(N, n, m, xr, f.astype(np.complex128) * w.astype(np.float64))
        # host form
        p.f[:] = f
        p.gridding(w)
        assert np.array_equal(p.f, f)
        e_host = rel(p.f_hat.astype(np.complex128), ref)
        # device form
        cd = torch.complex128 if precision == "double" else torch.complex64
        rd = torch.float64 if precision == "double" else torch.float32
        d_f = torch.from_numpy(f).to("cuda")
        d_w = torch.from_numpy(w).to("cuda")
        d_fh = torch.empty(p.p.N_total, dtype=cd, device="cuda")
        p.gridding_device(d_f.data_ptr(), d_w.data_ptr(), d_fh.data_ptr())
        p.synchronize()
        e_dev = rel(d_fh.cpu().numpy().astype(np.complex128), ref)
        assert torch.equal(d_f.cpu(), torch.from_numpy(f))
        # no weights: the plain adjoint
        p.gridding_device(d_f.data_ptr(), 0, d_fh.data_ptr())
        p.synchronize()
        plain = d_fh.cpu().numpy().astype(np.complex128)
        p.f[:] = f
        p.adjoint()
        assert rel(plain, p.f_hat.astype(np.complex128)) <= (1e-6 if precision == "single" else 1e-14)
        assert d_w.dtype == rd
    finally:
        p.finalize()
    print("%s %dD gridding rel l2: host %.2e device %.2e" % (precision, len(N), e_host, e_dev))
    assert e_host <= tol and e_dev <= tol


@pytest.mark.gpu
def test_gridding_with_density_weights_beats_unweighted():
    from nfft_amd import dcf_train as T
    from nfft_amd import density as D
    N = 256
    x, _ = T.trajectory(5001, N)  # held out from the model's training seeds
    p = T.make_plan(x, N)
    try:
        f_true = W.phantom(N, 20, seed=5001)
        p.f_hat[:] = f_true
        p.trafo()
        y = p.f.copy()
        errs = {}
        weights = {"plain": None, "pipe_menon": D.pipe_menon(p, 20)}
        if os.path.exists(T.DEFAULT_MODEL):
            model = D.DcfModel.load(T.DEFAULT_MODEL)
            try:
                weights["learned"] = model.weights(p).cpu().numpy()
            finally:
                model.close()
        for name, w in weights.items():
            p.f[:] = y
            p.gridding(w)
            errs[name] = scaled_err(p.f_hat.astype(np.complex128), f_true)
    finally:
        p.finalize()
    print("gridding image error (after LS scale): " + ", ".join("%s %.4f" % kv for kv in errs.items()))
    assert errs["pipe_menon"] < 0.6 * errs["plain"]
    if "learned" in errs:
        assert errs["learned"] < 0.6 * errs["plain"]
        assert errs["learned"] < 1.15 * errs["pipe_menon"]


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["double", "single"])
def test_direct_only_plan_reads_current_nodes(precision):
    """A plan too small for the fast path (N_t <= m) runs the NDFT, which
    reads the caller's x on every call (reference nfft.c:145-297 reads
    ths->x each time): changing x between two gridding calls, host and device
    forms, must change the result accordingly."""
    N = [4, 4]
    M = 40
    p = nf.Plan(N, M, n=[8, 8], m=4, flags=FLAGS, precision=precision)
    tol = 1e-12 if precision == "double" else 1e-5
    try:
        f = W.complex_normal(M, seed=61).astype(p.cdtype)
        w = (0.5 + np.random.default_rng(62).random(M)).astype(p.rdtype)
        cd = torch.complex128 if precision == "double" else torch.complex64
        d_f = torch.from_numpy(f).to("cuda")
        d_w = torch.from_numpy(w).to("cuda")
        d_fh = torch.empty(p.p.N_total, dtype=cd, device="cuda")
        for seed in (63, 64):
            p.x[:] = W.uniform(M, 2, seed=seed).astype(p.rdtype)
            xr = p.x.astype(np.float64)
            ref = O.ndft_adjoint(N, xr, f.astype(np.complex128) * w.astype(np.float64))
            p.f[:] = f
            p.gridding(w)
            assert rel(p.f_hat.astype(np.complex128), ref) <= tol
            p.gridding_device(d_f.data_ptr(), d_w.data_ptr(), d_fh.data_ptr())
            p.synchronize()
            assert rel(d_fh.cpu().numpy().astype(np.complex128), ref) <= tol
    finally:
        p.finalize()
