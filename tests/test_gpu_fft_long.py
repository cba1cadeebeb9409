"""FFT axes longer than one workgroup's LDS lines, and lengths with prime
factors above 13 (GPU, through the C ABI).

An oversampled axis of length n above 4096 (float) / 2048 (double) runs as
two four-step halves n = n1 n2 (fft.hip); FFTW, which the reference plans
with (kernel/nfft/nfft.c:6030-6031), takes any length.  Checked against the
FP64 oracle as the single-pass axes are:
  * 1-D FP32 n = 8192 and FP64 n = 4096, all nodes (rel l2 <= 1e-5 / 1e-12);
  * 2-D FP64 with one long axis (n = 4096 x 64) on a node sample;
  * the reference's own 1-D benchmark size N = M = 2^21 (n = 2^22,
    examples/nfft/nfft_benchomp.c:489) in both families on a 2000-node
    sample, FP32 against the FP64 plan over all nodes;
  * batched calls on such a plan (one right-hand side at a time) equal the
    single calls;
  * lengths with a prime factor above 13 (a generic-radix stage), a prime
    length, and a four-step split with such a factor, against the oracle.
"""
import numpy as np
import pytest

import nfft_amd as nf
import oracle as O
from nfft_amd import workloads as W

torch = pytest.importorskip("torch")

FLAGS = nf.PRE_PHI_HUT | nf.PRE_PSI | nf.MALLOC_X | nf.MALLOC_F_HAT | nf.MALLOC_F | nf.FFTW_INIT


def rel(a, b):
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def run(N, n, m, x, fh, f, precision):
    p = nf.Plan(N, x.shape[0], n=n, m=m, flags=FLAGS, precision=precision)
    try:
        p.x[:] = x
        p.precompute()
        p.f_hat[:] = fh
        p.trafo()
        t = p.f.astype(np.complex128).copy()
        p.f[:] = f
        p.adjoint()
        a = p.f_hat.astype(np.complex128).copy()
        return t, a, p.x.astype(np.float64).copy()
    finally:
        p.finalize()


@pytest.mark.gpu
@pytest.mark.parametrize("precision,N,n,tol", [("single", [4096], [8192], 1e-5),
                                               ("double", [2048], [4096], 1e-12),
                                               ("single", [3000], [12000], 1e-5)],
                         ids=["fp32_8192", "fp64_4096", "fp32_12000"])
def test_long_1d_vs_oracle(precision, N, n, tol):
    M = 5000
    x = W.uniform(M, 1, seed=21)
    cd = np.complex64 if precision == "single" else np.complex128
    fh = W.complex_normal(N[0], seed=22).astype(cd).astype(np.complex128)
    f = W.complex_normal(M, seed=23).astype(cd).astype(np.complex128)
    m = 4 if precision == "single" else 8
    t, a, xr = run(N, n, m, x, fh, f, precision)
    et = rel(t, O.nfft_trafo(N, n, m, xr, fh))
    ea = rel(a, O.nfft_adjoint(N, n, m, xr, f))
    print("%s N=%s n=%s: trafo %.2e adjoint %.2e" % (precision, N, n, et, ea))
    assert et < tol and ea < tol


@pytest.mark.gpu
def test_long_axis_2d_fp64_sample():
    N, n, m = [2048, 32], [4096, 64], 6
    M = 20000
    x = W.uniform(M, 2, seed=31)
    fh = W.complex_normal(int(np.prod(N)), seed=32)
    f = W.complex_normal(M, seed=33)
    t, a, xr = run(N, n, m, x, fh, f, "double")
    idx = np.random.default_rng(3).choice(M, size=1000, replace=False)
    g = O.fft(n, O.stage_D(N, n, m, fh), -1)
    et = rel(t[idx], O.stage_B(N, n, m, xr[idx], g))
    ea = rel(a, O.nfft_adjoint(N, n, m, xr, f))
    print("2-D n=%s: trafo (1000 nodes) %.2e adjoint %.2e" % (n, et, ea))
    assert et < 1e-12 and ea < 1e-12


@pytest.mark.gpu
def test_reference_benchmark_1d_2pow21():
    N, n, m = [1 << 21], [1 << 22], 4
    M = 1 << 21
    x = W.uniform(M, 1, seed=41)
    fh = W.complex_normal(N[0], seed=42).astype(np.complex64).astype(np.complex128)
    f = W.complex_normal(M, seed=43).astype(np.complex64).astype(np.complex128)
    t32, a32, xr = run(N, n, m, x, fh, f, "single")
    t64, a64, _ = run(N, n, m, xr, fh, f, "double")
    e32t, e32a = rel(t32, t64), rel(a32, a64)
    idx = np.random.default_rng(4).choice(M, size=2000, replace=False)
    O.set_threads(8)
    try:
        g = O.fft(n, O.stage_D(N, n, m, fh), -1)
        eot = rel(t64[idx], O.stage_B(N, n, m, xr[idx], g))
    finally:
        O.set_threads(1)
    print("1-D N=M=2^21: FP32 vs FP64 trafo %.2e adjoint %.2e; FP64 vs oracle trafo (2000 nodes) %.2e"
          % (e32t, e32a, eot))
    assert e32t < 1e-5 and e32a < 1e-5
    assert eot < 1e-10


@pytest.mark.gpu
def test_long_axis_batch_equals_single_calls():
    N, n = [4096], [8192]
    M = 3000
    p = nf.Plan(N, M, n=n, m=4, flags=FLAGS, precision="single")
    try:
        p.x[:] = W.uniform(M, 1, seed=51)
        p.precompute()
        R = 3
        fh = torch.from_numpy(W.complex_normal(R * N[0], seed=52)).to(torch.complex64).cuda()
        fo = torch.empty(R * M, dtype=torch.complex64, device="cuda")
        p.trafo_batch_device(fh.data_ptr(), fo.data_ptr(), R)
        fi = torch.from_numpy(W.complex_normal(R * M, seed=53)).to(torch.complex64).cuda()
        ho = torch.empty(R * N[0], dtype=torch.complex64, device="cuda")
        p.adjoint_batch_device(fi.data_ptr(), ho.data_ptr(), R)
        p.synchronize()
        for r in range(R):
            one = torch.empty(M, dtype=torch.complex64, device="cuda")
            p.trafo_device(fh[r * N[0]:].data_ptr(), one.data_ptr())
            oneh = torch.empty(N[0], dtype=torch.complex64, device="cuda")
            p.adjoint_device(fi[r * M:].data_ptr(), oneh.data_ptr())
            p.synchronize()
            assert torch.equal(one, fo[r * M:(r + 1) * M])
            # the 1-D adjoint adds with device atomics: order may vary
            assert float(torch.linalg.norm(oneh - ho[r * N[0]:(r + 1) * N[0]])
                         / torch.linalg.norm(oneh)) < 1e-6
    finally:
        p.finalize()


@pytest.mark.gpu
@pytest.mark.parametrize("precision,N,n,M", [("double", [16], [34], 300),
                                             ("double", [500], [1031], 2000),
                                             ("double", [16, 20], [38, 46], 1000),
                                             ("single", [1000], [2006], 3000),
                                             ("single", [4000], [8738], 3000)],
                         ids=["2x17", "prime1031", "2d_19_23", "fp32_2x17x59", "fp32_split_2x17x257"])
def test_prime_factors_vs_oracle(precision, N, n, M):
    d = len(N)
    x = W.uniform(M, d, seed=61)
    cd = np.complex64 if precision == "single" else np.complex128
    NN = int(np.prod(N))
    fh = W.complex_normal(NN, seed=62).astype(cd).astype(np.complex128)
    f = W.complex_normal(M, seed=63).astype(cd).astype(np.complex128)
    m = 4 if precision == "single" else 6
    t, a, xr = run(N, n, m, x, fh, f, precision)
    et = rel(t, O.nfft_trafo(N, n, m, xr, fh))
    ea = rel(a, O.nfft_adjoint(N, n, m, xr, f))
    tol = 1e-5 if precision == "single" else 1e-12
    print("%s N=%s n=%s: trafo %.2e adjoint %.2e" % (precision, N, n, et, ea))
    assert et < tol and ea < tol
