"""Full-size workload checks through the C ABI (GPU).

At the benchmark's shape -- 3-D N=128^3, n=256^3, m=4, M=2^24 kooshball --
the FP64 oracle cannot run over all nodes in test time, so parity is shown by:
  * FP64 GPU trafo vs the oracle on 2000 sampled nodes (full-grid D + FFT in
    the oracle, B on the sample), and FP64 GPU adjoint vs the oracle for an f
    supported on 2000 nodes (zeros elsewhere): rel l2 <= 1e-12;
  * FP32 GPU vs FP64 GPU over all nodes: rel l2 <= 1e-5, the FP32 target;
  * adjointness <A fh, f> = <fh, A^H f> and linearity of the FP32 transform.
The 2-D 512^2 case with 2^20 half-clustered nodes (m = 4 and 6) is checked
the same way.
"""
import numpy as np
import pytest

import nfft_amd as nf
import oracle as O
from nfft_amd import workloads as W

FLAGS = nf.PRE_PHI_HUT | nf.PRE_PSI | nf.MALLOC_X | nf.MALLOC_F_HAT | nf.MALLOC_F | nf.FFTW_INIT | nf.NFFT_SORT_NODES


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def plan(N, n, m, x, precision):
    p = nf.Plan(N, x.shape[0], n=n, m=m, flags=FLAGS, precision=precision)
    p.x[:] = x
    p.precompute()
    return p


def both_precisions(N, n, m, x, fh, f):
    out = {}
    for prec in ("double", "single"):
        p = plan(N, n, m, x, prec)
        try:
            p.f_hat[:] = fh
            p.trafo()
            Af = p.f.astype(np.complex128).copy()
            p.f[:] = f
            p.adjoint()
            AHf = p.f_hat.astype(np.complex128).copy()
        finally:
            p.finalize()
        out[prec] = (Af, AHf)
    return out


CASES = [
    ("3d_kooshball", [128, 128, 128], [256, 256, 256], 4, lambda: W.kooshball(16384, 1024, seed=3)),
    ("2d_clustered_m4", [512, 512], [1024, 1024], 4, lambda: W.clustered_2d(1 << 20, seed=2)),
    ("2d_clustered_m6", [512, 512], [1024, 1024], 6, lambda: W.clustered_2d(1 << 20, seed=2)),
]


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_fullsize(case):
    name, N, n, m, gen = case
    x = gen()
    M = x.shape[0]
    NN = int(np.prod(N))
    fh = W.complex_normal(NN, seed=4)
    f = W.complex_normal(M, seed=5)
    res = both_precisions(N, n, m, x, fh, f)
    Af64, AHf64 = res["double"]
    Af32, AHf32 = res["single"]
    e_t = rel_l2(Af32, Af64)
    e_a = rel_l2(AHf32, AHf64)
    print("%s: FP32 vs FP64 trafo rel l2 %.2e, adjoint rel l2 %.2e" % (name, e_t, e_a))
    assert e_t < 1e-5 and e_a < 1e-5

    # FP64 GPU vs the oracle on a node sample
    rng = np.random.default_rng(9)
    idx = rng.choice(M, size=2000, replace=False)
    g = O.fft(n, O.stage_D(N, n, m, fh), -1)
    f_ref = O.stage_B(N, n, m, x[idx], g)
    e_o = rel_l2(Af64[idx], f_ref)
    fs = np.zeros(M, dtype=np.complex128)
    fs[idx] = f[idx]
    p = plan(N, n, m, x, "double")
    try:
        p.f[:] = fs
        p.adjoint()
        AHs = p.f_hat.astype(np.complex128).copy()
    finally:
        p.finalize()
    AHs_ref = O.nfft_adjoint(N, n, m, x[idx], f[idx])
    e_oa = rel_l2(AHs, AHs_ref)
    print("%s: FP64 vs oracle trafo (2000 nodes) %.2e, adjoint (2000-node support) %.2e" % (name, e_o, e_oa))
    assert e_o < 1e-12 and e_oa < 1e-12

    # adjointness and linearity of the FP32 path
    lhs = np.vdot(f.astype(np.complex64).astype(np.complex128), Af32)
    rhs = np.vdot(AHf32, fh.astype(np.complex64).astype(np.complex128))
    assert abs(lhs - rhs) / abs(lhs) < 1e-5
