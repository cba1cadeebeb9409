"""Iterative inverse NFFT through the C ABI (GPU) against the solver oracle.

The oracle (oracle/solver.py, restating kernel/solver/solver.c and pinned
against least squares on CPU) runs the same recurrences with the FP64 NFFT
oracle as operator.  The library's device CGNR (plan is one of its own) and
its host path (the other methods) must follow the same iterates:
  double family: f_hat_iter and dot_* within 1e-9 relative for 8 steps
  float family:  within 2e-4 relative (FP32 operator, FP64 inner products;
                 1e-2 with the ill-conditioning Jackson damping factors)
"""
import numpy as np
import pytest

import nfft_amd as nf
import oracle as O
from oracle import solver as S
from nfft_amd import workloads as W

FLAGS = nf.PRE_PHI_HUT | nf.PRE_PSI | nf.MALLOC_X | nf.MALLOC_F_HAT | nf.MALLOC_F | nf.FFTW_INIT | nf.FFT_OUT_OF_PLACE


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def jackson_damping(N, precision):
    """w_hat[k0 N1 + k1] = prod_t modified_jackson2(N_t, k_t - N_t/2), the
    library's damping factors (reference kernel/util/damp.c:33-54) in the
    product form of examples/solver/glacier.c.in:81-86."""
    import ctypes
    L = nf.lib()
    fn = getattr(L, "nfft_modified_jackson2" if precision == "double" else "nfftf_modified_jackson2")
    fn.restype = ctypes.c_double if precision == "double" else ctypes.c_float
    fn.argtypes = [ctypes.c_ssize_t, ctypes.c_ssize_t]
    per = [np.array([fn(n, k - n // 2) for k in range(n)]) for n in N]
    return np.multiply.outer(per[0], per[1]).ravel()


def run_case(method, extra, precision, steps=8, seed=5, M=900, w_hat=None):
    N = [24, 20]
    m = 4
    x = W.clustered_2d(M, seed=seed)
    M = x.shape[0]
    NN = int(np.prod(N))
    p = nf.Plan(N, M, m=m, n=[48, 40], flags=FLAGS, precision=precision)
    p.x[:] = x
    p.precompute()
    xr = p.x.astype(np.float64).copy()
    n = p.n
    r = np.random.default_rng(seed)
    f_true = W.complex_normal(NN, seed=seed + 1)
    y = O.nfft_trafo(N, n, m, xr, f_true) + 0.01 * W.complex_normal(M, seed=seed + 2)
    flags = method | extra
    s = nf.Solver(p, flags)
    o = S.Solver(lambda f: O.nfft_trafo(N, n, m, xr, f), lambda v: O.nfft_adjoint(N, n, m, xr, v), M, NN, flags)
    try:
        s.y[:] = y
        o.y = s.y.astype(np.complex128).copy()
        s.f_hat_iter[:] = 0
        o.f_hat_iter = np.zeros(NN, complex)
        if flags & nf.PRECOMPUTE_WEIGHT:
            w = r.uniform(0.5, 2.0, M)
            s.w[:] = w
            o.w = s.w.astype(np.float64).copy()
        if flags & nf.PRECOMPUTE_DAMP:
            wh = r.uniform(0.5, 2.0, NN) if w_hat is None else w_hat
            s.w_hat[:] = wh
            o.w_hat = s.w_hat.astype(np.float64).copy()
        if flags & nf.LANDWEBER:
            s.s.alpha_iter = 0.5 / M
            o.alpha_iter = float(s.s.alpha_iter)
        s.before_loop()
        o.before_loop()
        out = []
        for it in range(steps):
            s.loop_one_step()
            o.loop_one_step()
            dr = np.linalg.norm(s.r_iter.astype(np.complex128) - o.r_iter)
            out.append((rel(s.f_hat_iter.astype(np.complex128), o.f_hat_iter),
                        abs(s.s.dot_r_iter - o.dot_r_iter) / abs(o.dot_r_iter),
                        dr / np.linalg.norm(o.r_iter), dr / np.linalg.norm(o.y)))
        return out
    finally:
        s.finalize()
        p.finalize()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("extra", [0, nf.PRECOMPUTE_WEIGHT, nf.PRECOMPUTE_WEIGHT | nf.PRECOMPUTE_DAMP])
@pytest.mark.parametrize("method", [nf.CGNR, nf.STEEPEST_DESCENT, nf.CGNE,
                                    nf.LANDWEBER | nf.NORMS_FOR_LANDWEBER])
def test_solver_vs_oracle(method, extra, precision):
    # CGNE (Craig's method) solves A W_hat A^H u = y: run it on an
    # underdetermined system (300 nodes, 480 unknowns), where it converges;
    # the other methods on an overdetermined one (900 nodes)
    M = 300 if method == nf.CGNE else 900
    res = run_case(method, extra, precision, steps=8, M=M)
    tol = 1e-9 if precision == "double" else 2e-4
    for it, (ef, er, err, _) in enumerate(res):
        assert ef < tol and er < tol and err < tol, (it, ef, er, err)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("method", [nf.CGNR, nf.CGNE])
def test_damped_solver_with_jackson_factors(method, precision):
    """Weighted and damped solve with the library's own damping factors
    (modified Jackson, per-dimension product) against the oracle."""
    M = 300 if method == nf.CGNE else 900
    wh = jackson_damping([24, 20], precision)
    assert wh.min() > 0 and wh.max() <= 4.0
    res = run_case(method, nf.PRECOMPUTE_WEIGHT | nf.PRECOMPUTE_DAMP, precision, steps=8, M=M,
                   w_hat=wh)
    # the Jackson factors span 5e-5 .. 3.8, so the damped normal operator
    # (W_hat enters squared) is ~1e9 times worse conditioned than with the
    # U[0.5, 2] factors above; the FP32 transforms' ~1e-6 relative error
    # grows through the recurrence to 0.5e-3 .. 4e-3 by step 8 (measured on
    # two runs), hence the looser float tolerance; the double family pins
    # the damped recurrences to 1e-9
    tol = 1e-9 if precision == "double" else 1e-2
    for it, (ef, er, err, err_y) in enumerate(res):
        # the FP32 residual, updated by recurrence, carries cancellation
        # error of the size of FP32 rounding of y: measured against |y|
        ok_r = err < tol if precision == "double" else err_y < tol
        assert ef < tol and er < tol and ok_r, (it, ef, er, err, err_y)


@pytest.mark.gpu
def test_device_cgnr_sync_levels():
    """Sync level 1 keeps f_hat_iter and the scalars current on the host."""
    N = [16, 16]
    x = W.uniform(500, 2, seed=71)
    p = nf.Plan(N, 500, m=4, n=[32, 32], flags=FLAGS, precision="double")
    p.x[:] = x
    p.precompute()
    s = nf.Solver(p, nf.CGNR)
    try:
        s.set_sync(1)
        s.y[:] = W.complex_normal(500, seed=72)
        s.f_hat_iter[:] = 0
        s.before_loop()
        d0 = s.s.dot_z_hat_iter
        for _ in range(5):
            s.loop_one_step()
        assert s.s.dot_z_hat_iter < d0
        assert np.linalg.norm(s.f_hat_iter) > 0
    finally:
        s.finalize()
        p.finalize()


@pytest.mark.gpu
def test_cgnr_step_graph_matches_eager(monkeypatch):
    """The device CGNR captures its step as a hipGraph on the first step and
    replays it afterwards; every replay (with the host copies and waits of
    sync_out between them) must equal the eager launches bit for bit.  Both
    solvers run on one plan: its node order (set by the binning's atomics)
    fixes the FP32 summation order of B^T, so separate plans may differ in
    the last bits."""
    N = [64, 64]
    x = W.golden_radial(60, 128)
    M = x.shape[0]
    y = W.complex_normal(M, seed=81)
    w = np.random.default_rng(82).uniform(0.5, 2.0, M)
    p = nf.Plan(N, M, m=4, n=[128, 128], flags=FLAGS, precision="single")
    p.x[:] = x
    p.precompute()
    runs = []
    try:
        for graph in ("0", "1", "0"):
            monkeypatch.setenv("NFFT_HIP_CGNR_GRAPH", graph)
            s = nf.Solver(p, nf.CGNR | nf.PRECOMPUTE_WEIGHT)
            try:
                s.y[:] = y
                s.w[:] = w
                s.f_hat_iter[:] = 0
                s.before_loop()
                steps = []
                for _ in range(6):
                    s.loop_one_step()
                    steps.append((s.f_hat_iter.copy(), s.r_iter.copy(), float(s.s.dot_z_hat_iter),
                                  float(s.s.alpha_iter), float(s.s.beta_iter)))
                runs.append(steps)
            finally:
                s.finalize()
    finally:
        p.finalize()
    for name, other in (("graph", runs[1]), ("eager again", runs[2])):
        for it, (a, b) in enumerate(zip(runs[0], other)):
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), (name, it)
            assert a[2:] == b[2:], (name, it, a[2:], b[2:])
