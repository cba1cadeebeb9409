"""The real-valued solver family solver[f]_*_double through the C ABI
(reference include/nfft3.h:786-819, kernel/solver/solver.c:392-714).

Its mv plan is the caller's own: here a ctypes struct with the reference's mv
prefix (N_total, M_total, f_hat, f, mv_trafo, mv_adjoint) whose callbacks
apply a dense real cosine matrix A[j, k] = cos(2 pi k x_j) (the shape of a
real nonequispaced cosine transform) and its transpose, reading and writing
whatever f_hat / f the solver has swapped in.  Every method, with and without
weights W and damping W_hat, is compared step by step with the solver oracle
(oracle/solver.py, pinned against numpy.linalg.lstsq) on the same system:
iterates, residuals and the scalars alpha, beta and the dots.  The solver
runs on the host around the callbacks, so no GPU is needed.
"""
import ctypes

import numpy as np
import pytest

import nfft_amd as nf
from oracle import solver as S

pytestmark = pytest.mark.filterwarnings("ignore::RuntimeWarning")


class RealOperator:
    """A ctypes mv_plan_double with numpy callbacks for A and A^T."""

    def __init__(self, A, precision):
        self.A = A
        self.R = ctypes.c_double if precision == "double" else ctypes.c_float
        self.dt = np.float64 if precision == "double" else np.float32
        T = nf.nfft_mv_plan_double if precision == "double" else nf.nfftf_mv_plan_double
        M, N = A.shape
        self.f_hat = np.zeros(N, self.dt)
        self.f = np.zeros(M, self.dt)
        self.mv = T()
        self.mv.M_total, self.mv.N_total = M, N
        P = ctypes.POINTER(self.R)
        self.mv.f_hat = self.f_hat.ctypes.data_as(P)
        self.mv.f = self.f.ctypes.data_as(P)
        self._t = nf._FN(self._trafo)
        self._a = nf._FN(self._adjoint)
        self.mv.mv_trafo = self._t
        self.mv.mv_adjoint = self._a
        self.calls = 0

    def _arr(self, ptr, n):
        return np.ctypeslib.as_array(ptr, shape=(n,))

    def _trafo(self, _p):
        M, N = self.A.shape
        fh = self._arr(self.mv.f_hat, N).astype(np.float64)
        self._arr(self.mv.f, M)[:] = (self.A @ fh).astype(self.dt)
        self.calls += 1

    def _adjoint(self, _p):
        M, N = self.A.shape
        f = self._arr(self.mv.f, M).astype(np.float64)
        self._arr(self.mv.f_hat, N)[:] = (self.A.T @ f).astype(self.dt)
        self.calls += 1


def system(seed, M, N):
    r = np.random.default_rng(seed)
    x = r.uniform(0.0, 0.5, size=M)
    A = np.cos(2 * np.pi * np.outer(x, np.arange(N)))
    f_true = r.standard_normal(N)
    y = A @ f_true + 0.01 * r.standard_normal(M)
    return A, y, r.uniform(0.5, 2.0, size=M), r.uniform(0.5, 2.0, size=N)


METHODS = [S.CGNR, S.CGNE, S.STEEPEST_DESCENT, S.LANDWEBER | S.NORMS_FOR_LANDWEBER, S.LANDWEBER]


@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("extra", [0, S.PRECOMPUTE_WEIGHT, S.PRECOMPUTE_DAMP,
                                   S.PRECOMPUTE_WEIGHT | S.PRECOMPUTE_DAMP])
@pytest.mark.parametrize("method", METHODS)
def test_real_solver_matches_oracle(method, extra, precision):
    flags = method | extra
    M, N = (24, 40) if method == S.CGNE else (60, 16)  # CGNE: underdetermined
    A, y, w, w_hat = system(7 + method, M, N)
    op = RealOperator(A, precision)
    L = nf.lib()
    pre = "solver_" if precision == "double" else "solverf_"
    Sp = nf.solver_plan_double if precision == "double" else nf.solverf_plan_double
    s = Sp()
    sp = ctypes.pointer(s)
    fn = {k: getattr(L, pre + k + "_double") for k in
          ("init_advanced", "before_loop", "loop_one_step", "finalize")}
    fn["init_advanced"].argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint]
    for k in ("before_loop", "loop_one_step", "finalize"):
        fn[k].argtypes = [ctypes.c_void_p]
    fn["init_advanced"](ctypes.cast(sp, ctypes.c_void_p), ctypes.cast(ctypes.pointer(op.mv), ctypes.c_void_p), flags)
    dt = op.dt
    view = lambda name, n: np.ctypeslib.as_array(getattr(s, name), shape=(n,))
    try:
        view("y", M)[:] = y.astype(dt)
        view("f_hat_iter", N)[:] = 0
        alpha = 1.0 / np.linalg.norm(A, 2) ** 2 * (0.5 if extra else 1.0)
        if flags & S.PRECOMPUTE_WEIGHT:
            view("w", M)[:] = w.astype(dt)
        if flags & S.PRECOMPUTE_DAMP:
            view("w_hat", N)[:] = w_hat.astype(dt)
        o = S.Solver(lambda v: A @ v, lambda v: A.T @ v, M, N, flags)
        o.y = y.astype(dt).astype(np.float64)
        o.f_hat_iter = np.zeros(N)
        if flags & S.PRECOMPUTE_WEIGHT:
            o.w = w.astype(dt).astype(np.float64)
        if flags & S.PRECOMPUTE_DAMP:
            o.w_hat = w_hat.astype(dt).astype(np.float64)
        if method & S.LANDWEBER:
            s.alpha_iter = alpha
            o.alpha_iter = float(dt(alpha))
        fn["before_loop"](ctypes.cast(sp, ctypes.c_void_p))
        o.before_loop()
        tol = 1e-10 if precision == "double" else 2e-4
        steps = 8
        for k in range(steps + 1):
            if k:
                fn["loop_one_step"](ctypes.cast(sp, ctypes.c_void_p))
                o.loop_one_step()
            fi = view("f_hat_iter", N).astype(np.float64)
            ri = view("r_iter", M).astype(np.float64)
            assert np.linalg.norm(fi - o.f_hat_iter) <= tol * max(1.0, np.linalg.norm(o.f_hat_iter)), k
            assert np.linalg.norm(ri - o.r_iter) <= tol * np.linalg.norm(o.r_iter), k
            if not (method & S.LANDWEBER) or (flags & S.NORMS_FOR_LANDWEBER):
                assert abs(s.dot_r_iter - o.dot_r_iter) <= tol * o.dot_r_iter, k
                if method != S.CGNE:
                    assert abs(s.dot_z_hat_iter - o.dot_z_hat_iter) <= tol * o.dot_z_hat_iter, k
            if k and method in (S.CGNR, S.CGNE):
                assert abs(s.alpha_iter - o.alpha_iter) <= tol * abs(o.alpha_iter), k
                assert abs(s.beta_iter - o.beta_iter) <= tol * abs(o.beta_iter), k
        # the caller's mv arrays are where they were (pointer swaps undone)
        assert ctypes.addressof(op.mv.f.contents) == op.f.ctypes.data
        assert ctypes.addressof(op.mv.f_hat.contents) == op.f_hat.ctypes.data
    finally:
        fn["finalize"](ctypes.cast(sp, ctypes.c_void_p))


def test_real_cgnr_converges_to_least_squares():
    """CGNR on the real system reaches numpy's least-squares solution."""
    A, y, _, _ = system(3, 80, 12)
    op = RealOperator(A, "double")
    L = nf.lib()
    s = nf.solver_plan_double()
    sp = ctypes.cast(ctypes.pointer(s), ctypes.c_void_p)
    L.solver_init_double.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    L.solver_init_double(sp, ctypes.cast(ctypes.pointer(op.mv), ctypes.c_void_p))
    try:
        assert s.flags == S.CGNR
        np.ctypeslib.as_array(s.y, shape=(80,))[:] = y
        np.ctypeslib.as_array(s.f_hat_iter, shape=(12,))[:] = 0
        L.solver_before_loop_double.argtypes = [ctypes.c_void_p]
        L.solver_loop_one_step_double.argtypes = [ctypes.c_void_p]
        L.solver_before_loop_double(sp)
        for _ in range(40):
            L.solver_loop_one_step_double(sp)
            if s.dot_z_hat_iter < 1e-28:
                break
        ref = np.linalg.lstsq(A, y, rcond=None)[0]
        got = np.ctypeslib.as_array(s.f_hat_iter, shape=(12,))
        assert np.linalg.norm(got - ref) <= 1e-8 * np.linalg.norm(ref)
    finally:
        L.solver_finalize_double.argtypes = [ctypes.c_void_p]
        L.solver_finalize_double(sp)
