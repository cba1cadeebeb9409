"""CPU pins of the solver oracle (oracle/solver.py).

CGNR, steepest descent, CGNE and Landweber restated from the reference
solver (kernel/solver/solver.c) are run on a small dense NDFT system built
with numpy; at convergence every method must reach the least-squares
solution of numpy.linalg.lstsq (weighted by W when PRECOMPUTE_WEIGHT is
set; damping W_hat changes the path, not the solution), and the CGNR
recurrences must keep r = y - A f_hat and z = A^H W r exact.
"""
import numpy as np
import pytest

from oracle import solver as S


def dense_ndft(N, x):
    ks = np.stack(np.meshgrid(*[np.arange(n) - n // 2 for n in N], indexing="ij"), -1).reshape(-1, len(N))
    return np.exp(-2j * np.pi * (x @ ks.T))


def setup(flags, seed=0, M=150, noise=0.01):
    r = np.random.default_rng(seed)
    N = [8, 6]
    x = r.uniform(-0.5, 0.5, size=(M, 2))
    A = dense_ndft(N, x)
    f_true = r.standard_normal(48) + 1j * r.standard_normal(48)
    y = A @ f_true + noise * (r.standard_normal(M) + 1j * r.standard_normal(M))
    s = S.Solver(lambda f: A @ f, lambda v: A.conj().T @ v, M, 48, flags)
    s.y = y.copy()
    if flags & S.PRECOMPUTE_WEIGHT:
        s.w = r.uniform(0.5, 2.0, size=M)
    if flags & S.PRECOMPUTE_DAMP:
        s.w_hat = r.uniform(0.5, 2.0, size=48)
    return A, y, s


def lstsq(A, y, w=None):
    if w is not None:
        sw = np.sqrt(w)
        return np.linalg.lstsq(sw[:, None] * A, sw * y, rcond=None)[0]
    return np.linalg.lstsq(A, y, rcond=None)[0]


@pytest.mark.parametrize("extra", [0, S.PRECOMPUTE_WEIGHT, S.PRECOMPUTE_DAMP,
                                   S.PRECOMPUTE_WEIGHT | S.PRECOMPUTE_DAMP])
@pytest.mark.parametrize("method", [S.CGNR, S.STEEPEST_DESCENT])
def test_converges_to_least_squares(method, extra):
    A, y, s = setup(method | extra)
    s.before_loop()
    iters = 60 if method == S.CGNR else 3000
    for _ in range(iters):
        s.loop_one_step()
        if s.dot_z_hat_iter < 1e-26:
            break
    ref = lstsq(A, y, s.w)
    assert np.linalg.norm(s.f_hat_iter - ref) < 1e-6 * np.linalg.norm(ref)


def test_cgne_reaches_minimum_norm_solution():
    """CGNE (Craig) on a consistent underdetermined system: residual to 0 and
    the iterate is the minimum-norm solution."""
    A, y, s = setup(S.CGNE, M=30, noise=0.0)
    s.before_loop()
    for _ in range(400):
        s.loop_one_step()
        if s.dot_r_iter < 1e-24 * np.vdot(y, y).real:
            break
    ref = np.linalg.lstsq(A, y, rcond=None)[0]
    assert np.linalg.norm(y - A @ s.f_hat_iter) < 1e-9 * np.linalg.norm(y)
    assert np.linalg.norm(s.f_hat_iter - ref) < 1e-6 * np.linalg.norm(ref)


def test_landweber_with_small_step():
    A, y, s = setup(S.LANDWEBER | S.NORMS_FOR_LANDWEBER)
    s.alpha_iter = 1.0 / np.linalg.norm(A, 2) ** 2
    s.before_loop()
    prev = s.dot_r_iter
    for _ in range(200):
        s.loop_one_step()
        assert s.dot_r_iter <= prev * (1 + 1e-12)
        prev = s.dot_r_iter


def test_cgnr_invariants():
    A, y, s = setup(S.CGNR | S.PRECOMPUTE_WEIGHT)
    s.before_loop()
    for _ in range(10):
        s.loop_one_step()
        assert np.allclose(s.r_iter, y - A @ s.f_hat_iter, atol=1e-10)
        assert np.allclose(s.z_hat_iter, A.conj().T @ (s.w * s.r_iter), atol=1e-9)
