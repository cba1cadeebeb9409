"""Iterative inverse NFFT oracle -- TEST INFRASTRUCTURE ONLY.

numpy restatement of the reference's solver recurrences
(kernel/solver/solver.c): before_loop 81-125, Landweber 128-174, steepest
descent 177-229, CGNR 232-292, CGNE 295-344, on complex128 vectors, with the
operator supplied as two callables A(f_hat) -> f and AH(f) -> f_hat.  Flags
as in include/nfft3.h:823-829.

Pins: CGNR run to convergence on a small dense NDFT system equals the
(weighted) least-squares solution of numpy.linalg.lstsq (tests/test_oracle_solver.py).
"""
import numpy as np

LANDWEBER = 1 << 0
STEEPEST_DESCENT = 1 << 1
CGNR = 1 << 2
CGNE = 1 << 3
NORMS_FOR_LANDWEBER = 1 << 4
PRECOMPUTE_WEIGHT = 1 << 5
PRECOMPUTE_DAMP = 1 << 6


def _dot(x, w=None):
    a = (x.conj() * x).real
    return float(np.sum(a if w is None else w * a))


class Solver:
    """State of solver_plan_complex (include/nfft3.h:758-780)."""

    def __init__(self, A, AH, M, N, flags=CGNR):
        self.A, self.AH = A, AH
        self.flags = flags
        self.y = np.zeros(M, complex)
        self.f_hat_iter = np.zeros(N, complex)
        self.w = np.ones(M) if flags & PRECOMPUTE_WEIGHT else None
        self.w_hat = np.ones(N) if flags & PRECOMPUTE_DAMP else None
        self.alpha_iter = 0.0
        self.beta_iter = 0.0
        self.dot_r_iter = self.dot_r_iter_old = 0.0
        self.dot_z_hat_iter = self.dot_z_hat_iter_old = 0.0
        self.dot_p_hat_iter = self.dot_v_iter = 0.0

    def _W(self, x):
        return x if self.w is None else self.w * x

    def _Wh(self, x):
        return x if self.w_hat is None else self.w_hat * x

    def before_loop(self):
        fl = self.flags
        self.r_iter = self.y - self.A(self.f_hat_iter)
        norms = (not fl & LANDWEBER) or (fl & NORMS_FOR_LANDWEBER)
        if norms:
            self.dot_r_iter = _dot(self.r_iter, self.w)
        self.z_hat_iter = self.AH(self._W(self.r_iter))
        if norms:
            self.dot_z_hat_iter = _dot(self.z_hat_iter, self.w_hat)
        if fl & CGNE:
            self.dot_p_hat_iter = self.dot_z_hat_iter
        self.p_hat_iter = self.z_hat_iter.copy()

    def loop_one_step(self):
        fl = self.flags
        if fl & LANDWEBER:
            self.f_hat_iter = self.f_hat_iter + self.alpha_iter * self._Wh(self.z_hat_iter)
            self.r_iter = self.y - self.A(self.f_hat_iter)
            if fl & NORMS_FOR_LANDWEBER:
                self.dot_r_iter = _dot(self.r_iter, self.w)
            self.z_hat_iter = self.AH(self._W(self.r_iter))
            if fl & NORMS_FOR_LANDWEBER:
                self.dot_z_hat_iter = _dot(self.z_hat_iter, self.w_hat)
        if fl & STEEPEST_DESCENT:
            self.v_iter = self.A(self._Wh(self.z_hat_iter))
            self.dot_v_iter = _dot(self.v_iter, self.w)
            self.alpha_iter = self.dot_z_hat_iter / self.dot_v_iter
            self.f_hat_iter = self.f_hat_iter + self.alpha_iter * self._Wh(self.z_hat_iter)
            self.r_iter = self.r_iter - self.alpha_iter * self.v_iter
            self.dot_r_iter = _dot(self.r_iter, self.w)
            self.z_hat_iter = self.AH(self._W(self.r_iter))
            self.dot_z_hat_iter = _dot(self.z_hat_iter, self.w_hat)
        if fl & CGNR:
            self.v_iter = self.A(self._Wh(self.p_hat_iter))
            self.dot_v_iter = _dot(self.v_iter, self.w)
            self.alpha_iter = self.dot_z_hat_iter / self.dot_v_iter
            self.f_hat_iter = self.f_hat_iter + self.alpha_iter * self._Wh(self.p_hat_iter)
            self.r_iter = self.r_iter - self.alpha_iter * self.v_iter
            self.dot_r_iter = _dot(self.r_iter, self.w)
            self.z_hat_iter = self.AH(self._W(self.r_iter))
            self.dot_z_hat_iter_old = self.dot_z_hat_iter
            self.dot_z_hat_iter = _dot(self.z_hat_iter, self.w_hat)
            self.beta_iter = self.dot_z_hat_iter / self.dot_z_hat_iter_old
            self.p_hat_iter = self.beta_iter * self.p_hat_iter + self.z_hat_iter
        if fl & CGNE:
            self.alpha_iter = self.dot_r_iter / self.dot_p_hat_iter
            self.f_hat_iter = self.f_hat_iter + self.alpha_iter * self._Wh(self.p_hat_iter)
            f = self.A(self._Wh(self.p_hat_iter))
            self.r_iter = self.r_iter - self.alpha_iter * f
            self.dot_r_iter_old = self.dot_r_iter
            self.dot_r_iter = _dot(self.r_iter, self.w)
            self.beta_iter = self.dot_r_iter / self.dot_r_iter_old
            self.p_hat_iter = self.beta_iter * self.p_hat_iter + self.AH(self._W(self.r_iter))
            self.dot_p_hat_iter = _dot(self.p_hat_iter, self.w_hat)
