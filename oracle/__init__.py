"""CPU FP64 oracle of the NFFT hot path -- TEST INFRASTRUCTURE ONLY.

Thin numpy wrappers over ``liboracle.so`` (built from ``nfft_oracle.c`` by
``make -C oracle``) plus the reference's error measures and bounds.  Only the
test suite, the smoke check and the CPU-baseline leg of the benchmark import
this package; the product library never does.

Complex vectors are numpy complex128 arrays (interleaved re/im in memory, the
layout of ``fftw_complex``).  Nodes are float64 arrays of shape (M, d).
"""
import ctypes
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


_LIB_PATH = None


def use_library(path):
    """Switch to another build of the oracle (the planted-mistake variants of
    the self-test); None restores liboracle.so."""
    global _LIB, _LIB_PATH
    _LIB = None
    _LIB_PATH = path


def lib():
    global _LIB
    if _LIB is None:
        path = _LIB_PATH or os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        lp = ctypes.POINTER(ctypes.c_long)
        L.oracle_bessel_i0.restype = ctypes.c_double
        L.oracle_bessel_i0.argtypes = [ctypes.c_double]
        L.oracle_phi.restype = ctypes.c_double
        L.oracle_phi.argtypes = [ctypes.c_long, ctypes.c_double, ctypes.c_long, ctypes.c_double]
        L.oracle_phi_hut.restype = ctypes.c_double
        L.oracle_phi_hut.argtypes = [ctypes.c_long, ctypes.c_long, ctypes.c_long, ctypes.c_double]
        L.oracle_kb_b.restype = ctypes.c_double
        L.oracle_kb_b.argtypes = [ctypes.c_double]
        L.oracle_next_power_of_2.restype = ctypes.c_long
        L.oracle_next_power_of_2.argtypes = [ctypes.c_long]
        L.oracle_c_phi_inv.argtypes = [ctypes.c_int, lp, lp, ctypes.c_long, dp]
        for name in ("oracle_ndft_trafo", "oracle_ndft_adjoint"):
            getattr(L, name).argtypes = [ctypes.c_int, lp, ctypes.c_long, dp, dp, dp]
        L.oracle_fft.argtypes = [ctypes.c_int, lp, dp, ctypes.c_int]
        L.oracle_set_threads.argtypes = [ctypes.c_int]
        L.oracle_get_threads.restype = ctypes.c_int
        for name in ("oracle_D", "oracle_DT"):
            getattr(L, name).argtypes = [ctypes.c_int, lp, lp, ctypes.c_long, dp, dp]
        for name in ("oracle_B", "oracle_BT", "oracle_nfft_trafo", "oracle_nfft_adjoint"):
            getattr(L, name).argtypes = [ctypes.c_int, lp, lp, ctypes.c_long, ctypes.c_long, dp, dp, dp]
        L.oracle_prepsi_3d.argtypes = [lp, lp, ctypes.c_long, ctypes.c_long, dp, dp, lp]
        for name in ("oracle_B_3d_prepsi", "oracle_BT_3d_prepsi"):
            getattr(L, name).argtypes = [lp, ctypes.c_long, ctypes.c_long, dp, lp, dp, dp]
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _lp(v):
    arr = (ctypes.c_long * len(v))(*[int(t) for t in v])
    return arr


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def _x(x, d):
    x = np.ascontiguousarray(x, dtype=np.float64)
    return x.reshape(-1, d)


# ---------------------------------------------------------------- scalars

def bessel_i0(x):
    return lib().oracle_bessel_i0(float(x))


def set_threads(n):
    """Threads of the oracle's FFT and B / B^T loops (1 by default)."""
    lib().oracle_set_threads(int(n))


def phi(n, x, m, b):
    return lib().oracle_phi(int(n), float(x), int(m), float(b))


def phi_hut(n, k, m, b):
    return lib().oracle_phi_hut(int(n), int(k), int(m), float(b))


def kb_b(sigma):
    return lib().oracle_kb_b(float(sigma))


def next_power_of_2(x):
    return lib().oracle_next_power_of_2(int(x))


def default_n(N):
    """n_t = 2 next_power_of_2(N_t) (kernel/nfft/nfft.c:6063-6066)."""
    return [2 * next_power_of_2(t) for t in N]


def c_phi_inv(N, n, m):
    out = np.empty(int(sum(N)), dtype=np.float64)
    lib().oracle_c_phi_inv(len(N), _lp(N), _lp(n), int(m), _dp(out))
    return out


# ---------------------------------------------------------------- transforms

def ndft_trafo(N, x, f_hat):
    d = len(N)
    x = _x(x, d)
    f_hat = _c128(f_hat).ravel()
    f = np.empty(x.shape[0], dtype=np.complex128)
    lib().oracle_ndft_trafo(d, _lp(N), x.shape[0], _dp(x), _dp(f_hat), _dp(f))
    return f


def ndft_adjoint(N, x, f):
    d = len(N)
    x = _x(x, d)
    f = _c128(f).ravel()
    f_hat = np.empty(int(np.prod(N)), dtype=np.complex128)
    lib().oracle_ndft_adjoint(d, _lp(N), x.shape[0], _dp(x), _dp(f), _dp(f_hat))
    return f_hat


def fft(n, g, sign):
    """Unnormalised d-D DFT of a row-major grid; sign -1 forward, +1 backward."""
    g = _c128(g).ravel().copy()
    lib().oracle_fft(len(n), _lp(n), _dp(g), int(sign))
    return g


def stage_D(N, n, m, f_hat):
    g = np.empty(int(np.prod(n)), dtype=np.complex128)
    lib().oracle_D(len(N), _lp(N), _lp(n), int(m), _dp(_c128(f_hat).ravel()), _dp(g))
    return g


def stage_DT(N, n, m, g_hat):
    f_hat = np.empty(int(np.prod(N)), dtype=np.complex128)
    lib().oracle_DT(len(N), _lp(N), _lp(n), int(m), _dp(_c128(g_hat).ravel()), _dp(f_hat))
    return f_hat


def stage_B(N, n, m, x, g):
    d = len(N)
    x = _x(x, d)
    g = _c128(g).ravel()
    f = np.empty(x.shape[0], dtype=np.complex128)
    lib().oracle_B(d, _lp(N), _lp(n), int(m), x.shape[0], _dp(x), _dp(g), _dp(f))
    return f


def stage_BT(N, n, m, x, f):
    d = len(N)
    x = _x(x, d)
    f = _c128(f).ravel()
    g = np.empty(int(np.prod(n)), dtype=np.complex128)
    lib().oracle_BT(d, _lp(N), _lp(n), int(m), x.shape[0], _dp(x), _dp(f), _dp(g))
    return g


class Prepsi3d:
    """3-D B / B^T in the shape of the reference's PRE_PSI path (window values
    precomputed per node, tensor loops, blockwise adjoint), for timing the
    CPU baseline; equal to stage_B / stage_BT to rounding."""

    def __init__(self, N, n, m, x):
        self.N, self.n, self.m = list(N), list(n), int(m)
        self.x = _x(x, 3)
        M = self.x.shape[0]
        w = 2 * self.m + 2
        self.psi = np.empty(M * 3 * w)
        self.u = np.empty(M * 3, dtype=np.int64)
        lib().oracle_prepsi_3d(_lp(N), _lp(n), self.m, M, _dp(self.x), _dp(self.psi),
                               self.u.ctypes.data_as(ctypes.POINTER(ctypes.c_long)))

    def B(self, g):
        g = _c128(g).ravel()
        f = np.empty(self.x.shape[0], dtype=np.complex128)
        lib().oracle_B_3d_prepsi(_lp(self.n), self.m, self.x.shape[0], _dp(self.psi),
                                 self.u.ctypes.data_as(ctypes.POINTER(ctypes.c_long)), _dp(g), _dp(f))
        return f

    def BT(self, f):
        f = _c128(f).ravel()
        g = np.empty(int(np.prod(self.n)), dtype=np.complex128)
        lib().oracle_BT_3d_prepsi(_lp(self.n), self.m, self.x.shape[0], _dp(self.psi),
                                  self.u.ctypes.data_as(ctypes.POINTER(ctypes.c_long)), _dp(f), _dp(g))
        return g


def nfft_trafo(N, n, m, x, f_hat):
    d = len(N)
    x = _x(x, d)
    f_hat = _c128(f_hat).ravel()
    f = np.empty(x.shape[0], dtype=np.complex128)
    lib().oracle_nfft_trafo(d, _lp(N), _lp(n), int(m), x.shape[0], _dp(x), _dp(f_hat), _dp(f))
    return f


def nfft_adjoint(N, n, m, x, f):
    d = len(N)
    x = _x(x, d)
    f = _c128(f).ravel()
    f_hat = np.empty(int(np.prod(N)), dtype=np.complex128)
    lib().oracle_nfft_adjoint(d, _lp(N), _lp(n), int(m), x.shape[0], _dp(x), _dp(f), _dp(f_hat))
    return f_hat


# ---------------------------------------------------------------- error measures

def err_trafo_measure(f_ref, f, f_hat):
    """||f - f_ref||_inf / ||f_hat||_1  (tests/nfft.c:739-759)."""
    num = np.max(np.abs(np.asarray(f_ref) - np.asarray(f))) if len(f) else 0.0
    den = np.sum(np.abs(f_hat))
    return 0.0 if num == 0.0 else float(num / den)


def err_adjoint_measure(f_hat_ref, f_hat, f):
    """||f_hat - f_hat_ref||_inf / ||f||_1  (tests/nfft.c:761-781)."""
    num = np.max(np.abs(np.asarray(f_hat_ref) - np.asarray(f_hat)))
    den = np.sum(np.abs(f))
    return 0.0 if num == 0.0 else float(num / den)


def eps_of(precision):
    """2^(1-MANT_DIG) (kernel/util/float.c:37-41)."""
    return 2.0 ** (1 - (53 if precision == "double" else 24))


def err_bound_direct(precision):
    """48 eps (tests/nfft.c:211-215)."""
    return 48.0 * eps_of(precision)


def err_bound_fast(m, precision):
    """Kaiser-Bessel bound max(a err_KB, b eps, 48 eps) (tests/nfft.c:217-284).

    err_KB = pi (sqrt(m) + m) (1/2)^(1/4) exp(-2 pi m sqrt(1/2)); (a, b) =
    (0.3, 2100) in double and (0.4, 2000) in single.
    """
    eps = eps_of(precision)
    a, b = (0.3, 2100.0) if precision == "double" else (0.4, 2000.0)
    err = math.pi * (math.sqrt(m) + m) * math.sqrt(math.sqrt(0.5)) * math.exp(-2 * math.pi * m * math.sqrt(0.5))
    return max(a * err, b * eps, 48.0 * eps)


# ---------------------------------------------------------------- golden files

def read_golden(path):
    """Parse a tests/data/nfft_*.txt file (format: tests/nfft.c:379-448).

    Returns dict(d, N, M, x (M, d), f_hat, f).
    """
    toks = open(path).read().split()
    pos = 0
    d = int(toks[pos]); pos += 1
    N = [int(toks[pos + i]) for i in range(d)]; pos += d
    M = int(toks[pos]); pos += 1
    x = np.array([float(t) for t in toks[pos:pos + M * d]], dtype=np.float64).reshape(M, d)
    pos += M * d
    NN = int(np.prod(N))
    v = np.array([float(t) for t in toks[pos:pos + 2 * NN]], dtype=np.float64)
    f_hat = v[0::2] + 1j * v[1::2]
    pos += 2 * NN
    v = np.array([float(t) for t in toks[pos:pos + 2 * M]], dtype=np.float64)
    f = v[0::2] + 1j * v[1::2]
    return dict(d=d, N=N, M=M, x=x, f_hat=f_hat, f=f)


def nfft_check(N, n, x):
    """Validation strings of nfft_check (kernel/nfft/nfft.c:6169-6207), for the
    geometry part (x range, oversampling, even N); None if fine."""
    x = np.asarray(x)
    if np.any(x < -0.5) or np.any(x >= 0.5):
        return "ths->x out of range [-0.5,0.5)"
    for Nt, nt in zip(N, n):
        if nt / Nt <= 1:
            return "Oversampling factor too small"
        if Nt % 2 == 1:
            return "polynomial degree N has to be even"
    return None
