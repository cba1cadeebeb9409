"""Damping factors nfft[f]_modified_* (reference kernel/util/damp.c:25-142,
include/infft.h:1506-1520) through the C ABI, on the CPU.

Each "modified" factor of index kk is the sum of a kernel's Fourier
coefficients at kk and kk + 1.  Pins fixed by the mathematics, independent
of the closed forms the library evaluates:
  * Fejer: N^2 * factor = F(kk) + F(kk+1), F the coefficients of the Fejer
    kernel (sin(n x/2)/sin(x/2))^2 with n = N/2, normalised to F(0) = 1;
  * Jackson (power 4, n = (N/2+1)/2) and generalised Jackson (power 8,
    n = (N/2+3)/4): the coefficients of those kernels, computed by an FFT of
    the kernel sampled on 4096 points, for every kk in [-N/2, N/2) and N with
    integer n (where the kernel is a trigonometric polynomial);
  * Sobolev and multiquadric: values worked out by hand.
A planted mistake (Jackson coefficients at kk and kk - 1) must fail.
"""
import ctypes

import numpy as np
import pytest

import nfft_amd as nf


def kernel_coeffs(n, power, K=4096):
    x = 2 * np.pi * np.arange(K) / K
    with np.errstate(all="ignore"):
        v = (np.sin(n * x / 2) / np.sin(x / 2)) ** power
    v[0] = float(n) ** power
    c = np.fft.fft(v).real / K
    return c / c[0]


def fn(name, precision):
    L = nf.lib()
    pre = "nfft_" if precision == "double" else "nfftf_"
    f = getattr(L, pre + name)
    R = ctypes.c_double if precision == "double" else ctypes.c_float
    f.restype = R
    if name in ("modified_fejer", "modified_jackson2", "modified_jackson4"):
        f.argtypes = [ctypes.c_ssize_t, ctypes.c_ssize_t]
    elif name == "modified_sobolev":
        f.argtypes = [R, ctypes.c_ssize_t]
    else:
        f.argtypes = [R, R, ctypes.c_ssize_t]
    return f


def modified(c, kk):
    return c[abs(kk)] + c[abs(kk + 1)]


@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("N", [4, 8, 16, 64])
def test_fejer(precision, N):
    f = fn("modified_fejer", precision)
    c = kernel_coeffs(N // 2, 2)
    tol = 1e-12 if precision == "double" else 2e-6
    for kk in range(-N // 2, N // 2):
        assert abs(N * N * f(N, kk) - modified(c, kk)) <= tol, kk


@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("N", [2, 6, 10, 14, 30, 62])
def test_jackson2(precision, N):
    n = (N // 2 + 1) // 2
    assert 2 * n == N // 2 + 1
    f = fn("modified_jackson2", precision)
    c = kernel_coeffs(n, 4)
    tol = 1e-12 if precision == "double" else 2e-6
    for kk in range(-N // 2, N // 2):
        assert abs(f(N, kk) - modified(c, kk)) <= tol, kk


@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("N", [2, 10, 18, 34, 66])
def test_jackson4(precision, N):
    n = (N // 2 + 3) // 4
    assert 4 * n == N // 2 + 3
    f = fn("modified_jackson4", precision)
    c = kernel_coeffs(n, 8)
    tol = 1e-11 if precision == "double" else 5e-5
    for kk in range(-N // 2, N // 2):
        assert abs(f(N, kk) - modified(c, kk)) <= tol, kk


def test_jackson_planted_mistake_fails():
    """Coefficients at kk - 1 and kk instead of kk and kk + 1: off the kernel."""
    N = 14
    n = (N // 2 + 1) // 2
    f = fn("modified_jackson2", "double")
    c = kernel_coeffs(n, 4)
    wrong = [c[abs(kk - 1)] + c[abs(kk)] for kk in range(-N // 2, N // 2)]
    got = [f(N, kk) for kk in range(-N // 2, N // 2)]
    assert max(abs(a - b) for a, b in zip(wrong, got)) > 1e-3


@pytest.mark.parametrize("precision", ["double", "single"])
def test_sobolev_and_multiquadric_by_hand(precision):
    s = fn("modified_sobolev", precision)
    q = fn("modified_multiquadric", precision)
    tol = 1e-14 if precision == "double" else 1e-6
    # |k|^(-2 mu) with 1 at k = 0, summed over kk, kk + 1
    assert abs(s(1.0, 0) - 2.0) <= tol              # 1 + 1
    assert abs(s(1.0, 1) - 1.25) <= tol             # 1 + 1/4
    assert abs(s(0.5, -1) - 2.0) <= tol             # 1/1 + 1
    assert abs(s(1.5, 2) - (1 / 8 + 1 / 27)) <= tol
    assert abs(s(1.0, -3) - (1 / 9 + 1 / 4)) <= tol
    # (k^2 + c^2)^(-mu)
    assert abs(q(1.0, 1.0, 0) - 1.5) <= tol         # 1/1 + 1/2
    assert abs(q(0.5, 3.0, 3) - (1 / np.sqrt(18) + 1 / 5)) <= tol
    assert abs(q(2.0, 0.5, -1) - (1 / 1.25 ** 2 + 1 / 0.25 ** 2)) <= tol * 16
