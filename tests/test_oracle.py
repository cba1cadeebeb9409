"""CPU tests that pin the FP64 oracle before it is trusted as the checker.

Pins, per oracle function:
  bessel_i0        100-entry known-answer table of the reference (tests/bessel.c:29-131)
                   at the reference's bound 4 eps (tests/bessel.c:134-141)
  ndft_*           Mathematica 64-digit golden files (tests/data/nfft_*.txt) at 48 eps
  nfft_*, stages   the same golden files under the reference's fast-transform bound
                   (tests/nfft.c:217-284) for m = 4, 6, 8; adjointness
  fft              numpy.fft and a brute-force DFT
  next_power_of_2  the naive loop of tests/util.c:100-152
  phi (sin branch) the analytic continuation of the sinh branch, evaluated in
                   complex arithmetic (the branch lies below the golden files'
                   resolution)
and a planted-mistake test: deliberately wrong builds of the oracle
(truncated window, missing deconvolution factor, shifted tap window, flipped
FFT sign) must FAIL the golden checks.
"""
import glob
import os
import subprocess

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")
FILES = sorted(glob.glob(os.path.join(GOLD, "nfft", "nfft_*.txt")))


@pytest.fixture(scope="module", autouse=True)
def built():
    subprocess.run(["make", "-s", "-C", os.path.join(O._HERE), "liboracle.so"], check=True)


def test_fixture_count():
    assert len(FILES) == 66


def test_bessel_i0_kat():
    vals = [float(v) for v in open(os.path.join(GOLD, "bessel_i0.txt")).read().split()]
    assert len(vals) == 100
    bound = 4.0 * O.eps_of("double")
    for x, ref in enumerate(vals):
        y = O.bessel_i0(float(x))
        err = abs(y - ref) / abs(ref)
        assert err < bound, (x, y, ref, err)


def test_next_power_of_2_kat():
    for x in range(0, 5000):
        naive = 1
        while naive < x:
            naive <<= 1
        expect = x + 1 if x < 2 else naive
        assert O.next_power_of_2(x) == expect
    assert O.next_power_of_2(-3) == -1


def _golden_cases():
    return [os.path.basename(f) for f in FILES]


@pytest.mark.parametrize("name", _golden_cases())
def test_ndft_golden(name):
    g = O.read_golden(os.path.join(GOLD, "nfft", name))
    N = g["N"]
    bound = O.err_bound_direct("double")
    if "adjoint" in name:
        fh = O.ndft_adjoint(N, g["x"], g["f"])
        err = O.err_adjoint_measure(g["f_hat"], fh, g["f"])
    else:
        f = O.ndft_trafo(N, g["x"], g["f_hat"])
        err = O.err_trafo_measure(g["f"], f, g["f_hat"])
    assert err < bound, err


@pytest.mark.parametrize("m", [4, 6, 8])
@pytest.mark.parametrize("name", _golden_cases())
def test_nfft_golden(name, m):
    g = O.read_golden(os.path.join(GOLD, "nfft", name))
    N = g["N"]
    n = O.default_n(N)
    if O.nfft_check(N, n, g["x"]) is not None:
        pytest.skip("nfft_check rejects this case (the reference counts it as OK)")
    bound = O.err_bound_fast(m, "double")
    if "adjoint" in name:
        fh = O.nfft_adjoint(N, n, m, g["x"], g["f"])
        err = O.err_adjoint_measure(g["f_hat"], fh, g["f"])
    else:
        f = O.nfft_trafo(N, n, m, g["x"], g["f_hat"])
        err = O.err_trafo_measure(g["f"], f, g["f_hat"])
    assert err < bound, err


@pytest.mark.parametrize("n", [[2], [8], [16], [12], [10], [30], [4, 8], [6, 10], [8, 4, 6], [2, 3, 5]])
def test_fft_vs_numpy(n):
    r = np.random.default_rng(7)
    a = r.standard_normal(int(np.prod(n))) + 1j * r.standard_normal(int(np.prod(n)))
    fw = O.fft(n, a, -1)
    bw = O.fft(n, a, +1)
    ref_fw = np.fft.fftn(a.reshape(n)).ravel()
    ref_bw = (np.fft.ifftn(a.reshape(n)) * np.prod(n)).ravel()
    assert np.max(np.abs(fw - ref_fw)) < 1e-12 * np.max(np.abs(ref_fw))
    assert np.max(np.abs(bw - ref_bw)) < 1e-12 * np.max(np.abs(ref_bw))


def test_fft_bruteforce_tiny():
    n = [6]
    a = np.arange(6) + 1j * np.arange(6)[::-1]
    k = np.arange(6)
    ref = np.array([np.sum(a * np.exp(-2j * np.pi * kk * k / 6)) for kk in k])
    assert np.allclose(O.fft(n, a, -1), ref, atol=1e-13)


@pytest.mark.parametrize("N,M,m", [([32], 50, 4), ([16, 12], 60, 5), ([8, 10, 6], 40, 3)])
def test_adjointness(N, M, m):
    r = np.random.default_rng(11)
    x = r.uniform(-0.5, 0.5, size=(M, len(N)))
    n = O.default_n(N)
    fh = r.standard_normal(int(np.prod(N))) + 1j * r.standard_normal(int(np.prod(N)))
    f = r.standard_normal(M) + 1j * r.standard_normal(M)
    Af = O.nfft_trafo(N, n, m, x, fh)
    AHf = O.nfft_adjoint(N, n, m, x, f)
    lhs = np.vdot(f, Af)
    rhs = np.vdot(AHf, fh)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_phi_branches_closed_form():
    # |y| = m gives b/pi exactly; the sin branch is sin(b s)/(pi s)
    n, m = 64, 4
    b = O.kb_b(2.0)
    assert O.phi(n, m / n, m, b) == pytest.approx(b / np.pi, rel=1e-15)
    y = m + 0.5
    s = np.sqrt(y * y - m * m)
    assert O.phi(n, y / n, m, b) == pytest.approx(np.sin(b * s) / (np.pi * s), rel=1e-13)
    # phi_hut(k=0) = I0(m b)
    assert O.phi_hut(n, 0, m, b) == pytest.approx(O.bessel_i0(m * b), rel=1e-15)


def _phi_continuation(y, m, b):
    """sinh(b s) / (pi s), s = sqrt(m^2 - y^2), with complex s: the function
    is entire in s^2 (sinh(b s)/s is even in s), so for |y| > m it continues
    analytically to sin(b t) / (pi t), t = sqrt(y^2 - m^2) -- the second
    branch of the reference's window (include/infft.h:209-215) is fixed by
    the first."""
    s = np.sqrt(np.complex128(m * m - y * y))
    v = np.sinh(b * s) / (np.pi * s)
    assert abs(v.imag) <= 1e-12 * max(1.0, abs(v.real))
    return float(v.real)


@pytest.mark.parametrize("m", [2, 4, 6, 8])
@pytest.mark.parametrize("sigma", [1.5, 2.0, 3.0])
def test_phi_sin_branch_is_analytic_continuation(m, sigma):
    """Pins the window's sin branch (|y| > m, the two outermost taps of every
    node) against the analytic continuation of its sinh branch, evaluated
    in complex arithmetic by numpy -- and shows that a build that truncates
    the window at |y| = m (planted mistake 1) fails this pin."""
    n = 64
    b = O.kb_b(sigma)
    ys = np.concatenate([np.linspace(m + 1e-3, m + 1.0, 41), -np.linspace(m + 1e-3, m + 1.0, 41)])
    ref = np.array([_phi_continuation(y, m, b) for y in ys])
    got = np.array([O.phi(n, y / n, m, b) for y in ys])
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))
    assert np.max(np.abs(ref)) > 0.05  # the branch is not negligible in absolute terms
    # inside the support the same expression is the sinh branch itself
    yi = np.linspace(-m + 1e-3, m - 1e-3, 33)
    gi = np.array([O.phi(n, y / n, m, b) for y in yi])
    ri = np.array([_phi_continuation(y, m, b) for y in yi])
    assert np.max(np.abs(gi - ri) / np.abs(ri)) <= 1e-13
    subprocess.run(["make", "-s", "-C", O._HERE, "liboracle_plant1.so"], check=True)
    O.use_library(os.path.join(O._HERE, "liboracle_plant1.so"))
    try:
        bad = np.array([O.phi(n, y / n, m, b) for y in ys])
    finally:
        O.use_library(None)
    assert np.max(np.abs(bad - ref)) > 1e3 * 1e-13 * np.max(np.abs(ref))


def _golden_pass_rate(libname):
    """Fraction of fast-path golden checks (m = 6, double) a build passes."""
    O.use_library(os.path.join(O._HERE, libname))
    try:
        ok = tot = 0
        for path in FILES:
            g = O.read_golden(path)
            N = g["N"]
            n = O.default_n(N)
            if O.nfft_check(N, n, g["x"]) is not None or min(N) <= 6:
                continue
            bound = O.err_bound_fast(6, "double")
            if "adjoint" in path:
                fh = O.nfft_adjoint(N, n, 6, g["x"], g["f"])
                err = O.err_adjoint_measure(g["f_hat"], fh, g["f"])
            else:
                f = O.nfft_trafo(N, n, 6, g["x"], g["f_hat"])
                err = O.err_trafo_measure(g["f"], f, g["f_hat"])
            tot += 1
            ok += err < bound
        return ok / tot
    finally:
        O.use_library(None)


@pytest.mark.parametrize("plant", [2, 3, 4])
def test_planted_mistakes_are_caught(plant):
    subprocess.run(["make", "-s", "-C", O._HERE, "liboracle_plant%d.so" % plant], check=True)
    assert _golden_pass_rate("liboracle.so") == 1.0
    rate = _golden_pass_rate("liboracle_plant%d.so" % plant)
    assert rate < 0.5, "planted mistake %d passed %.0f%% of golden checks" % (plant, 100 * rate)


def test_window_outer_taps_below_golden_resolution():
    """Documents a limit of the golden files: zeroing the two outermost taps
    (the sin branch of include/infft.h:209-215) changes results by less than
    they resolve (relative ~1e-8 at m=4, ~1e-13 at m=6).  The branch is
    pinned instead by test_phi_sin_branch_is_analytic_continuation."""
    subprocess.run(["make", "-s", "-C", O._HERE, "liboracle_plant1.so"], check=True)
    assert _golden_pass_rate("liboracle_plant1.so") == 1.0


@pytest.mark.parametrize("threads", [1, 4])
def test_prepsi_3d_equals_oracle_b(threads):
    """The reference-shaped 3-D PRE_PSI B / B^T (the CPU baseline's timing
    path) equal the oracle's own B / B^T, edge nodes and wrap included."""
    N, n, m = [8, 10, 12], [16, 20, 24], 3
    rng = np.random.default_rng(7)
    x = rng.uniform(-0.5, 0.5, size=(400, 3))
    x[:5] = -0.5
    x[5:10] = 0.49999
    g = rng.standard_normal(int(np.prod(n))) + 1j * rng.standard_normal(int(np.prod(n)))
    f = rng.standard_normal(400) + 1j * rng.standard_normal(400)
    O.set_threads(threads)
    try:
        P = O.Prepsi3d(N, n, m, x)
        fb = P.B(g)
        gb = P.BT(f)
    finally:
        O.set_threads(1)
    fr = O.stage_B(N, n, m, x, g)
    gr = O.stage_BT(N, n, m, x, f)
    assert np.linalg.norm(fb - fr) / np.linalg.norm(fr) < 1e-13
    assert np.linalg.norm(gb - gr) / np.linalg.norm(gr) < 1e-13
