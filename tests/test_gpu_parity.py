"""GPU library vs the FP64 CPU oracle at workload-shaped sizes, through the C ABI.

Every case states relative l2 and l-infinity errors against the oracle run on
the same inputs (for the float family, the inputs rounded to float first, so
only arithmetic error is measured).  Tolerances:
  double family: rel l2 <= 1e-12 vs the oracle's identical algorithm
  float family:  rel l2 <= 1e-5 (target of the FP32 path) and the reference's
                 own fast-transform bound under its error measure
Full-size cases (the 3-D kooshball at M = 2^24) are checked through
size-independent properties: adjointness, linearity, and float vs double
agreement.
"""
import numpy as np
import pytest

import nfft_amd as nf
import oracle as O
from nfft_amd import workloads as W

FLAGS = nf.PRE_PHI_HUT | nf.PRE_PSI | nf.MALLOC_X | nf.MALLOC_F_HAT | nf.MALLOC_F | nf.FFTW_INIT | nf.FFT_OUT_OF_PLACE


def rel_errs(a, ref):
    a = np.asarray(a, dtype=np.complex128)
    ref = np.asarray(ref, dtype=np.complex128)
    l2 = np.linalg.norm(a - ref) / np.linalg.norm(ref)
    linf = np.max(np.abs(a - ref)) / np.max(np.abs(ref))
    return l2, linf


def make_plan(N, x, m, precision, n=None):
    n = n or [2 * O.next_power_of_2(t) for t in N]
    p = nf.Plan(N, x.shape[0], n=n, m=m, flags=FLAGS, precision=precision)
    p.x[:] = x
    p.precompute()
    return p


def run_pair(N, x, fh, f, m, precision, n=None):
    """GPU trafo and adjoint; oracle on the same (rounded) inputs."""
    p = make_plan(N, x, m, precision, n)
    try:
        xr = p.x.astype(np.float64).copy()
        p.f_hat[:] = fh
        fhr = p.f_hat.astype(np.complex128).copy()
        p.trafo()
        f_gpu = p.f.astype(np.complex128).copy()
        p.f[:] = f
        fr = p.f.astype(np.complex128).copy()
        p.adjoint()
        fh_gpu = p.f_hat.astype(np.complex128).copy()
        nn = p.n
    finally:
        p.finalize()
    f_ref = O.nfft_trafo(N, nn, m, xr, fhr)
    fh_ref = O.nfft_adjoint(N, nn, m, xr, fr)
    return (f_gpu, f_ref, fhr), (fh_gpu, fh_ref, fr)


CASES = [
    # name, N, node generator, m
    ("1d_cfg1", [1024], lambda: W.uniform(4096, 1, seed=11), 6),
    ("1d_small", [64], lambda: W.uniform(300, 1, seed=12), 4),
    ("2d_clustered_m4", [64, 64], lambda: W.clustered_2d(8192, seed=13), 4),
    ("2d_clustered_m6", [64, 48], lambda: W.clustered_2d(6000, seed=14), 6),
    ("2d_radial", [64, 64], lambda: W.golden_radial(40, 128), 4),
    ("3d_kooshball_m4", [16, 16, 16], lambda: W.kooshball(64, 64, seed=15), 4),
    ("3d_uniform_m3", [12, 16, 20], lambda: W.uniform(2000, 3, seed=16), 3),
    ("3d_m2", [8, 8, 8], lambda: W.uniform(500, 3, seed=17), 2),
]


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_vs_oracle(case, precision):
    name, N, gen, m = case
    x = gen()
    NN = int(np.prod(N))
    fh = W.complex_normal(NN, seed=21)
    f = W.complex_normal(x.shape[0], seed=22)
    (fg, fr, fhr), (hg, hr, frr) = run_pair(N, x, fh, f, m, precision)
    l2t, lit = rel_errs(fg, fr)
    l2a, lia = rel_errs(hg, hr)
    print("%s %s trafo l2 %.2e linf %.2e | adjoint l2 %.2e linf %.2e" % (name, precision, l2t, lit, l2a, lia))
    if precision == "double":
        assert l2t < 1e-12 and l2a < 1e-12
    else:
        assert l2t < 1e-5 and l2a < 1e-5
        bound = O.err_bound_fast(m, "single")
        assert O.err_trafo_measure(fr, fg, fhr) < bound
        assert O.err_adjoint_measure(hr, hg, frr) < bound


@pytest.mark.gpu
@pytest.mark.parametrize("variant", ["lane", "wave", "priv", "rows", "runs2sbc", "runs2rs", "runs2d"])
@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("case", [c for c in CASES if len(c[1]) >= 2], ids=[c[0] for c in CASES if len(c[1]) >= 2])
def test_b_kernel_variants(case, precision, variant, monkeypatch):
    """Both shapes of the B / B^T kernels (lane per node, wave per node)."""
    monkeypatch.setenv("NFFT_HIP_INTERP", variant)
    monkeypatch.setenv("NFFT_HIP_SPREAD", variant)
    name, N, gen, m = case
    x = gen()
    fh = W.complex_normal(int(np.prod(N)), seed=23)
    f = W.complex_normal(x.shape[0], seed=24)
    (fg, fr, _), (hg, hr, _) = run_pair(N, x, fh, f, m, precision)
    tol = 1e-12 if precision == "double" else 1e-5
    assert rel_errs(fg, fr)[0] < tol and rel_errs(hg, hr)[0] < tol


@pytest.mark.gpu
def test_cfg1_vs_ndft():
    """1-D N=1024, M=4096, sigma=2, m=6, FP64 against the direct NDFT under the
    reference's bound (tests/nfft.c:217-284)."""
    N = [1024]
    x = W.uniform(4096, 1, seed=11)
    fh = W.complex_normal(1024, seed=21)
    f = W.complex_normal(4096, seed=22)
    (fg, _, fhr), (hg, _, frr) = run_pair(N, x, fh, f, 6, "double")
    f_nd = O.ndft_trafo(N, x, fh)
    fh_nd = O.ndft_adjoint(N, x, f)
    bound = O.err_bound_fast(6, "double")
    assert O.err_trafo_measure(f_nd, fg, fh) < bound
    assert O.err_adjoint_measure(fh_nd, hg, f) < bound


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["double", "single"])
def test_adjointness_and_direct(precision):
    N = [32, 24]
    x = W.clustered_2d(3000, seed=31)
    p = make_plan(N, x, 4, precision)
    try:
        fh = W.complex_normal(p.p.N_total, seed=32)
        f = W.complex_normal(p.M, seed=33)
        p.f_hat[:] = fh
        p.trafo()
        Af = p.f.astype(np.complex128).copy()
        p.trafo_direct()
        Af_direct = p.f.astype(np.complex128).copy()
        p.f[:] = f
        p.adjoint()
        AHf = p.f_hat.astype(np.complex128).copy()
        fhr = np.asarray(fh, dtype=p.cdtype).astype(np.complex128)
        frr = np.asarray(f, dtype=p.cdtype).astype(np.complex128)
        lhs = np.vdot(frr, Af)
        rhs = np.vdot(AHf, fhr)
        tol = 1e-12 if precision == "double" else 2e-6
        assert abs(lhs - rhs) <= tol * abs(lhs)
        # fast vs direct under the reference bound
        assert O.err_trafo_measure(Af_direct, Af, fhr) < O.err_bound_fast(4, precision)
    finally:
        p.finalize()


@pytest.mark.gpu
def test_empty_and_edge_sizes():
    # M = 0: trafo is a no-op, adjoint yields zeros
    p = nf.Plan([16, 16], 0, n=[32, 32], m=4, flags=FLAGS, precision="single")
    try:
        p.precompute()
        p.f_hat[:] = 1.0
        p.trafo()
        p.adjoint()
        assert np.all(p.f_hat == 0)
    finally:
        p.finalize()
    # all nodes in one cell (worst bin imbalance) and nodes at the -0.5 edge
    x = np.zeros((5000, 2))
    x[::2] = -0.5
    fh = W.complex_normal(16 * 16, seed=41)
    f = W.complex_normal(5000, seed=42)
    (fg, fr, _), (hg, hr, _) = run_pair([16, 16], x, fh, f, 4, "double")
    assert rel_errs(fg, fr)[0] < 1e-12 and rel_errs(hg, hr)[0] < 1e-12
    # direct fallback when N <= m
    (fg, fr, _), (hg, hr, _) = run_pair([4, 6], W.uniform(50, 2, seed=43), W.complex_normal(24, seed=44),
                                        W.complex_normal(50, seed=45), 6, "double")
    assert rel_errs(fg, fr)[0] < 1e-12 and rel_errs(hg, hr)[0] < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("variant,d,bins", [("runs2sbc", 3, None), ("runs2rs", 3, None),
                                            ("runs2sbc", 3, "8,8,7"), ("runs2rs", 3, "8,8,5"),
                                            ("runs2sbc", 3, "4,8,8"), ("runs2sbc", 3, "8,4,6"),
                                            ("runs2sbc", 3, "8,8,3"), ("runs2sbc", 3, "8,8,2"),
                                            ("runs2rs", 3, "8,8,1"), ("runs2d", 2, None)])
def test_spreading_runs_edge_cases(variant, d, bins, monkeypatch):
    """FP32 run-merging B^T kernels on work items that stress their node
    loops: 9001 nodes in one cell (runs across many 64-node chunks and
    several work items, streams of unequal length, partial last chunks), a
    lone node in a bin of its own (an item of one node: a stream without
    nodes), nodes on the -0.5 edge, and uniform nodes in one corner.  In 3-D
    with the default 8^3 bins (runs of three cells along dimension 2, the
    last group of a bin row clamped to cells 5..7) and with other bin edges:
    7, 5 and 3 in dimension 2 (other clamped groups; bins that do not divide
    the grid, so tiles go to scratch without the zone order), 2 and 1 (runs
    of one cell), and other shapes (other plane pitches and lane tables)."""
    monkeypatch.setenv("NFFT_HIP_SPREAD", variant)
    if bins:
        monkeypatch.setenv("NFFT_HIP_BINS", bins)
    rng = np.random.default_rng(51)
    N = [16] * d
    x = np.concatenate([np.zeros((9001, d)), np.full((7, d), -0.5),
                        rng.uniform(-0.5, -0.1, (513, d)), np.full((1, d), 0.37)])
    fh = W.complex_normal(int(np.prod(N)), seed=52)
    f = W.complex_normal(x.shape[0], seed=53)
    (fg, fr, _), (hg, hr, _) = run_pair(N, x, fh, f, 4, "single")
    print("%s %dD edge cases: adjoint rel l2 %.2e, trafo %.2e" % (variant, d, rel_errs(hg, hr)[0], rel_errs(fg, fr)[0]))
    assert rel_errs(hg, hr)[0] < 1e-5 and rel_errs(fg, fr)[0] < 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("m,bins", [(5, None), (6, None), (7, None), (8, None), (5, "4,4,3"),
                                    (3, "8,8,5")])
def test_wide_and_odd_3d_fp32_windows(m, bins, monkeypatch):
    """3-D FP32 B and B^T (run-merging spreading, slot interpolation) with
    windows of 8 to 18 taps: above 10 taps one node per window pass (3 W >
    32 lanes), two or three lane slots per wavefront from the plan's lane
    table, interpolation pairs instead of quads and one-cell runs; with the
    default bins and with odd dimension-2 bin edges, against the oracle."""
    if bins:
        monkeypatch.setenv("NFFT_HIP_BINS", bins)
    N = [16, 16, 16]
    x = W.kooshball(64, 48, seed=91)
    fh = W.complex_normal(int(np.prod(N)), seed=92)
    f = W.complex_normal(x.shape[0], seed=93)
    (fg, fr, _), (hg, hr, _) = run_pair(N, x, fh, f, m, "single")
    print("3-D FP32 m=%d bins %s: trafo rel l2 %.2e adjoint %.2e" % (m, bins, rel_errs(fg, fr)[0], rel_errs(hg, hr)[0]))
    assert rel_errs(fg, fr)[0] < 1e-5 and rel_errs(hg, hr)[0] < 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("d,N,n,m", [(5, [4] * 5, [8] * 5, 2), (6, [4] * 6, [8] * 6, 2),
                                     (5, [4, 6, 4, 4, 4], [8, 12, 8, 8, 8], 2),
                                     (5, [4] * 5, [8] * 5, 3)])
@pytest.mark.parametrize("precision", ["double", "single"])
def test_five_and_six_dimensions(d, N, n, m, precision):
    """d = 5 and 6 (runtime-window B / B^T kernels with their tiles in HBM,
    one band-pruned FFT pass per dimension, a length-12 axis among them)
    against the oracle; at m = 3 with n_t = 2m + 2 the plan runs the direct
    NDFT, as the reference does."""
    M = 300
    x = W.uniform(M, d, seed=70 + d)
    fh = W.complex_normal(int(np.prod(N)), seed=80 + d)
    f = W.complex_normal(M, seed=90 + d)
    (fg, fr, _), (hg, hr, _) = run_pair(N, x, fh, f, m, precision, n=n)
    tol = 1e-12 if precision == "double" else 1e-5
    print("d=%d N=%s m=%d %s: trafo rel l2 %.2e adjoint %.2e"
          % (d, N, m, precision, rel_errs(fg, fr)[0], rel_errs(hg, hr)[0]))
    assert rel_errs(fg, fr)[0] < tol and rel_errs(hg, hr)[0] < tol


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["double", "single"])
def test_non_power_of_two_grid(precision):
    # oversampled lengths with factors 3, 5 exercise the mixed-radix FFT
    N = [20, 12]
    x = W.uniform(700, 2, seed=51)
    fh = W.complex_normal(240, seed=52)
    f = W.complex_normal(700, seed=53)
    (fg, fr, _), (hg, hr, _) = run_pair(N, x, fh, f, 4, precision, n=[60, 30])
    tol = 1e-12 if precision == "double" else 1e-5
    assert rel_errs(fg, fr)[0] < tol and rel_errs(hg, hr)[0] < tol


@pytest.mark.gpu
def test_pointer_swap_between_calls():
    """The solver swaps f / f_hat pointers between calls; host pointers are
    authoritative on every call."""
    N = [32, 32]
    x = W.uniform(1000, 2, seed=61)
    p = make_plan(N, x, 4, "double")
    try:
        a = np.ascontiguousarray(W.complex_normal(1024, seed=62))
        out = np.zeros(1000, dtype=np.complex128)
        import ctypes
        saved_fh, saved_f = p.p.f_hat, p.p.f
        p.p.f_hat = ctypes.cast(a.ctypes.data, type(p.p.f_hat))
        p.p.f = ctypes.cast(out.ctypes.data, type(p.p.f))
        p.trafo()
        p.p.f_hat, p.p.f = saved_fh, saved_f
        ref = O.nfft_trafo(N, p.n, 4, x, a)
        assert rel_errs(out, ref)[0] < 1e-12
    finally:
        p.finalize()


@pytest.mark.gpu
def test_c_caller_runs(tmp_path):
    """tests/c/simple_test.c, a C caller of the plan API written like the
    reference's examples, compiled against include/nfft3.h and run."""
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = tmp_path / "simple_test"
    subprocess.run(["gcc", "-std=c99", "-O1", "-I", os.path.join(root, "include"),
                    os.path.join(root, "tests", "c", "simple_test.c"), "-o", str(exe),
                    "-L", os.path.join(root, "nfft_amd"), "-lnfft3_hip", "-lm",
                    "-Wl,-rpath," + os.path.join(root, "nfft_amd")], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0 and "OK" in r.stdout, r.stdout + r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("case", [c for c in CASES if len(c[1]) >= 2], ids=[c[0] for c in CASES if len(c[1]) >= 2])
def test_measured_plan_vs_oracle(case, precision):
    """FFTW_MEASURE plans time the B / B^T kernel shapes at precompute and keep
    the fastest; results stay within the same tolerances."""
    name, N, gen, m = case
    x = gen()
    p = nf.Plan(N, x.shape[0], n=[2 * O.next_power_of_2(t) for t in N], m=m, flags=FLAGS,
                precision=precision, fftw_flags=nf.FFTW_MEASURE)
    try:
        p.x[:] = x
        p.precompute()
        xr = p.x.astype(np.float64).copy()
        p.f_hat[:] = W.complex_normal(int(np.prod(N)), seed=71)
        fhr = p.f_hat.astype(np.complex128).copy()
        p.trafo()
        fg = p.f.astype(np.complex128).copy()
        p.f[:] = W.complex_normal(x.shape[0], seed=72)
        fr = p.f.astype(np.complex128).copy()
        p.adjoint()
        hg = p.f_hat.astype(np.complex128).copy()
        nn = p.n
    finally:
        p.finalize()
    tol = 1e-12 if precision == "double" else 1e-5
    assert rel_errs(fg, O.nfft_trafo(N, nn, m, xr, fhr))[0] < tol
    assert rel_errs(hg, O.nfft_adjoint(N, nn, m, xr, fr))[0] < tol


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("m,N,n", [(2, [6, 4, 6, 4], [16, 8, 16, 8]), (3, [6, 6, 4, 6], [12, 12, 10, 12])])
def test_four_dimensional(precision, m, N, n):
    """d = 4 goes through the generic-d kernels (tap odometer, reference
    kernel/nfft/nfft.c:630-1100) and four pruned FFT passes."""
    x = W.uniform(400, 4, seed=81)
    fh = W.complex_normal(int(np.prod(N)), seed=82)
    f = W.complex_normal(400, seed=83)
    (fg, fr, fhr), (hg, hr, frr) = run_pair(N, x, fh, f, m, precision, n=n)
    l2t, _ = rel_errs(fg, fr)
    l2a, _ = rel_errs(hg, hr)
    print("4-D m=%d %s trafo l2 %.2e adjoint l2 %.2e" % (m, precision, l2t, l2a))
    tol = 1e-12 if precision == "double" else 1e-5
    assert l2t < tol and l2a < tol


@pytest.mark.gpu
@pytest.mark.parametrize("pairs,slot,order", [("0", "2", "1"), ("1", "2", "1"), ("1", "4", "1"),
                                              ("1", "4", "0")])
@pytest.mark.parametrize("pitch,plane", [("0", "0"), ("1", "0"), ("22", "3")])
def test_interp_pairs_on_and_off(pairs, slot, order, pitch, plane, monkeypatch):
    """3-D FP32 interpolation without same-cell slots, with slots of up to two
    and of up to four nodes (dense kooshball centre: most nodes share their
    cell; the outer shell mixes cells of one to three nodes, so the single,
    pair, triple and quad code all run), each item's slots in count and bank
    order or in node order (NFFT_HIP_SLOTSORT), against the oracle, with the
    interpolation tile's rows at their natural length, padded to the default
    pitch, and at 22 cells with 3 more cells between planes
    (NFFT_HIP_IPITCH, NFFT_HIP_IPLANE)."""
    monkeypatch.setenv("NFFT_HIP_PAIRS", pairs)
    monkeypatch.setenv("NFFT_HIP_SLOT", slot)
    monkeypatch.setenv("NFFT_HIP_SLOTSORT", order)
    monkeypatch.setenv("NFFT_HIP_IPITCH", pitch)
    monkeypatch.setenv("NFFT_HIP_IPLANE", plane)
    N = [16, 16, 16]
    x = W.kooshball(96, 64, seed=5)
    fh = W.complex_normal(int(np.prod(N)), seed=81)
    f = W.complex_normal(x.shape[0], seed=82)
    (fg, fr, _), (hg, hr, _) = run_pair(N, x, fh, f, 4, "single")
    assert rel_errs(fg, fr)[0] < 1e-5 and rel_errs(hg, hr)[0] < 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("N,gen", [([64, 64], lambda: W.clustered_2d(20000, seed=71)),
                                   ([16, 16, 16], lambda: W.kooshball(256, 96, seed=72))],
                         ids=["2d_clustered", "3d_kooshball"])
def test_fp32_results_reproducible(N, gen):
    """The node binning orders each cell's nodes by node index (a stable
    radix sort, as the reference's), so a fresh plan on the same nodes runs
    the same FP32 sums: trafo and adjoint bit-identical across plans (the
    FP32 kernels add in a fixed order, with no atomics)."""
    x = gen()
    fh = W.complex_normal(int(np.prod(N)), seed=73)
    f = W.complex_normal(x.shape[0], seed=74)
    outs = []
    for _ in range(2):
        p = make_plan(N, x, 4, "single")
        try:
            p.f_hat[:] = fh
            p.trafo()
            ft = p.f.copy()
            p.f[:] = f
            p.adjoint()
            outs.append((ft, p.f_hat.copy()))
        finally:
            p.finalize()
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1])
