"""The allocation-free batched adjoint (nfft[f]_adjoint_batch_device_noalloc)
through the C ABI against adjoint_batch_device, the entry point it mirrors.

Its batch grids and spreading tiles live in the library's static 64 MiB
per-GPU scratch array instead of plan allocations.  Checks:
  * two calls in a row, the second with a smaller input (a smaller plan and
    fewer right-hand sides), each equal to adjoint_batch_device bit for bit
    in the float family (deterministic kernels) -- the first call's scratch
    contents must not leak into the second;
  * a batch larger than one scratch chunk (several chunks per call);
  * free device memory does not change across calls;
  * the double family (not batched: one right-hand side at a time on the
    plan's own buffers) within 1e-13, ds_add_f64 order varying;
  * a plan whose right-hand side needs more than the array is refused with
    -4 and nothing written; a plan without node data with -2.
"""
import ctypes

import numpy as np
import pytest

import nfft_amd as nf
from nfft_amd import workloads as W

torch = pytest.importorskip("torch")

FLAGS = nf.PRE_PHI_HUT | nf.PRE_PSI | nf.MALLOC_X | nf.MALLOC_F_HAT | nf.MALLOC_F | nf.FFTW_INIT


def noalloc_fn(precision):
    L = nf.lib()
    fn = getattr(L, ("nfftf_" if precision == "single" else "nfft_") + "adjoint_batch_device_noalloc")
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    fn.restype = ctypes.c_int
    return fn


def make_plan(N, x, precision, precompute=True):
    p = nf.Plan(N, x.shape[0], n=[2 * t for t in N], m=4, flags=FLAGS, precision=precision)
    p.x[:] = x
    if precompute:
        p.precompute()
    return p


def cdt(precision):
    return torch.complex64 if precision == "single" else torch.complex128


def run_pair(p, nrhs, seed, precision):
    """(no-alloc result, adjoint_batch_device result, return code)."""
    M, NN = p.M, int(np.prod(p.N))
    f = torch.from_numpy(W.complex_normal(nrhs * M, seed=seed)).to(cdt(precision)).cuda()
    out_na = torch.full((nrhs * NN,), complex(np.nan, np.nan), dtype=cdt(precision), device="cuda")
    out_ref = torch.empty(nrhs * NN, dtype=cdt(precision), device="cuda")
    torch.cuda.synchronize()
    rc = noalloc_fn(precision)(ctypes.addressof(p.p), f.data_ptr(), out_na.data_ptr(), nrhs)
    p.synchronize()
    p.adjoint_batch_device(f.data_ptr(), out_ref.data_ptr(), nrhs)
    p.synchronize()
    return out_na, out_ref, rc


@pytest.mark.gpu
def test_noalloc_matches_batched_adjoint_two_calls():
    # first call: the 32-coil radial plan at N = 256^2, 12 right-hand sides
    # (more than one scratch chunk); second: a smaller plan, 3 right-hand sides
    big = make_plan([256, 256], W.golden_radial(403, 512), "single")
    small = make_plan([64, 64], W.uniform(5000, 2, seed=8), "single")
    try:
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        a, b, rc = run_pair(big, 12, 1, "single")
        assert rc == 0
        assert torch.equal(a, b)
        a, b, rc = run_pair(small, 3, 2, "single")
        assert rc == 0
        assert torch.equal(a, b)
        # the plan-allocated batch buffers of adjoint_batch_device are kept
        # with the plans; repeated no-alloc calls change nothing
        torch.cuda.synchronize()
        free1 = torch.cuda.mem_get_info()[0]
        f = torch.from_numpy(W.complex_normal(12 * big.M, seed=3)).to(torch.complex64).cuda()
        out = torch.empty(12 * 256 * 256, dtype=torch.complex64, device="cuda")
        torch.cuda.synchronize()
        free2 = torch.cuda.mem_get_info()[0]
        for _ in range(3):
            assert noalloc_fn("single")(ctypes.addressof(big.p), f.data_ptr(), out.data_ptr(), 12) == 0
        big.synchronize()
        assert torch.cuda.mem_get_info()[0] == free2
        print("free device memory: before %d, after the pairs %d" % (free0, free1))
    finally:
        big.finalize()
        small.finalize()


@pytest.mark.gpu
def test_noalloc_double_family():
    p = make_plan([32, 32], W.golden_radial(50, 64), "double")
    try:
        a, b, rc = run_pair(p, 4, 5, "double")
        assert rc == 0
        assert float(torch.linalg.norm(a - b) / torch.linalg.norm(b)) <= 1e-13
    finally:
        p.finalize()


@pytest.mark.gpu
def test_noalloc_refuses_what_does_not_fit():
    # 3-D n = 128^3: one right-hand side's grid alone is 16 MiB and its
    # spreading tiles (17^3 cells per 8^3 bin) about 160 MiB
    p = make_plan([64, 64, 64], W.uniform(200000, 3, seed=9), "single")
    try:
        f = torch.zeros(2 * p.M, dtype=torch.complex64, device="cuda")
        out = torch.full((2 * 64 ** 3,), 7.0, dtype=torch.complex64, device="cuda")
        rc = noalloc_fn("single")(ctypes.addressof(p.p), f.data_ptr(), out.data_ptr(), 2)
        p.synchronize()
        assert rc == -4
        assert bool(torch.all(out == 7.0))
    finally:
        p.finalize()
    q = make_plan([32, 32], W.uniform(3000, 2, seed=10), "single", precompute=False)
    try:
        f = torch.zeros(2 * q.M, dtype=torch.complex64, device="cuda")
        out = torch.zeros(2 * 32 * 32, dtype=torch.complex64, device="cuda")
        assert noalloc_fn("single")(ctypes.addressof(q.p), f.data_ptr(), out.data_ptr(), 2) == -2
    finally:
        q.finalize()
