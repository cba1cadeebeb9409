"""Batched transforms through the C ABI: many right-hand sides on one node
set (trafo_batch_device / adjoint_batch_device) and several plans in one
hipGraph (graph_trafo / graph_adjoint).  Both run the same kernels as the
single calls, so results must equal them: bit for bit where the kernels are
deterministic (the float family), to 1e-13 relative l2 where B^T adds with
LDS atomics (the double family's ds_add_f64 order varies from run to run);
graph replays are checked one by one, with eager work and synchronisations
between replays and new input values written before each replay."""
import numpy as np
import pytest

import nfft_amd as nf
from nfft_amd import workloads as W

torch = pytest.importorskip("torch")

FLAGS = nf.PRE_PHI_HUT | nf.PRE_PSI | nf.MALLOC_X | nf.MALLOC_F_HAT | nf.MALLOC_F | nf.FFTW_INIT | nf.FFT_OUT_OF_PLACE


def plan(N, x, precision="single", m=4):
    p = nf.Plan(N, x.shape[0], n=[2 * t for t in N], m=m, flags=FLAGS, precision=precision)
    p.x[:] = x
    p.precompute()
    return p


def cdt(precision):
    return torch.complex64 if precision == "single" else torch.complex128


def same(a, b, precision):
    if precision == "single":
        assert torch.equal(a, b)
    else:
        assert float(torch.linalg.norm(a - b) / torch.linalg.norm(b)) <= 1e-13


def rand_dev(n, seed, precision):
    return torch.from_numpy(W.complex_normal(n, seed=seed)).to(cdt(precision)).cuda()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("N,gen", [([32, 32], lambda: W.golden_radial(50, 64)),
                                   ([16, 16, 16], lambda: W.kooshball(64, 64, seed=3))],
                         ids=["2d", "3d"])
def test_batch_matches_single_calls(precision, N, gen):
    """FP32 plans with the default kernels run the batch as one launch per
    kernel (grid y = right-hand side); the others loop.  Either way each
    right-hand side equals its single call."""
    x = gen()
    p = plan(N, x, precision)
    try:
        nrhs, NN, M = 5, int(np.prod(N)), p.M
        fh = rand_dev(nrhs * NN, 1, precision)
        f = torch.empty(nrhs * M, dtype=cdt(precision), device="cuda")
        p.trafo_batch_device(fh.data_ptr(), f.data_ptr(), nrhs)
        fin = rand_dev(nrhs * M, 2, precision)
        fho = torch.empty(nrhs * NN, dtype=cdt(precision), device="cuda")
        p.adjoint_batch_device(fin.data_ptr(), fho.data_ptr(), nrhs)
        p.synchronize()
        for r in range(nrhs):
            one = torch.empty(M, dtype=cdt(precision), device="cuda")
            p.trafo_device(fh[r * NN:].data_ptr(), one.data_ptr())
            oneh = torch.empty(NN, dtype=cdt(precision), device="cuda")
            p.adjoint_device(fin[r * M:].data_ptr(), oneh.data_ptr())
            p.synchronize()
            assert torch.equal(one, f[r * M:(r + 1) * M])
            same(oneh, fho[r * NN:(r + 1) * NN], precision)
    finally:
        p.finalize()


@pytest.mark.gpu
@pytest.mark.parametrize("m,nrhs", [(2, 2), (4, 8), (4, 9), (6, 3), (8, 5)])
def test_2d_batch_several_rhs_per_workgroup(m, nrhs):
    """2-D FP32 batched B^T takes two right-hand sides per workgroup (shared
    node work, one LDS tile each; odd counts end with a group of one), B one
    per workgroup: every right-hand side bit-identical to its single call,
    for windows of 6 to 18 taps (one or two lane slots)."""
    x = W.golden_radial(60, 64)
    N = [32, 32]
    p = plan(N, x, "single", m=m)
    try:
        NN, M = int(np.prod(N)), p.M
        fh = rand_dev(nrhs * NN, 4, "single")
        f = torch.empty(nrhs * M, dtype=torch.complex64, device="cuda")
        p.trafo_batch_device(fh.data_ptr(), f.data_ptr(), nrhs)
        fin = rand_dev(nrhs * M, 5, "single")
        fho = torch.empty(nrhs * NN, dtype=torch.complex64, device="cuda")
        p.adjoint_batch_device(fin.data_ptr(), fho.data_ptr(), nrhs)
        p.synchronize()
        for r in range(nrhs):
            one = torch.empty(M, dtype=torch.complex64, device="cuda")
            p.trafo_device(fh[r * NN:].data_ptr(), one.data_ptr())
            oneh = torch.empty(NN, dtype=torch.complex64, device="cuda")
            p.adjoint_device(fin[r * M:].data_ptr(), oneh.data_ptr())
            p.synchronize()
            assert torch.equal(one, f[r * M:(r + 1) * M]), r
            assert torch.equal(oneh, fho[r * NN:(r + 1) * NN]), r
    finally:
        p.finalize()


@pytest.mark.gpu
@pytest.mark.parametrize("adjoint", [False, True])
def test_graph_of_many_plans_replays(adjoint):
    sizes = [(16, 16), (32, 32), (24, 40)]
    Ms = [700, 5000, 2100]
    plans = [plan(list(N), W.uniform(M, 2, seed=10 + i)) for i, (N, M) in enumerate(zip(sizes, Ms))]
    try:
        nrhs = 3
        NNs = [int(np.prod(N)) for N in sizes]
        nin = [(Ms[i] if adjoint else NNs[i]) * nrhs for i in range(3)]
        nout = [(NNs[i] if adjoint else Ms[i]) * nrhs for i in range(3)]
        ins = [rand_dev(nin[i], 20 + i, "single") for i in range(3)]
        outs = [torch.zeros(nout[i], dtype=torch.complex64, device="cuda") for i in range(3)]
        g = nf.Graph(plans, [t.data_ptr() for t in ins], [t.data_ptr() for t in outs], nrhs, adjoint)
        try:
            for rep in range(4):
                for i in range(3):  # new input values before every replay
                    ins[i].copy_(rand_dev(nin[i], 100 * rep + i, "single"))
                torch.cuda.synchronize()
                g.launch()
                g.synchronize()
                for i, p in enumerate(plans):
                    ref = torch.empty_like(outs[i])
                    if adjoint:
                        p.adjoint_batch_device(ins[i].data_ptr(), ref.data_ptr(), nrhs)
                    else:
                        p.trafo_batch_device(ins[i].data_ptr(), ref.data_ptr(), nrhs)
                    p.synchronize()
                    assert torch.equal(ref, outs[i]), (rep, i)
        finally:
            g.destroy()
    finally:
        for p in plans:
            p.finalize()


@pytest.mark.gpu
def test_2d_split_zone_reduction_non_square_core():
    """2-D n = 1024 x 768: bins of 32 x 24 cells (a core of 768 cells, summed
    by four workgroups per bin of which the last quarter gets no full row),
    FP32, against the oracle on a node sample, and batched calls (the zone
    reduction's batched launch, grid x = bins x parts, y = right-hand side)
    equal to single calls."""
    import oracle as O
    N, n, m = [512, 384], [1024, 768], 4
    x = W.clustered_2d(200000, seed=12)
    M = x.shape[0]
    p = nf.Plan(N, M, n=n, m=m, flags=FLAGS, precision="single")
    try:
        p.x[:] = x
        p.precompute()
        xr = p.x.astype(np.float64).copy()
        NN = int(np.prod(N))
        R = 3
        fh = rand_dev(R * NN, 31, "single")
        fi = rand_dev(R * M, 32, "single")
        t_b = torch.empty(R * M, dtype=torch.complex64, device="cuda")
        a_b = torch.empty(R * NN, dtype=torch.complex64, device="cuda")
        p.trafo_batch_device(fh.data_ptr(), t_b.data_ptr(), R)
        p.adjoint_batch_device(fi.data_ptr(), a_b.data_ptr(), R)
        p.synchronize()
        for r in range(R):
            one = torch.empty(M, dtype=torch.complex64, device="cuda")
            oneh = torch.empty(NN, dtype=torch.complex64, device="cuda")
            p.trafo_device(fh[r * NN:].data_ptr(), one.data_ptr())
            p.adjoint_device(fi[r * M:].data_ptr(), oneh.data_ptr())
            p.synchronize()
            assert torch.equal(one, t_b[r * M:(r + 1) * M]), r
            assert torch.equal(oneh, a_b[r * NN:(r + 1) * NN]), r
        # right-hand side 0 against the oracle (FP32 inputs, FP64 oracle)
        fh0 = fh[:NN].cpu().numpy().astype(np.complex128)
        fi0 = fi[:M].cpu().numpy().astype(np.complex128)
        t0 = t_b[:M].cpu().numpy().astype(np.complex128)
        a0 = a_b[:NN].cpu().numpy().astype(np.complex128)
        # The oracle's D, B, B^T and D^T around numpy's FFT: the oracle's own
        # FFT is a direct DFT on the 768-point axis (minutes of CPU time at
        # this size) and equals numpy's to rounding (test_oracle).
        idx = np.random.default_rng(3).choice(M, size=2000, replace=False)
        g = np.fft.fftn(O.stage_D(N, n, m, fh0).reshape(n)).ravel()
        tr = O.stage_B(N, n, m, xr[idx], g)
        et = float(np.linalg.norm(t0[idx] - tr) / np.linalg.norm(tr))
        gt = (np.fft.ifftn(O.stage_BT(N, n, m, xr, fi0).reshape(n)) * np.prod(n)).ravel()
        ar = O.stage_DT(N, n, m, gt)
        ea = float(np.linalg.norm(a0 - ar) / np.linalg.norm(ar))
        print("2-D 1024 x 768: trafo (2000 nodes) rel l2 %.2e, adjoint rel l2 %.2e" % (et, ea))
        assert et < 1e-5 and ea < 1e-5
    finally:
        p.finalize()
