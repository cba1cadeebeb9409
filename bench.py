#!/usr/bin/env python3
"""Throughput of the NFFT hot path on one MI355X.

Workload (one "step" = one nfftf_trafo + one nfftf_adjoint through the C ABI,
device-resident inputs):
  3-D NFFT, FP32 spreading / FFT with FP64 phi_hut and coordinate reduction,
  N = 128^3 Fourier coefficients (iid complex N(0,1)), sigma = 2 (n = 256^3),
  Kaiser-Bessel m = 4, M = 2^24 nodes on a 3-D radial "kooshball" (16384
  spokes x 1024 samples, directions uniform on the sphere, radii -1/2 + k/1024).
Metric: nodes/s = M / (t_trafo + t_adjoint), with precompute (node binning)
excluded and reported separately.

Also reported:
  roofline      the dominant kernel (B interpolation or B^T spreading): its
                algorithmic HBM bytes per launch (n_total c + M d r + M c)
                over its average duration from HIP events on the plan's stream
                inside the timed region, against 8 TB/s
  step_hbm      the whole step's algorithmic bytes (D + FFT + B both ways)
                over the step time
  stages        per stage: algorithmic HBM bytes and GB/s against 8 TB/s;
                for B and B^T also the on-chip rates that bound them: taps
                (2m+2)^d M per launch, 8 bytes of LDS grid data per tap
                (algorithmic: one complex value per tap; the 3-D kernel's
                same-cell node pairs read about 55% of that) against
                256 CUs x 128 B/clk x 2.4 GHz = 78.6 TB/s, and 4 flops per tap
                (complex value times real weight, two FMAs) against the
                78.6 TFLOP/s of unpacked FP32 FMA issue (256 CUs x 4 SIMDs x
                16 lanes x 2 flops x 2.4 GHz; the published 157.3 TFLOP/s
                counts packed FP32, which this library does not emit).  B^T's
                HBM figure includes the tile reduction, its on-chip figures
                cover the spreading kernel alone
  cpu_baseline  the FP64 CPU oracle on a bounded sample of the same workload
                (full-grid D + FFT, B on 2^22 of the nodes, scaled to M), on
                up to 16 host cores (OpenMP)
  batched       32 coil right-hand sides on one 2-D plan and 64 plans of
                uneven size, eager and as hipGraph replays
  other_configs the 1-D FP64 and 2-D (512^2, FP32 / FP64, m 4 / 6) single-plan
                workloads: ms per trafo+adjoint (with --no-batch skipped too)
  cgnr          the metric's second part on held-out 2-D trajectories
                (N = 256^2, sigma 2, m = 4, FP32; golden-angle radial and
                variable-density spiral sets; phantom of random ellipses):
                unweighted vs Pipe-Menon weights vs the learned weights of
                nfft_amd/data/dcf_model.npy, wall time including the weight
                computation, to normal-equation residual 1e-6 (noise-free
                data) and to an image-error target (1% complex noise)

python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline] [--no-cgnr] [--no-batch]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "nodes/s 3-D trafo+adjoint (GB/s vs roofline); CGNR time-to-1e-6 learned vs plain"
HBM_PEAK_GBS = 8000.0
LDS_PEAK_GBS = 256 * 128 * 2.4  # 78643 GB/s
VALU_FP32_TFLOPS = 256 * 4 * 16 * 2 * 2.4e9 / 1e12  # 78.6, unpacked FMA
FP32_PEAK_TFLOPS = 157.3  # MI355X published FP32 vector peak (packed)


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def algorithmic_bytes(N, n, M, d, c, r):
    """Algorithmic HBM bytes of trafo and adjoint stages (D, F, B)."""
    Nt = int(np.prod(N))
    nt = int(np.prod(n))
    D = Nt * c + nt * c
    F = d * 2 * nt * c
    B = nt * c + M * d * r + M * c
    DT = 2 * Nt * c
    return dict(D=D, F=F, B=B, DT=DT, trafo=D + F + B, adjoint=B + F + DT)


def err_stats(a, ref):
    """Relative l2 and l-infinity error of a against ref."""
    a = np.asarray(a, dtype=np.complex128)
    ref = np.asarray(ref, dtype=np.complex128)
    d = a - ref
    return {"rel_l2": float(np.linalg.norm(d) / np.linalg.norm(ref)),
            "rel_linf": float(np.max(np.abs(d)) / np.max(np.abs(ref)))}


def oracle_sample_errors(N, n, m, x, fh, f, t64, a64, sample, seed):
    """FP64 GPU results against the FP64 oracle: the trafo on `sample` random
    nodes (the oracle's full-grid D + FFT, B on the sample) and the adjoint of
    f restricted to those nodes (a64 = the GPU adjoint of that restriction)."""
    import oracle as O
    idx = np.random.default_rng(seed).choice(x.shape[0], size=sample, replace=False)
    g = O.fft(n, O.stage_D(N, n, m, fh), -1)
    return {"trafo": err_stats(t64[idx], O.stage_B(N, n, m, x[idx], g)),
            "adjoint": err_stats(a64, O.nfft_adjoint(N, n, m, x[idx], f[idx])),
            "sample_nodes": sample}, idx


def cpu_baseline(N, n, m, x, fh, f, sample, threads):
    """FP64 CPU path on the host's cores (OpenMP, `threads` threads), shaped as
    the reference's PRE_PSI trafo / blockwise adjoint (oracle.Prepsi3d: window
    values precomputed per node and not timed, as nfft_benchomp times
    precompute separately; tensor loops; the adjoint split into grid slabs
    without atomics), with the oracle's full-grid D + FFT both ways; B and
    B^T on `sample` nodes, their time scaled to all nodes."""
    import oracle as O
    O.set_threads(threads)
    M = x.shape[0]
    idx = np.random.default_rng(5).choice(M, size=sample, replace=False)
    xs = np.ascontiguousarray(x[idx])
    tp = time.perf_counter()
    P = O.Prepsi3d(N, n, m, xs)
    t_pre = time.perf_counter() - tp
    t0 = time.perf_counter()
    g = O.stage_D(N, n, m, fh)
    g = O.fft(n, g, -1)
    t1 = time.perf_counter()
    P.B(g)
    t2 = time.perf_counter()
    gt = P.BT(f[idx])
    t3 = time.perf_counter()
    gt = O.fft(n, gt, +1)
    O.stage_DT(N, n, m, gt)
    t4 = time.perf_counter()
    O.set_threads(1)
    fixed = (t1 - t0) + (t4 - t3)
    per_node = ((t2 - t1) + (t3 - t2)) / sample
    est = fixed + per_node * M
    return dict(value=M / est, unit="nodes/s", cores=threads, kind="port",
                sample="FP64 CPU path shaped as the reference's PRE_PSI trafo and blockwise adjoint "
                       "(oracle.Prepsi3d; the reference itself needs FFTW, absent here) on %d OpenMP "
                       "threads, same workload: full 256^3 D+FFT for trafo and adjoint, B and B^T on "
                       "%d random nodes of the kooshball, B time scaled to M=%d; B %.2f s, B^T %.2f s, "
                       "D+FFT both ways %.2f s, untimed window precompute %.2f s"
                       % (threads, sample, M, t2 - t1, t3 - t2, fixed, t_pre))


def cgnr_leg(args):
    """CGNR on held-out trajectories: unweighted vs Pipe-Menon vs learned W.

    Criteria, each per method and summed over the cases:
      residual_1e-6     iterations and wall time (weights included) until the
                        normal-equation residual sqrt(z^H z / z_0^H z_0)
                        (each solver in its own norm) falls below 1e-6, on
                        the workload's noisy data (1% complex noise); not
                        reached within --cgnr-max-iter steps is reported as
                        such (total_s null)
      residual_1e-6_noise_free  the same on noise-free data y = A f_true
      to_image_target   on the noisy data: iterations and wall time until
                        the image error |Re f - f_true| / |f_true| is within
                        10% of the best any of the three attains in 300 steps
                        (weighted CGNR reaches its best image in a few steps
                        and then semi-converges; unweighted CGNR needs more
                        steps to the same image quality)
    """
    import torch
    from nfft_amd import dcf_train as T
    from nfft_amd import density as D
    from nfft_amd import workloads as W

    N, tol = 256, 1e-6
    model_path = T.DEFAULT_MODEL
    if not os.path.exists(model_path):
        log("cgnr: no trained model at %s, skipped" % model_path)
        return {"skipped": "no trained model"}
    model = D.DcfModel.load(model_path)
    names = ("plain", "pipe_menon", "learned")
    # one-time warm-up outside the timed calls (first launches of the
    # density, MLP and solver kernels), like loading the model itself
    xw, _ = T.trajectory(4999, N)
    pw = T.make_plan(xw, N)
    model.weights(pw)
    D.pipe_menon(pw, 1)
    T.cgnr_time_to(pw, np.ones(pw.M, dtype=np.complex128), model.weights(pw).cpu().numpy(), max_iter=2)
    pw.finalize()
    cases = []
    for seed in (5001, 5003, 5006, 5008):  # held out: training uses seeds 0..1999
        x, desc = T.trajectory(seed, N)
        p = T.make_plan(x, N)
        f_true = W.phantom(N, 20, seed=seed)
        p.f_hat[:] = f_true
        p.trafo()
        y0 = p.f.astype(np.complex128).copy()
        y = y0 + 0.01 * np.sqrt(np.mean(np.abs(y0) ** 2)) * W.complex_normal(p.M, seed=seed + 1)
        res = {"trajectory": list(desc), "M": int(p.M)}
        traces = {}
        for name in names:
            # weight computation time: median of 3 (device work and the copy
            # of w to the host, where the solver's W lives)
            w, tws = None, [0.0]
            if name != "plain":
                tws = []
                for _ in range(3):
                    torch.cuda.synchronize()
                    t0 = time.perf_counter()
                    w = D.pipe_menon(p, 20) if name == "pipe_menon" else model.weights(p).cpu().numpy()
                    tws.append(time.perf_counter() - t0)
            t_w = float(np.median(tws))
            # the stated criterion on the noisy measurements, and the same
            # criterion on noise-free data (a separately named figure)
            it, t_s, r, _ = T.cgnr_time_to(p, y, w, tol=tol, max_iter=args.cgnr_max_iter)
            it0, t_s0, r0, _ = T.cgnr_time_to(p, y0, w, tol=tol, max_iter=args.cgnr_max_iter)
            traces[name] = T.image_error_trace(p, y, f_true, w, 300)
            # gridding: one weighted adjoint of the noisy data
            p.f[:] = y
            p.synchronize()
            t0 = time.perf_counter()
            p.gridding(w)
            t_g = time.perf_counter() - t0
            img = p.f_hat.astype(np.complex128).ravel()
            a = np.vdot(img, f_true.ravel()) / np.vdot(img, img)
            g_err = float(np.linalg.norm(a * img - f_true.ravel()) / np.linalg.norm(f_true))
            res[name] = {"weights_s": t_w,
                         "residual_1e-6": {"iterations": it, "solve_s": t_s, "residual": r,
                                           "reached": bool(r < tol),
                                           "total_s": (t_s + t_w) if r < tol else None},
                         "residual_1e-6_noise_free": {
                             "iterations": it0, "solve_s": t_s0, "residual": r0,
                             "reached": bool(r0 < tol),
                             "total_s": (t_s0 + t_w) if r0 < tol else None},
                         "s_per_iteration": t_s / max(1, it),
                         "gridding": {"rel_err_scaled": g_err, "total_s": t_g + t_w}}
            log("cgnr %s %s: noisy %d it to residual %.1e (%s), %.1f ms; noise-free %d it to %.1e "
                "(%s), %.1f ms (+%.1f ms weights)"
                % (desc[0], name, it, r, "reached" if r < tol else "cap", t_s * 1e3, it0, r0,
                   "reached" if r0 < tol else "cap", t_s0 * 1e3, t_w * 1e3))
        target = 1.10 * min(min(t) for t in traces.values())
        res["image_target_rel_err"] = target
        for name in names:
            e = traces[name]
            k = next((i + 1 for i, v in enumerate(e) if v <= target), None)
            spi = res[name]["s_per_iteration"]
            res[name]["to_image_target"] = {
                "iterations": k, "total_s": None if k is None else k * spi + res[name]["weights_s"],
                "best_rel_err": min(e), "rel_err_after": {str(j): e[j - 1] for j in (1, 3, 5, 10, 20)}}
            log("cgnr %s %s: image error <= %.4f after %s it (%s ms), own best %.4f"
                % (desc[0], name, target, k,
                   "-" if k is None else "%.1f" % (1e3 * res[name]["to_image_target"]["total_s"]),
                   min(e)))
        cases.append(res)
        p.finalize()
    model.close()

    def total(name, crit):
        v = [c[name][crit]["total_s"] for c in cases]
        return None if any(t is None for t in v) else sum(v)

    out = {
        "workload": "2-D N=256^2, sigma=2, m=4, FP32; held-out golden-angle radial (2) and "
                    "variable-density spiral (2) trajectories; ellipse phantom; CGNR from 0; "
                    "gridding = one weighted adjoint (host arrays), error after the "
                    "least-squares complex scale",
        "tol": tol,
        "cases": cases,
    }
    for crit in ("residual_1e-6", "residual_1e-6_noise_free", "to_image_target"):
        t = {k: total(k, crit) for k in names}
        out[crit] = {"total_s": t,
                     "speedup_learned_vs_plain": (t["plain"] / t["learned"]
                                                  if t["plain"] and t["learned"] else None)}
    return out


def configs_leg(args):
    """The other single-plan workloads, device-resident, precompute excluded:
    1-D FP64 (N=1024, sigma 2, m=6, M=4096 uniform nodes); 2-D N=512^2
    (sigma 2, M=2^20: half uniform, half Gaussian sd 0.05 at the origin) in
    FP32 and FP64 with m = 4 and 6; and the reference's own benchmark sizes
    (examples/nfft/nfft_benchomp.c:485-575, sigma 2, m = 4, iid uniform
    nodes): 1-D N = M = 2^21, 2-D N = 1024^2 with M = 2^20, 3-D N = 128^3
    with M = 2^21, FP32 and FP64.  Per case: ms per trafo+adjoint, nodes/s
    and the algorithmic HBM bytes of both directions over that time, and the
    relative l2 / l-infinity error of its results: FP32 against the FP64 plan
    on the same inputs (all nodes), FP64 against the FP64 oracle (all nodes
    in the small 1-D case, 2000 sampled nodes otherwise).  Plans use
    FFTW_MEASURE (FFTW_ESTIMATE with --estimate)."""
    import torch
    import nfft_amd as nf
    import oracle as O
    from nfft_amd import workloads as W
    flags = nf.PRE_PHI_HUT | nf.PRE_PSI | nf.MALLOC_X | nf.MALLOC_F_HAT | nf.MALLOC_F | nf.FFTW_INIT
    r = np.random.default_rng(2)
    M2 = 1 << 20
    x2 = np.concatenate([r.random((M2 // 2, 2)) - 0.5, 0.05 * r.standard_normal((M2 // 2, 2))])
    x2 = x2 - np.round(x2)
    x2[x2 >= 0.5] -= 1.0
    cases = [("1d_fp64_m6", "1d", [1024], [2048], 6, W.uniform(4096, 1, seed=11), "double")]
    for prec in ("double", "single"):
        for m in (4, 6):
            cases.append(("2d_%s_m%d" % ("fp32" if prec == "single" else "fp64", m), "2d512",
                          [512, 512], [1024, 1024], m, x2, prec))
    refs = [("ref_1d_2pow21", [1 << 21], [1 << 22], W.uniform(1 << 21, 1, seed=13)),
            ("ref_2d_1024", [1024, 1024], [2048, 2048], W.uniform(1 << 20, 2, seed=14)),
            ("ref_3d_128", [128, 128, 128], [256, 256, 256], W.uniform(1 << 21, 3, seed=15))]
    for prec in ("double", "single"):
        for name, N, n, x in refs:
            cases.append(("%s_%s" % (name, "fp32" if prec == "single" else "fp64"), name, N, n, 4, x,
                          prec))
    # nodes FP32-representable for both precisions, as the inputs below: an
    # FP32 plan stores x in float, and at n = 2^22 rounding x to float moves
    # a node by up to 1/16 cell (a phase error of up to 0.1 rad at the
    # highest frequency), which is the input's error, not the transform's
    cases = [c[:5] + (c[5].astype(np.float32).astype(np.float64),) + c[6:] for c in cases]
    O.set_threads(min(16, len(os.sched_getaffinity(0))))
    out = {}
    res64 = {}
    # planned like the headline (FFTW_MEASURE unless --estimate): the tuner
    # picks B / B^T kernels, bins and item sizes per workload (FP64 3-D
    # interpolation, for one, is 2x faster on uniform nodes with one
    # wavefront per node, 2.4x slower with it on the kooshball)
    fftw_flags = (nf.FFTW_ESTIMATE if args.estimate else nf.FFTW_MEASURE) | nf.FFTW_DESTROY_INPUT
    for name, group, N, n, m, x, prec in cases:
        M = x.shape[0]
        p = nf.Plan(N, M, n=n, m=m, flags=flags, precision=prec, fftw_flags=fftw_flags)
        p.x[:] = x
        p.set_stream(torch.cuda.current_stream().cuda_stream)
        p.precompute()
        ct = torch.complex64 if prec == "single" else torch.complex128
        NN = int(np.prod(N))
        # the same FP32-representable inputs for both precisions
        fh_np = W.complex_normal(NN, seed=40 + m).astype(np.complex64).astype(np.complex128)
        f_np = W.complex_normal(M, seed=50 + m).astype(np.complex64).astype(np.complex128)
        fh = torch.from_numpy(fh_np).to(ct).cuda()
        f = torch.from_numpy(f_np).to(ct).cuda()
        fo = torch.empty(M, dtype=ct, device="cuda")
        ho = torch.empty(NN, dtype=ct, device="cuda")

        def step():
            p.trafo_device(fh.data_ptr(), fo.data_ptr())
            p.adjoint_device(f.data_ptr(), ho.data_ptr())
        for _ in range(3):
            step()
        torch.cuda.synchronize()
        K = 20
        t0 = time.perf_counter()
        for _ in range(K):
            step()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / K
        c = 8 if prec == "single" else 16
        ab = algorithmic_bytes(N, n, M, len(N), c, c // 2)
        out[name] = {"ms": dt * 1e3, "nodes_per_s": M / dt, "M": M, "N": N, "m": m,
                     "algorithmic_GBs": (ab["trafo"] + ab["adjoint"]) / dt / 1e9}
        t_res = fo.cpu().numpy().astype(np.complex128)
        a_res = ho.cpu().numpy().astype(np.complex128)
        xr = p.x.astype(np.float64).copy()
        if prec == "double":
            res64[(group, m)] = (t_res, a_res)
            if group == "1d":
                err = {"vs": "FP64 oracle, all nodes",
                       "trafo": err_stats(t_res, O.nfft_trafo(N, n, m, xr, fh_np)),
                       "adjoint": err_stats(a_res, O.nfft_adjoint(N, n, m, xr, f_np))}
            else:
                idx = np.random.default_rng(9).choice(M, size=2000, replace=False)
                fs = np.zeros(M, dtype=np.complex128)
                fs[idx] = f_np[idx]
                d_fs = torch.from_numpy(fs).cuda()
                p.adjoint_device(d_fs.data_ptr(), ho.data_ptr())
                torch.cuda.synchronize()
                a_s = ho.cpu().numpy()
                err, _ = oracle_sample_errors(N, n, m, xr, fh_np, f_np, t_res, a_s, 2000, 9)
                err["vs"] = "FP64 oracle on 2000 sampled nodes"
        else:
            t64, a64 = res64[(group, m)]
            err = {"vs": "FP64 GPU plan, all nodes", "trafo": err_stats(t_res, t64),
                   "adjoint": err_stats(a_res, a64)}
        out[name]["error"] = err
        log("config %s: %.3f ms per trafo+adjoint, trafo rel l2 %.2e (%s)"
            % (name, dt * 1e3, err["trafo"]["rel_l2"], err["vs"]))
        p.finalize()
    O.set_threads(1)
    return out


def batch_leg(args):
    """Batched FP32 transforms (the two batched workloads): (a) 32 coil
    right-hand sides on one 2-D N=256^2 plan (sigma 2, m 4) of 403 x 512
    golden-angle radial nodes, trafo + adjoint of all 32, eager
    (trafo_batch_device / adjoint_batch_device) and as one hipGraph replay;
    (b) 64 plans of N=128^2 with M log-uniform in [1e4, 1e6] (iid uniform
    nodes), trafo + adjoint of every plan as two hipGraph replays.  Inputs
    resident on the device; nodes/s counts every right-hand side's nodes."""
    import torch
    import nfft_amd as nf
    from nfft_amd import workloads as W
    flags = (nf.PRE_PHI_HUT | nf.PRE_PSI | nf.MALLOC_X | nf.MALLOC_F_HAT | nf.MALLOC_F
             | nf.FFTW_INIT)
    out = {}

    def timed(fn, reps=10):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps

    # (a) 32 right-hand sides, one node set
    x = W.golden_radial(403, 512)
    p = nf.Plan([256, 256], x.shape[0], n=[512, 512], m=4, flags=flags, precision="single")
    p.x[:] = x
    p.precompute()
    M, NN, R = p.M, 256 * 256, 32
    stream = torch.cuda.current_stream()
    p.set_stream(stream.cuda_stream)
    fh = torch.randn(R * NN, dtype=torch.complex64, device="cuda")
    f = torch.randn(R * M, dtype=torch.complex64, device="cuda")
    fo = torch.empty_like(f)
    ho = torch.empty_like(fh)

    def eager():
        p.trafo_batch_device(fh.data_ptr(), fo.data_ptr(), R)
        p.adjoint_batch_device(f.data_ptr(), ho.data_ptr(), R)
    t_eager = timed(eager)
    gt = nf.Graph([p], [fh.data_ptr()], [fo.data_ptr()], nrhs=R)
    ga = nf.Graph([p], [f.data_ptr()], [ho.data_ptr()], nrhs=R, adjoint=True)

    def graphs():
        gt.launch()
        ga.launch()
        ga.synchronize()
    t_graph = timed(graphs)
    gt.destroy()
    ga.destroy()
    p.finalize()
    out["coils_32"] = {"M": M, "nrhs": R, "eager_ms": t_eager * 1e3, "graph_ms": t_graph * 1e3,
                       "nodes_per_s_graph": R * M / t_graph}
    log("batch: 32 coils, M=%d: eager %.2f ms, graph %.2f ms per trafo+adjoint of all"
        % (M, t_eager * 1e3, t_graph * 1e3))

    # (b) 64 plans of uneven node count
    r = np.random.default_rng(4)
    Ms = np.round(10 ** r.uniform(4, 6, 64)).astype(int)
    plans, ins, outs, ains, aouts = [], [], [], [], []
    for i, Mi in enumerate(Ms):
        q = nf.Plan([128, 128], int(Mi), n=[256, 256], m=4, flags=flags, precision="single")
        q.x[:] = W.uniform(int(Mi), 2, seed=100 + i)
        q.precompute()
        plans.append(q)
        ins.append(torch.randn(128 * 128, dtype=torch.complex64, device="cuda"))
        outs.append(torch.empty(int(Mi), dtype=torch.complex64, device="cuda"))
        ains.append(torch.randn(int(Mi), dtype=torch.complex64, device="cuda"))
        aouts.append(torch.empty(128 * 128, dtype=torch.complex64, device="cuda"))
    gt = nf.Graph(plans, [t.data_ptr() for t in ins], [t.data_ptr() for t in outs])
    ga = nf.Graph(plans, [t.data_ptr() for t in ains], [t.data_ptr() for t in aouts], adjoint=True)

    def graphs64():
        gt.launch()
        ga.launch()
        ga.synchronize()
    t64 = timed(graphs64)
    gt.destroy()
    ga.destroy()
    for q in plans:
        q.finalize()
    out["plans_64"] = {"M_total": int(Ms.sum()), "graph_ms": t64 * 1e3,
                       "nodes_per_s": float(Ms.sum()) / t64}
    log("batch: 64 plans, %d nodes: graph %.2f ms per trafo+adjoint of all" % (Ms.sum(), t64 * 1e3))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--spokes", type=int, default=16384)
    ap.add_argument("--samples", type=int, default=1024)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-cgnr", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--estimate", action="store_true",
                    help="plan the headline workload with FFTW_ESTIMATE (built-in kernel "
                         "choices) instead of FFTW_MEASURE (timed at precompute)")
    ap.add_argument("--no-error", action="store_true",
                    help="skip the accuracy figures (FP64 plan and oracle sample)")
    ap.add_argument("--cgnr-max-iter", type=int, default=1000)
    ap.add_argument("--cpu-sample", type=int, default=0,
                    help="nodes of the CPU baseline's B / B^T sample (0: 2^22 with 8 or more "
                         "threads, else 2^18)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run a plain device-to-device copy of this many bytes "
                         "(calibrates the HBM byte counters of a profiler run)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    import nfft_amd as nf
    from nfft_amd import workloads as W

    if args.copy_probe:
        n = args.copy_probe // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n * 4, "copies": 3}))
        return

    N = [128, 128, 128]
    n = [256, 256, 256]
    m = 4
    d = 3
    log("generating kooshball nodes (%d x %d)" % (args.spokes, args.samples))
    x = W.kooshball(args.spokes, args.samples, seed=3)
    M = x.shape[0]
    fh = W.complex_normal(int(np.prod(N)), seed=4)
    f_in = W.complex_normal(M, seed=5)

    flags = (nf.PRE_PHI_HUT | nf.PRE_PSI | nf.MALLOC_X | nf.MALLOC_F_HAT | nf.MALLOC_F
             | nf.FFTW_INIT | nf.NFFT_SORT_NODES)
    # FFTW_MEASURE: precompute times the bin edges, B / B^T kernel shapes,
    # spreading item size and FFT lines per workgroup on these nodes and
    # keeps the fastest (included in precompute_ms, outside the timed steps)
    fftw_flags = nf.FFTW_ESTIMATE if args.estimate else nf.FFTW_MEASURE
    p = nf.Plan(N, M, n=n, m=m, flags=flags, precision="single",
                fftw_flags=fftw_flags | nf.FFTW_DESTROY_INPUT)
    p.x[:] = x
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    p.set_stream(stream.cuda_stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    p.precompute()
    p.synchronize()
    precompute_ms = (time.perf_counter() - t0) * 1e3
    log("precompute %.1f ms, %d work items" % (precompute_ms, p.num_subproblems()))

    d_fh = torch.from_numpy(fh.astype(np.complex64)).to(dev)
    d_f = torch.empty(M, dtype=torch.complex64, device=dev)
    d_fin = torch.from_numpy(f_in.astype(np.complex64)).to(dev)
    d_fh_out = torch.empty(int(np.prod(N)), dtype=torch.complex64, device=dev)

    def step():
        p.trafo_device(d_fh.data_ptr(), d_f.data_ptr())
        p.adjoint_device(d_fin.data_ptr(), d_fh_out.data_ptr())

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    log("warmup done")

    # timed region: K steps, HIP events around every stage on the plan stream
    p.timing_begin()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    t_wall = time.perf_counter()
    for _ in range(args.steps):
        step()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall_s = time.perf_counter() - t_wall
    st = p.timing_end()
    ms_total = ev0.elapsed_time(ev1)
    ms_step = ms_total / args.steps
    log("timed %d steps: %.3f ms/step (wall %.3f ms/step)" % (args.steps, ms_step, wall_s * 1e3 / args.steps))

    K = args.steps
    stages_ms = {k: v * 1e3 / K for k, v in st.items() if not k.startswith("n_")}
    ab = algorithmic_bytes(N, n, M, d, 8, 4)
    b_trafo_ms = stages_ms["trafo_B"]
    b_adj_ms = stages_ms["adjoint_B"]
    if b_adj_ms >= b_trafo_ms:
        dom_name, dom_ms, dom_sym = "B^T spreading (k_spread_runs, adjoint)", b_adj_ms, "k_spread_runs"
    else:
        dom_name, dom_ms, dom_sym = "B interpolation (k_interp_tile, trafo)", b_trafo_ms, "k_interp_tile"
    # HBM bytes per launch of that kernel measured by rocprofv3 FETCH_SIZE /
    # WRITE_SIZE (copy-calibrated) and committed with the profile summaries
    # (the profile names the exact kernel each stage ran in its run)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_kooshball3d.json")
    if os.path.exists(tpath):
        with open(tpath) as fh_:
            tj = json.load(fh_)
        stage = "adjoint_B" if dom_sym == "k_spread_runs" else "trafo_B"
        kname = tj.get("stage_kernels", {}).get(stage)
        if kname in tj.get("kernels", {}):
            traffic = tj["kernels"][kname]["traffic_bytes"]
    achieved = ab["B"] / (dom_ms * 1e-3) / 1e9
    step_bytes = ab["trafo"] + ab["adjoint"]
    taps = (2 * m + 2) ** d * M

    def hbm(nbytes, ms):
        gbs = nbytes / (ms * 1e-3) / 1e9
        return {"algorithmic_bytes": nbytes, "ms": ms, "GBs": gbs, "hbm_frac": gbs / HBM_PEAK_GBS}

    def onchip(ms):
        lds = 8.0 * taps / (ms * 1e-3) / 1e9
        tf = 4.0 * taps / (ms * 1e-3) / 1e12
        return {"taps": taps, "lds_GBs": lds, "lds_frac": lds / LDS_PEAK_GBS,
                "fma_TFLOPs": tf, "valu_frac": tf / VALU_FP32_TFLOPS}

    stages = {
        "trafo_D_FFT": hbm(ab["D"] + ab["F"], stages_ms["trafo_D"] + stages_ms["trafo_FFT"]),
        "trafo_B": dict(hbm(ab["B"], b_trafo_ms), **onchip(b_trafo_ms)),
        "adjoint_BT_reduce": dict(hbm(ab["B"], b_adj_ms + stages_ms["adjoint_Bred"]),
                                  **onchip(b_adj_ms)),
        "adjoint_FFT_DT": hbm(ab["F"] + ab["DT"], stages_ms["adjoint_FFT"] + stages_ms["adjoint_D"]),
    }

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle, bounded sample) ...")
        threads = min(16, len(os.sched_getaffinity(0)))
        sample = min(args.cpu_sample if args.cpu_sample else (1 << 22 if threads >= 8 else 1 << 18), M)
        cpu = cpu_baseline(N, n, m, x, fh, f_in, sample, threads)
        log("cpu baseline %.3g nodes/s" % cpu["value"])

    out = {
        "metric": METRIC,
        "value": M / (ms_step * 1e-3),
        "unit": "nodes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "3-D NFFT trafo+adjoint, FP32, N=128^3, sigma=2 (n=256^3), Kaiser-Bessel m=4, "
                        "M=2^24 kooshball nodes (16384 spokes x 1024 samples), device-resident",
            "N": N, "n": n, "m": m, "M": M, "parallelism": "single GPU",
            "planning": "FFTW_ESTIMATE" if args.estimate else "FFTW_MEASURE",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom_name,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": "profiles/traffic_kooshball3d.json (rocprofv3 FETCH_SIZE + WRITE_SIZE "
                              "per dispatch, scaled by a 1 GiB copy kernel's readings)",
            "avg_launch_ms": dom_ms,
            "algorithmic_bytes_per_launch": ab["B"],
            # the same launch against the on-chip peak that bounds it: 4
            # flops per tap (complex value times real weight), (2m+2)^d taps
            # per node, against the published FP32 vector peak (157.3 TFLOP/s,
            # which counts packed FP32) and the unpacked FMA issue rate this
            # library's code runs at (78.6 TFLOP/s)
            "onchip": {
                "bound": "valu",
                "achieved": 4.0 * taps / (dom_ms * 1e-3) / 1e12,
                "peak": FP32_PEAK_TFLOPS,
                "unit": "TFLOP/s",
                "frac": 4.0 * taps / (dom_ms * 1e-3) / 1e12 / FP32_PEAK_TFLOPS,
                "frac_of_unpacked_fma": 4.0 * taps / (dom_ms * 1e-3) / 1e12 / VALU_FP32_TFLOPS,
                "flops_per_launch": 4.0 * taps,
            },
            # which of the two fractions is higher, i.e. the roof this kernel
            # is nearer to
            "limiter": ("valu" if 4.0 * taps / (dom_ms * 1e-3) / 1e12 / FP32_PEAK_TFLOPS
                        > achieved / HBM_PEAK_GBS else "hbm"),
        },
        "step_hbm": {
            "algorithmic_bytes": step_bytes,
            "achieved_GBs": step_bytes / (ms_step * 1e-3) / 1e9,
            "frac": step_bytes / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
        },
        "stages_ms": stages_ms,
        "stages": stages,
        "precompute_ms": precompute_ms,
        "cpu_baseline": cpu,
    }
    # accuracy of the timed FP32 results: against the FP64 plan on the same
    # (FP32-representable) inputs over all nodes, and the FP64 plan against
    # the FP64 oracle on 2000 sampled nodes
    t32 = d_f.cpu().numpy().astype(np.complex128)
    a32 = d_fh_out.cpu().numpy().astype(np.complex128)
    xr = p.x.astype(np.float64).copy()
    p.finalize()
    if not args.no_error:
        log("accuracy: FP64 plan and oracle sample ...")
        fh64 = fh.astype(np.complex64).astype(np.complex128)
        f64 = f_in.astype(np.complex64).astype(np.complex128)
        q = nf.Plan(N, M, n=n, m=m, flags=flags, precision="double")
        q.x[:] = xr
        q.precompute()
        q.f_hat[:] = fh64
        q.trafo()
        t64 = q.f.astype(np.complex128).copy()
        q.f[:] = f64
        q.adjoint()
        a64 = q.f_hat.astype(np.complex128).copy()
        idx = np.random.default_rng(9).choice(M, size=2000, replace=False)
        fs = np.zeros(M, dtype=np.complex128)
        fs[idx] = f64[idx]
        q.f[:] = fs
        q.adjoint()
        a64s = q.f_hat.astype(np.complex128).copy()
        q.finalize()
        import oracle as O
        O.set_threads(min(16, len(os.sched_getaffinity(0))))
        osamp, _ = oracle_sample_errors(N, n, m, xr, fh64, f64, t64, a64s, 2000, 9)
        O.set_threads(1)
        out["error"] = {
            "fp32_vs_fp64": {"trafo": err_stats(t32, t64), "adjoint": err_stats(a32, a64),
                             "vs": "FP64 GPU plan, all nodes, same inputs"},
            "fp64_vs_oracle": dict(osamp, vs="FP64 oracle on 2000 sampled nodes"),
            "target_rel_l2": 1e-5,
        }
        log("accuracy: FP32 vs FP64 trafo rel l2 %.2e linf %.2e, adjoint rel l2 %.2e linf %.2e"
            % (out["error"]["fp32_vs_fp64"]["trafo"]["rel_l2"],
               out["error"]["fp32_vs_fp64"]["trafo"]["rel_linf"],
               out["error"]["fp32_vs_fp64"]["adjoint"]["rel_l2"],
               out["error"]["fp32_vs_fp64"]["adjoint"]["rel_linf"]))
    if not args.no_batch:
        out["batched"] = batch_leg(args)
        out["other_configs"] = configs_leg(args)
    if not args.no_cgnr:
        out["cgnr"] = cgnr_leg(args)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
