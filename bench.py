"""Benchmark: qLogNEI value+grad evals/s (n=2048, d=12, b=512, q=4, S=512) on one
MI355X, mixed precision vs the plain FP64 GPU path, plus MLL-grad/s (n=4096, d=16).

One "step" is one batched qLogNEI value + d/dX over all 512 restarts (one
L-BFGS-B step with every restart active), inputs resident in HBM.  ``value`` is
steps/s of the mixed path.  Baseline legs (FP64 oracle on PyTorch-ROCm, FP64
oracle on the host CPU) use the test oracle as the comparator only.

Prints ONE JSON line.  Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import torch

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "qLogNEI value+grad evals/s (n=2048,d=12), mixed vs FP64 GPU; MLL-grad/s"
PEAK_TFLOPS = {"f32": 157.3, "f64": 78.6, "f16": 2500.0}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _timed(fn, steps, warmup):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def _timed_median(fn, steps, warmup, reps=3):
    """Median over reps back-to-back timings of steps calls (after warmup)."""
    ts = [_timed(fn, steps, warmup if r == 0 else 0) for r in range(reps)]
    return sorted(ts)[len(ts) // 2]


def build_headline(precision: str, n=2048, d=12, b=512, q=4, S=512):
    """The product acquisition of the headline workload (inputs on the GPU)."""
    import botorch_amd as B
    from botorch_amd import workloads

    w = workloads.gp_headline(n=n, d=d, b=b, q=q, S=S)
    dev = "cuda"
    base = B.MaternKernel(ard_num_dims=d)
    base.lengthscale = w["lengthscale"]
    covar = B.ScaleKernel(base)
    covar.outputscale = 1.0
    lik = B.GaussianLikelihood()
    lik.noise = w["noise"]
    gp = B.SingleTaskGP(w["X"].to(dev), w["Y"].to(dev), likelihood=lik, covar_module=covar)
    with torch.no_grad():
        Xb = B.prune_inferior_points(
            gp, w["X"].to(dev), sampler=B.SobolQMCNormalSampler(torch.Size([2048]), seed=123))
    sampler = B.SobolQMCNormalSampler(torch.Size([S]), seed=0)
    acq = B.qLogNoisyExpectedImprovement(gp, Xb, sampler=sampler, prune_baseline=False,
                                         precision=precision)
    return acq, w, Xb


def build_comparator(w, Xb, S, dev="cuda"):
    """FP64 oracle on the same inputs: the accuracy checker and the plain FP64
    PyTorch baseline leg."""
    from oracle import gp_oracle as O
    orc = O.OracleGP(w["X"].to(dev), w["Y"].to(dev), O.MATERN52, w["lengthscale"].to(dev), 1.0,
                     w["noise"])
    return O.OracleQLogNEI(orc, Xb, num_samples=S, seed=0)


def headline_flops(n, d, b, q, n_b, S):
    """Algorithmic FLOPs of one value+grad step (triangular contraction with
    L^-1 forward and adjoint, kernel rows, Sigma_qb, samples, reductions)."""
    m = n_b + q
    return (2 * b * q * n * n + 3 * b * q * n * (2 * d + 8) + 6 * b * q * n * n_b
            + 6 * b * q * m * S + 40 * b * q * S)


def gemm_fwd_roofline(acq, X, steps, precision):
    """Times the forward contraction GEMM of a step on its own (same arguments
    the step issues) with events on the launch stream."""
    from botorch_amd import _lib
    L = _lib.lib()
    cache = acq._keep["cache"]
    n_pad = cache["n_pad"]
    b, q, d = X.shape
    R_pad = _lib.round_up(b * q, 128)
    dt = torch.float64 if precision == "fp64" else torch.float32
    vt = torch.empty(R_pad, n_pad, dtype=dt, device="cuda")
    linv = acq._keep["linv"]
    if precision == "mixed":
        # K_x planes (values in [0, 1)) and V planes out, as the step issues it
        kx = torch.rand(2, R_pad, n_pad, device="cuda").half()
        vh = torch.empty(2, R_pad, n_pad, dtype=torch.float16, device="cuda")
        pl = linv[0].numel() * 2   # bytes per L^-1 plane
        lscale = acq._state.linv_scale

        def launch():
            _lib.check(L.bta_gemm_f16x3(_lib.GEMM_B_LOWER, R_pad, n_pad, n_pad, 1.0,
                                        _lib.ptr(kx[0]), _lib.ptr(kx[1]), n_pad, None, 1.0,
                                        _lib.ptr(linv[0]), _lib.ptr(linv[1]), n_pad, lscale,
                                        _lib.ptr(vt), n_pad, _lib.ptr(vh[0]), _lib.ptr(vh[1]),
                                        1.0, 0, _lib.stream()))
    else:
        kx = torch.rand(R_pad, n_pad, dtype=dt, device="cuda")
        flags = {"fp64": 0, "mixed_f32": _lib.GEMM_F32 | _lib.GEMM_ACC64}[precision]

        def launch():
            _lib.check(L.bta_gemm(flags | _lib.GEMM_B_LOWER, R_pad, n_pad, n_pad, 1.0,
                                  _lib.ptr(kx), n_pad, _lib.ptr(linv), n_pad, 0.0, None, 0,
                                  _lib.ptr(vt), n_pad, 1, 0, 0, 0, 0, _lib.stream()))

    for _ in range(3):
        launch()
    s0, s1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s0.record()
    for _ in range(steps):
        launch()
    s1.record()
    torch.cuda.synchronize()
    avg_ms = s0.elapsed_time(s1) / steps
    # algorithmic flops: K_x L^-T over the lower triangle of L^-1
    n = cache["model"].n
    flops = 2.0 * b * q * (n * (n + 1) / 2)
    return avg_ms, flops


TRAFFIC_FILE = os.path.join(HERE, "profiles", "gemm_h2_kernel_qlognei_n2048_b512_q4_pmc.txt")


def committed_traffic():
    """HBM bytes per launch of the forward contraction, from the committed
    FETCH_SIZE / WRITE_SIZE passes (line ``traffic_bytes_per_launch = N``)."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch"):
                    return int(float(line.split("=")[1]))
    except OSError:
        pass
    return None


def roofline_only(reps=10):
    """Counter passes: a plain copy of known size, then the forward contraction
    of the mixed step, each launched ``reps`` times after the set-up work."""
    acq, w, _ = build_headline("mixed")
    X = w["Xq"].to("cuda")
    acq._launch(X, True)
    nbytes = 256 << 20
    src = torch.rand(nbytes // 4, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(reps):
        torch.neg(src, out=dst)
    torch.cuda.synchronize()
    gemm_fwd_roofline(acq, X, reps, "mixed")
    torch.cuda.synchronize()
    log(f"copy kernel: {nbytes} bytes read + {nbytes} bytes written per launch, {reps} launches")


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline_headline(w, Xb, budget_s=15.0):
    """FP64 oracle (PyTorch CPU, the reference's op sequence) timed on full
    512-restart value+grad steps of the headline workload on the host cores."""
    from oracle import gp_oracle as O
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    torch.set_num_threads(cores)
    orc = O.OracleGP(w["X"], w["Y"], O.MATERN52, w["lengthscale"], 1.0, w["noise"])
    ref = O.OracleQLogNEI(orc, Xb.cpu(), num_samples=w["S"], seed=0)
    X = w["Xq"]
    ref.value_and_grad(X)  # warm-up
    times = []
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s or len(times) < 3:
        t1 = time.perf_counter()
        ref.value_and_grad(X)
        times.append(time.perf_counter() - t1)
    times.sort()
    med = times[len(times) // 2]
    b = X.shape[0]
    return {"value": 1.0 / med, "unit": "value+grad evals/s", "cores": torch.get_num_threads(),
            "kind": "port", "cpu_model": _cpu_model(),
            "sample": f"{len(times)} full {b}-restart value+grad steps "
                      f"({sum(times):.1f}s), median {med * 1e3:.1f} ms per step"}


def bench_batched(steps=20, warmup=3):
    """The batched configuration: 64 uneven-n Matern-5/2 GPs, each scoring 256
    restarts x q = 2 with 256 samples (value + grad, mixed), as one batched step
    (five launches for all models), that step replayed from a HIP graph, and the
    same models stepped one after another on one stream."""
    import botorch_amd as B
    from botorch_amd import workloads
    ws = workloads.gp_batch()
    acqs = []
    for w in ws:
        d = w["X"].shape[-1]
        base = B.MaternKernel(ard_num_dims=d)
        base.lengthscale = w["lengthscale"]
        covar = B.ScaleKernel(base)
        covar.outputscale = w["outputscale"]
        lik = B.GaussianLikelihood()
        lik.noise = w["noise"]
        gp = B.SingleTaskGP(w["X"].to("cuda"), w["Y"].to("cuda"), likelihood=lik,
                            covar_module=covar)
        with torch.no_grad():
            Xb = B.prune_inferior_points(
                gp, w["X"].to("cuda"),
                sampler=B.SobolQMCNormalSampler(torch.Size([2048]), seed=123))
        acqs.append(B.qLogNoisyExpectedImprovement(
            gp, Xb, sampler=B.SobolQMCNormalSampler(torch.Size([w["S"]]), seed=0),
            prune_baseline=False, precision="mixed"))
    Xs = [w["Xq"].to("cuda").contiguous() for w in ws]
    batch = B.BatchedQLogNEI(acqs, 256, q=2)
    for buf, x in zip(batch.X, Xs):
        buf.copy_(x)
    t_batch = _timed(batch.launch, steps, warmup)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        batch.launch()
    t_graph = _timed(graph.replay, steps, warmup)

    def sequential():
        for acq, x in zip(acqs, Xs):
            acq._launch(x, True)

    t_seq = _timed(sequential, max(5, steps // 4), 2)
    ns = [w["X"].shape[0] for w in ws]
    nb = sum(len(w["Xq"]) for w in ws)
    log(f"batched 64 models: {t_batch * 1e3:.3f} ms/step (graph {t_graph * 1e3:.3f}), "
        f"one model after another {t_seq * 1e3:.3f} ms")
    return {"workload": "64 Matern-5/2 GPs, n log-uniform in [128, 4096] (here "
                        f"{min(ns)}..{max(ns)}), d in 4..20, 256 restarts x q=2 x S=256 each, "
                        "value+grad, mixed",
            "models": len(ws), "ms_per_step": t_batch * 1e3, "graph_ms_per_step": t_graph * 1e3,
            "sequential_ms_per_step": t_seq * 1e3, "speedup_vs_sequential": t_seq / t_batch,
            "restart_evals_per_s": nb / t_batch}


def bench_optimize(acq, X, maxiter=50, host_iters=3):
    """gen_candidates_scipy on the headline acquisition from the headline's
    restarts: the device L-BFGS-B (all restarts' optimiser state on the GPU, one
    optimiser kernel per round) for maxiter iterations, and the host loop of
    SciPy's setulb per restart (the reference's optimiser) for host_iters
    iterations; both per round of one value+grad step for all restarts."""
    from botorch_amd.optim import fmin_l_bfgs_b_device, gen_candidates_scipy
    import numpy as np
    b, q, d = X.shape

    def evaluate(x_eval):
        return acq._launch(x_eval.view(b, q, d), True)

    bounds = (np.zeros(q * d), np.ones(q * d))
    x0 = X.reshape(b, -1)
    fmin_l_bfgs_b_device(evaluate, x0, bounds, fscale=-1.0, maxiter=3)  # warm-up
    torch.cuda.synchronize()
    t = time.perf_counter()
    x, f, info = fmin_l_bfgs_b_device(evaluate, x0, bounds, fscale=-1.0, maxiter=maxiter)
    torch.cuda.synchronize()
    t_dev = time.perf_counter() - t
    rounds = max(i["nfev"] for i in info)
    t = time.perf_counter()
    fmin_l_bfgs_b_device(evaluate, x0, bounds, fscale=-1.0, maxiter=maxiter, use_graph=True)
    torch.cuda.synchronize()
    t_graph = time.perf_counter() - t
    host_rounds = [0]

    class Counting:  # the same acquisition, counting the host loop's rounds
        def __init__(self, a):
            self.a = a
            self.X_baseline = a.X_baseline

        def value_and_grad(self, Xq):
            host_rounds[0] += 1
            return self.a.value_and_grad(Xq)

        def __call__(self, Xq):
            return self.a(Xq)

    torch.cuda.synchronize()
    t = time.perf_counter()
    gen_candidates_scipy(X, Counting(acq), 0.0, 1.0,
                         options={"maxiter": host_iters, "device_lbfgsb": False})
    torch.cuda.synchronize()
    t_host = time.perf_counter() - t
    log(f"optimize: device L-BFGS-B {t_dev * 1e3:.1f} ms for {rounds} rounds "
        f"({t_dev / rounds * 1e3:.3f} ms/round, graph {t_graph / rounds * 1e3:.3f}); "
        f"host SciPy loop {t_host / host_rounds[0] * 1e3:.2f} ms/round")
    return {"workload": f"gen_candidates_scipy L-BFGS-B from the {b} headline restarts "
                        f"(q={q}, d={d}, n_vars={q * d}), maxiter={maxiter}",
            "device_ms": t_dev * 1e3, "rounds": rounds,
            "device_ms_per_round": t_dev / rounds * 1e3,
            "device_graph_ms_per_round": t_graph / rounds * 1e3,
            "mean_iterations": float(np.mean([i["nit"] for i in info])),
            "host_scipy_ms_per_round": t_host / host_rounds[0] * 1e3,
            "host_rounds_timed": host_rounds[0],
            "speedup_device_vs_host_loop": (t_host / host_rounds[0]) / (t_dev / rounds)}


def build_mll(n=4096, d=16):
    from botorch_amd import workloads
    w = workloads.gp_mll(n=n, d=d)
    X = w["X"].to("cuda")
    f = workloads.prior_draw(X, w["z"].to("cuda"), w["lengthscale"])
    y = f + w["noise_sd"] * w["eps"].to("cuda")
    y = (y - y.mean()) / y.std()
    return X.contiguous(), y.contiguous()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-baselines", action="store_true")
    ap.add_argument("--no-mll", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-optim", action="store_true")
    ap.add_argument("--roofline-only", action="store_true",
                    help="launch only a copy and the forward contraction (counter passes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported here "
              "(restarts shard across GPUs by design, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(2)
    if args.roofline_only:
        roofline_only()
        return
    from botorch_amd import _lib
    from oracle import gp_oracle as O

    n, d, b, q, S = 2048, 12, 512, 4, 512
    out = {}
    # ---------------- mixed-precision HIP path (the measured value) --------------
    log("building headline model / acquisition (mixed)")
    acq_m, w, Xb = build_headline("mixed", n, d, b, q, S)
    X = w["Xq"].to("cuda")
    n_b = Xb.shape[0]
    # the measured tuner picks the split products' plans for this shape first
    plans = acq_m.autotune(X)
    log(f"autotuned plans {plans}")

    def step_mixed():
        v, g, st = acq_m._launch(X, True)
        return v, g

    log("timing mixed steps")
    t_mixed = _timed(step_mixed, args.steps, args.warmup)
    v_m, g_m, st_m = acq_m._launch(X, True)
    acq_m.check_status(st_m)
    log(f"mixed: {t_mixed * 1e3:.3f} ms/step")
    # the same step captured once in a HIP graph and replayed (inputs copied
    # into the captured buffer by the caller between steps)
    Xg = X.clone()
    acq_m._launch(Xg, True)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        acq_m._launch(Xg, True)
    t_graph = _timed(graph.replay, args.steps, args.warmup)
    log(f"mixed, graph replay: {t_graph * 1e3:.3f} ms/step")
    # ---------------- FP64 and FP32-contraction HIP paths -------------------------
    acq_d, *_ = build_headline("fp64", n, d, b, q, S)
    t_fp64 = _timed(lambda: acq_d._launch(X, True), max(10, args.steps // 2), args.warmup)
    v_d, g_d, st_d = acq_d._launch(X, True)
    acq_d.check_status(st_d)
    log(f"fp64: {t_fp64 * 1e3:.3f} ms/step")
    acq_f, *_ = build_headline("mixed_f32", n, d, b, q, S)
    t_f32 = _timed(lambda: acq_f._launch(X, True), max(10, args.steps // 2), args.warmup)
    v_f, g_f, st_f = acq_f._launch(X, True)
    acq_f.check_status(st_f)
    log(f"mixed_f32: {t_f32 * 1e3:.3f} ms/step")
    # ---------------- accuracy vs the FP64 oracle (same inputs, on device) ---------
    ref = build_comparator(w, Xb, S)
    v_ref, g_ref = ref.value_and_grad(X)

    def rel(a, r):
        return float(((a - r).abs().max() / r.abs().max()).item())

    acc = {"mixed_value": rel(v_m, v_ref), "mixed_grad": rel(g_m, g_ref),
           "mixed_f32_value": rel(v_f, v_ref), "mixed_f32_grad": rel(g_f, g_ref),
           "fp64_value": rel(v_d, v_ref), "fp64_grad": rel(g_d, g_ref)}
    log(f"accuracy {acc}")
    # ---------------- plain FP64 GPU path (oracle through PyTorch-ROCm) -----------
    t_torch = None
    if not args.no_baselines:
        t_torch = _timed(lambda: ref.value_and_grad(X), 5, 2)
        log(f"torch fp64 GPU: {t_torch * 1e3:.3f} ms/step")
    # ---------------- roofline of the dominant kernel ------------------------------
    gemm_ms, gemm_flops = gemm_fwd_roofline(acq_m, X, 50, "mixed")
    achieved = gemm_flops / (gemm_ms * 1e-3) / 1e12
    # split-FP16: each algorithmic FMA costs three FP16 MFMA FMAs
    peak = PEAK_TFLOPS["f16"] / 3.0
    roof = {"bound": "mfma", "achieved": round(achieved, 3), "peak": round(peak, 1),
            "unit": "TFLOP/s", "frac": round(achieved / peak, 4), "traffic": committed_traffic(),
            "kernel": "gemm_h2_kernel V = K_x L^-T (split-FP16 planes, 3 products)",
            "peak_basis": "dense FP16 MFMA 2500 TFLOP/s / 3 products",
            "avg_launch_ms": round(gemm_ms, 5), "flops_per_launch": gemm_flops}
    gemm_ms32, _ = gemm_fwd_roofline(acq_f, X, 30, "mixed_f32")
    roof_f32 = {"achieved": round(gemm_flops / (gemm_ms32 * 1e-3) / 1e12, 3),
                "peak": PEAK_TFLOPS["f32"], "avg_launch_ms": round(gemm_ms32, 5),
                "kernel": "gemm_kernel<float,...,ACC64> V = K_x L^-T"}
    step_flops = headline_flops(n, d, b, q, n_b, S)
    # ---------------- MLL gradient ----------------------------------------------
    mll = None
    if not args.no_mll:
        import ctypes
        log("MLL n=4096 d=16")
        Xm, ym = build_mll()
        nm, dm = Xm.shape
        L = _lib.lib()
        ws = _lib.workspace(L.bta_mll_workspace_size(nm, dm), "cuda")
        ls = [0.4] * dm
        lsh = (ctypes.c_double * dm)(*ls)
        outb = (ctypes.c_double * (4 + dm))()

        def mll_step():
            _lib.check(L.bta_mll_value_grad(0, _lib.ptr(Xm), _lib.ptr(ym), nm, dm, lsh, 1.0,
                                            0.05**2, 0.0, 6, outb, _lib.ptr(ws), ws.numel(),
                                            _lib.stream()))

        t_mll = _timed_median(mll_step, 10, 5)
        g_hip = torch.tensor(list(outb), dtype=torch.float64)
        log(f"MLL hip: {t_mll * 1e3:.2f} ms")
        lsr = torch.full((dm,), 0.4, dtype=torch.float64, device="cuda", requires_grad=True)
        osr = torch.tensor(1.0, dtype=torch.float64, device="cuda", requires_grad=True)
        nr = torch.tensor(0.05**2, dtype=torch.float64, device="cuda", requires_grad=True)
        cr = torch.tensor(0.0, dtype=torch.float64, device="cuda", requires_grad=True)

        def torch_mll():
            for p in (lsr, osr, nr, cr):
                p.grad = None
            ll = O.exact_mll(O.MATERN52, Xm, ym, lsr, osr, nr, cr) * nm
            ll.backward()
            return ll

        t_mll_torch = None
        ll = torch_mll()
        g_ref_m = torch.cat([torch.stack([ll.detach(), osr.grad, nr.grad, cr.grad]),
                             lsr.grad]).cpu()
        if not args.no_baselines:
            t_mll_torch = _timed_median(torch_mll, 3, 1)
            log(f"MLL torch fp64: {t_mll_torch * 1e3:.2f} ms")
        mll = {"workload": "exact MLL value+grad, n=4096, d=16, Matern-5/2 ARD",
               "mll_grad_per_s": 1.0 / t_mll, "ms": t_mll * 1e3,
               "torch_fp64_gpu_per_s": (1.0 / t_mll_torch) if t_mll_torch else None,
               "speedup_vs_fp64_gpu": (t_mll_torch / t_mll) if t_mll_torch else None,
               "max_rel_err_grad": float(((g_hip[1:] - g_ref_m[1:]).abs().max()
                                          / g_ref_m[1:].abs().max()).item()),
               "rel_err_loglik": float(abs(g_hip[0] - g_ref_m[0]) / abs(g_ref_m[0])),
               "flops": nm**3 / 3 + 2 * nm**3 / 3 + nm * nm * (5 * dm + 18) / 2}
    # ---------------- batched uneven models -------------------------------------------
    batched = None
    if not args.no_batch:
        log("batched uneven models")
        batched = bench_batched()
    # ---------------- optimiser rounds -------------------------------------------------
    optim = None
    if not args.no_optim:
        log("optimize rounds")
        optim = bench_optimize(acq_m, X)
    # ---------------- CPU baseline --------------------------------------------------
    cpu = None
    if not args.no_cpu:
        log("cpu baseline")
        cpu = cpu_baseline_headline(w, Xb)
    res = {
        "metric": METRIC, "value": 1.0 / t_mixed, "unit": "value+grad evals/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t_mixed * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 split into 2x f16 (MFMA contraction, f64 sums) + f64 (factors, MC)",
        "data": "synthetic",
        "config": {"workload": "qLogNEI value+grad, SingleTaskGP Matern-5/2 ARD n=2048 d=12, "
                               "b=512 restarts x q=4, S=512 Sobol samples",
                   "n": n, "d": d, "restarts": b, "q": q, "mc_samples": S, "n_baseline": n_b,
                   "precision": "mixed"},
        "restart_evals_per_s": b / t_mixed,
        "mixed_graph_replay_ms_per_step": t_graph * 1e3,
        "fp64_hip_steps_per_s": 1.0 / t_fp64,
        "mixed_f32_hip_steps_per_s": 1.0 / t_f32,
        "roofline_mixed_f32_gemm": roof_f32,
        "torch_fp64_gpu_steps_per_s": (1.0 / t_torch) if t_torch else None,
        "speedup_mixed_vs_fp64_gpu": (t_torch / t_mixed) if t_torch else None,
        "speedup_fp64hip_vs_fp64_gpu": (t_torch / t_fp64) if t_torch else None,
        "step_tflops_mixed": step_flops / t_mixed / 1e12,
        "accuracy_max_rel_err_vs_fp64_oracle": acc,
        "roofline": roof, "cpu_baseline": cpu, "mll": mll, "batched_uneven": batched,
        "optimize_acqf_rounds": optim,
        "autotuned_plans": plans,
    }
    print(json.dumps(res))


if __name__ == "__main__":
    main()
