"""Benchmark of the factor+solve path on one MI355X.

Default workload (the single-GPU configuration of the metric): 3D 7-point
Laplacian on a 100^3 grid (n = 1e6, FP64 SPD), LL^t with grid nested
dissection, one RHS b = A x_true with x_true ~ U[-1,1] (numpy default_rng,
seed 0).

A step is one numerical pass with the matrix and b already resident in HBM:
coeftab fill from the device CSC + factorization + solve + FP64 GMRES
refinement to ||r||_2/||b||_2 <= 1e-13 (1e-14 for convdiff).  The timed steps run the FP32-factor
+ FP64-GMRES path (the headline); the plain FP64 path is timed the same way
alongside.  value = factorization flops of the reference model (THFLOPS, 1e9
scale) per second of a full FP32+IR step.

Other workloads (--workload):
  convdiff  3D 27-point convection-diffusion 80^3, LU with static pivoting,
            FP32+GMRES vs FP64 (same step and line shape);
  batch     256 uneven 2D Laplacians (n_i ~ U{64..512}, seed 1) factored
            and solved (16 RHS each, FP64) in one pass;
  anisodiff forward + backward of the differentiable solve on the
            anisotropic diffusion 64^3 (FP64 LL^t).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload W] [--grid N]
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"fp64": 78.6e12, "fp32": 157.3e12}
HBM_PEAK = 8.0e12


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_instance(pa, A, grid, mixed, eps_refine, factotype, order=None, refine=True, tune=True):
    ip, dp = pa.initParam()
    ip[pa.iparm.factorization] = factotype
    ip[pa.iparm.verbose] = pa.verbose.Not
    ip[pa.iparm.refinement] = pa.refine.GMRES
    # GMRES stops on ||r||_2/||b||_2 (reference refinement/z_refine_gmres.c:169);
    # 1e-13 keeps the normwise inf-norm backward error below 1e-12 at n = 1e6
    dp[pa.dparm.epsilon_refinement] = eps_refine
    if mixed:
        ip[pa.iparm.mixed] = 1
    h = pa.init(ip, dp)
    mt = pa.mtxtype.General if factotype == pa.factotype.LU else pa.mtxtype.Symmetric
    spm = pa.spmatrix(A, mtxtype=mt)
    o = order if order is not None else pa.order.grid(*grid)
    t0 = time.perf_counter()
    pa.task_analyze(h, spm, o)
    t_an = time.perf_counter() - t0
    tuning = None
    if isinstance(tune, (list, tuple)):
        # a schedule chosen by an earlier tuning run, in place before any factorization
        pa.set_tuning(h, tune)
        tuning = {"params": list(tune) + [0] * (4 - len(tune)), "given": True}
    t0 = time.perf_counter()
    pa.task_numfact(h, spm)          # CSC permutation + upload + first factorization
    t_nf = time.perf_counter() - t0
    if tune is True:
        # the schedule autotuner: candidate tile classes / kernels timed on this matrix
        t0 = time.perf_counter()
        tuning = pa.tune(h)
        tuning["seconds"] = time.perf_counter() - t0
        log("tuned schedule", tuning)
    return dict(h=h, ip=ip, dp=dp, spm=spm, order=o, t_analyze=t_an, t_first_numfact=t_nf, refine=refine,
                tuning=tuning)


def step(pa, lib, inst, d_b, d_x, stages=None):
    """One numerical pass on HBM-resident data; returns refinement iterations.
    d_b / d_x are (n) or column-major (nrhs, n) device tensors.
    With `stages` (a dict), the stages are separated by device syncs and
    timed one by one."""
    import torch
    h = inst["h"]
    n = d_b.shape[-1]
    nrhs = 1 if d_b.dim() == 1 else d_b.shape[0]

    def mark(name, t0):
        if stages is None:
            return t0
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        stages[name] = stages.get(name, 0.0) + (t1 - t0) * 1e3
        return t1
    t0 = mark("start", time.perf_counter())
    rc = lib.pastix_subtask_bcsc2ctab(h)
    assert rc == 0, rc
    t0 = mark("fill", t0)
    rc = lib.pastix_subtask_sopalin(h)
    assert rc == 0, rc
    t0 = mark("factor", t0)
    d_x.copy_(d_b)
    rc = lib.pastix_amd_solve_dev(h, pa.trans.NoTrans, nrhs, C.c_void_p(d_x.data_ptr()), n)
    assert rc == 0, rc
    t0 = mark("solve", t0)
    it = 0
    if inst["refine"]:
        it = lib.pastix_amd_refine_dev(h, nrhs, C.c_void_p(d_b.data_ptr()), n, C.c_void_p(d_x.data_ptr()), n)
        assert it >= 0, it
        mark("refine", t0)
    return it


def timed(torch, pa, lib, inst, d_b, d_x, steps, warmup):
    """K timed steps (factorization and solves replayed from hipGraphs), then
    K more steps with the per-level HIP events of the kernel classes recorded
    on the library stream (eager launches), then one step with device syncs
    between the stages."""
    for _ in range(warmup):
        step(pa, lib, inst, d_b, d_x)
    torch.cuda.synchronize()
    iters = []
    t0 = time.perf_counter()
    for _ in range(steps):
        iters.append(step(pa, lib, inst, d_b, d_x))
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / steps

    import numpy as np
    lib.pastix_amd_profile(inst["h"], 1, None)
    step(pa, lib, inst, d_b, d_x)                  # eager warm-up of the profiled path
    torch.cuda.synchronize()
    lib.pastix_amd_profile(inst["h"], 1, None)     # reset the accumulators
    t0 = time.perf_counter()
    for _ in range(steps):
        step(pa, lib, inst, d_b, d_x)
    torch.cuda.synchronize()
    t_prof = (time.perf_counter() - t0) / steps
    prof = np.zeros(9)
    lib.pastix_amd_profile(inst["h"], -1, C.c_void_p(prof.ctypes.data))
    lib.pastix_amd_profile(inst["h"], 0, None)
    stages = {}
    step(pa, lib, inst, d_b, d_x, stages)
    stages.pop("start", None)
    return t, iters, prof, stages, t_prof


def measured_traffic(schedule):
    """HBM bytes per launch of the level GEMM class, measured with rocprofv3
    FETCH_SIZE / WRITE_SIZE passes of this workload and scaled by the same
    counters on a plain copy of known size; the figures and the commands
    that produced them are committed beside the file.  Used only when the
    measurement was taken with the same launch schedule."""
    path = os.path.join(ROOT, "profiles", "gemm_level_fp32_laplacian100_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    if schedule is None or list(d.get("schedule", [])) != (list(schedule) + [0] * 4)[:4]:
        return None
    return {"bytes_per_launch": d["hbm_bytes_per_launch"], "source": os.path.relpath(path, ROOT)}


def host_cpu():
    """Host processor facts of the machine the baseline ran on."""
    info = {"nproc": os.cpu_count()}
    try:
        info["affinity"] = len(os.sched_getaffinity(0))
    except AttributeError:
        info["affinity"] = info["nproc"]
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    info["model"] = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return info


def cpu_baseline(pa, problems, family="laplacian", seconds_target=20.0, full_grid=None, full_thflops=None):
    """The restated PaStiX CPU path with the reference's default static
    scheduling: the analysis' proportional mapping and simulation give each
    host thread its ordered cblk list (SolverMatrix ttsktab), and the oracle's
    static-scheduled factorization runs them, each thread waiting on the
    incoming-update counters (sopalin/sequential_zpotrf.c:104-145), the dense
    per-task kernels on SciPy's BLAS with one BLAS thread per task (the
    reference's sequential per-task MKL).  Timed like DPARM_FACT_TIME: the
    factorization alone, FP64.  A bounded sample: the largest cube of the
    workload's stencil that takes about the time budget; the full workload's
    factorization time is extrapolated at the sample's rate and labelled so."""
    import numpy as np
    import scipy.sparse as sp
    from oracle import pyoracle as orc
    lu = family == "convdiff"
    host = host_cpu()
    threads = max(1, min(16, host["affinity"]))   # the CPU share of one GPU on the box
    sizes = (36, 44, 52, 60, 68) if lu else (50, 60, 70, 80, 90)
    res = None
    for N in sizes:
        A = problems.convection_diffusion_27(N, N, N) if lu else problems.laplacian(N, N, N)
        ip, dp = pa.initParam()
        ip[pa.iparm.factorization] = pa.factotype.LU if lu else pa.factotype.LLT
        ip[pa.iparm.verbose] = pa.verbose.Not
        ip[pa.iparm.thread_nbr] = threads
        h = pa.init(ip, dp)
        spm = pa.spmatrix(A, mtxtype=pa.mtxtype.General if lu else pa.mtxtype.Symmetric)
        pa.task_analyze(h, spm, pa.order.grid(N, N, N))
        cblk, blok, perm, invp, dims = pa.get_structure(h)
        tptr, ttsk = pa.get_schedule(h)
        th = float(dp[pa.dparm.fact_thflops])
        pa.finalize(h)
        S = orc.Structure(cblk, blok)
        crit = np.finfo(np.float64).eps / 2 * sp.linalg.norm(A)
        _, _, _, t = orc.factorize_static(S, A, perm, crit, tptr, ttsk, use_blas=True, factotype=2 if lu else 0)
        rate = th / t
        res = {"value": rate / 1e9, "unit": "GFLOP/s", "cores": threads, "kind": "port",
               "host": host,
               "sample": "restated PaStiX CPU path, static scheduling (proportional mapping + simulation, "
                         "%d thread lists), %s factorization only, FP64, per-task SciPy BLAS (1 thread each), "
                         "%s %d^3 (n=%d): %.3e flops in %.2f s" %
                         (threads, "LU" if lu else "LL^t",
                          "27-point convection-diffusion" if lu else "7-point Laplacian", N, N ** 3, th, t)}
        if full_thflops:
            res["full_workload_factor_s_extrapolated"] = full_thflops / rate
            res["extrapolation"] = "%s^3 factorization time at the sample's rate (not measured)" % full_grid
        log("cpu baseline", res["sample"])
        if t > seconds_target / 4:
            break
    return res


def gemm_roofline(prof, steps, precision):
    """Roofline figures of the level scatter-GEMM class from the per-level HIP
    events (prof: the pastix_amd_profile accumulators)."""
    gemm_ms = prof[2] / steps
    gemm_launches = prof[5] / steps
    gemm_flops = prof[8]
    achieved = gemm_flops / (gemm_ms * 1e-3)
    tname = "float" if precision == "fp32" else "double"
    return {"bound": "mfma",
            "kernel": "k_gemm_level<%s,TM,TN> (level scatter-GEMM, all tile classes of the schedule)" % tname,
            "achieved": achieved / 1e12, "peak": PEAK[precision] / 1e12, "unit": "TFLOP/s",
            "frac": achieved / PEAK[precision], "traffic": None,
            "launches_per_step": gemm_launches, "avg_launch_ms": gemm_ms / max(gemm_launches, 1),
            "algorithmic_flops_per_launch": gemm_flops / max(gemm_launches, 1),
            "algorithmic_flops_per_step": gemm_flops}


def solve_roofline(r, word):
    """HBM roofline of the step's first solve (graph-replayed, timed between
    device syncs): the forward and the backward sweep each read every factor
    entry once (the diagonal blocks as their stored inverses, the same
    count), 2 * coefnbr * word bytes; the right-hand-side traffic (a few
    n-vectors) is left out."""
    ms = r["stages"]["solve"]
    nbytes = 2.0 * r["coefnbr"] * word
    achieved = nbytes / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": achieved / (HBM_PEAK / 1e9), "algorithmic_bytes": nbytes, "ms": ms}


def kernel_ms(prof, steps):
    return {"diag": prof[0] / steps, "trsm": prof[1] / steps, "gemm": prof[2] / steps}


def run_pair(args, torch, pa, lib, A, grid, factotype, d_b, d_x, x_true=None):
    """FP32+IR and plain FP64 (unless --no-fp64) on the same matrix and RHS."""
    import numpy as np
    from pastix_amd import problems
    results = {}
    b = d_b.cpu().numpy()
    for mode in (["fp32ir"] if args.no_fp64 else ["fp32ir", "fp64"]):
        log("mode", mode, ": analysis + first factorization")
        sched = args.schedule if mode == "fp32ir" else args.schedule64
        eps = args.eps if args.eps is not None else 1e-13
        inst = make_instance(pa, A, grid, mixed=(mode == "fp32ir"), eps_refine=eps, factotype=factotype,
                             tune=sched if sched else not args.no_tune)
        th = float(inst["dp"][pa.dparm.fact_thflops])
        dims = np.zeros(8, dtype=np.int64)
        assert lib.pastix_amd_get_dims(inst["h"], C.c_void_p(dims.ctypes.data)) == 0
        t, iters, prof, stages, t_prof = timed(torch, pa, lib, inst, d_b, d_x, args.steps, args.warmup)
        x = d_x.cpu().numpy()
        berr = problems.backward_error(A, x, b)
        ferr = np.abs(x - x_true).max() / np.abs(x_true).max() if x_true is not None else None
        results[mode] = dict(t=t, iters=iters, prof=prof, th=th, berr=berr, ferr=ferr, stages=stages, t_prof=t_prof,
                             t_analyze=inst["t_analyze"], fact_s=float(inst["dp"][pa.dparm.fact_time]),
                             tuning=inst["tuning"], coefnbr=int(dims[3]),
                             x=x if mode == "fp64" else None)
        log("mode %s: %.1f ms/step, berr %.2e, ferr %s, iters %s, fact %.1f ms" %
            (mode, t * 1e3, berr, ferr, iters, results[mode]["fact_s"] * 1e3))
        if mode == "fp32ir":
            results[mode]["x32"] = x
        pa.finalize(inst["h"])
    return results


def pair_line(args, pa, results, metric_config, family):
    """The contract JSON line of an FP32+IR vs FP64 workload."""
    import numpy as np
    r = results["fp32ir"]
    steps = args.steps
    roofline = gemm_roofline(r["prof"], steps, "fp32")
    roofline["events_timed_ms_per_step"] = r["t_prof"] * 1e3
    schedule = r["tuning"]["params"] if r["tuning"] else None
    if family == "laplacian":
        traffic = measured_traffic(schedule)
        if traffic:
            roofline["traffic"] = traffic["bytes_per_launch"]
            roofline["traffic_source"] = traffic["source"]
    line = {
        "metric": "factor+solve time & GFLOP/s, FP32+IR vs FP64 at backward error <=1e-12",
        "value": r["th"] / r["t"] / 1e9,
        "unit": "GFLOP/s",
        "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": r["t"] * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": metric_config,
        "berr": r["berr"], "ferr": r["ferr"], "refine_iters": r["iters"][-1],
        "kernel_ms_per_step": kernel_ms(r["prof"], steps),
        "stage_ms": r["stages"],
        "analysis_s": r["t_analyze"],
        "tuning": r["tuning"],
        "roofline": roofline,
        "solve_roofline": solve_roofline(r, 4),
    }
    if "fp64" in results:
        r64 = results["fp64"]
        line["fp64_ms_per_step"] = r64["t"] * 1e3
        line["fp64_gflops"] = r64["th"] / r64["t"] / 1e9
        line["fp64_berr"] = r64["berr"]
        line["fp64_ferr"] = r64["ferr"]
        line["speedup_fp32ir_vs_fp64"] = r64["t"] / r["t"]
        line["fp64_kernel_ms_per_step"] = kernel_ms(r64["prof"], steps)
        line["fp64_roofline"] = gemm_roofline(r64["prof"], steps, "fp64")
        line["fp64_stage_ms"] = r64["stages"]
        line["fp64_tuning"] = r64["tuning"]
        # forward error of the FP32+IR solution against the FP64 factorization's
        x64 = r64["x"]
        line["ferr_vs_fp64_solution"] = float(np.abs(r["x32"] - x64).max() / np.abs(x64).max())
    if not args.no_cpu:
        from pastix_amd import problems
        cb = cpu_baseline(pa, problems, family, full_grid=metric_config.get("grid"), full_thflops=r["th"])
        line["cpu_baseline"] = cb
        # like for like: factorization-only rates (GPU factorization events
        # of the timed steps against the CPU's factorization-only sample)
        line["gpu_factor_gflops"] = {"fp32": r["th"] / (r["stages"]["factor"] * 1e-3) / 1e9}
        if "fp64" in results:
            line["gpu_factor_gflops"]["fp64"] = results["fp64"]["th"] / (results["fp64"]["stages"]["factor"] * 1e-3) / 1e9
            line["fp64_factor_speedup_vs_cpu_baseline"] = line["gpu_factor_gflops"]["fp64"] / cb["value"]
        line["fp32_factor_speedup_vs_cpu_baseline"] = line["gpu_factor_gflops"]["fp32"] / cb["value"]
    return line


def workload_laplacian(args, torch, pa, lib, dev):
    from pastix_amd import problems
    N = args.grid or 100
    grid = (N, N, N)
    log("generating 7-point Laplacian %d^3" % N)
    A = problems.laplacian(*grid)
    n = A.shape[0]
    xt = problems.uniform(n, seed=0)
    d_b = torch.from_numpy(A @ xt).to(dev)
    d_x = torch.empty_like(d_b)
    results = run_pair(args, torch, pa, lib, A, grid, pa.factotype.LLT, d_b, d_x, x_true=xt)
    cfg = {"workload": "3D 7-point Laplacian %d^3 (n=%d), LL^t, grid nested dissection, "
                       "1 RHS b=A*x_true, x_true~U[-1,1] seed 0" % (N, n),
           "grid": N, "thflops": results["fp32ir"]["th"], "parallelism": "single GPU"}
    return pair_line(args, pa, results, cfg, "laplacian")


def workload_convdiff(args, torch, pa, lib, dev):
    from pastix_amd import problems
    N = args.grid or 80
    grid = (N, N, N)
    log("generating 27-point convection-diffusion %d^3" % N)
    A = problems.convection_diffusion_27(*grid)
    n = A.shape[0]
    # the normwise 2-norm stop at 1e-13 can leave the inf-norm backward error
    # of this nonsymmetric matrix just under 1e-12 after one GMRES iteration:
    # 1e-14 keeps a margin
    if args.eps is None:
        args.eps = 1e-14
    xt = problems.uniform(n, seed=0)
    d_b = torch.from_numpy(A @ xt).to(dev)
    d_x = torch.empty_like(d_b)
    results = run_pair(args, torch, pa, lib, A, grid, pa.factotype.LU, d_b, d_x, x_true=xt)
    cfg = {"workload": "3D 27-point convection-diffusion %d^3 (n=%d, nnz=%d), diffusion 1, upwind velocity "
                       "(0.3,0.2,0.1), LU with static pivoting, grid nested dissection, GMRES, "
                       "1 RHS b=A*x_true, x_true~U[-1,1] seed 0" % (N, n, A.nnz),
           "grid": N, "thflops": results["fp32ir"]["th"], "parallelism": "single GPU"}
    return pair_line(args, pa, results, cfg, "convdiff")


def workload_batch(args, torch, pa, lib, dev):
    """Many uneven SPD systems factored and solved in one pass (FP64, 16 RHS
    each): the block-diagonal forest of 2D Laplacians n_i x n_i, n_i ~
    U{64..512}, diagonal shift ~ U[0,1], seed 1."""
    import numpy as np
    from pastix_amd import batch, problems
    count = args.count or 256
    t0 = time.perf_counter()
    mats, grids = batch.uneven_laplacians(count=count, low=64, high=512, seed=1)
    log("generated %d systems, sum n = %d (%.1f s)" % (count, sum(m.shape[0] for m in mats),
                                                       time.perf_counter() - t0))
    t0 = time.perf_counter()
    s = batch.BatchedSolver(mats, grids, mixed=False)
    t_an = time.perf_counter() - t0
    log("analysis %.1f s" % t_an)
    n = int(s.offsets[-1])
    nrhs = 16
    rng = np.random.default_rng(1)
    # column-major (n x nrhs): row r of the (nrhs, n) tensor is right-hand side r
    B = np.empty((nrhs, n))
    for i, m in enumerate(mats):
        B[:, s.offsets[i]:s.offsets[i + 1]] = rng.uniform(-1, 1, (mats[i].shape[0], nrhs)).T
    d_b = torch.from_numpy(B).to(dev)
    d_x = torch.empty_like(d_b)
    del B
    s.factorize()
    inst = dict(h=s.h, refine=False)
    t, iters, prof, stages, t_prof = timed(torch, pa, lib, inst, d_b, d_x, args.steps, args.warmup)
    th = float(s.dparm[pa.dparm.fact_thflops])
    X = d_x.cpu().numpy()
    Bh = d_b.cpu().numpy()
    berr = 0.0
    for i, m in enumerate(mats):
        sl = slice(s.offsets[i], s.offsets[i + 1])
        berr = max(berr, problems.backward_error(m, X[:, sl].T, Bh[:, sl].T))
    log("batch: %.1f ms/pass, max berr %.2e" % (t * 1e3, berr))
    steps = args.steps
    line = {
        "metric": "factor+solve time & GFLOP/s, FP64 batched pass of many uneven systems",
        "value": th / t / 1e9, "unit": "GFLOP/s",
        "n_gpus": 1, "steps": steps, "warmup": args.warmup, "ms_per_step": t * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "%d 2D 5-point Laplacians n_i x n_i, n_i~U{64..512} + shift~U[0,1] seed 1 "
                               "(sum n=%d), LL^t FP64, 16 RHS~U[-1,1] each, one batched pass" % (count, n),
                   "thflops": th, "parallelism": "single GPU"},
        "berr_max": berr, "systems": count, "nrhs": nrhs,
        "kernel_ms_per_step": kernel_ms(prof, steps), "stage_ms": stages, "analysis_s": t_an,
        "roofline": gemm_roofline(prof, steps, "fp64"),
    }
    s.close()
    return line


def workload_anisodiff(args, torch, pa, lib, dev):
    """Differentiable solve: a step is one forward (device value update,
    FP64 LL^t factorization, solve, refinement) and one backward (adjoint
    solve, sampled outer product on the pattern) of loss = 0.5||x - x_t||^2
    for the anisotropic diffusion 64^3, theta = exp(N(0,1)) seed 2, ratio 10
    in z, b ~ U[-1,1] seed 3, x_t ~ U[-1,1] seed 4."""
    import numpy as np
    import scipy.sparse as sp
    from pastix_amd import problems
    from pastix_amd.autograd import SparseSolver, solve
    N = args.grid or 64
    grid = (N, N, N)
    n = N ** 3
    theta = problems.anisotropic_conductivity(*grid, seed=2)
    P, J = problems.anisotropic_diffusion_jacobian(*grid, ratio_z=10.0)
    b = problems.uniform(n, seed=3)
    xt = problems.uniform(n, seed=4)
    t0 = time.perf_counter()
    s = SparseSolver(P, symmetric=True, grid=grid)
    t_an = time.perf_counter() - t0
    vals0 = torch.from_numpy(J @ theta).to(dev)
    bt = torch.from_numpy(b).to(dev)
    xtt = torch.from_numpy(xt).to(dev)
    s.factor(vals0)                      # first factorization: builds the permuted CSC on the host

    def one_step():
        v = vals0.clone().requires_grad_(True)
        bb = bt.clone().requires_grad_(True)
        x = solve(s, v, bb)
        loss = 0.5 * ((x - xtt) ** 2).sum()
        loss.backward()
        return x, v.grad, bb.grad

    for _ in range(args.warmup):
        one_step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        x, gv, gb = one_step()
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / args.steps
    th = float(s.dparm[pa.dparm.fact_thflops])
    x0 = x.detach().cpu().numpy()
    A = sp.csc_matrix((J @ theta, P.indices, P.indptr), shape=(n, n))
    lam = gb.cpu().numpy()
    adj_res = float(np.abs(A.T @ lam - (x0 - xt)).max() / np.abs(x0 - xt).max())
    berr = problems.backward_error(A, x0, b)
    # the pattern gradient is -lambda_i x_j on every stored entry
    C = sp.csc_matrix(P)
    C.sort_indices()
    cols = np.repeat(np.arange(n), np.diff(C.indptr))
    ref = -lam[C.indices] * x0[cols]
    g = gv.cpu().numpy()
    grad_err = float(np.abs(g - ref).max() / np.abs(ref).max())
    log("anisodiff: %.1f ms/step, berr %.2e, adjoint residual %.2e, grad err %.2e" % (t * 1e3, berr, adj_res, grad_err))
    s.close()
    return {
        "metric": "forward+backward time of the differentiable solve (FP64 LL^t)",
        "value": th / t / 1e9, "unit": "GFLOP/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": t * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "3D 7-point anisotropic diffusion %d^3 (n=%d), theta=exp(N(0,1)) seed 2, ratio 10 in z, "
                               "LL^t FP64; loss 0.5||x-x_t||^2, b~U[-1,1] seed 3, x_t~U[-1,1] seed 4; step = forward "
                               "(device value update, factor, solve, refine) + backward (adjoint solve, sampled "
                               "outer product)" % (N, n),
                   "thflops": th, "parallelism": "single GPU"},
        "berr": berr, "adjoint_residual": adj_res, "pattern_grad_rel_err": grad_err, "analysis_s": t_an,
    }


WORKLOADS = {"laplacian": workload_laplacian, "convdiff": workload_convdiff, "batch": workload_batch,
             "anisodiff": workload_anisodiff}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="laplacian")
    ap.add_argument("--grid", type=int, default=0, help="cube side (default: 100 laplacian, 80 convdiff)")
    ap.add_argument("--count", type=int, default=0, help="systems of the batch workload (default 256)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="keep the default schedule (no autotuning)")
    sched = lambda v: [int(x) for x in v.split(",")]
    ap.add_argument("--schedule", type=sched, default=None,
                    help="tm_min,narrow,small[,tall64] of the FP32 run instead of autotuning (as printed by a tuning run)")
    ap.add_argument("--schedule64", type=sched, default=None, help="the same for the FP64 run")
    ap.add_argument("--eps", type=float, default=None,
                    help="refinement stop on ||r||_2/||b||_2 (default 1e-13 laplacian, 1e-14 convdiff)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; --gpus %d is not supported "
              "(the multi-GPU sharding is a design, not a build)" % args.gpus)
        sys.exit(2)

    import torch  # torch's HIP runtime is loaded first and shared by the engine
    import pastix_amd as pa
    from pastix_amd import _lib
    lib = _lib.load()
    assert torch.cuda.is_available()
    dev = torch.device("cuda:0")
    line = WORKLOADS[args.workload](args, torch, pa, lib, dev)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
