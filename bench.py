"""Benchmark of the factor+solve path on one MI355X.

Workload (the single-GPU configuration of the metric): 3D 7-point Laplacian
on a 100^3 grid (n = 1e6, FP64 SPD), LL^t with grid nested dissection, one RHS
b = A x_true with x_true ~ U[-1,1] (numpy default_rng, seed 0).

A step is one numerical pass with the matrix and b already resident in HBM:
coeftab fill from the device CSC + factorization + solve + refinement to
||r||_2/||b||_2 <= 1e-12.  The timed steps run the FP32-factor + FP64-GMRES
path (the headline); the plain FP64 path is timed the same way alongside.
value = factorization flops of the reference model (THFLOPS, 1e9 scale) per
second of a full FP32+IR step.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--grid 100]
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


def make_instance(pa, A, grid, mixed, eps_refine):
    ip, dp = pa.initParam()
    ip[pa.iparm.factorization] = pa.factotype.LLT
    ip[pa.iparm.verbose] = pa.verbose.Not
    ip[pa.iparm.refinement] = pa.refine.GMRES
    # GMRES stops on ||r||_2/||b||_2 (reference refinement/z_refine_gmres.c:169);
    # 1e-13 keeps the normwise inf-norm backward error below 1e-12 at n = 1e6
    dp[pa.dparm.epsilon_refinement] = eps_refine
    if mixed:
        ip[pa.iparm.mixed] = 1
    h = pa.init(ip, dp)
    spm = pa.spmatrix(A, mtxtype=pa.mtxtype.Symmetric)
    o = pa.order.grid(*grid)
    t0 = time.perf_counter()
    pa.task_analyze(h, spm, o)
    t_an = time.perf_counter() - t0
    t0 = time.perf_counter()
    pa.task_numfact(h, spm)          # CSC permutation + upload + first factorization
    t_nf = time.perf_counter() - t0
    return dict(h=h, ip=ip, dp=dp, spm=spm, order=o, t_analyze=t_an, t_first_numfact=t_nf)


def step(pa, lib, inst, d_b, d_x, stages=None):
    """One numerical pass on HBM-resident data; returns refinement iterations.
    With `stages` (a dict), the stages are separated by device syncs and
    timed: fill, factor, solve, refine (ms)."""
    import torch
    h = inst["h"]

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
    rc = lib.pastix_amd_solve_dev(h, pa.trans.NoTrans, 1, C.c_void_p(d_x.data_ptr()), d_x.numel())
    assert rc == 0, rc
    t0 = mark("solve", t0)
    it = lib.pastix_amd_refine_dev(h, 1, C.c_void_p(d_b.data_ptr()), d_b.numel(),
                                   C.c_void_p(d_x.data_ptr()), d_x.numel())
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


def measured_traffic():
    """HBM bytes per launch of the level GEMM class, measured with rocprofv3
    FETCH_SIZE / WRITE_SIZE passes of this workload and scaled by the same
    counters on a plain copy of known size; the figures and the commands
    that produced them are committed beside the file."""
    path = os.path.join(ROOT, "profiles", "gemm_level_fp32_laplacian100_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return {"bytes_per_launch": d["hbm_bytes_per_launch"], "source": os.path.relpath(path, ROOT)}


def cpu_baseline(pa, problems, seconds_target=20.0):
    """The oracle's sequential LL^t (restated PaStiX CPU path, one core, its
    own dense loops) on a bounded sample of the same workload family: the
    7-point Laplacian on the largest cube that fits the time budget."""
    import numpy as np
    import scipy.sparse as sp
    from oracle import pyoracle as orc
    res = None
    for N in (30, 40, 50):
        A = problems.laplacian(N, N, N)
        ip, dp = pa.initParam()
        ip[pa.iparm.factorization] = pa.factotype.LLT
        ip[pa.iparm.verbose] = pa.verbose.Not
        h = pa.init(ip, dp)
        spm = pa.spmatrix(A, mtxtype=pa.mtxtype.Symmetric)
        pa.task_analyze(h, spm, pa.order.grid(N, N, N))
        cblk, blok, perm, invp, dims = pa.get_structure(h)
        th = float(dp[pa.dparm.fact_thflops])
        pa.finalize(h)
        S = orc.Structure(cblk, blok)
        crit = np.finfo(np.float64).eps / 2 * sp.linalg.norm(A)
        t0 = time.perf_counter()
        orc.factorize(S, A, perm, 0, crit)
        t = time.perf_counter() - t0
        res = {"value": th / t / 1e9, "unit": "GFLOP/s", "cores": 1, "kind": "port",
               "sample": "oracle sequential LL^t (fill+factor) of the 7-point Laplacian %d^3, "
                         "n=%d, %.3e flops, %.2f s" % (N, N ** 3, th, t)}
        log("cpu baseline", res["sample"])
        if t > seconds_target / 4:
            break
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--eps", type=float, default=1e-13)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; --gpus %d is not supported "
              "(the multi-GPU sharding is a design, not a build)" % args.gpus)
        sys.exit(2)

    import torch  # torch's HIP runtime is loaded first and shared by the engine
    import numpy as np
    import pastix_amd as pa
    from pastix_amd import problems, _lib
    lib = _lib.load()
    assert torch.cuda.is_available()
    dev = torch.device("cuda:0")

    N = args.grid
    grid = (N, N, N)
    log("generating 7-point Laplacian %d^3" % N)
    A = problems.laplacian(*grid)
    n = A.shape[0]
    xt = problems.uniform(n, seed=0)
    b = A @ xt
    d_b = torch.from_numpy(b).to(dev)
    d_x = torch.empty_like(d_b)

    results = {}
    for mode in (["fp32ir"] if args.no_fp64 else ["fp32ir", "fp64"]):
        log("mode", mode, ": analysis + first factorization")
        inst = make_instance(pa, A, grid, mixed=(mode == "fp32ir"), eps_refine=args.eps)
        th = float(inst["dp"][pa.dparm.fact_thflops])
        t, iters, prof, stages, t_prof = timed(torch, pa, lib, inst, d_b, d_x, args.steps, args.warmup)
        x = d_x.cpu().numpy()
        berr = problems.backward_error(A, x, b)
        ferr = np.abs(x - xt).max() / np.abs(xt).max()
        results[mode] = dict(t=t, iters=iters, prof=prof, th=th, berr=berr, ferr=ferr, stages=stages, t_prof=t_prof,
                             t_analyze=inst["t_analyze"], fact_s=float(inst["dp"][pa.dparm.fact_time]),
                             refine_res=float(inst["dp"][pa.dparm.relative_error]))
        log("mode %s: %.1f ms/step, berr %.2e, ferr %.2e, iters %s, fact %.1f ms" %
            (mode, t * 1e3, berr, ferr, iters, results[mode]["fact_s"] * 1e3))
        pa.finalize(inst["h"])

    r = results["fp32ir"]
    prof = r["prof"]
    steps = args.steps
    gemm_ms = prof[2] / steps
    gemm_launches = prof[5] / steps
    gemm_flops = prof[8]
    achieved = gemm_flops / (gemm_ms * 1e-3)
    traffic = measured_traffic()
    roofline = {"bound": "mfma",
                "kernel": "k_gemm_level<float,128> + k_gemm_level<float,64> (level scatter-GEMM)",
                "achieved": achieved / 1e12, "peak": PEAK["fp32"] / 1e12, "unit": "TFLOP/s",
                "frac": achieved / PEAK["fp32"],
                "traffic": traffic["bytes_per_launch"] if traffic else None,
                "launches_per_step": gemm_launches, "avg_launch_ms": gemm_ms / max(gemm_launches, 1),
                "algorithmic_flops_per_launch": gemm_flops / max(gemm_launches, 1),
                "algorithmic_flops_per_step": gemm_flops,
                "events_timed_ms_per_step": r["t_prof"] * 1e3}
    if traffic:
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
        "config": {"workload": "3D 7-point Laplacian %d^3 (n=%d), LL^t, grid nested dissection, "
                               "1 RHS b=A*x_true, x_true~U[-1,1] seed 0" % (N, n),
                   "thflops": r["th"], "parallelism": "single GPU"},
        "berr": r["berr"], "ferr": r["ferr"], "refine_iters": r["iters"][-1],
        "kernel_ms_per_step": {"diag": prof[0] / steps, "trsm": prof[1] / steps, "gemm": prof[2] / steps},
        "stage_ms": r["stages"],
        "roofline": roofline,
    }
    if "fp64" in results:
        r64 = results["fp64"]
        line["fp64_ms_per_step"] = r64["t"] * 1e3
        line["fp64_gflops"] = r64["th"] / r64["t"] / 1e9
        line["fp64_berr"] = r64["berr"]
        line["speedup_fp32ir_vs_fp64"] = r64["t"] / r["t"]
        p64 = r64["prof"]
        line["fp64_kernel_ms_per_step"] = {"diag": p64[0] / steps, "trsm": p64[1] / steps,
                                           "gemm": p64[2] / steps}
        line["fp64_gemm_tflops"] = p64[8] / (p64[2] / steps * 1e-3) / 1e12
        line["fp64_stage_ms"] = r64["stages"]
    if not args.no_cpu:
        cb = cpu_baseline(pa, problems)
        line["cpu_baseline"] = cb
        line["speedup_vs_cpu_baseline"] = line["value"] / cb["value"]
    print(json.dumps(line))


if __name__ == "__main__":
    main()
