"""Benchmark: voxel fits/s of the batched LM fitter on 1,048,576-voxel IVIM.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload (16 b-values linspace(0, 800) s/mm^2, 128x128x64 = 1,048,576
voxels, float32, Rician SNR ~ U[20, 100] per voxel, seeded): one step = the
default initial guess (segmented log-linear fits, one kernel) + state reset +
the fused LM fit (one kernel, at most 100 iterations), inputs resident in HBM.
Prints one JSON line with the roofline of the LM kernel (algorithmic FP32
FLOPs of the iterations actually executed / kernel time, events on the launch
stream) and the CPU oracle timed on a sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "voxel fits/sec at matched final residual: MLP-init LM vs plain LM, 1M-voxel IVIM"
PEAK_FP32 = 157.3e12   # MI355X FP32 vector (packed) FLOP/s
PEAK_HBM = 8.0e12


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(w, p0, lb, ub, gpu_params, gpu_ssr, sample, min_seconds=10.0):
    """Oracle (FP32 restatement, OpenMP) on a sample of the workload."""
    import numpy as np
    import oracle
    threads = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit():
        threads = min(threads, int(omp))
    y = np.ascontiguousarray(w["data"][:, :sample])
    p0s, lbs, ubs = (np.ascontiguousarray(a[:, :sample]) for a in (p0, lb, ub))
    runs, t0 = 0, time.perf_counter()
    while True:
        ref = oracle.lm_run("ivim", w["consts"], y, lbs, ubs, p0s, 100, nthreads=threads)
        runs += 1
        el = time.perf_counter() - t0
        if el >= min_seconds or runs >= 50:
            break
    ssr = oracle.ssr("ivim", w["consts"], y, ref["params"], nthreads=threads)
    rel_f = np.abs(gpu_ssr[:sample] - ssr) / np.abs(ssr)
    conv = (ref["step_type"] == 0)
    rel_p = (np.abs(gpu_params[:, :sample] - ref["params"]) /
             np.maximum(np.abs(ref["params"]), 1e-7)).max(axis=0)
    return ({"value": sample * runs / el, "unit": "voxel-fits/s", "cores": threads,
             "kind": "port",
             "sample": f"first {sample} voxels of the same workload, {runs} full fits "
                       f"(FP32 restatement of the reference iteration, OpenMP)"},
            {"sample": sample,
             "final_ssr_rel_le_1e-4": float(np.mean(rel_f <= 1e-4)),
             "final_ssr_rel_p99": float(np.percentile(rel_f, 99)),
             "params_rel_le_1e-4_converged": float(np.mean(rel_p[conv] <= 1e-4)),
             "oracle_converged": float(conv.mean())})


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--voxels", type=int, default=1 << 20)
    ap.add_argument("--cpu-sample", type=int, default=65536)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): voxel fits need no "
              "exchange, so N GPUs would each fit an independent shard; not measured here.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from pycompute_amd import fit, synth

    dev = torch.device("cuda:0")
    w = synth.ivim_c2(args.voxels, seed=2)
    prob = fit.make_problem("ivim", w["consts"], w["data"], device=dev)
    N, M = prob.N, prob.M
    solver = fit.make_solver(prob)
    pars = torch.empty((4, N), dtype=torch.float32, device=dev)
    fit.default_init(prob, out=pars)
    solver.setup(pars, prob.consts_t, prob.data_t, prob.lb_t, prob.ub_t)
    stream = torch.cuda.current_stream(dev)
    flops = solver.program.flops

    def step(ev=None):
        fit.default_init(prob, out=pars)
        solver.lam_t.fill_(5.0)
        solver.step_type_t.fill_(1)
        solver.n_iter_t.zero_()
        if ev is not None:
            ev[0].record(stream)
        solver.run(100)
        if ev is not None:
            ev[1].record(stream)

    log(f"workload: IVIM, N={N}, M={M}, warmup {args.warmup}, steps {args.steps}")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    events = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(args.steps)]
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(events[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = 1e3 * wall / args.steps
    k_ms = [a.elapsed_time(b) for a, b in events]
    k_avg = sum(k_ms) / len(k_ms)

    n_iter = solver.n_iter_t[0].to(torch.int64)
    conv = (solver.step_type_t[0] == 0).to(torch.int64)
    it_sum = int(n_iter.sum())
    conv_sum = int(conv.sum())
    alg_flops = (it_sum - conv_sum) * flops["F_it"] + conv_sum * flops["F_conv"]
    achieved = alg_flops / (k_avg * 1e-3)
    log(f"LM kernel {k_avg:.3f} ms avg, {it_sum / N:.2f} iterations/voxel, converged "
        f"{conv_sum / N:.4f}, {achieved / 1e12:.2f} TFLOP/s")

    result = {
        "metric": METRIC, "value": N / (ms_step * 1e-3), "unit": "voxel-fits/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded IVIM, Rician SNR~U[20,100])",
        "config": {"workload": "IVIM 1,048,576 voxels (128x128x64), 16 b-values 0-800, plain LM "
                               "(reference default init), max 100 iterations",
                   "voxels": N, "measurements": M, "max_iters": 100, "method": "plain LM",
                   "init": "default segmented log-linear"},
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": PEAK_FP32 / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32, "traffic": None,
                     "kernel": "pca_lm_fit", "kernel_ms": k_avg,
                     "flops_per_iteration": flops["F_it"], "exp_per_iteration": flops["exp_it"],
                     "iterations_per_voxel": it_sum / N, "converged_frac": conv_sum / N},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        fit.default_init(prob, out=pars)
        p0 = pars.cpu().numpy()
        step()
        torch.cuda.synchronize()
        gp = solver.pars_t.cpu().numpy()
        gf = solver.final_f().cpu().numpy()[0]
        cb, par = cpu_baseline(w, p0, prob.lb_t.cpu().numpy(), prob.ub_t.cpu().numpy(), gp, gf,
                               min(args.cpu_sample, N))
        result["cpu_baseline"] = cb
        result["parity"] = par
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
