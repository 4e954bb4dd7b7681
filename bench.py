"""Benchmark: contour-steps/s of the SCFT iteration at 128^3, Ns = 200.

Workload (the configuration the metric is quoted on): AB diblock melt,
f_A = 0.35, chiN = 18, cubic cell a = 4.2, 128^3 grid, Ns = 200
(Ms = [71, 131]), RQM4 propagator, Simpson quadrature, Anderson mixing
(history 20, lam 0.1), seeded double-gyroid template + noise.  One "step" is
one full SCFT iteration (Model_AB::update: four propagators, Q, densities,
Anderson mixing) with every input resident in HBM; it advances 2*Ns = 400
contour steps.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision fp64|fp32]

Prints one JSON line.  value = whole-job contour-steps/s of the plain FP64
path; the FP32-bulk form is timed in the same run and reported beside it.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "traffic_rqm4_step_128cubed.json")


def run_form(precision, steps, warmup, L=128, Ns=200):
    from polystring_amd import capi, workloads
    wl = workloads.cfg2(L, Ns)
    ctx = capi.Context(0, precision)
    m = capi.ModelAB(ctx, wl.wA, wl.wB, **wl.model_kwargs())
    m.update(warmup)
    ctx.synchronize()
    m.set_profiling(True)
    ctx.synchronize()
    t0 = time.perf_counter()
    m.update(steps)
    ctx.synchronize()
    t1 = time.perf_counter()
    prof = m.profile()
    obs = m.observe()
    m.set_profiling(False)
    wall = t1 - t0
    cps = m.contour_steps()
    p = 8 if precision == "fp64" else 4
    B_step = workloads.algorithmic_bytes_per_step(wl.shape, p)
    # dominant kernel: the propagator contour step (k_plane + k_xpass launches of
    # one lockstep step, which advances two chains by one contour step each)
    n_launch = prof["iters"] * prof["lockstep_steps"]
    step_ms = prof["prop_ms"] / n_launch
    achieved = 2 * B_step / (step_ms * 1e-3) / 1e9
    out = dict(
        precision=precision, wall_s=wall, iters=steps, contour_steps=cps * steps,
        contour_steps_per_s=cps * steps / wall, iters_per_s=steps / wall,
        propagator_ms=prof["prop_ms"], density_anderson_ms=prof["rest_ms"],
        lockstep_step_us=step_ms * 1e3, achieved_GBs=achieved, bytes_per_launch=2 * B_step,
        H=obs["H"], residual=obs["residual"], Q=obs["Q"],
    )
    del m
    ctx.close()
    return out


def cpu_baseline(seconds_target=15.0):
    """The FP64 oracle (a C restatement of the reference's RQM4 dataflow, C2C
    transforms) timed on this host: RQM4 contour steps of one propagator at
    128^3, on all host threads the oracle's OpenMP uses."""
    import numpy as np

    from oracle import oracle as O
    from polystring_amd import workloads
    wl = workloads.cfg2()
    l, g = O.cell_shape(3, "CUBIC", 4.2)
    k2 = O.k2_orthogonal(3, l, wl.shape)
    head = np.ones(wl.shape)
    ds = 1.0 / wl.Ns
    O.rqm4_solve(k2, ds, wl.wA, head, 2)  # plan + page-in
    t0 = time.perf_counter()
    n = 2
    O.rqm4_solve(k2, ds, wl.wA, head, n + 1)
    dt = time.perf_counter() - t0
    per = dt / n
    nsteps = max(2, min(40, int(seconds_target / per)))
    t0 = time.perf_counter()
    O.rqm4_solve(k2, ds, wl.wA, head, nsteps + 1)
    dt = time.perf_counter() - t0
    return dict(value=nsteps / dt, unit="contour-steps/s", cores=O.num_threads(), kind="port",
                sample=f"{nsteps} RQM4 contour steps of one propagator at 128^3 FP64 (oracle C restatement, "
                       f"complex-to-complex FFTs as the reference), {dt:.1f} s on {O.num_threads()} threads")


def copy_probe(nbytes=1 << 30, reps=3):
    """Plain device-to-device copy of a known size (torch), the calibration
    kernel for rocprofv3 FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    x = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": nbytes, "copy_probe_bytes_written": nbytes, "reps": reps}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--copy-probe", action="store_true", help="run only the traffic-calibration copy")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (each GPU would run "
              "independent solves / string images); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    f64 = run_form("fp64", args.steps, args.warmup)
    print(f"[bench] fp64 {json.dumps(f64)}", file=sys.stderr, flush=True)
    f32 = None
    if not args.no_fp32:
        f32 = run_form("fp32", args.steps, args.warmup)
        print(f"[bench] fp32 {json.dumps(f32)}", file=sys.stderr, flush=True)

    traffic = None
    if os.path.exists(PROFILE_TRAFFIC):
        with open(PROFILE_TRAFFIC) as fh:
            traffic = json.load(fh).get("hbm_bytes_per_lockstep_step_fp64")
    cpu = None if args.no_cpu_baseline else cpu_baseline()

    line = {
        "metric": "contour-steps/s at 128^3 Ns=200: mixed+adjoint vs FP64 plain, |dH/H|<=1e-6",
        "value": f64["contour_steps_per_s"],
        "unit": "contour-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * f64["wall_s"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded double-gyroid template + Philox Gaussian noise)",
        "config": {"workload": "AB diblock SCFT, RQM4 + Simpson + Anderson(20), 128^3, Ns=200, "
                               "f=0.35, chiN=18, cubic a=4.2 Rg, FP64 plain",
                   "grid": [128, 128, 128], "Ns": 200, "Ms": [71, 131], "step": "one SCFT iteration "
                   "(2*Ns contour steps)", "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": f64["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": f64["achieved_GBs"] / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "RQM4 contour step (k_plane + k_xpass launches of one lockstep step)",
                     "bytes_per_launch": f64["bytes_per_launch"], "avg_launch_us": f64["lockstep_step_us"]},
        "cpu_baseline": cpu,
        "iters_per_s": f64["iters_per_s"],
        "H_final": f64["H"],
    }
    if f32 is not None:
        line["fp32_bulk"] = {"contour_steps_per_s": f32["contour_steps_per_s"], "iters_per_s": f32["iters_per_s"],
                             "speedup_vs_fp64": f32["contour_steps_per_s"] / f64["contour_steps_per_s"],
                             "dH_over_H_vs_fp64": abs(f32["H"] - f64["H"]) / abs(f64["H"]),
                             "achieved_GBs": f32["achieved_GBs"], "lockstep_step_us": f32["lockstep_step_us"]}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
