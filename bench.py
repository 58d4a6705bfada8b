"""Throughput of the xSPDE ensemble integration path on one MI355X.

Workload (one "step" = one full ensemble pass of the hot path):
  2D damped complex Ginzburg-Landau SPDE on a 128x128 periodic lattice,
  da = [(1+i) lap a + a - (1+0.5i)|a|^2 a] dt + 0.05 dW, RK4 interaction
  picture with the FFT propagator, 1024 trajectories (ensembles [64,16,1]),
  101 stored points x 20 steps (dt = 0.002 coarse, 2000 coarse steps) with
  checks on: every trajectory runs the fine (dt/2) and the coarse path,
  6000 trajectory-steps each.  FP32 fields, FP64 moments.  Synthetic input:
  random-phase initial fields and counter-based noise, generated on device.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
Prints one JSON line (rank 0, N=1).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12          # B/s, MI355X HBM3E


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(seconds=12.0):
    """FP64 oracle (numpy restatement of the reference path), 1 core, bounded sample."""
    import numpy as np
    from oracle import forms
    from oracle import xspde as X
    from xspde_matlab_amd.configs import cgl2d, shrink
    E = 8
    p = shrink(cgl2d("fp64"), E=E, e2=1)
    q = X.xpreferences(forms.apply_forms(p))
    q.dtr = q.dt / (2 * q.steps[0])
    prop = X.xpropfactor(q)
    v = X.xrandom(q, 1)
    a = X.initial_field(q, v)
    n = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        w = X.noise_draw(q, 1, n)
        a = X.RK4(a, w, q, prop)
        n += 1
    dt = time.perf_counter() - t0
    return dict(value=n * E / dt, unit="trajectory-steps/s", cores=1, kind="port",
                sample="oracle RK4IP fp64 numpy, 128x128 CGL, %d trajectories x %d steps (%.1f s)" % (E, n, dt))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--precision", default="fp32")
    ap.add_argument("--boundaries", default="periodic", choices=["periodic", "neumann"],
                    help="periodic (FFT) or Neumann-Neumann (DCT-I) lattice")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs a single MI355X (--gpus %d not supported)"
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    from xspde_matlab_amd import Problem
    from xspde_matlab_amd.configs import cgl2d

    p = cgl2d(args.precision)
    if args.boundaries == "neumann":
        p["boundaries"] = [[-1, -1], [-1, -1]]
    log("bench: creating problem %s %s %s" % (p["name"], p["points"], p["gpu"]["precision"]))
    prob = Problem(p)
    info = prob.info
    ns = info.nsites
    esize = 8 if args.precision == "fp32" else 16
    traj_steps = info.traj_steps
    for w in range(args.warmup):
        prob.handle.run_ensemble(1)
        ms = prob.handle.synchronize()
        log("bench: warmup %d done (%.1f ms kernel)" % (w, ms))
    kms = []
    t0 = time.perf_counter()
    for s in range(args.steps):
        prob.handle.run_ensemble(1)
        kms.append(prob.handle.synchronize())
        log("bench: step %d kernel %.1f ms" % (s, kms[-1]))
    wall = time.perf_counter() - t0
    ms_step = 1e3 * wall / args.steps
    kavg = sum(kms) / len(kms)
    rate = traj_steps / (ms_step / 1e3)
    alg_bytes = traj_steps * 2 * ns * esize            # state read + write per trajectory-step
    achieved = alg_bytes / (kavg / 1e3)
    # a sanity read of the moments so the run is checked, outside the timed region
    av = prob.handle.read_averages()
    import numpy as np
    finite = bool(np.isfinite(av).all())
    prob.close()
    cpu = None
    if not args.no_cpu:
        log("bench: timing the CPU oracle sample")
        cpu = cpu_baseline()
    line = {
        "metric": "trajectory-steps/s (2D 128^2 RK4IP, FP32 vs FP64) + wall time to fixed error bar",
        "value": rate, "unit": "trajectory-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32" if args.precision == "fp32" else "f64", "data": "synthetic",
        "config": {"workload": "2D CGL SPDE 128x128 %s, RK4IP+%s, 1024 trajectories, "
                               "2000 coarse steps with checks (6000 traj-steps/trajectory)"
                               % (args.boundaries, "FFT" if args.boundaries == "periodic" else "DCT-I"),
                   "lattice": [128, 128], "trajectories": 1024, "ensembles": [64, 16, 1],
                   "points": 101, "steps": 20, "checks": 1, "traj_steps_per_step": traj_steps,
                   "precision": args.precision, "moments": "fp64", "finite": finite},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": None},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
