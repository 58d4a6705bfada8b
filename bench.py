"""Throughput of the xSPDE ensemble integration path on one MI355X.

Workload (one "step" = one full ensemble pass of the hot path):
  2D damped complex Ginzburg-Landau SPDE on a 128x128 periodic lattice,
  da = [(1+i) lap a + a - (1+0.5i)|a|^2 a] dt + 0.05 dW, RK4 interaction
  picture with the FFT propagator, 1024 trajectories (ensembles [64,16,1]),
  101 stored points x 20 steps (dt = 0.002 coarse, 2000 coarse steps) with
  checks on: every trajectory runs the fine (dt/2) and the coarse path,
  6000 trajectory-steps each.  FP32 fields, FP64 moments.  Synthetic input:
  random-phase initial fields and counter-based noise, generated on device.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-fp64]
                       [--precision fp32|fp64] [--boundaries periodic|neumann]
       python bench.py --workload tw     (control variate vs plain ensemble)
       python bench.py --workload copy   (256 MiB device copies: the reference
                                          kernel that scales rocprofv3's
                                          FETCH_SIZE / WRITE_SIZE to bytes)
Prints one JSON line (rank 0, N=1).  The default FP32 line also carries the
FP64 rate of the same workload (the other half of the metric) in "fp64",
measured after the timed region (1 warmup + 2 launches).

--workload tw measures the second half of the headline metric, the wall time
to a fixed error bar: on-site g2 of a truncated-Wigner 2D Bose lattice
(128x128 periodic, coherent n0 = 100 plus vacuum noise, RK4IP, 1000 steps of
dt = 0.001), plain 65536-trajectory ensemble against the control-variate
estimate (4096 training trajectories, GPU-trained network, 10^6 cheap samples).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12          # B/s, MI355X HBM3E
VALU_PEAK = {"fp32": 157.3e12, "fp64": 78.6e12}   # FLOP/s, MI355X vector
MFMA_PEAK = {"fp32": 157.3e12, "fp64": 78.6e12}   # FLOP/s, MI355X dense matrix cores


def step_flops(ns, transforms=8, drift_evals=4, c_drift=20):
    """Algorithmic flops of one RK4IP trajectory-step on an ns-site periodic
    lattice: 8 FFTs of 5 ns log2 ns, 4 propagator products (6 ns), 4 drift
    evaluations (c_drift ns) and the RK combinations (10 ns)."""
    import math
    return transforms * 5 * ns * math.log2(ns) + 4 * 6 * ns + drift_evals * c_drift * ns + 10 * ns


def dense_flops(lattice, transforms=8, split=True):
    """Flops of the dense real transform passes of one RK4IP trajectory-step on
    a non-periodic lattice: per transform and dimension d, a real N_d x N_d
    matrix times the complex plane, 4 N_d ns flops; a DCT-I / DST-I dimension
    with N_d % 64 == 0 runs as its even/odd split (two N_d/2 x N_d/2 products,
    half of that)."""
    ns = 1
    for n in lattice:
        ns *= n
    return transforms * sum((2 if split and n % 64 == 0 else 4) * n * ns for n in lattice)


def measured_traffic(kernel, workload):
    """HBM bytes per launch from the committed rocprofv3 PMC summary, if it was
    taken for this kernel and workload (else None)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f).get(workload)
    except (OSError, ValueError):
        return None
    if not rec or kernel not in rec.get("kernel", ""):
        return None
    return rec.get("bytes_per_launch")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _cpu_worker(args):
    """One single-threaded worker of the CPU baseline (like one parfor worker)."""
    seconds, seed = args
    os.environ["OMP_NUM_THREADS"] = "1"
    from oracle import forms
    from oracle import xspde as X
    from xspde_matlab_amd.configs import cgl2d, shrink
    E = 8
    p = shrink(cgl2d("fp64"), E=E, e2=1)
    q = X.xpreferences(forms.apply_forms(p))
    q.dtr = q.dt / (2 * q.steps[0])
    prop = X.xpropfactor(q)
    v = X.xrandom(q, 1 + seed)
    a = X.initial_field(q, v)
    n = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        w = X.noise_draw(q, 1 + seed, n)
        a = X.RK4(a, w, q, prop)
        n += 1
    return n * E, time.perf_counter() - t0


def cpu_baseline(seconds=12.0, workers=None):
    """FP64 oracle (numpy restatement of the reference path) on the host cores:
    one single-threaded worker process per core, each stepping its own 8
    trajectories for a bounded time, as the reference's parfor over
    ensembles(3) would (xsim.m:51)."""
    import multiprocessing as mp
    workers = workers or min(16, os.cpu_count() or 1)
    ctx = mp.get_context("spawn")
    with ctx.Pool(workers) as pool:
        res = pool.map(_cpu_worker, [(seconds, w) for w in range(workers)])
    rate = sum(n / dt for n, dt in res)
    steps = sum(n for n, _ in res)
    return dict(value=rate, unit="trajectory-steps/s", cores=workers, kind="port",
                sample="oracle RK4IP fp64 numpy, 128x128 CGL periodic, %d single-threaded workers x 8 "
                       "trajectories, %d trajectory-steps in %.0f s" % (workers, steps, seconds))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--precision", default="fp32")
    ap.add_argument("--boundaries", default="periodic", choices=["periodic", "neumann"],
                    help="periodic (FFT) or Neumann-Neumann (DCT-I) lattice")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 companion measurement")
    ap.add_argument("--observe", default=None,
                    help="comma-separated observe forms replacing the workload's (e.g. abs2_ave): "
                         "traffic attribution runs")
    ap.add_argument("--workload", default="cgl2d", choices=["cgl2d", "nls1d", "tw", "batch", "copy"])
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs a single MI355X (--gpus %d not supported)"
              % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.workload == "copy":
        return copy_calibration()
    if args.workload == "tw":
        return bench_tw(args)
    if args.workload == "batch":
        return bench_batch(args)

    from xspde_matlab_amd import Problem
    from xspde_matlab_amd.configs import cgl2d, nls1d

    if args.workload == "nls1d":
        p = nls1d(args.precision)
        if args.boundaries == "neumann":
            p["boundaries"] = [[-1, -1]]
    else:
        p = cgl2d(args.precision)
        if args.boundaries == "neumann":
            p["boundaries"] = [[-1, -1], [-1, -1]]
    if args.observe:
        p["gpu"] = dict(p["gpu"], observe=args.observe.split(","))
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
    flops = traj_steps * step_flops(ns) / (kavg / 1e3)
    kernel = ("spde_path", "spde_stream", "spde_path", "sde_path")[info.kernel_kind]
    wl = "%s-%s-%s" % ("nls1d256" if args.workload == "nls1d" else "cgl2d128", args.precision, args.boundaries)
    if args.observe:
        wl += "-" + args.observe.replace(",", "+")
    # a sanity read of the moments so the run is checked, outside the timed region
    av = prob.handle.read_averages()
    import numpy as np
    finite = bool(np.isfinite(av).all())
    prob.close()
    fp64 = None
    if args.workload == "cgl2d" and args.precision == "fp32" and not args.no_fp64:
        fp64 = fp64_companion(args.boundaries)
    cpu = None
    if not args.no_cpu:
        log("bench: timing the CPU oracle sample")
        cpu = cpu_baseline()
    tname = "FFT" if args.boundaries == "periodic" else "DCT-I"
    if args.workload == "nls1d":
        metric = "trajectory-steps/s (1D 256 RK4IP stochastic NLS)"
        config = {"workload": "1D stochastic NLS 256 sites %s, RK4IP+%s, 4096 trajectories, 1000 coarse "
                              "steps with checks (3000 traj-steps/trajectory)" % (args.boundaries, tname),
                  "lattice": [256], "trajectories": 4096, "ensembles": p["ensembles"], "points": 101,
                  "steps": 10, "checks": 1}
    else:
        metric = "trajectory-steps/s (2D 128^2 RK4IP, FP32 vs FP64) + wall time to fixed error bar"
        config = {"workload": "2D CGL SPDE 128x128 %s, RK4IP+%s, 1024 trajectories, 2000 coarse steps "
                              "with checks (6000 traj-steps/trajectory)" % (args.boundaries, tname),
                  "lattice": [128, 128], "trajectories": 1024, "ensembles": [64, 16, 1],
                  "points": 101, "steps": 20, "checks": 1}
    config.update(traj_steps_per_step=traj_steps, precision=args.precision, moments="fp64", finite=finite,
                  kernel=kernel, threads=info.threads, traj_per_block=info.traj_per_block)
    line = {
        "metric": metric,
        "value": rate, "unit": "trajectory-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32" if args.precision == "fp32" else "f64", "data": "synthetic",
        "config": config,
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": measured_traffic(kernel, wl),
                     "kernel": kernel, "kernel_ms": kavg,
                     "valu": {"achieved": flops / 1e12, "peak": VALU_PEAK[args.precision] / 1e12,
                              "unit": "TFLOP/s", "frac": flops / VALU_PEAK[args.precision],
                              "flops_per_traj_step": step_flops(ns)}},
        "cpu_baseline": cpu,
    }
    if args.boundaries == "neumann":
        lat = config["lattice"]
        dfl = traj_steps * dense_flops(lat) / (kavg / 1e3)
        line["roofline"]["mfma"] = {"achieved": dfl / 1e12, "peak": MFMA_PEAK[args.precision] / 1e12,
                                    "unit": "TFLOP/s", "frac": dfl / MFMA_PEAK[args.precision],
                                    "flops_per_traj_step": dense_flops(lat),
                                    "what": "dense DCT-I passes on the matrix cores (even/odd split: two 64 x 64 products per 128-point pass)"}
    if fp64 is not None:
        fp64["fp32_over_fp64"] = rate / fp64["value"]
        line["fp64"] = fp64
    print(json.dumps(line), flush=True)


def fp64_companion(boundaries, warmup=1, launches=2):
    """The same 128x128 workload with FP64 fields: whole-ensemble launches timed
    with HIP events, after the FP32 timed region (the FP64 half of the metric)."""
    from xspde_matlab_amd import Problem
    from xspde_matlab_amd.configs import cgl2d
    p = cgl2d("fp64")
    if boundaries == "neumann":
        p["boundaries"] = [[-1, -1], [-1, -1]]
    prob = Problem(p)
    info = prob.info
    for w in range(warmup):
        prob.handle.run_ensemble(1)
        log("bench fp64: warmup %d done (%.1f ms kernel)" % (w, prob.handle.synchronize()))
    kms = []
    t0 = time.perf_counter()
    for s in range(launches):
        prob.handle.run_ensemble(1)
        kms.append(prob.handle.synchronize())
        log("bench fp64: launch %d kernel %.1f ms" % (s, kms[-1]))
    wall = (time.perf_counter() - t0) / launches
    prob.close()
    kernel = ("spde_path", "spde_stream", "spde_path", "sde_path")[info.kernel_kind]
    kavg = sum(kms) / len(kms)
    return {"value": info.traj_steps / wall, "unit": "trajectory-steps/s", "dtype": "f64", "launches": launches,
            "warmup": warmup, "ms_per_step": 1e3 * wall, "kernel": kernel, "kernel_ms": kavg,
            "hbm_frac": info.traj_steps * 2 * info.nsites * 16 / (kavg / 1e3) / HBM_PEAK,
            "traffic": measured_traffic(kernel, "cgl2d128-fp64-%s" % boundaries),
            "mfma_frac": (info.traj_steps * dense_flops([128, 128]) / (kavg / 1e3) / MFMA_PEAK["fp64"]
                          if boundaries == "neumann" else None)}


def copy_calibration(n=64 * 1024 * 1024, reps=3):
    """Three plain device copies of 256 MiB (n float32), the kernel of known
    traffic (262,144 KiB read + 262,144 KiB written per copy) whose FETCH_SIZE and
    WRITE_SIZE readings scale the counters of the path kernels to bytes."""
    import torch
    x = torch.ones(n, device="cuda")
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"workload": "copy", "bytes_read_per_copy": 4 * n, "bytes_written_per_copy": 4 * n,
                      "copies": reps}), flush=True)


# attractive lattice whose modulational-instability band (hop k^2 < 2 |g| n0)
# holds the 32 feature modes: vacuum noise in those modes seeds the growth
TW_G, TW_HOP = -0.035, 270.0


def bench_tw(args):
    """Wall time to the plain ensemble's error bar: plain vs control variate.

    Plain: 65536 trajectories.  Control variate: the network is trained on
    4096 trajectories, its mean is taken over 10^6 cheap device-drawn noise
    samples, and then only as many independent trajectories are run as the
    residual variance needs to reach the plain ensemble's error bar.  Its wall
    time counts the training trajectories, the training, the cheap samples and
    the estimation trajectories."""
    import math
    import numpy as np
    from xspde_matlab_amd import cv
    from xspde_matlab_amd.configs import tw2d

    E, M, NA, NP = 1024, 10 ** 6, 4096, 65536

    def mk(e2, seed):
        return tw2d("fp32", E=E, e2=e2, seed=seed, g=TW_G, hop=TW_HOP)

    tr = cv.feature_translations(mk(1, 0))
    log("bench tw: warmup")
    cv.run_samples(mk(1, 77))
    log("bench tw: %d training trajectories" % NA)
    (oA, zA, kA), wA = cv.timed(cv.run_samples, mk(NA // E, 0))
    log("bench tw: plain ensemble (%d trajectories)" % NP)
    (oP, zP, kP), wP = cv.timed(cv.run_samples, mk(NP // E, 1000))
    log("bench tw: plain %.2f s (kernel %.2f s), training set %.2f s" % (wP, kP, wA))
    spt = wP / NP
    out = {}
    for name in ("g2", "lowk"):
        if name == "g2":
            g_plain, yP = cv.g2_pseudo(*cv.g2_parts(oP))
            _, yA = cv.g2_pseudo(*cv.g2_parts(oA))
        else:
            yP, yA = oP[:, 2], oA[:, 2]
            g_plain = float(np.mean(yP))
        sig_target = float(np.std(yP, ddof=1) / math.sqrt(NP))
        model = cv.CVModel(seed=0, batch_max=8192)
        info, wfit = cv.timed(model.fit, zA, yA, translations=tr)
        (mu_f, var_f), wcheap = cv.timed(model.cheap_moments, M, 7)
        # residual variance measured over the whole plain ensemble (independent of training)
        var_res = float(np.var(yP - model.predict(zP), ddof=1))
        n_need = cv.trajectories_to_error_bar(var_res, var_f, M, sig_target)
        nB = int(min(NP, max(E, math.ceil(n_need / E) * E)))
        # the estimate itself, on the first nB plain trajectories
        if name == "g2":
            _, yB = cv.g2_pseudo(*cv.g2_parts(oP[:nB]))
        else:
            yB = yP[:nB]
        r = cv.estimate(model, zP[:nB], yB, m_cheap=M)
        model.close()
        t_cv = wA + wfit + wcheap + nB * spt
        out[name] = dict(plain_mean=g_plain, plain_sigma=sig_target, cv_mean=r["cv_mean"], cv_sigma=r["cv_sigma"],
                         var_ratio=float(np.var(yP, ddof=1) / var_res), beta=info["beta"],
                         estimate_trajectories=nB, cv_seconds=t_cv, plain_seconds=wP, speedup=wP / t_cv,
                         fit_seconds=wfit, cheap_seconds=wcheap, training_seconds=wA)
        log("bench tw: %s var ratio %.2f, %d estimate trajectories, cv %.2f s vs plain %.2f s (x%.2f)"
            % (name, out[name]["var_ratio"], nB, t_cv, wP, wP / t_cv))
    line = {
        "metric": "wall time to fixed error bar (on-site g2, truncated-Wigner 2D 128^2)",
        "value": out["g2"]["cv_seconds"], "unit": "s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": 1e3 * out["g2"]["cv_seconds"], "higher_is_better": False, "scaling": "strong",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "truncated-Wigner GP 128x128 periodic, n0=100, g=%g, hop=%g, RK4IP 1000 x "
                               "dt=0.001; plain 65536 trajectories vs control variate (network trained on "
                               "4096 trajectories, 64-256-256-1 GELU on bf16 MFMA, 10^6 cheap samples)"
                               % (TW_G, TW_HOP),
                   "plain_seconds": wP, "plain_kernel_seconds": kP,
                   "g2": out["g2"], "low_mode_occupation": out["lowk"]},
    }
    print(json.dumps(line), flush=True)


def _batch_time(ps, args, serial):
    import numpy as np
    from xspde_matlab_amd.xsim import BatchProblem

    t0 = time.perf_counter()
    bp = BatchProblem(ps, autotune=True, serial=serial)
    t_tune = time.perf_counter() - t0
    total = sum(info.traj_steps for info in bp.infos)
    for _ in range(args.warmup):
        bp.run(1)
        bp.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        bp.run(1)
        bp.synchronize()
    wall = (time.perf_counter() - t0) / args.steps
    finite = all(bool(np.isfinite(bp.read_averages(i)).all()) for i in range(len(ps)))
    plan = bp.plan()
    bp.close()
    return total, wall, t_tune, finite, plan


def _batch_run(ps, args):
    """Autotuned graph runs of the instances ps (families concurrent, then
    serial) against one launch per instance; the faster graph is reported."""
    from xspde_matlab_amd import Problem

    total, wall_c, t_tune, finite, plan = _batch_time(ps, args, serial=False)
    _, wall_s, _, finite_s, _ = _batch_time(ps, args, serial=True)
    wall = min(wall_c, wall_s)
    finite = finite and finite_s
    probs = [Problem(p) for p in ps]
    for pr in probs:                       # warm every kernel once
        pr.handle.run_ensemble(1)
    for pr in probs:
        pr.handle.synchronize()
    t0 = time.perf_counter()
    for pr in probs:
        pr.handle.run_ensemble(1)
        pr.handle.synchronize()
    seq = time.perf_counter() - t0
    for pr in probs:
        pr.close()
    return dict(traj_steps_per_run=total, batch_seconds=wall, concurrent_seconds=wall_c, serial_seconds=wall_s,
                sequential_seconds=seq, speedup_vs_sequential=seq / wall, tuning_seconds=t_tune, plan=plan,
                finite=finite)


def bench_batch(args):
    """64 uneven GP instances (1D 32..512, 2D 32^2..128^2, E 256..4096, 500 RK4IP
    steps, FP32) in one graph launch, autotuned, against the same instances
    launched one after another on their own handles; plus a batch of small
    lattices only (1D, 32^2), where one launch per instance leaves the GPU idle."""
    from xspde_matlab_amd.configs import batch_instances

    log("bench batch: 64 instances, all lattice families")
    full = _batch_run(batch_instances(count=64, seed=0, steps=500), args)
    log("bench batch: %.2f s per run, sequential %.2f s" % (full["batch_seconds"], full["sequential_seconds"]))
    log("bench batch: 64 small-lattice instances")
    small = _batch_run(batch_instances(count=64, seed=1, steps=500, sizes_2d=(32,)), args)
    log("bench batch: small %.3f s per run, sequential %.3f s" % (small["batch_seconds"], small["sequential_seconds"]))
    line = {
        "metric": "trajectory-steps/s (batched uneven ensembles, one graph launch)",
        "value": full["traj_steps_per_run"] / full["batch_seconds"], "unit": "trajectory-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * full["batch_seconds"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "64 GP instances, 1D {32..512} / 2D {32^2..128^2} by seed, gamma U[0.01,0.5], "
                               "g U[0.1,2], E {256..4096}, 500 RK4IP steps, FP32, autotuned",
                   "full": full,
                   "small_lattices": dict(small, rate=small["traj_steps_per_run"] / small["batch_seconds"])},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
