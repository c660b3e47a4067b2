#!/usr/bin/env python
"""Throughput of the OpenCFD-SCU right-hand-side path on one MI355X.

Workload (N=1): the 1024x384x256 swept compression ramp (24-degree ramp
swept 30 degrees in z, 100.7 M points) - hybrid_auto WENO7/UP7 inviscid
fluxes on the curvilinear (tanh-stretched) grid, CD8 viscous terms, IBC 108
inflow/isothermal wall, non-reflecting outflow/top, the shock-capturing
filter every 10 steps, SSP-RK3 - plain FP64.  Synthetic seeded inputs.  A
step is one RK3 time step (three right-hand sides and stages); the metric is
grid-point RK-stage updates per second.  --nx/--ny/--nz/--sweep select
another ramp (e.g. 512x128x64 unswept).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints one JSON line.  `roofline` covers the kernel with the largest share
of the timed region, timed with HIP events on the library's stream inside the
timed region; `cpu_baseline` times the CPU oracle (the C restatement of
the same routines, OpenMP over the point loops) on --cpu-sample-steps full
steps of the same grid (default 1 on grids above 50 M points, where a step
takes about 17 s on 16 threads, else 15).
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "opencfd-scu_amd"))

PEAK_HBM_GBS = 8000.0

# Algorithmic HBM bytes per interior point per launch: each array the kernel
# must read or write, once (FP64 = 8 B).
ALGO_BYTES = {
    "sweep_x": (5 + 3 + 5) * 8,           # rho,u,v,w,T + 3 cofactors; write du(5) (Ajac applied by the stage)
    "sweep_y": (5 + 3 + 5 + 5) * 8,       # same + read-modify-write du(5)
    "sweep_z": (5 + 3 + 5 + 5) * 8,
    "visc_flux": (4 + 10 + 12) * 8,       # u,v,w,T, 9 cofactors + Ajac; write 12 Ev (mu formed from T)
    "visc_div": (2 * 3 + 2) * 8,          # two flux components per launch: their 6 Ev; write 2 dv
    "stage": (5 + 4 + 1 + 5 + 5 + 5 + 5) * 8,  # du, dv, Ajac, fn, f; write f, 5 primitives (stages 2-3; p in 3)
    "sensor": 2 * 8 + 3,                  # rho, T; write 3 int8 flags
}


def log(msg):
    """Progress on stderr (the JSON line stays alone on stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--nx", type=int, default=1024)
    ap.add_argument("--ny", type=int, default=384)
    ap.add_argument("--nz", type=int, default=256)
    ap.add_argument("--sweep", type=float, default=30.0, help="ramp sweep angle in degrees (0 for the unswept ramp)")
    ap.add_argument("--sensor-params", default=str(REPO / "opencfd-scu_amd" / "pyocfd" / "sensor_params.npy"),
                    help="learned-sensor weights for the mixed + learned-sensor leg (skipped if absent)")
    ap.add_argument("--cpu-sample-steps", type=int, default=0,
                    help="full steps of the CPU oracle (0: 1 on grids above 50 M points, else 15)")
    ap.add_argument("--profile-steps", type=int, default=10, help="steps of the per-kernel timing pass")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision leg")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain device copy of 1 GiB five times (the reference "
                         "kernel for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE) and exit")
    return ap.parse_args()


def copy_calibration():
    import torch
    n = (1 << 30) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "reps": 5}))


def traffic_for(kernel, workload):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary."""
    p = REPO / "profiles" / "traffic.json"
    if not p.exists():
        return None
    try:
        d = json.loads(p.read_text())
        return d.get(workload, {}).get(kernel)
    except Exception:
        return None


def cpu_baseline(case, steps):
    """The CPU oracle (test infrastructure, used here only as the baseline)
    on `steps` full time steps of the same grid."""
    sys.path.insert(0, str(REPO / "oracle"))
    import oracle as O
    o = case.load(O.Oracle(case.cfg))
    threads = O.lib().orc_num_threads()
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    rate = case.npoints * 3 * steps / dt
    return {"value": rate, "unit": "grid-point RK-stage updates/s", "cores": threads, "kind": "port",
            "sample": f"{steps} RK3 step(s) of the full {case.cfg.nx}x{case.cfg.ny}x{case.cfg.nz} ramp, "
                      f"C oracle with {threads} OpenMP threads, {dt:.1f} s"}


def mixed_leg(case, solver64, args, value64, sensor=None):
    """The same workload on the mixed-precision path (FP32 bulk; FP64 metrics,
    WENO weights, Jacobian factors and RK state) from the same initial state
    and for the same warmup + timed steps, optionally with the learned hybrid
    sensor: its rate, its speedup over the FP64 path and the relative L2
    error of density and pressure against the FP64 path's state at the same
    step."""
    import numpy as np
    from pyocfd import capi
    cfg = capi.OcfdConfig.from_buffer_copy(case.cfg)
    cfg.precision = capi.PREC_MIXED
    m = case.load(capi.Solver(cfg, device=0))
    if sensor is not None:
        m.set_sensor(sensor, 0.2, 0.8)
    if args.warmup:
        m.step(args.warmup)
    m.sync()
    t0 = time.perf_counter()
    m.step(args.steps)
    m.sync()
    wall = time.perf_counter() - t0
    ms = m.last_step_ms()
    secs = ms * 1e-3 if ms > 0 else wall
    rate = case.npoints * 3 * args.steps / secs
    # catch up with the FP64 path's kernel-timing steps before comparing
    (p64, i64, _) = solver64.get_flow()
    m.step(i64 - m.get_flow()[1])
    (p32, i32, _) = m.get_flow()
    assert i64 == i32
    g = case.cfg.gamma * case.cfg.ama ** 2
    pr64, pr32 = p64[0] * p64[4] / g, p32[0] * p32[4] / g

    def l2(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))

    out = {"value": rate, "ms_per_step": secs * 1e3 / args.steps, "speedup": rate / value64,
           "l2_rel_rho": l2(p32[0], p64[0]), "l2_rel_p": l2(pr32, pr64),
           "max_rel_rho": float(np.abs(p32[0] - p64[0]).max() / np.abs(p64[0]).max()),
           "after_steps": int(i64), "dtype": "f32 bulk / f64 metrics, weights, RK state"}
    if sensor is not None:
        flags = m.get_flags()
        n = sum(f.size for f in flags)
        out["sensor"] = ("learned (BF16-MFMA MLP inside the x sweep of the first RK stage, "
                         "confidence band [0.2, 0.8])")
        out["weno_fraction"] = round(sum(int((f > 1).sum()) for f in flags) / n, 4)
    m.close()
    return out


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported: the solver runs on one GPU "
              "(the multi-GPU halo exchange is a design only)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    from pyocfd import capi, cases

    workload = f"ramp{args.nx}x{args.ny}x{args.nz}" + (f"_sweep{args.sweep:g}" if args.sweep else "")
    log(f"generating {workload}")
    case = cases.ramp(args.nx, args.ny, args.nz, sweep=args.sweep, nsteps=args.warmup + args.steps,
                      progress=lambda m: log(f"case: {m}"))
    log("loading the FP64 context")
    solver = case.load(capi.Solver(case.cfg, device=0))
    log("timing the FP64 path")
    npts = case.npoints
    if args.warmup:
        solver.step(args.warmup)
    solver.sync()
    t0 = time.perf_counter()
    solver.step(args.steps)
    solver.sync()
    wall = time.perf_counter() - t0
    ev_ms = solver.last_step_ms()
    secs = ev_ms * 1e-3 if ev_ms > 0 else wall
    value = npts * 3 * args.steps / secs
    # per-kernel times from a few more steps with every launch bracketed by
    # events (this runs the kernels in order on one stream, without the
    # overlap of the viscous fluxes with the sweeps that the timed steps use)
    log("timing the kernels")
    solver.profile(True)
    solver.profile_reset()
    solver.step(args.profile_steps)
    solver.sync()
    prof = solver.profile_entries()
    solver.profile(False)

    # dominant kernel of the timed region
    kname, (launches, tot_ms) = max(prof.items(), key=lambda kv: kv[1][1])
    avg_s = tot_ms * 1e-3 / launches
    roof = None
    if kname in ALGO_BYTES:
        bytes_per_launch = ALGO_BYTES[kname] * npts
        achieved = bytes_per_launch / avg_s / 1e9
        roof = {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
                "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
                "traffic": traffic_for(kname, workload), "algo_bytes_per_launch": bytes_per_launch,
                "avg_launch_ms": round(avg_s * 1e3, 4)}
    share = {k: round(v[1] / sum(x[1] for x in prof.values()), 4) for k, v in sorted(prof.items())}
    flags = solver.get_flags()
    cnt = np.bincount(np.concatenate([f.ravel() for f in flags]), minlength=4)[1:4]
    mixed = learned = None
    if not args.no_mixed:
        log("timing the mixed-precision path")
        mixed = mixed_leg(case, solver, args, value)
        if Path(args.sensor_params).exists():
            log("timing the mixed-precision path with the learned sensor")
            learned = mixed_leg(case, solver, args, value, sensor=np.load(args.sensor_params))
    cpu = None
    if not args.no_cpu_baseline and int(os.environ.get("RANK", "0")) == 0:
        log("timing the CPU oracle")
        cpu_steps = args.cpu_sample_steps or (1 if npts > 50_000_000 else 15)
        cpu = cpu_baseline(case, cpu_steps)
    out = {
        "metric": "grid-point RK-stage updates/sec on 1 MI355X",
        "value": value,
        "unit": "grid-point RK-stage updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": secs * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator: tanh-stretched 24-degree ramp%s, Blasius-like inflow)" % (
            f", swept {args.sweep:g} degrees" if args.sweep else ""),
        "config": {"workload": workload, "points": npts, "scheme": "hybrid_auto WENO7/UP7 + CD8",
                   "ibc": 108, "filter_every": 10, "precision": "fp64",
                   "hybrid_flag_fractions": {"UP7": round(cnt[0] / cnt.sum(), 4),
                                             "WENO7": round(cnt[1] / cnt.sum(), 4),
                                             "WENO5": round(cnt[2] / cnt.sum(), 4)}},
        "wall_s": wall,
        "kernel_share": share,
        "roofline": roof,
        "mixed_precision": mixed,
        "mixed_learned_sensor": learned,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
