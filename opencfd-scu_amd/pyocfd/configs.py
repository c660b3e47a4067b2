"""Throughput of the other configurations on the FP64 and mixed paths.

    python -m pyocfd.configs [c1 c3 c4]

C1: Taylor-Green vortex 64^3, WENO7 (periodic, identity metrics);
C3: shock-turbulence interaction 512x256x256, WENO7;
C4: compression ramp 1024x256x128, hybrid_auto.
Each runs a few warm-up steps and then a timed run on the FP64 and on the
mixed path from the same state (HIP events on the library's stream around
the timed steps), and prints one JSON line: ms per step, G grid-point
RK-stage updates/s, the mixed path's speedup and its relative L2 error of
rho and p against the FP64 path after the same steps.
"""
import json
import sys

import numpy as np

from . import capi, cases


def measure(name, case, steps, warm=5):
    out = {"case": name, "points": case.npoints, "steps": warm + steps}
    flows = {}
    for prec, tag in ((capi.PREC_FP64, "fp64"), (capi.PREC_MIXED, "mixed")):
        cfg = capi.OcfdConfig.from_buffer_copy(case.cfg)
        cfg.precision = prec
        s = case.load(capi.Solver(cfg, device=0))
        s.step(warm)
        s.sync()
        s.step(steps)
        s.sync()
        ms = s.last_step_ms() / steps
        out[tag + "_ms_per_step"] = round(ms, 4)
        out[tag + "_G_pt_stage_per_s"] = round(case.npoints * 3 / ms / 1e6, 3)
        flows[tag] = s.get_flow()[0]
        s.close()
        print(f"[configs] {name} {tag}: {ms:.3f} ms/step", file=sys.stderr, flush=True)
    a, b = flows["mixed"], flows["fp64"]
    g = case.cfg.gamma * case.cfg.ama ** 2

    def l2(x, y):
        return float(np.linalg.norm(x - y) / np.linalg.norm(y))

    out["mixed_speedup"] = round(out["fp64_ms_per_step"] / out["mixed_ms_per_step"], 3)
    out["mixed_l2_rel_rho"] = l2(a[0], b[0])
    out["mixed_l2_rel_p"] = l2(a[0] * a[4] / g, b[0] * b[4] / g)
    print(json.dumps(out), flush=True)
    return out


CONFIGS = {
    "c1": lambda: measure("C1 Taylor-Green 64^3 WENO7", cases.tgv(64, nsteps=300), 200),
    "c3": lambda: measure("C3 shock-turbulence 512x256x256 WENO7", cases.shock_turbulence(512, 256, 256, nsteps=200),
                          100),
    "c4": lambda: measure("C4 ramp 1024x256x128 hybrid_auto", cases.ramp(1024, 256, 128, nsteps=200), 100),
}

if __name__ == "__main__":
    for key in sys.argv[1:] or list(CONFIGS):
        CONFIGS[key]()
