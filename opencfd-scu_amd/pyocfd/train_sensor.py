"""Train the learned hybrid sensor on seeded ramp flows and save its weights.

    python -m pyocfd.train_sensor [--nx 1024 --ny 256 --nz 128] [--stride 2]
                                  [--steps 60] [--out sensor_params.npy]

(the defaults: the 1024x256x128 ramp, every second node in each direction).

Training data are the plain FP64 solver's own runs of the compression ramp
over Mach {2.5, 2.9, 3.5} x angle {18, 24, 30} degrees, leaving Mach 2.9 /
24 degrees out as the held-out flow: each run is advanced `steps` steps and
its interior nodes give 24 features and the plain style-2 decision (WENO if
kp > 1).  Training (FP32, Adam, positives weighted by the class ratio) and
the collection of samples run in the library's own kernels
(ocfd_mlp_collect / ocfd_mlp_train).  The script prints the held-out
balanced accuracy of the untrained and the trained network and writes the
1889 parameters, and the report (sample counts, losses, held-out accuracy)
as JSON beside them (same name, .json).
"""
from __future__ import annotations

import argparse
import json
from pathlib import Path

import numpy as np

from . import capi, cases

GRID_MA = (2.5, 2.9, 3.5)
GRID_ANGLE = (18.0, 24.0, 30.0)
HELD_OUT = (2.9, 24.0)


def balanced_accuracy(prob, y):
    pos, neg = y > 0.5, y <= 0.5
    tpr = ((prob > 0.5) & pos).sum() / max(pos.sum(), 1)
    tnr = ((prob <= 0.5) & neg).sum() / max(neg.sum(), 1)
    return float(0.5 * (tpr + tnr)), float(tpr), float(tnr)


def collect(trainer, ma, angle, nx, ny, nz, steps, stride):
    case = cases.ramp(nx, ny, nz, ma=ma, angle=angle, nsteps=steps + 1)
    s = case.load(capi.Solver(case.cfg))
    s.step(steps)
    trainer.collect(s, stride)
    s.close()


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--nx", type=int, default=1024)
    ap.add_argument("--ny", type=int, default=256)
    ap.add_argument("--nz", type=int, default=128)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--stride", type=int, default=2)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--out", default=str(Path(__file__).resolve().parent / "sensor_params.npy"))
    a = ap.parse_args(argv)
    per = ((a.nx + a.stride - 1) // a.stride) * ((a.ny + a.stride - 1) // a.stride) * \
        ((a.nz + a.stride - 1) // a.stride)
    train = capi.Trainer(per * len(GRID_MA) * len(GRID_ANGLE))
    held = capi.Trainer(per)
    for ma in GRID_MA:
        for ang in GRID_ANGLE:
            if (ma, ang) == HELD_OUT:
                collect(held, ma, ang, a.nx, a.ny, a.nz, a.steps, a.stride)
            else:
                collect(train, ma, ang, a.nx, a.ny, a.nz, a.steps, a.stride)
            print(f"collected Ma {ma} angle {ang}: train {train.n} held-out {held.n}", flush=True)
    _, y = train.get()
    pos = float(y.mean())
    _, yh = held.get()
    held.set_params(train.get_params())
    before = balanced_accuracy(held.predict(), yh)
    losses = train.train(a.epochs, a.batch, a.lr, wpos=(1 - pos) / max(pos, 1e-12), seed=a.seed)
    params = train.get_params()
    held.set_params(params)
    after = balanced_accuracy(held.predict(), yh)
    np.save(a.out, params)
    report = {"samples": int(train.n), "positive_fraction": pos, "epoch_losses": [float(x) for x in losses],
              "held_out": {"flow": {"ma": HELD_OUT[0], "angle": HELD_OUT[1]}, "samples": int(held.n),
                           "balanced_accuracy_untrained": before[0], "balanced_accuracy_trained": after[0],
                           "tpr": after[1], "tnr": after[2]},
              "grid": [a.nx, a.ny, a.nz], "stride": a.stride, "steps": a.steps,
              "epochs": a.epochs, "batch": a.batch, "lr": a.lr, "seed": a.seed}
    Path(a.out).with_suffix(".json").write_text(json.dumps(report))
    print(json.dumps(report), flush=True)
    return report


if __name__ == "__main__":
    main()
