"""The learned hybrid sensor on the GPU against the NumPy restatement
(oracle/mlp_sensor.py): features, the FP32 training gradients, BF16-MFMA
inference with the confidence-band fallback, and what training gains.

Bars: features to FP32 rounding (1e-6 absolute; they are O(1) and the
kernel evaluates them in FP32 with the hardware reciprocal); gradients to
1e-4 of each parameter block's largest entry (FP32 sums of a 256-sample
batch against FP64); inference probabilities against a NumPy forward that
rounds inputs, weights and hidden activations to BF16 as the kernel does:
mean difference 2e-5 (FP32 accumulation order), largest 5e-4 (a value that
sits on a BF16 rounding boundary may round the other way in the kernel,
whose features and GELU differ from FP64 in the last FP32 bits: one BF16 ulp
of one hidden unit moves the probability by about 1e-4 with these
weights); flags exactly by the flag rule on the kernel's own
probabilities.
"""
import numpy as np
import pytest

import mlp_sensor as M
from pyocfd import capi, cases

pytestmark = pytest.mark.gpu


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    b = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32)


def wrap_lap(a):
    """Interior (nz, ny, nx) field of a fully periodic case -> LAP box."""
    return np.pad(a, 4, mode="wrap")


def periodic_fields(case, solver):
    d, u, v, w, T = [wrap_lap(a) for a in solver.get_flow()[0]]
    met = solver.get_metrics()
    A = np.stack([wrap_lap(met[3 + m]) for m in range(9)])
    return d, u, v, w, T, A


def test_features_and_labels_match_oracle():
    case = cases.wavy_box(20, 16, 12, scheme="SCHEME_HYBRIDAUTO")
    g = case.load(capi.Solver(case.cfg))
    t = capi.Trainer(10000)
    stride = 3
    t.collect(g, stride)
    X, y = t.get()
    nodes = [(i, j, k) for k in range(0, 12, stride) for j in range(0, 16, stride) for i in range(0, 20, stride)]
    assert len(nodes) == t.n
    d, u, v, w, T, A = periodic_fields(case, g)
    p00 = 1.0 / (case.cfg.gamma * case.cfg.ama ** 2)
    F = M.features_field(d, u, v, w, T, A, p00, 1.0 / case.cfg.ama, nodes)
    assert np.abs(X - F).max() <= 1e-6
    kp = M.jameson_kp(p00 * d[3:-3, 3:-3, 3:-3] * T[3:-3, 3:-3, 3:-3], *case.cfg.hy_dp_intv)
    lab = np.array([kp[k, j, i] > 1 for (i, j, k) in nodes], np.float32)
    assert np.array_equal(y, lab)


def test_training_gradients_match_oracle():
    rng = np.random.default_rng(7)
    n = 256
    X = rng.standard_normal((n, 24)).astype(np.float32) * 0.3
    y = (rng.random(n) < 0.3).astype(np.float32)
    p = M.init_params(seed=3, dtype=np.float32)
    t = capi.Trainer(n)
    t.set_params(capi.mlp_pack(p))
    t.put(X, y)
    loss, g = t.grad(0, n, wpos=2.5)
    p64 = {k: v.astype(np.float64) for k, v in p.items()}
    loss_ref, g_ref = M.loss_and_grads(p64, X.astype(np.float64), y.astype(np.float64), wpos=2.5)
    assert abs(loss - loss_ref) <= 1e-5 * abs(loss_ref)
    gg = capi.mlp_unpack(g)
    for k in g_ref:
        err = np.abs(gg[k] - g_ref[k]).max() / max(np.abs(g_ref[k]).max(), 1e-30)
        print(f"[mlp] grad {k}: rel err {err:.2e}", flush=True)
        assert err <= 1e-4, (k, err)
    prob = t.predict(0, n)
    prob_ref, _ = M.forward(p64, X.astype(np.float64))
    assert np.abs(prob - prob_ref).max() <= 1e-5


def test_inference_matches_bf16_oracle_and_flag_rule():
    case = cases.wavy_box(20, 16, 12, scheme="SCHEME_HYBRIDAUTO", amp=0.08)
    g = case.load(capi.Solver(case.cfg))
    p = M.init_params(seed=11, dtype=np.float32)
    p["W3"] = p["W3"] * 20.0  # spread the probabilities over (0, 1)
    g.set_sensor(capi.mlp_pack(p), 0.2, 0.8)
    g.rhs(1)
    # a confidence band with a third of the nodes below, inside and above it
    lo, hi = [float(x) for x in np.quantile(g.get_sensor_prob(), [1 / 3, 2 / 3])]
    g.set_sensor(capi.mlp_pack(p), lo, hi)
    g.rhs(1)
    prob = g.get_sensor_prob()
    fx, fy, fz = g.get_flags()
    d, u, v, w, T, A = periodic_fields(case, g)
    p00 = 1.0 / (case.cfg.gamma * case.cfg.ama ** 2)
    nz, ny, nx = 12, 16, 20
    nodes = [(i, j, k) for k in range(nz) for j in range(ny) for i in range(nx)]
    F = M.features_field(d, u, v, w, T, A, p00, 1.0 / case.cfg.ama, nodes).astype(np.float32)
    # forward with the kernel's BF16 roundings
    W1, W2 = bf16(p["W1"]), bf16(p["W2"])
    a1 = bf16(M.gelu(bf16(F) @ W1.T + p["b1"]).astype(np.float32))
    a2 = M.gelu(a1 @ W2.T + p["b2"])
    ref = 1.0 / (1.0 + np.exp(-(a2 @ p["W3"] + p["b3"][0])))
    got = prob.reshape(-1)
    err = np.abs(got - ref)
    print(f"[mlp] inference |prob - ref| max {err.max():.2e} mean {err.mean():.2e}", flush=True)
    assert err.max() <= 5e-4 and err.mean() <= 2e-5
    kp = M.jameson_kp(p00 * d[3:-3, 3:-3, 3:-3] * T[3:-3, 3:-3, 3:-3], *case.cfg.hy_dp_intv).reshape(-1)
    want = M.flag_rule(got, kp, lo, hi).reshape(nz, ny, nx)
    assert np.array_equal(fx[:, :, 1:], want)
    assert np.array_equal(fy[:, 1:, :], want)
    assert np.array_equal(fz[1:, :, :], want)
    assert ((got > lo) & (got < hi)).any() and ((got < lo) | (got > hi)).any()
    # switching back restores the plain sensor
    g.set_sensor(None)
    g.rhs(1)
    fx2 = g.get_flags()[0]
    assert np.array_equal(fx2[:, :, 1:], kp.reshape(nz, ny, nx))


def test_training_learns_the_plain_sensor():
    """Seeded ramp flows (two Mach numbers, two angles) give the samples;
    training on them lowers the loss and the trained network agrees with the
    plain sensor on a held-out flow (third angle) much better than the
    untrained one does."""
    t = capi.Trainer(2_000_000)
    for ma, ang in [(2.5, 18.0), (3.5, 30.0)]:
        case = cases.ramp(96, 48, 8, ma=ma, angle=ang, nsteps=40, filt=False)
        g = case.load(capi.Solver(case.cfg))
        g.step(30)
        t.collect(g, 1)
        g.close()
    X, y = t.get()
    pos = y.mean()
    print(f"[mlp] samples {t.n}, positive fraction {pos:.4f}", flush=True)
    assert 0.0 < pos < 0.5
    held = cases.ramp(96, 48, 8, ma=2.9, angle=24.0, nsteps=40, filt=False)
    hg = held.load(capi.Solver(held.cfg))
    hg.step(30)
    h = capi.Trainer(200_000)
    h.collect(hg, 1)
    Xh, yh = h.get()
    def balanced(prob):
        tpr = ((prob > 0.5) & (yh > 0.5)).sum() / max((yh > 0.5).sum(), 1)
        tnr = ((prob <= 0.5) & (yh <= 0.5)).sum() / max((yh <= 0.5).sum(), 1)
        return 0.5 * (tpr + tnr), tpr, tnr

    h.set_params(t.get_params())
    b0 = balanced(h.predict())
    bal = float((1 - pos) / pos)
    losses = t.train(epochs=6, batch=2048, lr=3e-3, wpos=bal, seed=5)
    print("[mlp] epoch losses " + " ".join(f"{x:.4f}" for x in losses), flush=True)
    assert losses[-1] < 0.7 * losses[0]
    h.set_params(t.get_params())
    b1 = balanced(h.predict())
    print(f"[mlp] held-out balanced accuracy untrained {b0[0]:.4f} trained {b1[0]:.4f} "
          f"(TPR {b1[1]:.3f}, TNR {b1[2]:.3f})", flush=True)
    assert b1[0] > b0[0] + 0.2 and b1[0] > 0.85


def test_trained_sensor_gain_on_the_held_out_ramp():
    """What the shipped network gains, on the flow it was not trained on
    (the Mach 2.9, 24-degree ramp) on a 256x64x32 grid, 200 steps on the
    mixed path from the same initial state, each run judged against the
    plain FP64 solution (hybrid_auto with the plain sensor): relative L2 of
    density and pressure, and the measured time per step.

      no sensor          UP7 everywhere (the hybrid scheme without a sensor)
      untrained network  the network's seeded initial weights deciding every
                         node (confidence band [0.5, 0.5]: no fallback)
      trained network    the shipped weights deciding every node
      trained + band     the shipped weights with the [0.2, 0.8] band that
                         falls back to the plain sensor
      plain sensor       the mixed path with the plain sensor (the FP32
                         rounding alone)

    The trained network alone must come within the 1e-4 error bar and
    closer to the plain FP64 solution than no sensor and than the untrained
    network; with the band, within the bar as well."""
    from pathlib import Path
    import pyocfd
    params = np.load(Path(pyocfd.__file__).parent / "sensor_params.npy")
    untrained = capi.mlp_pack(M.init_params(seed=7, dtype=np.float32))
    nsteps = 200
    case = cases.ramp(256, 64, 32, ma=2.9, angle=24.0, nsteps=nsteps, filt=False)
    ref = case.load(capi.Solver(case.cfg))
    ref.step(nsteps)
    (p64, _, _) = ref.get_flow()
    ref.close()
    g = case.cfg.gamma * case.cfg.ama ** 2
    pr64 = p64[0] * p64[4] / g

    def run(tag, scheme=None, sensor=None, lo=0.2, hi=0.8):
        cfg = capi.OcfdConfig.from_buffer_copy(case.cfg)
        cfg.precision = capi.PREC_MIXED
        if scheme is not None:
            cfg.scheme_invis = scheme
        s = case.load(capi.Solver(cfg))
        if sensor is not None:
            s.set_sensor(sensor, lo, hi)
        s.step(nsteps)
        s.sync()
        ms = s.last_step_ms() / nsteps   # the time of the whole step() call
        (p, _, _) = s.get_flow()
        weno = None
        if cfg.scheme_invis == capi.HYBRIDAUTO:
            fl = s.get_flags()
            weno = sum(int((f > 1).sum()) for f in fl) / sum(f.size for f in fl)
        s.close()
        e_rho = float(np.linalg.norm(p[0] - p64[0]) / np.linalg.norm(p64[0]))
        e_p = float(np.linalg.norm(p[0] * p[4] / g - pr64) / np.linalg.norm(pr64))
        w = "-" if weno is None else f"{weno:.4f}"
        print(f"[mlp] gain {tag:18s} L2 rho {e_rho:.2e} p {e_p:.2e}  WENO faces {w}  {ms:.3f} ms/step",
              flush=True)
        return max(e_rho, e_p)

    e_none = run("no sensor", scheme=capi.UP7)
    e_untr = run("untrained network", sensor=untrained, lo=0.5, hi=0.5)
    e_tr = run("trained network", sensor=params, lo=0.5, hi=0.5)
    e_band = run("trained + band", sensor=params)
    e_plain = run("plain sensor")
    print(f"[mlp] gain: trained network {e_none / e_tr:.1f}x closer than no sensor, "
          f"{e_untr / e_tr:.1f}x closer than the untrained network; plain sensor {e_plain:.2e}", flush=True)
    assert e_tr <= 1e-4 and e_band <= 1e-4
    assert e_tr < e_none and e_tr < e_untr
