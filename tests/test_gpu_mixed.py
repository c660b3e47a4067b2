"""Mixed-precision path (PRECISION = MIXED): FP32 primitives, split fluxes,
reconstructions, derivatives and viscous fluxes; FP64 metrics, WENO weight
arithmetic, Jacobian factors, du accumulation and RK state.

Its reference is the FP64 oracle on the same inputs.  Bars (stated per test):
the right-hand side to 1e-4 relative max-norm per variable, the state after
a few steps to 1e-6 relative L2 per variable, and mass conservation on the
periodic box to 1e-7 relative (FP32 fluxes telescope only to FP32 round-off
and the sums are taken in FP64).
"""
import numpy as np
import pytest

import oracle as O
from pyocfd import capi, cases

pytestmark = pytest.mark.gpu


def rel_max(a, b):
    a = np.asarray(a).reshape(a.shape[0], -1)
    b = np.asarray(b).reshape(b.shape[0], -1)
    return np.array([np.abs(a[v] - b[v]).max() / max(np.abs(b[v]).max(), 1e-300) for v in range(a.shape[0])])


def rel_l2(a, b):
    a = np.asarray(a).reshape(a.shape[0], -1)
    b = np.asarray(b).reshape(b.shape[0], -1)
    return np.array([np.linalg.norm(a[v] - b[v]) / max(np.linalg.norm(b[v]), 1e-300) for v in range(a.shape[0])])


def mixed_pair(case):
    cfg64 = case.cfg
    cfg32 = capi.OcfdConfig.from_buffer_copy(cfg64)
    cfg32.precision = capi.PREC_MIXED
    g = case.load(capi.Solver(cfg32))
    o = case.load(O.Oracle(cfg64))
    return g, o


@pytest.mark.parametrize("scheme", ["WENO7", "UP7", "SCHEME_HYBRIDAUTO", "OMP6_HR", "WENO7_SYMBO"])
def test_mixed_rhs_wavy_box(scheme):
    case = cases.wavy_box(20, 16, 12, scheme=scheme)
    g, o = mixed_pair(case)
    g.rhs(1)
    o.rhs(1)
    e = rel_max(g.get_du(), o.get_du())
    print(f"[mixed] rhs wavy {scheme}: " + " ".join(f"{x:.2e}" for x in e), flush=True)
    assert (e <= 1e-4).all(), e


def test_mixed_steps_ramp():
    case = cases.ramp(48, 32, 10, nsteps=6, filter_every=2)
    g, o = mixed_pair(case)
    g.step(6)
    o.step(6)
    e = rel_l2(g.get_state(), o.get_state())
    print("[mixed] 6 steps ramp L2: " + " ".join(f"{x:.2e}" for x in e), flush=True)
    assert (e <= 1e-6).all(), e
    (pg, ig, _), (po, io, _) = g.get_flow(), o.get_flow()
    assert ig == io == 6
    assert rel_l2(np.stack(pg), np.stack(po)).max() <= 1e-6


def test_mixed_tgv_mass():
    case = cases.tgv(32, nsteps=10)
    cfg = capi.OcfdConfig.from_buffer_copy(case.cfg)
    cfg.precision = capi.PREC_MIXED
    g = case.load(capi.Solver(cfg))
    m0 = g.get_state()[0].sum()
    g.step(10)
    f = g.get_state()
    assert np.isfinite(f).all()
    assert abs(f[0].sum() - m0) <= 1e-7 * abs(m0)


@pytest.mark.parametrize("zclosed", [False, True])
def test_mixed_across_viscous_plane_chunks(zclosed):
    """The FP32 gradient/flux kernel and the FP32 divergence march through z
    in chunks of 16 planes; a 40-plane box crosses two chunk seams (and, with
    closed z faces, the one-sided closures meet a chunk boundary).  Momentum
    and energy carry the viscous terms: right-hand side to 1e-4 relative
    max, state after 4 steps to 1e-6 relative L2, against the FP64 oracle."""
    case = cases.wavy_box(20, 12, 40, scheme="WENO7", periodic=(1, 1, 0) if zclosed else (1, 1, 1))
    if zclosed:
        case.cfg.dif_bound[4] = case.cfg.dif_bound[5] = 1
    g, o = mixed_pair(case)
    g.rhs(1)
    o.rhs(1)
    e = rel_max(g.get_du(), o.get_du())
    print(f"[mixed] rhs wavy 20x12x40 z {'closed' if zclosed else 'periodic'}: " + " ".join(f"{x:.2e}" for x in e),
          flush=True)
    assert (e <= 1e-4).all(), e
    g.step(4)
    o.step(4)
    e = rel_l2(g.get_state(), o.get_state())
    print("[mixed] 4 steps L2: " + " ".join(f"{x:.2e}" for x in e), flush=True)
    assert (e <= 1e-6).all(), e
