"""CHARTERIC = 1 (reconstruction in characteristic space,
src/OCFD_flux_charteric.cu) on the GPU against the oracle.

Bar: the parity bar of the physical path, relative max-norm per conserved
variable <= 1e-12 (one right-hand side on a curvilinear box; several steps of
the ramp with its closures, non-reflecting faces, IBC 108 and the filter);
FP32 bulk against the FP64 oracle to 1e-4 relative (the mixed-precision bar).
"""
import numpy as np
import pytest

import oracle as O
from pyocfd import capi, cases

pytestmark = pytest.mark.gpu
TOL = 1e-12


def relerr(a, b):
    a = np.asarray(a).reshape(a.shape[0], -1)
    b = np.asarray(b).reshape(b.shape[0], -1)
    return np.array([np.abs(a[v] - b[v]).max() / max(np.abs(b[v]).max(), 1e-300) for v in range(a.shape[0])])


@pytest.mark.parametrize("scheme", ["WENO7_SYMBO", "WENO7_SYMBO_LIM", "SCHEME_HYBRIDAUTO"])
def test_char_rhs_wavy_box(scheme):
    case = cases.wavy_box(20, 16, 12, scheme=scheme, amp=0.08)
    case.cfg.charteric = 1
    g = case.load(capi.Solver(case.cfg))
    o = case.load(O.Oracle(case.cfg))
    g.rhs(1)
    o.rhs(1)
    e = relerr(g.get_du(), o.get_du())
    print(f"[char] rhs wavy {scheme}: " + " ".join(f"{x:.2e}" for x in e), flush=True)
    assert (e <= TOL).all(), e
    # the projection changes the result (the test would pass vacuously otherwise)
    case.cfg.charteric = 0
    p = case.load(capi.Solver(case.cfg))
    p.rhs(1)
    assert relerr(p.get_du(), o.get_du()).max() > 1e-8


@pytest.mark.parametrize("bounds,nonref", [((1, 0, 0, 1, 1, 1), (1, 0, 0, 1, 0, 0)),
                                           ((0, 1, 1, 0, 0, 0), (0, 1, 1, 1, 0, 0))])
def test_char_rhs_closures(bounds, nonref):
    """Closures on one face (low or high) and on both, with non-reflecting
    faces: the boundary nodes take the physical-space differences."""
    case = cases.wavy_box(20, 16, 12, scheme="WENO7_SYMBO", amp=0.08, periodic=(0, 0, 0))
    case.cfg.charteric = 1
    for q in range(6):
        case.cfg.dif_bound[q] = bounds[q]
        case.cfg.non_ref[q] = nonref[q]
    g = case.load(capi.Solver(case.cfg))
    o = case.load(O.Oracle(case.cfg))
    g.rhs(1)
    o.rhs(1)
    e = relerr(g.get_du(), o.get_du())
    print(f"[char] rhs closures {bounds} {nonref}: " + " ".join(f"{x:.2e}" for x in e), flush=True)
    assert (e <= TOL).all(), e


def test_char_ramp_steps():
    case = cases.ramp(48, 32, 10, nsteps=6, filter_every=2)
    case.cfg.charteric = 1
    g = case.load(capi.Solver(case.cfg))
    o = case.load(O.Oracle(case.cfg))
    g.step(6)
    o.step(6)
    e = relerr(g.get_state(), o.get_state())
    print("[char] 6 steps ramp: " + " ".join(f"{x:.2e}" for x in e), flush=True)
    assert (e <= TOL).all(), e
    fg, fo = g.get_flags(), o.get_flags()
    assert all(np.array_equal(a, b) for a, b in zip(fg, fo))


def test_char_mixed_precision():
    c64 = cases.wavy_box(20, 16, 12, scheme="WENO7_SYMBO", amp=0.08)
    c64.cfg.charteric = 1
    c32 = cases.wavy_box(20, 16, 12, scheme="WENO7_SYMBO", amp=0.08, precision=capi.PREC_MIXED)
    c32.cfg.charteric = 1
    o = c64.load(O.Oracle(c64.cfg))
    g = c32.load(capi.Solver(c32.cfg))
    o.rhs(1)
    g.rhs(1)
    e = relerr(g.get_du(), o.get_du())
    print("[char] mixed rhs: " + " ".join(f"{x:.2e}" for x in e), flush=True)
    assert (e <= 1e-4).all(), e


def test_char_rejects_other_schemes():
    case = cases.wavy_box(20, 16, 12, scheme="WENO7")
    case.cfg.charteric = 1
    with pytest.raises(RuntimeError, match="CHARTERIC"):
        capi.Solver(case.cfg)
