"""Steps replayed from a captured HIP graph (ocfd_step with 4 or more steps,
when no launch argument changes between steps) against the same steps
launched one by one (OCFD_GRAPH=0 at ocfd_create).

Bar: bitwise the same conserved state, primitives, step counter and time,
after every call of a sequence that mixes graph replays (4 and 6 steps) with
single eager steps and downloads between them.
"""
from pathlib import Path

import numpy as np
import pytest

from pyocfd import capi, cases

pytestmark = pytest.mark.gpu
PARAMS = Path(__file__).resolve().parents[1] / "opencfd-scu_amd" / "pyocfd" / "sensor_params.npy"


def _ramp_steady_wall():
    c = cases.ramp(48, 32, 10, nsteps=40, filt=False)
    c.cfg.bc_rpara[4] = 0.0  # no time-dependent wall disturbance: the steps replay
    return c


CASES = {
    "tgv_weno7": lambda: cases.tgv(24, nsteps=40),
    "wavy_hybrid": lambda: cases.wavy_box(20, 16, 12, scheme="SCHEME_HYBRIDAUTO"),
    "ramp_ibc108": _ramp_steady_wall,
    "tgv_mixed": lambda: cases.tgv(24, nsteps=40, precision=capi.PREC_MIXED),
    "ramp_mixed_learned": lambda: _ramp_steady_wall(),
}


@pytest.mark.parametrize("name", list(CASES))
def test_graph_replay_matches_eager(name, monkeypatch):
    case = CASES[name]()
    if name == "ramp_mixed_learned":
        case.cfg.precision = capi.PREC_MIXED
    sols = []
    for graph in ("0", "1"):
        monkeypatch.setenv("OCFD_GRAPH", graph)
        s = case.load(capi.Solver(case.cfg))
        if name == "ramp_mixed_learned":
            s.set_sensor(np.load(PARAMS), 0.2, 0.8)
        sols.append(s)
    eager, graph = sols
    for n in (4, 1, 6, 5):
        eager.step(n)
        graph.step(n)
        a, b = eager.get_state(), graph.get_state()
        (pa, ia, ta), (pb, ib, tb) = eager.get_flow(), graph.get_flow()
        print(f"[graph] {name}: after {ia} steps max |diff| = {np.abs(a - b).max():.3e}", flush=True)
        assert ia == ib and ta == tb
        assert np.array_equal(a, b)
        assert all(np.array_equal(x, y) for x, y in zip(pa, pb))
