"""The reference's own benchmark deck through the drop-in driver, and restart
through Opencfd.msg.

* tests/golden/opencfd-scu.in and flow1d-inlet.dat are the reference's
  benchmark_comp/ input files, byte for byte (tests/golden/SHA256SUMS.json).
  The deck asks for a 25x240x20 ramp, WENO7_SYMBO + CD8, IBC 108 with the
  inlet read from flow1d-inlet.dat, z periodic, DT 0.01 up to END_TIME 10
  (1000 steps), a flow file every 1000 steps.  The mesh and the initial field
  it also needs come from cases.deck_ramp (the reference's own generators are
  separate MPI programs).  Bar: the OCFD00001000.dat the driver writes equals
  the oracle's primitives after 1000 steps to a relative max-norm of 1e-12
  per variable, or, where rounding alone moves this long run further, to
  three times what it moves the oracle itself: the same oracle built with
  contracted multiply-adds (oracle/liboracle_fma.so, which is what the GPU
  compiler does, in other places) differs from the plain build by a few
  1e-12 after 1000 steps of this deck (WENO7-SYMBO's nonlinear weights and
  TV limiter switch amplify last-bit differences), and the GPU differs from
  the oracle by the same amount.
* Restart (src/OCFD_IO.c:37-91 reads Opencfd.msg as one int32, then
  OCFD%08d.dat; src/OCFD_NS_Solver.cu:63-167 runs while tt < END_TIME): a run
  restarted from the flow file of step k reaches the same state at step N as
  the run that went through.
"""
import hashlib
import json
import shutil
import struct
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import oracle as O
from pyocfd import capi, cases

REPO = Path(__file__).resolve().parents[1]
GOLDEN = REPO / "tests" / "golden"
EXE = REPO / "opencfd-scu_amd" / "opencfd-scu.out"
DECK = GOLDEN / "opencfd-scu.in"
INLET = GOLDEN / "flow1d-inlet.dat"


def relerr(a, b):
    return [float(np.abs(x - y).max() / np.abs(y).max()) for x, y in zip(a, b)]


def run_driver(d, tag):
    print(f"[deck] {tag}: driver in {d}", flush=True)
    r = subprocess.run([str(EXE)], cwd=d, capture_output=True, text=True, timeout=300)
    sys.stdout.write(r.stdout[-1500:])
    sys.stderr.write(r.stderr[-1500:])
    assert r.returncode == 0
    assert "OK! opencfd is finished" in r.stdout
    return r.stdout


# ---------------------------------------------------------------- CPU checks
def test_fixtures_are_the_reference_files():
    sums = json.loads((GOLDEN / "SHA256SUMS.json").read_text())
    for name, h in sums.items():
        assert hashlib.sha256((GOLDEN / name).read_bytes()).hexdigest() == h, name


def test_deck_parses_to_the_benchmark_configuration():
    c = capi.read_config(DECK)
    assert (c.nx, c.ny, c.nz) == (25, 240, 20)
    assert c.scheme_invis == capi.WENO7_SYMBO and c.scheme_vis == capi.CD8
    assert list(c.iperiodic) == [0, 0, 1]
    assert list(c.dif_bound) == [0, 1, 0, 1, 0, 0] and list(c.non_ref) == [0, 1, 0, 1, 0, 0]
    assert c.ibc == 108 and list(c.bc_npara[:4]) == [10, 5, 1, 0]
    assert list(c.bc_rpara[:8]) == [2.84, 0.2, -320.0, -300.0, 0.1, -400.0, -50.0, 14.0]
    assert (c.re, c.ama, c.gamma, c.pr, c.t_ref) == (5581.4, 2.9, 1.4, 0.7, 108.1)
    assert (c.dt, c.end_time, c.kstep_save) == (0.01, 10.0, 1000)
    assert c.nfilter == 0     # the deck's filter lines are commented out


def test_deck_mesh_is_a_valid_grid_on_the_inlet_rows():
    case = cases.deck_ramp(DECK, INLET)
    x, y, z = case.mesh
    # the generated wall-normal distribution is the one the inlet profile is
    # given on (8 decimals in the file)
    assert np.abs(y[0, :, 0] - case.extra["inlet_y"]).max() < 1e-7
    # right-handed, non-degenerate cells: the 2-D cell areas are positive
    ax = np.diff(x[0], axis=1)[:-1] * np.diff(y[0], axis=0)[:, :-1] - \
        np.diff(y[0], axis=1)[:-1] * np.diff(x[0], axis=0)[:, :-1]
    assert (ax > 0).all()
    assert np.allclose(z[:, 0, 0], np.linspace(0.0, 14.0, 20))
    d, u, v, w, T = case.prims
    assert np.array_equal(d[3, :, 7], case.inlet[0]) and np.array_equal(T[5, :, 11], case.inlet[3])
    assert not w.any()


# ---------------------------------------------------------------- GPU checks
@pytest.mark.gpu
def test_reference_deck_1000_steps_through_the_driver(tmp_path):
    case = cases.deck_ramp(DECK, INLET)
    shutil.copyfile(DECK, tmp_path / "opencfd-scu.in")
    shutil.copyfile(INLET, tmp_path / "flow1d-inlet.dat")
    capi.write_records(tmp_path / "OCFD3d-Mesh.dat", list(case.mesh))
    capi.write_flow_file(tmp_path / "opencfd.dat", case.prims, 0, 0.0)
    out = run_driver(tmp_path, "benchmark deck 25x240x20, 1000 steps")
    assert "write data file: OCFD00001000.dat" in out
    prims, istep, tt = capi.read_flow_file(tmp_path / "OCFD00001000.dat", 25, 240, 20)
    assert istep == 1000
    print("[deck] oracle, 1000 steps", flush=True)
    o = case.load(O.Oracle(case.cfg))
    o.step(1000)
    po, io, to = o.get_flow()
    assert io == 1000 and tt == to
    print("[deck] oracle with contracted multiply-adds, 1000 steps", flush=True)
    f = case.load(O.Oracle(case.cfg, fma=True))
    f.step(1000)
    pf = f.get_flow()[0]
    # w is zero at the start and stays at the round-off of the blowing (z
    # periodic, nothing drives it): measured against the scale of u
    def dev(a, b):
        e = relerr(a, b)
        e[3] = float(np.abs(a[3] - b[3]).max() / np.abs(b[1]).max())
        return e

    e, noise = dev(prims, po), dev(pf, po)
    print("[deck] flow file after 1000 steps vs oracle: " + " ".join(f"{x:.2e}" for x in e), flush=True)
    print("[deck] oracle (fma) vs oracle:               " + " ".join(f"{x:.2e}" for x in noise), flush=True)
    assert all(np.isfinite(a).all() for a in prims)
    for v in range(5):
        assert e[v] <= max(1e-12, 3.0 * noise[v]), (v, e, noise)


@pytest.mark.gpu
def test_restart_through_opencfd_msg(tmp_path):
    """Continuous: 6 steps with a flow file every 3.  Restarted: Opencfd.msg
    names step 3, the run reads OCFD00000003.dat and goes on to END_TIME.
    The filter runs every 2 steps, so step 3's primitives carry no lag of it.

    The restart rebuilds the conservative state from the saved primitives
    (a prim -> cons round trip), so it differs from the continuous run by
    what three more steps make of last-bit differences.  Bars: the
    restarted GPU run equals the oracle restarted from the same flow file
    to 1e-12 per variable (the same inputs and round trip on both sides);
    it differs from the continuous GPU run by at most three times what the
    restart moves the oracle's own run (or 1e-12).  w, which only the
    seeded wall blowing drives, is measured against the scale of u."""
    n, k = 6, 3
    case = cases.ramp(48, 32, 10, nsteps=n, filter_every=2)
    case.cfg.kstep_save = k
    a, b = tmp_path / "through", tmp_path / "restart"
    case.write(a)
    run_driver(a, f"ramp 48x32x10, {n} steps, a flow file every {k}")
    case.write(b)
    shutil.copyfile(a / f"OCFD{k:08d}.dat", b / f"OCFD{k:08d}.dat")
    (b / "Opencfd.msg").write_bytes(struct.pack("<i", k))
    (b / "opencfd.dat").unlink()         # the restart must not fall back to it
    out = run_driver(b, f"restart from step {k}")
    assert f"OCFD{k:08d}.dat" in out
    pa, ia, ta = capi.read_flow_file(a / f"OCFD{n:08d}.dat", 48, 32, 10)
    pb, ib, tb = capi.read_flow_file(b / f"OCFD{n:08d}.dat", 48, 32, 10)
    assert ia == ib == n and ta == tb

    def dev(x, y):
        e = relerr(x, y)
        e[3] = float(np.abs(x[3] - y[3]).max() / np.abs(y[1]).max())
        return e

    # the oracle, continuous and restarted from the same flow file
    o = case.load(O.Oracle(case.cfg))
    o.step(n)
    po = o.get_flow()[0]
    pk, ik, tk = capi.read_flow_file(a / f"OCFD{k:08d}.dat", 48, 32, 10)
    r = O.Oracle(case.cfg)
    r.load_mesh(*case.mesh)
    r.load_flow(*pk, istep=ik, tt=tk)
    r.set_inlet_profile(case.inlet)
    r.init()
    r.step(n - k)
    pr = r.get_flow()[0]
    e_gpu, e_ref, noise = dev(pb, pa), dev(pb, pr), dev(pr, po)
    print("[deck] restarted vs continuous (GPU):       " + " ".join(f"{x:.2e}" for x in e_gpu), flush=True)
    print("[deck] restarted GPU vs restarted oracle:   " + " ".join(f"{x:.2e}" for x in e_ref), flush=True)
    print("[deck] restarted vs continuous (oracle):    " + " ".join(f"{x:.2e}" for x in noise), flush=True)
    assert max(e_ref) <= 1e-12, e_ref
    for v in range(5):
        assert e_gpu[v] <= max(1e-12, 3.0 * noise[v]), (v, e_gpu, noise)
