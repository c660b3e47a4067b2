"""The drop-in driver (driver/opencfd.c: read_parameters, init,
NS_solver_real on libocfd_hip.so) run on the reference's own input files,
against the oracle advanced from the same inputs.

Bar: the flow file it writes after N steps matches the oracle's primitives
to the parity bar (relative max-norm 1e-12 per variable); the analysis
events named in the input file run at their steps and leave their output.
"""
import os
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import oracle as O
from pyocfd import capi, cases

pytestmark = pytest.mark.gpu
EXE = Path(__file__).resolve().parents[1] / "opencfd-scu_amd" / "opencfd-scu.out"


def test_driver_ramp_with_analysis_events(tmp_path):
    nsteps = 4
    case = cases.ramp(48, 32, 10, nsteps=nsteps, filter_every=2)
    case.cfg.kstep_save = nsteps
    case.cfg.kstep_show = 2
    capi.set_analysis(case.cfg, [(100, 1), (107, 2), (104, nsteps, [1, 1, 5]), (101, 1)])
    case.write(tmp_path)
    print(f"[driver] ramp 48x32x10, {nsteps} steps, in {tmp_path}", flush=True)
    r = subprocess.run([str(EXE)], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    sys.stdout.write(r.stdout[-3000:])
    sys.stderr.write(r.stderr[-3000:])
    assert r.returncode == 0
    assert "OK! opencfd is finished" in r.stdout
    assert r.stdout.count("The first type scheme of Hybrid schemes in direction X") == nsteps // 2
    assert r.stdout.count("It is analyzing NAN") == nsteps
    prims, istep, tt = capi.read_flow_file(tmp_path / f"OCFD{nsteps:08d}.dat", 48, 32, 10)
    o = case.load(O.Oracle(case.cfg))
    o.step(nsteps)
    po, io, to = o.get_flow()
    assert istep == io == nsteps and tt == to
    e = [np.abs(a - b).max() / np.abs(b).max() for a, b in zip(prims, po)]
    print("[driver] flow file vs oracle: " + " ".join(f"{x:.2e}" for x in e), flush=True)
    assert max(e) <= 1e-12
    # time average over the steps, written at KSTEP_SAVE
    avg, na, ta = capi.read_flow_file(tmp_path / f"OCFD{nsteps:08d}.average", 48, 32, 10)
    assert na == nsteps
    assert os.path.getsize(tmp_path / "Savedata-XY000.dat") == 4 + 12 + 4 + 5 * (4 + 48 * 32 * 8 + 4)


def test_driver_liftbody_reads_boundary_tables(tmp_path):
    """IBC 124 through the driver: flow-inlet-section.dat and
    flow-outboundary.dat read by init, symmetry planes, the adiabatic wall with
    the multi-wave disturbance."""
    nsteps = 3
    case = cases.liftbody(40, 20, 16, nsteps=nsteps, mzmax=3, tw=-1.0, upper=1, withleading=1)
    case.cfg.kstep_save = nsteps
    case.write(tmp_path)
    print(f"[driver] liftbody 40x20x16, {nsteps} steps, in {tmp_path}", flush=True)
    r = subprocess.run([str(EXE)], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    sys.stdout.write(r.stdout[-2000:])
    sys.stderr.write(r.stderr[-2000:])
    assert r.returncode == 0
    assert "read inlet boundary data" in r.stdout and "read upper boundary data" in r.stdout
    prims, istep, tt = capi.read_flow_file(tmp_path / f"OCFD{nsteps:08d}.dat", 40, 20, 16)
    o = case.load(O.Oracle(case.cfg))
    o.step(nsteps)
    po, io, to = o.get_flow()
    e = [np.abs(a - b).max() / np.abs(b).max() for a, b in zip(prims, po)]
    print("[driver] liftbody flow file vs oracle: " + " ".join(f"{x:.2e}" for x in e), flush=True)
    assert istep == io == nsteps and max(e) <= 1e-12
