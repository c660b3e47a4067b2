"""Analysis events (OCFD_ana, src/OCFD_ana.cu:588-619) on the GPU.

Q criterion against the oracle (relative max-norm 1e-12, the parity bar);
the running time average against the mean of the snapshots it saw (1e-13
relative: the running formula rounds once per sample) and through its file
round trip; the plane, scheme and Q files against the arrays they are made
of; the hybrid proportions against counts of the flags; the NaN / negative-T
guard on healthy and on broken fields.  Files are written to the working
directory, as the reference does, so every test runs in its own tmp_path.
"""
import struct

import numpy as np
import pytest

import oracle as O
from pyocfd import capi, cases

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def in_tmp(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    return tmp_path


def read_records(path):
    """All Fortran records of a file as bytes."""
    out = []
    with open(path, "rb") as fh:
        while True:
            h = fh.read(4)
            if not h:
                return out
            n = struct.unpack("<i", h)[0]
            out.append(fh.read(n))
            assert struct.unpack("<i", fh.read(4))[0] == n


@pytest.mark.parametrize("scheme_vis,steps", [("CD6", 3), ("CD8", 0)])
def test_q_matches_oracle(scheme_vis, steps):
    case = cases.wavy_box(20, 16, 12, scheme_vis=scheme_vis)
    g = case.load(capi.Solver(case.cfg))
    o = case.load(O.Oracle(case.cfg))
    g.step(steps)
    o.step(steps)
    qg, qo = g.get_q(), o.get_q()
    err = np.abs(qg - qo).max() / np.abs(qo).max()
    print(f"[ana] Q wavy {scheme_vis} after {steps} steps: rel err {err:.2e}", flush=True)
    assert err <= 1e-12


def test_q_ramp_closures_and_files():
    case = cases.ramp(48, 32, 10, nsteps=3)
    g = case.load(capi.Solver(case.cfg))
    o = case.load(O.Oracle(case.cfg))
    g.step(2)
    o.step(2)
    qg, qo = g.get_q(), o.get_q()
    err = np.abs(qg - qo).max() / np.abs(qo).max()
    print(f"[ana] Q ramp after 2 steps: rel err {err:.2e}", flush=True)
    assert err <= 1e-12
    g.analysis(103)
    q = capi.read_records("Q.dat", 48, 32, 10, 1)
    assert np.array_equal(q.reshape(qg.shape), qg)
    lines = open("Q000000.dat").read().splitlines()
    assert lines[0] == "variables=x,y,z,Q" and lines[1] == "zone i=48 ,j=32 ,k=10"
    assert len(lines) == 2 + qg.size
    x, y, z = case.mesh
    for p in (0, 777, qg.size - 1):
        vals = [float(lines[2 + p][15 * c:15 * (c + 1)]) for c in range(4)]
        want = [x.reshape(-1)[p], y.reshape(-1)[p], z.reshape(-1)[p], qg.reshape(-1)[p]]
        assert np.allclose(vals, want, rtol=0, atol=5e-7 + 1e-6 * np.abs(want)), (p, vals, want)


def test_q_mixed_precision_close_to_fp64():
    c64 = cases.wavy_box(20, 16, 12)
    c32 = cases.wavy_box(20, 16, 12, precision=capi.PREC_MIXED)
    q64 = c64.load(capi.Solver(c64.cfg)).get_q()
    q32 = c32.load(capi.Solver(c32.cfg)).get_q()
    assert np.abs(q32 - q64).max() <= 1e-5 * np.abs(q64).max()


def test_time_average_running_mean_and_files():
    case = cases.tgv(16, nsteps=8)
    case.cfg.kstep_save = 3
    capi.set_analysis(case.cfg, [(101, 1)])
    g = case.load(capi.Solver(case.cfg))
    snaps = []
    for s in range(4):
        g.step(1)
        snaps.append(np.stack(g.get_flow()[0]))
        g.analysis(101, 0)
    avg, n, tt = g.get_time_average()
    assert n == 4 and abs(tt - 4 * case.cfg.dt) <= 1e-15
    want = np.mean(snaps, axis=0)
    err = np.abs(np.stack(avg) - want).max() / np.abs(want).max()
    print(f"[ana] time average of 4 snapshots: rel err {err:.2e}", flush=True)
    assert err <= 1e-13
    # written at Istep = 3 (KSTEP_SAVE), named by the sample count
    prims, istep, t = capi.read_flow_file("OCFD00000003.average", 16, 16, 16)
    assert istep == 3 and abs(t - 3 * case.cfg.dt) <= 1e-15
    assert np.abs(np.stack(prims) - np.mean(snaps[:3], axis=0)).max() <= 1e-13 * np.abs(want).max()
    # a new run picks the average up from ./opencfd.average and continues it
    import shutil
    shutil.copy("OCFD00000003.average", "opencfd.average")
    g2 = case.load(capi.Solver(case.cfg))
    g2.step(1)
    x = np.stack(g2.get_flow()[0])
    g2.analysis(101, 0)
    avg2, n2, tt2 = g2.get_time_average()
    assert n2 == 4 and abs(tt2 - 4 * case.cfg.dt) <= 1e-15
    want2 = (3 * np.stack(prims) + x) / 4
    assert np.abs(np.stack(avg2) - want2).max() <= 1e-13 * np.abs(want2).max()


def test_plane_saves():
    nx, ny, nz = 20, 16, 12
    case = cases.wavy_box(nx, ny, nz)
    capi.set_analysis(case.cfg, [(104, 1, [2, 2, 1, 7]), (105, 1, [1, 1, 19]), (106, 1, [1, 3, 0])])
    g = case.load(capi.Solver(case.cfg))
    flows = []
    for s in range(2):
        g.step(1)
        flows.append(g.get_flow())
        for ev, style in enumerate((104, 105, 106)):
            g.analysis(style, ev)
    checks = [("XY", 0, 2, [(1, 2), (7, 2)]), ("YZ", 1, 1, [(19, 1)]), ("XZ", 2, 1, [(0, 3)])]
    for tag, ev, count, bands in checks:
        for p, (first, band) in enumerate(bands):
            recs = read_records(f"Savedata-{tag}{p:03d}.dat")
            assert len(recs) == 2 * (1 + 5 * band)  # two calls appended
            for call in range(2):
                prims, istep, tt = flows[call]
                r = recs[call * (1 + 5 * band):(call + 1) * (1 + 5 * band)]
                assert struct.unpack("<id", r[0]) == (istep, tt)
                for b in range(band):
                    for m in range(5):
                        a = np.frombuffer(r[1 + 5 * b + m], dtype="<f8")
                        f = prims[m]
                        if tag == "XY":
                            want = f[first + b].reshape(-1)
                        elif tag == "YZ":
                            want = f[:, :, first + b].reshape(-1)
                        else:
                            want = f[:, first + b, :].reshape(-1)
                        assert np.array_equal(a, want), (tag, p, call, b, m)


def test_scheme_file_and_proportions():
    nx, ny, nz = 20, 16, 12
    case = cases.wavy_box(nx, ny, nz, scheme="SCHEME_HYBRIDAUTO", amp=0.08)
    # thresholds at the 30% and 70% points of the sensor value, so that all
    # three flag types occur (periodic box: neighbours by np.roll)
    P = case.prims[0] * case.prims[4]
    dp = sum(np.abs(-np.roll(P, 1, a) + 2 * P - np.roll(P, -1, a)) / (np.roll(P, 1, a) + 2 * P + np.roll(P, -1, a))
             for a in range(3))
    case.cfg.hy_dp_intv[0], case.cfg.hy_dp_intv[1] = np.quantile(dp, [0.3, 0.7])
    g = case.load(capi.Solver(case.cfg))
    g.rhs(1)  # flags of the loaded state; ghosts hold the periodic copies
    fx, fy, fz = g.get_flags()
    counts = [np.bincount(f.reshape(-1), minlength=4)[1:4] for f in (fx, fy, fz)]
    assert all((c > 0).all() for c in counts), counts
    fr = g.hybrid_proportions()
    other = [ny * nz, nx * nz, nx * ny]
    ln = [nx + 1, ny + 1, nz + 1]
    for d in range(3):
        assert np.allclose(fr[d], counts[d] / other[d] / ln[d], rtol=1e-15, atol=0)
    g.analysis(107)  # prints the same numbers
    g.analysis(102)
    (d, u, v, w, T), istep, _ = g.get_flow()
    p00 = 1.0 / (case.cfg.gamma * case.cfg.ama ** 2)
    K = nz // 2
    flat = [fx.reshape(-1), fy.reshape(-1), fz.reshape(-1)]
    rowlen = [nx + 1, nx, nx]
    rows = [ny, ny + 1, ny]
    for dd, tag in enumerate("xyz"):
        lines = open(f"Scheme_{tag}{istep:08d}.dat").read().splitlines()
        assert lines[0] == "variables=scheme" and lines[1] == f"zone i={nx}, j={ny}"
        assert len(lines) == 2 + nx * ny
        for j in range(ny):
            for i in range(nx):
                s = lines[2 + j * nx + i]
                flag, pv = int(s[:8]), float(s[8:])
                assert flag == flat[dd][(K * rows[dd] + j) * rowlen[dd] + i + 1]
                ii = (i + 1) % nx  # x = nx is the periodic ghost copy of x = 0
                assert abs(pv - p00 * d[K, j, ii] * T[K, j, ii]) <= 5e-7


def test_nan_and_negative_t_guard():
    n = 16
    case = cases.tgv(n, nsteps=2)
    g = case.load(capi.Solver(case.cfg))
    g.step(1)
    g.analysis(100)  # healthy: no error
    d, u, v, w, T = [a.copy() for a in case.prims]
    T.reshape(-1)[:5] = -1.0
    g.load_flow(d, u, v, w, T)
    g.analysis(100)  # five negative points: printed, not fatal
    T.reshape(-1)[:12] = -1.0
    g.load_flow(d, u, v, w, T)
    with pytest.raises(RuntimeError, match="status 7"):
        g.analysis(100)
    T = case.prims[4].copy()
    T[3, 4, 5] = np.nan
    g.load_flow(d, u, v, w, T)
    with pytest.raises(RuntimeError, match="status 7"):
        g.analysis(100)
