"""Parity of the HIP path (libocfd_hip.so on the MI355X) with the CPU oracle.

Bar (from the plain FP64 target): relative max-norm per conserved variable
    max|gpu - oracle| / max|oracle|  <=  1e-12
after the stated number of steps; hybrid flags bit-exact; metrics to 1e-13.
The two sides differ only in floating-point contraction (the GPU fuses
multiply-adds), so agreement is at the level of a few ulps per operation.
"""
import numpy as np
import pytest

import oracle as O
from pyocfd import capi, cases

pytestmark = pytest.mark.gpu
TOL = 1e-12


def relerr(a, b):
    """Per-variable relative max-norm of a (test) against b (oracle)."""
    a = np.asarray(a).reshape(a.shape[0], -1)
    b = np.asarray(b).reshape(b.shape[0], -1)
    return np.array([np.abs(a[v] - b[v]).max() / max(np.abs(b[v]).max(), 1e-300) for v in range(a.shape[0])])


def pair(case):
    g = case.load(capi.Solver(case.cfg))
    o = case.load(O.Oracle(case.cfg))
    return g, o


def report(tag, e):
    print(f"[parity] {tag}: " + " ".join(f"{x:.2e}" for x in e), flush=True)


@pytest.mark.parametrize("scheme", ["WENO7", "UP7", "WENO5", "WENO7_SYMBO", "WENO7_SYMBO_LIM", "NND2", "OMP6_HR",
                                    "OMP6_LD", "OMP6_CD8", "SCHEME_HYBRIDAUTO"])
def test_rhs_wavy_box_all_schemes(scheme):
    case = cases.wavy_box(20, 16, 12, scheme=scheme, scheme_vis="CD8" if scheme == "UP7" else "CD6")
    g, o = pair(case)
    assert np.allclose(g.get_metrics()[3:], o.get_metrics()[3:], rtol=1e-13, atol=1e-13)
    g.rhs(1)
    o.rhs(1)
    e = relerr(g.get_du(), o.get_du())
    report(f"rhs wavy {scheme}", e)
    assert (e <= TOL).all(), e


@pytest.mark.parametrize("scheme", ["WENO7", "UP7", "WENO5", "WENO7_SYMBO", "NND2", "OMP6_HR", "SCHEME_HYBRIDAUTO"])
@pytest.mark.parametrize("bounds,nonref", [((1, 0, 0, 1, 1, 1), (1, 0, 0, 1, 0, 0)),
                                           ((0, 1, 1, 0, 0, 0), (0, 1, 1, 1, 0, 0))])
def test_rhs_closures_low_high_both(scheme, bounds, nonref):
    """Boundary closures on the low face, the high face and both, with
    non-reflecting faces, on a non-periodic curvilinear box."""
    case = cases.wavy_box(20, 16, 12, scheme=scheme, amp=0.08, periodic=(0, 0, 0))
    for q in range(6):
        case.cfg.dif_bound[q] = bounds[q]
        case.cfg.non_ref[q] = nonref[q]
    g, o = pair(case)
    g.rhs(1)
    o.rhs(1)
    e = relerr(g.get_du(), o.get_du())
    report(f"rhs closures {scheme} {bounds}", e)
    assert (e <= TOL).all(), e


@pytest.mark.parametrize("intv", [(0.01, 0.03), (0.005, 0.02), (0.003, 0.01), (0.0, 0.0), (0.0, 1e30)])
@pytest.mark.parametrize("closed", [False, True])
def test_hybrid_flag_mixes(intv, closed):
    """hybrid_auto with HY_DP_INTV chosen for about 4 %, 31 %, 50 % and all of
    the faces on WENO7 or WENO5 (64x24x12 wavy box, periodic or closed on
    every face): waves whose faces mix UP7, WENO7 and WENO5 run each branch
    for their own lanes; every mix and the closures must give the oracle's
    flags and du; the FP32 bulk path is held to 1e-4."""
    case = cases.wavy_box(64, 24, 12, scheme="SCHEME_HYBRIDAUTO", amp=0.08,
                          periodic=(0, 0, 0) if closed else (1, 1, 1))
    case.cfg.hy_dp_intv[0], case.cfg.hy_dp_intv[1] = intv
    if closed:
        for q in range(6):
            case.cfg.dif_bound[q] = 1
    g, o = pair(case)
    g.rhs(1)
    o.rhs(1)
    for a, b in zip(g.get_flags(), o.get_flags()):
        assert np.array_equal(a, b)
    e = relerr(g.get_du(), o.get_du())
    report(f"rhs hybrid intv={intv} closed={closed}", e)
    assert (e <= TOL).all(), e
    case.cfg.precision = capi.PREC_MIXED
    m = case.load(capi.Solver(case.cfg))
    m.rhs(1)
    e = relerr(m.get_du(), o.get_du())
    report(f"rhs hybrid mixed intv={intv} closed={closed}", e)
    assert (e <= 1e-4).all(), e


@pytest.mark.parametrize("closed", [False, True])
def test_hybrid_tall_z(closed):
    """A 20x12x40 wavy box: the sensor's blocks (16 nodes along z each) and
    the gradient/flux and divergence kernels' blocks (32 planes on the FP64
    path, 16 on the FP32 path) meet inside the grid, with and
    without closures on every face; the oracle's flags bit for bit and its du
    to the bar."""
    case = cases.wavy_box(20, 12, 40, scheme="SCHEME_HYBRIDAUTO", amp=0.08,
                          periodic=(0, 0, 0) if closed else (1, 1, 1))
    case.cfg.hy_dp_intv[0], case.cfg.hy_dp_intv[1] = 0.005, 0.02
    if closed:
        for q in range(6):
            case.cfg.dif_bound[q] = 1
    g, o = pair(case)
    g.rhs(1)
    o.rhs(1)
    for a, b in zip(g.get_flags(), o.get_flags()):
        assert np.array_equal(a, b)
    e = relerr(g.get_du(), o.get_du())
    report(f"rhs hybrid tall z closed={closed}", e)
    assert (e <= TOL).all(), e
    case.cfg.precision = capi.PREC_MIXED
    m = case.load(capi.Solver(case.cfg))
    m.rhs(1)
    e = relerr(m.get_du(), o.get_du())
    report(f"rhs hybrid tall z mixed closed={closed}", e)
    assert (e <= 1e-4).all(), e


@pytest.mark.parametrize("kw", [{}, {"mzmax": 3, "tw": -1.0, "upper": 1, "withleading": 1},
                                {"mzmax": -2, "scheme": "SCHEME_HYBRIDAUTO"}])
def test_liftbody_ibc124_steps(kw):
    """IBC 124 on half a body of revolution: symmetry planes (metrics and
    primitives), inlet from the flow or a table, free-stream or tabulated
    upper face with its disturbance, isothermal or adiabatic wall with the
    seeded or multi-wave disturbance, JAC_BOUND 0 across the symmetry
    planes."""
    case = cases.liftbody(40, 20, 16, nsteps=5, **kw)
    g, o = pair(case)
    assert np.allclose(g.get_metrics()[3:], o.get_metrics()[3:], rtol=1e-13, atol=1e-13)
    g.step(5)
    o.step(5)
    e = relerr(g.get_state(), o.get_state())
    report(f"5 steps liftbody {kw}", e)
    assert (e <= TOL).all(), e


@pytest.mark.parametrize("scheme", ["WENO7", "SCHEME_HYBRIDAUTO", "OMP6_HR", "WENO7_SYMBO"])
def test_steps_wavy_box(scheme):
    case = cases.wavy_box(20, 16, 12, scheme=scheme)
    g, o = pair(case)
    g.step(6)
    o.step(6)
    e = relerr(g.get_state(), o.get_state())
    report(f"6 steps wavy {scheme}", e)
    assert (e <= TOL).all(), e
    (pg, ig, tg), (po, io, to) = g.get_flow(), o.get_flow()
    assert ig == io == 6 and tg == to
    assert relerr(np.stack(pg), np.stack(po)).max() <= TOL


def test_tgv_steps():
    case = cases.tgv(24, nsteps=10)
    g, o = pair(case)
    g.step(10)
    o.step(10)
    e = relerr(g.get_state(), o.get_state())
    report("10 steps tgv24", e)
    assert (e <= TOL).all(), e


def test_ramp_flags_rhs_and_steps():
    """IBC 108, DIF_BOUND closures, non-reflecting faces, hybrid flags, CD8,
    and the filter every 2 steps (x, y and the z-call of the y kernel)."""
    case = cases.ramp(48, 32, 10, nsteps=6, filter_every=2)
    g, o = pair(case)
    assert np.allclose(g.get_metrics()[3:], o.get_metrics()[3:], rtol=1e-13, atol=1e-13)
    assert np.array_equal(g.get_amu(), o.get_amu()) or np.allclose(g.get_amu(), o.get_amu(), rtol=1e-15)
    g.rhs(1)
    o.rhs(1)
    for a, b in zip(g.get_flags(), o.get_flags()):
        assert np.array_equal(a, b)
    e = relerr(g.get_du(), o.get_du())
    report("rhs ramp", e)
    assert (e <= TOL).all(), e
    g2, o2 = pair(case)
    g2.step(6)
    o2.step(6)
    e = relerr(g2.get_state(), o2.get_state())
    report("6 steps ramp (filter every 2)", e)
    assert (e <= TOL).all(), e
    for a, b in zip(g2.get_flags(), o2.get_flags()):
        assert np.array_equal(a, b)
    # the pressure of the last stage (the filter's input), faces included
    pg, po = g2.get_pressure(), o2.get_pressure()
    assert np.abs(pg - po).max() / np.abs(po).max() <= TOL


def test_ramp_long_run():
    """60 steps (180 right-hand sides, 6 filter passes) of the ramp with its
    shock, closures, IBC 108 and non-reflecting faces: the rounding
    differences (FMA contraction) must not grow past the bar."""
    case = cases.ramp(48, 32, 10, nsteps=60, filter_every=10)
    g, o = pair(case)
    for n in (20, 20, 20):
        g.step(n)
        o.step(n)
        e = relerr(g.get_state(), o.get_state())
        report(f"ramp after {g.get_flow()[1]} steps", e)
        assert (e <= TOL).all(), e
    for a, b in zip(g.get_flags(), o.get_flags()):
        assert np.array_equal(a, b)


def test_ramp_wall_and_inlet_boundary_values():
    case = cases.ramp(40, 24, 8, nsteps=3, filt=False)
    g, o = pair(case)
    g.step(3)
    o.step(3)
    (pg, _, _), (po, _, _) = g.get_flow(), o.get_flow()
    tw = case.cfg.bc_rpara[0]
    assert np.all(pg[4][:, 0, :] == tw)           # isothermal wall
    assert np.all(pg[1][:, 0, :] == 0.0)          # no slip
    # inlet column x=0 above the wall is the inflow profile exactly; the wall
    # point itself gets rho = p_w/T_w from the extrapolated wall pressure
    assert np.array_equal(pg[0][:, 1:, 0], np.broadcast_to(case.inlet[0][1:], pg[0][:, 1:, 0].shape))
    assert np.array_equal(pg[0][:, 1:, 0], po[0][:, 1:, 0])
    assert relerr(np.stack(pg), np.stack(po)).max() <= TOL


@pytest.mark.parametrize("variant", ["adiabatic", "not_normal", "x_begin", "split"])
def test_ibc108_variants(variant, monkeypatch):
    """IBC 108 paths beside the isothermal ramp: adiabatic wall (TW <= 0),
    wall velocity only (BC_NPARA(4) != 0), a top row that starts at x_begin
    > 0 (XUB), and the three separate boundary launches (OCFD_BC108_SPLIT=1)
    instead of the merged one: 3 steps against the oracle, primitives and
    state."""
    case = cases.ramp(40, 24, 8, nsteps=3, filt=False)
    if variant == "adiabatic":
        case.cfg.bc_rpara[0] = -1.0
    elif variant == "not_normal":
        case.cfg.bc_npara[3] = 1
    elif variant == "x_begin":
        case.cfg.bc_rpara[6] = 0.3
    elif variant == "split":
        monkeypatch.setenv("OCFD_BC108_SPLIT", "1")
    g, o = pair(case)
    g.step(3)
    o.step(3)
    (pg, _, _), (po, _, _) = g.get_flow(), o.get_flow()
    e = relerr(np.stack(pg), np.stack(po))
    report(f"ibc108 {variant} prims", e)
    assert (e <= TOL).all(), e
    e = relerr(g.get_state(), o.get_state())
    report(f"ibc108 {variant} state", e)
    assert (e <= TOL).all(), e


def test_full_size_tgv_mass_conservation():
    """64^3 periodic TGV (the C1 grid): total mass is conserved to round-off
    over 20 steps (the inviscid and viscous fluxes telescope and the filter
    is off), and the solution stays finite."""
    case = cases.tgv(64, nsteps=20)
    g = case.load(capi.Solver(case.cfg))
    m0 = g.get_state()[0].sum()
    g.step(20)
    f = g.get_state()
    assert np.isfinite(f).all()
    assert abs(f[0].sum() - m0) <= 1e-12 * abs(m0)


def test_full_size_ramp_runs_finite():
    """512x128x64 ramp (the C2 grid): 12 steps with the filter every 10 are
    finite, the wall stays isothermal and no flag is outside 1..3."""
    case = cases.ramp(512, 128, 64, nsteps=12)
    g = case.load(capi.Solver(case.cfg))
    g.step(12)
    (p, istep, _) = g.get_flow()
    assert istep == 12
    assert all(np.isfinite(a).all() for a in p)
    assert np.all(p[4][:, 0, :] == case.cfg.bc_rpara[0])
    for fl in g.get_flags():
        inner = fl[1:-1, 1:-1, 1:-1]
        assert inner.min() >= 1 and inner.max() <= 3


def test_repeatable_bitwise():
    """No atomics: two runs give identical bytes."""
    case = cases.ramp(48, 32, 10, nsteps=4, filter_every=2)
    a = case.load(capi.Solver(case.cfg))
    b = case.load(capi.Solver(case.cfg))
    a.step(4)
    b.step(4)
    assert np.array_equal(a.get_state(), b.get_state())


def test_jacobi_writer_round_trip(tmp_path):
    """OCFD3d-Jacobi.dat written from a mesh-derived context reloads to the
    same metrics and the same right-hand side, bit for bit."""
    case = cases.wavy_box(20, 16, 12, scheme="WENO7")
    a = case.load(capi.Solver(case.cfg))
    path = tmp_path / "OCFD3d-Jacobi.dat"
    a.write_jacobi(path)
    n = case.npoints
    arr = capi.read_records(path, case.cfg.nx, case.cfg.ny, case.cfg.nz, 13)
    assert np.array_equal(arr[:3], np.stack(case.mesh))
    b = capi.Solver(case.cfg)
    b.load_jacobi(arr)
    b.load_flow(*case.prims, istep=0, tt=0.0)
    b.init()
    assert np.array_equal(a.get_metrics(), b.get_metrics())
    a.rhs(1)
    b.rhs(1)
    assert np.array_equal(a.get_du(), b.get_du())
    assert n == arr[0].size


def test_jacobi_file_layout_and_oracle_metrics(tmp_path):
    """OCFD3d-Jacobi.dat parsed byte by byte as the reference reads it
    (src/OCFD_Comput_Jacobian3d.cu:148-215: 13 arrays Axx, Ayy, Azz, Akx ..
    Asz, Ajac, each NZ Fortran records int32 n | n bytes | int32 n of NX*NY
    float64, x fastest), on a curvilinear closed box whose metrics use the
    derivative closures: the coordinates as loaded, the ten metric arrays
    equal to the oracle's from the same mesh to 1e-13, nothing after them."""
    import struct
    case = cases.wavy_box(20, 16, 12, scheme="WENO7", periodic=(0, 0, 0))
    nx, ny, nz = case.cfg.nx, case.cfg.ny, case.cfg.nz
    a = case.load(capi.Solver(case.cfg))
    path = tmp_path / "OCFD3d-Jacobi.dat"
    a.write_jacobi(path)
    raw = path.read_bytes()
    off, arrays = 0, []
    for _ in range(13):
        planes = []
        for _ in range(nz):
            (nb,) = struct.unpack_from("<i", raw, off)
            assert nb == nx * ny * 8
            planes.append(np.frombuffer(raw, "<f8", nx * ny, off + 4).reshape(ny, nx))
            (nb2,) = struct.unpack_from("<i", raw, off + 4 + nb)
            assert nb2 == nb
            off += nb + 8
        arrays.append(np.stack(planes))
    assert off == len(raw)
    for m in range(3):
        assert np.array_equal(arrays[m], case.mesh[m])
    om = case.load(O.Oracle(case.cfg)).get_metrics()
    for m in range(3, 13):
        e = np.abs(arrays[m] - om[m]).max() / np.abs(om[m]).max()
        assert e <= 1e-13, (m, e)


def test_step_diagnostics():
    """Means of f[4] and T over the interior, NaN / negative counts of T over
    the field with its ghost layers (the KSTEP_SHOW report)."""
    case = cases.wavy_box(20, 16, 12, scheme="WENO7")
    g = case.load(capi.Solver(case.cfg))
    g.step(2)
    e, t, nn, ng = g.diagnostics()
    f = g.get_state()
    T = g.get_flow()[0][4]
    assert abs(e - f[4].mean()) <= 1e-13 * abs(f[4].mean())
    assert abs(t - T.mean()) <= 1e-13 * abs(T.mean())
    assert nn == 0 and ng == 0
    # three negative temperatures and one NaN away from the periodic faces
    d, u, v, w, T0 = [a.copy() for a in case.prims]
    T0[5, 6, 7] = T0[6, 7, 8] = T0[5, 8, 9] = -1.0
    T0[6, 6, 6] = np.nan
    h = capi.Solver(case.cfg)
    h.load_mesh(*case.mesh)
    h.load_flow(d, u, v, w, T0, istep=0, tt=0.0)
    h.init()
    e, t, nn, ng = h.diagnostics()
    assert nn == 1 and ng == 3 and np.isnan(e)
