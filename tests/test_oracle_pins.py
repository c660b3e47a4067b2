"""Pins of the CPU oracle against what the mathematics fixes.

The reference ships no tests, golden vectors or fixtures for this path, so the
oracle restatement is pinned by closed forms and invariants chosen so that a
plausible slip (a wrong coefficient, a shifted stencil, a mirrored index, a
dropped term) fails one of them:

* polynomial exactness of UP7 / CD6 / CD8 and of the D0 closures,
* observed convergence order of WENO7, WENO7-SYMBO, WENO5, OMP6 on smooth data,
* mirror symmetry between the P and M half-node fluxes,
* linear-data values of the inviscid boundary closures,
* Steger-Warming consistency (f+ + f- = Euler flux) and upwinding,
* the SSP-RK3 stability polynomial,
* telescoping conservation of the periodic RHS and exact freestream
  preservation on a uniform grid,
* metric identities on an affine mesh, Sutherland mu(1) = 1/Re,
* Jameson sensor zeros on linear pressure,
* the IBC 108 / IBC 124 wall extrapolations on polynomial profiles, the
  inlet / top / upper tables and the mirrored ghost layers,
* the filter stencil with the shock sensor forced to 1.
"""
import math

import numpy as np
import pytest

import oracle as O
from pyocfd import capi, cases

P, M = 0, 1


def nodes(n, h, x0=0.0):
    """x of nodes -4..n+3."""
    return x0 + (np.arange(n + 8) - 4) * h


@pytest.mark.parametrize("side", [P, M])
@pytest.mark.parametrize("deg", range(0, 8))
def test_up7_exact_for_degree_le_7(side, deg):
    n, h = 12, 0.05
    x = nodes(n, h, 0.3)
    f = x ** deg
    du = O.line_flux_derivative(capi.UP7, side, f, np.ones(n + 8), h)
    exact = -deg * x[4:-4] ** max(deg - 1, 0) if deg > 0 else np.zeros(n)
    assert np.allclose(du, exact, rtol=1e-9, atol=1e-9)


def test_up7_not_exact_for_degree_8():
    n, h = 12, 0.05
    x = nodes(n, h, 0.3)
    du = O.line_flux_derivative(capi.UP7, P, x ** 8, np.ones(n + 8), h)
    assert np.abs(du + 8 * x[4:-4] ** 7).max() > 1e-9


@pytest.mark.parametrize("svis,deg_max", [(capi.CD6, 6), (capi.CD8, 8)])
def test_central_exact_interior(svis, deg_max):
    n, h = 14, 0.07
    x = nodes(n, h, 0.2)
    for deg in range(deg_max + 1):
        out = O.line_central(svis, x ** deg, h)
        exact = deg * x[4:-4] ** max(deg - 1, 0) if deg else 0 * x[4:-4]
        assert np.allclose(out, exact, rtol=1e-9, atol=1e-9), deg
    out = O.line_central(svis, x ** (deg_max + 1), h)
    assert np.abs(out - (deg_max + 1) * x[4:-4] ** deg_max).max() > 1e-9


@pytest.mark.parametrize("svis", [capi.CD6, capi.CD8])
def test_d0_closures_orders(svis):
    """Closure points 0,1,2,3 are exact for degree 1,2,4,6 (and mirrored)."""
    n, h = 16, 0.1
    x = nodes(n, h, 0.5)
    for pt, deg in [(0, 1), (1, 2), (2, 4), (3, 6)]:
        f = x ** deg
        out = O.line_central(svis, f, h, bound_l=1, bound_r=1)
        ex = deg * x[4:-4] ** (deg - 1)
        assert abs(out[pt] - ex[pt]) < 1e-9 * max(1, abs(ex[pt])), (pt, deg)
        assert abs(out[n - 1 - pt] - ex[n - 1 - pt]) < 1e-9 * max(1, abs(ex[n - 1 - pt])), (pt, deg)
        if deg < 6:  # not exact one degree higher
            f2 = x ** (deg + 1)
            o2 = O.line_central(svis, f2, h, bound_l=1, bound_r=1)
            e2 = (deg + 1) * x[4:-4] ** deg
            assert abs(o2[pt] - e2[pt]) > 1e-9


def _order(scheme, side, n0=16, n1=24, f=lambda x: np.exp(1.5 * x) + 0.2 * np.sin(x),
           df=lambda x: 1.5 * np.exp(1.5 * x) + 0.2 * np.cos(x)):
    """Observed order between n0 and n1 points on smooth data without critical
    points (where the nonlinear weights of WENO-type schemes approach the
    linear ones), before round-off dominates."""
    errs = []
    for n in (n0, n1):
        h = 1.0 / n
        x = nodes(n, h)
        du = O.line_flux_derivative(scheme, side, f(x), np.ones(n + 8), h)
        errs.append(np.abs(du + df(x[4:-4])).max())
    return math.log(errs[0] / errs[1], n1 / n0), errs


# Formal orders: UP7/WENO7 7, OMP6 (HR, LD) 6, OMP6 with the CD8 parameters 8,
# WENO5 5, NND2 2.  WENO7-SYMBO's bandwidth-optimised linear weights
# (src/OCFD_Schemes.cu:415-419) make it 4th order.
@pytest.mark.parametrize("scheme,min_order,max_order", [
    (capi.WENO7, 6.5, 7.5), (capi.WENO7_SYMBO, 3.7, 4.3), (capi.UP7, 6.6, 7.5), (capi.WENO5, 4.5, 5.5),
    (capi.OMP6_HR, 5.5, 6.5), (capi.OMP6_LD, 5.5, 6.5), (capi.OMP6_CD8, 7.5, 8.5), (capi.NND2, 1.7, 2.3)])
@pytest.mark.parametrize("side", [P, M])
def test_convergence_order(scheme, min_order, max_order, side):
    order, errs = _order(scheme, side)
    assert min_order < order < max_order, (order, errs)


@pytest.mark.parametrize("scheme,smooth", [
    (capi.WENO7, 0), (capi.UP7, 0), (capi.WENO5, 0), (capi.WENO7_SYMBO, 0), (capi.NND2, 0),
    # The OMP6 M limiter is not the mirror of the P limiter: it uses
    # stencil[4] - stencil[3] where the mirror of the P side would use
    # stencil[4] - stencil[5] (src/OCFD_Schemes.cu:1051 vs :959).  On smooth
    # data the limiter is inactive and the linear parts are mirrors.
    (capi.OMP6_HR, 1), (capi.OMP6_CD8, 1)])
def test_m_flux_is_mirror_of_p_flux(scheme, smooth):
    """du_M(f)[i] == -du_P(reversed f)[n-1-i]: pins the M-side stencil mapping."""
    rng = np.random.default_rng(5)
    n, h = 20, 0.1
    if smooth:
        f = np.exp(0.1 * np.arange(n + 8))
    else:
        f = np.cumsum(rng.standard_normal(n + 8)) * 0.3 + 2.0
    duM = O.line_flux_derivative(scheme, M, f, np.ones(n + 8), h)
    duP = O.line_flux_derivative(scheme, P, f[::-1].copy(), np.ones(n + 8), h)
    assert np.allclose(duM, -duP[::-1], rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("hy,base", [(1, capi.UP7), (2, capi.WENO7), (3, capi.WENO5)])
@pytest.mark.parametrize("side", [P, M])
def test_hybrid_flags_select_schemes(hy, base, side):
    rng = np.random.default_rng(7)
    n, h = 16, 0.1
    f = np.cumsum(rng.standard_normal(n + 8))
    du_h = O.line_flux_derivative(capi.HYBRIDAUTO, side, f, np.ones(n + 8), h, flags=np.full(n + 1, hy))
    du_b = O.line_flux_derivative(base, side, f, np.ones(n + 8), h)
    assert np.array_equal(du_h, du_b)


@pytest.mark.parametrize("scheme", [capi.WENO7, capi.UP7, capi.HYBRIDAUTO, capi.WENO7_SYMBO])
def test_inviscid_closures_on_linear_data(scheme):
    """With DIF_BOUND on both faces and f = a + b x: the P flux at the first
    face is the first node value (first order) so du_P(0) = -b/2 and every
    other node is exact; the M flux at the last face is the last node value so
    du_M(n-1) = -b/2 and every other node is exact."""
    n, h, b = 20, 0.1, 1.7
    x = nodes(n, h)
    f = 0.4 + b * x
    flags = np.full(n + 1, 2) if scheme == capi.HYBRIDAUTO else None
    duP = O.line_flux_derivative(scheme, P, f, np.ones(n + 8), h, bound_l=1, bound_r=1, flags=flags)
    duM = O.line_flux_derivative(scheme, M, f, np.ones(n + 8), h, bound_l=1, bound_r=1, flags=flags)
    assert abs(duP[0] + b / 2) < 1e-12
    assert np.allclose(duP[1:], -b, atol=1e-12)
    assert abs(duM[-1] + b / 2) < 1e-12
    assert np.allclose(duM[:-1], -b, atol=1e-12)
    # ghosts must not be read when both closures are on
    f2 = f.copy()
    f2[:4] = np.nan
    f2[-4:] = np.nan
    assert np.array_equal(O.line_flux_derivative(scheme, P, f2, np.ones(n + 8), h, 1, 1, flags=flags), duP)
    assert np.array_equal(O.line_flux_derivative(scheme, M, f2, np.ones(n + 8), h, 1, 1, flags=flags), duM)


def test_nonreflecting_drops_boundary_term():
    n, h = 12, 0.1
    rng = np.random.default_rng(1)
    f = rng.standard_normal(n + 8)
    a = O.line_flux_derivative(capi.UP7, P, f, np.ones(n + 8), h, nonref=1)
    b = O.line_flux_derivative(capi.UP7, P, f, np.ones(n + 8), h, nonref=0)
    assert a[0] == 0.0 and np.array_equal(a[1:], b[1:])
    a = O.line_flux_derivative(capi.UP7, M, f, np.ones(n + 8), h, nonref=1)
    b = O.line_flux_derivative(capi.UP7, M, f, np.ones(n + 8), h, nonref=0)
    assert a[-1] == 0.0 and np.array_equal(a[:-1], b[:-1])


def euler_flux(d, u, v, w, c, A, gamma):
    p = d * c * c / gamma
    U = A[0] * u + A[1] * v + A[2] * w
    E = p / (gamma - 1) + 0.5 * d * (u * u + v * v + w * w)
    return np.array([d * U, d * u * U + p * A[0], d * v * U + p * A[1], d * w * U + p * A[2], (E + p) * U])


def test_steger_warming_consistency_and_upwinding():
    rng = np.random.default_rng(3)
    for _ in range(200):
        d = rng.uniform(0.2, 3)
        u, v, w = rng.uniform(-2, 2, 3)
        c = rng.uniform(0.2, 3)
        A = rng.uniform(-2, 2, 3)
        fp, fm = O.split_point(d, u, v, w, c, *A, gamma=1.4, eps=0.0)
        F = euler_flux(d, u, v, w, c, A, 1.4)
        assert np.allclose(fp + fm, F, rtol=1e-12, atol=1e-12 * np.abs(F).max())
    # supersonic along A: f- vanishes, f+ is the whole flux
    A = np.array([0.6, 0.8, 0.0])
    fp, fm = O.split_point(1.3, 3.0, 2.0, 0.1, 1.0, *A, gamma=1.4, eps=0.0)
    assert np.all(fm == 0.0)
    assert np.allclose(fp, euler_flux(1.3, 3.0, 2.0, 0.1, 1.0, A, 1.4), rtol=1e-13)


def test_rk3_stability_polynomial():
    lam = np.array([-1.3, 0.7, -0.05, 2.0])
    u0 = np.array([1.0, -2.0, 0.5, 3.0])
    dt = 0.37
    f1 = O.rk_stage(1, dt, np.zeros(4), u0, lam * u0)
    f2 = O.rk_stage(2, dt, f1, u0, lam * f1)
    f3 = O.rk_stage(3, dt, f2, u0, lam * f2)
    z = lam * dt
    assert np.allclose(f3, u0 * (1 + z + z * z / 2 + z ** 3 / 6), rtol=1e-14, atol=1e-15)


def test_periodic_rhs_telescopes():
    c = cases.tgv(16, nsteps=1)
    o = c.load(O.Oracle(c.cfg))
    o.rhs(1)
    du = o.get_du()
    for v in range(5):
        assert abs(du[v].sum()) <= 1e-11 * np.abs(du[v]).sum() + 1e-13, v


def test_uniform_flow_is_preserved_exactly():
    c = cases.tgv(12, nsteps=1)
    n = 12
    prims = [np.full((n, n, n), 1.2), np.full((n, n, n), 0.3), np.full((n, n, n), -0.2),
             np.full((n, n, n), 0.1), np.full((n, n, n), 1.1)]
    c.prims = prims
    for scheme in (capi.WENO7, capi.UP7, capi.WENO7_SYMBO, capi.OMP6_HR):
        c.cfg.scheme_invis = scheme
        o = c.load(O.Oracle(c.cfg))
        o.rhs(1)
        assert np.all(o.get_du() == 0.0), scheme
    # hybrid: the sensor at a ghost-layer node reads edge ghosts that no
    # exchange writes (p = 0 there), which flags WENO5 on the first/last face of
    # boundary rows while the neighbouring faces use UP7; both reproduce a
    # constant only to round-off.
    c.cfg.scheme_invis = capi.HYBRIDAUTO
    o = c.load(O.Oracle(c.cfg))
    o.rhs(1)
    du = o.get_du()
    assert np.abs(du).max() < 1e-10  # flux magnitude ~1e2, h = 1/11
    fx, fy, fz = o.get_flags()
    assert (fx[1:-1, 1:-1, 1:-1] == 1).all() and (fx == 3).any()


def test_metrics_of_affine_mesh():
    nx, ny, nz = 10, 9, 8
    cfg = capi.make_config(nx=nx, ny=ny, nz=nz, iperiodic=(0, 0, 0), jac_bound=(1, 1, 1), re=100.0, ama=0.5,
                           dt=1e-3)
    rng = np.random.default_rng(11)
    B = np.eye(3) + 0.3 * rng.standard_normal((3, 3))
    hx, hy, hz = 1 / (nx - 1), 1 / (ny - 1), 1 / (nz - 1)
    k, j, i = np.meshgrid(np.arange(nz) * hz, np.arange(ny) * hy, np.arange(nx) * hx, indexing="ij")
    X = B[0, 0] * i + B[0, 1] * j + B[0, 2] * k + 0.2
    Y = B[1, 0] * i + B[1, 1] * j + B[1, 2] * k - 0.1
    Z = B[2, 0] * i + B[2, 1] * j + B[2, 2] * k
    o = O.Oracle(cfg)
    o.load_mesh(X, Y, Z)
    m = o.get_metrics()
    cof = np.linalg.inv(B) * np.linalg.det(B)  # rows: grad xi, grad eta, grad zeta times det
    names = [(3, 0, 0), (4, 0, 1), (5, 0, 2), (6, 1, 0), (7, 1, 1), (8, 1, 2), (9, 2, 0), (10, 2, 1),
             (11, 2, 2)]
    for idx, r, cc in names:
        assert np.allclose(m[idx], cof[r, cc], rtol=1e-10, atol=1e-10), idx
    assert np.allclose(m[12], 1.0 / np.linalg.det(B), rtol=1e-10)


@pytest.mark.parametrize("svis", [capi.CD6, capi.CD8])
def test_q_criterion_of_linear_velocity(svis):
    """Q of u = G x on an affine mesh is the sum of the principal 2x2 minors
    of G (the second invariant of the velocity gradient), exactly up to
    rounding: the derivatives and their one-sided closures are exact on linear
    data and the cofactors times Ajac invert the mesh map."""
    nx, ny, nz = 11, 10, 9
    cfg = capi.make_config(nx=nx, ny=ny, nz=nz, iperiodic=(0, 0, 0), jac_bound=(1, 1, 1),
                           dif_bound=(1, 1, 1, 1, 1, 1), re=100.0, ama=0.5, dt=1e-3, scheme_vis=svis)
    rng = np.random.default_rng(5)
    B = np.eye(3) + 0.3 * rng.standard_normal((3, 3))
    G = rng.standard_normal((3, 3))
    hx, hy, hz = 1 / (nx - 1), 1 / (ny - 1), 1 / (nz - 1)
    k, j, i = np.meshgrid(np.arange(nz) * hz, np.arange(ny) * hy, np.arange(nx) * hx, indexing="ij")
    X = [B[r, 0] * i + B[r, 1] * j + B[r, 2] * k for r in range(3)]
    vel = [G[r, 0] * X[0] + G[r, 1] * X[1] + G[r, 2] * X[2] for r in range(3)]
    o = O.Oracle(cfg)
    o.load_mesh(*X)
    one = np.ones((nz, ny, nx))
    o.load_flow(one, *vel, one)
    q = o.get_q()
    want = (G[0, 0] * G[1, 1] - G[0, 1] * G[1, 0] + G[0, 0] * G[2, 2] - G[0, 2] * G[2, 0]
            + G[1, 1] * G[2, 2] - G[1, 2] * G[2, 1])
    assert np.abs(q - want).max() <= 1e-10 * max(1.0, abs(want))
    # a rotation about z has Q = w^2 > 0, a pure strain Q < 0
    vel = [-X[1], X[0], 0 * X[2]]
    o.load_flow(one, *vel, one)
    assert np.allclose(o.get_q(), 1.0, rtol=0, atol=1e-10)
    vel = [X[0], -X[1], 0 * X[2]]
    o.load_flow(one, *vel, one)
    assert np.allclose(o.get_q(), -1.0, rtol=0, atol=1e-10)


def _euler_flux(U, nvec, gamma):
    rho, mu, mv, mw, E = U
    u, v, w = mu / rho, mv / rho, mw / rho
    p = (gamma - 1) * (E - 0.5 * rho * (u * u + v * v + w * w))
    un = u * nvec[0] + v * nvec[1] + w * nvec[2]
    return np.array([rho * un, mu * un + p * nvec[0], mv * un + p * nvec[1], mw * un + p * nvec[2], (E + p) * un])


@pytest.mark.parametrize("seed", range(6))
def test_char_matrices_are_inverse_eigenvectors(seed):
    """The characteristic projection of a face: L R = I, and L J R =
    diag(un, un, un, un - c, un + c) for J the Jacobian of the Euler flux
    along n = A/|A| at the face state (u, v, w, c) - every row and column of
    both matrices is pinned by these."""
    rng = np.random.default_rng(seed)
    gamma = 1.4
    s0 = np.concatenate([rng.normal(0, 0.6, 3), [rng.uniform(0.5, 2.0)], rng.normal(0, 1, 3)])
    s1 = np.concatenate([rng.normal(0, 0.6, 3), [rng.uniform(0.5, 2.0)], rng.normal(0, 1, 3)])
    if seed % 2:  # exercise both choices of l
        s0[6], s1[6] = 3.0, 2.5
    L, R = O.char_matrices(s0, s1, gamma)
    assert np.allclose(L @ R, np.eye(5), rtol=0, atol=1e-13)
    assert np.allclose(R @ L, np.eye(5), rtol=0, atol=1e-13)
    st = 0.5 * (s0 + s1)
    u, v, w, c = st[:4]
    nvec = st[4:] / np.linalg.norm(st[4:])
    rho = 1.3
    p = rho * c * c / gamma
    U = np.array([rho, rho * u, rho * v, rho * w, p / (gamma - 1) + 0.5 * rho * (u * u + v * v + w * w)])
    J = np.empty((5, 5))
    for k in range(5):
        e = np.zeros(5)
        e[k] = 1e-6 * max(1.0, abs(U[k]))
        J[:, k] = (_euler_flux(U + e, nvec, gamma) - _euler_flux(U - e, nvec, gamma)) / (2 * e[k])
    un = u * nvec[0] + v * nvec[1] + w * nvec[2]
    D = L @ J @ R
    assert np.allclose(D, np.diag([un, un, un, un - c, un + c]), rtol=0, atol=1e-6 * (abs(un) + c))


@pytest.mark.parametrize("scheme", [capi.WENO7_SYMBO, capi.WENO7_SYMBO_LIM, capi.HYBRIDAUTO])
@pytest.mark.parametrize("side", [0, 1])
@pytest.mark.parametrize("dirn", [0, 1, 2])
@pytest.mark.parametrize("bounds", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("nonref", [0, 1])
def test_char_line_is_physical_line_on_linear_fluxes(scheme, side, dirn, bounds, nonref):
    """On split fluxes linear along the line every reconstruction and closure
    is a fixed linear combination of the stencil, so projecting with L and
    mapping back with R changes nothing: the CHARTERIC = 1 derivative equals
    the physical-space one at every node except the boundary nodes where the
    reference takes a physical-space difference instead, which on linear data
    is the exact slope (or 0 under the non-reflecting rule of that path)."""
    n = 16
    bl, br = bounds
    rng = np.random.default_rng(7 + 3 * side + dirn)
    q = np.arange(-4, n + 4, dtype=float)
    a = rng.normal(0, 1, 5)
    b = rng.normal(0, 0.3, 5)
    F = a[:, None] + b[:, None] * q[None, :]
    prim = np.empty((7, n + 8))
    prim[0:3] = rng.normal(0, 0.4, (3, 1)) + 0.05 * np.sin(q)[None, :]
    prim[3] = 1.0 + 0.1 * np.cos(q)
    prim[4:7] = rng.normal(0, 1, (3, 1)) + 0.05 * np.cos(0.7 * q)[None, :]
    ajac = 1.0 + 0.2 * np.sin(0.3 * q)
    h = 0.1
    flags = rng.integers(1, 4, n + 1) if scheme == capi.HYBRIDAUTO else None
    got = O.line_flux_char(scheme, side, dirn, F, prim, ajac, h, bl, br, nonref, flags)
    want = np.stack([O.line_flux_derivative(scheme, side, F[m], ajac, h, bl, br, 0, flags) for m in range(5)])
    fy = 7 if bl and br else 1 if bl else 4 if br else 0
    exact = -ajac[4:n + 4][None, :] * b[:, None] / h
    if side == 0:
        if fy == 1:
            want[:, 0] = 0.0 if nonref else exact[:, 0]
        if fy == 4:
            want[:, n - 1] = exact[:, n - 1]
    else:
        if fy == 4:
            want[:, n - 1] = 0.0 if nonref else exact[:, n - 1]
        if fy == 1 and dirn != 1:
            want[:, 0] = exact[:, 0]
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


def test_symmetry_plane_metrics_match_a_mirrored_mesh():
    """IBC 124 symmetry planes: the metric ghosts the mirror rule writes
    (y odd; the y row of the cofactors odd) and the metrics of the points
    next to the planes (which read the mirrored coordinate ghosts) equal the
    metrics of a mesh extended by the real mirror image of the grid."""
    nx, ny, nz = 10, 12, 9
    c = cases.liftbody(nx, ny, nz)
    x, y, z = c.mesh
    cfg = c.cfg
    o = O.Oracle(cfg)
    o.load_mesh(x, y, z)
    half = [o.get_metric_lap(k) for k in range(13)]
    # the grid plus its mirror images, 8 rows beyond each symmetry plane (so
    # that the rows compared are 4 or more from the extended grid's edges)
    ext = [np.concatenate([a[:, 8:0:-1, :] * s_, a, a[:, -2:-10:-1, :] * s_], axis=1)
           for a, s_ in ((x, 1.0), (y, -1.0), (z, 1.0))]
    cfg2 = capi.make_config(nx=nx, ny=ny + 16, nz=nz, iperiodic=(0, 0, 0), jac_bound=(1, 1, 1), re=100.0, ama=0.5,
                            dt=1e-3)
    o2 = O.Oracle(cfg2)
    o2.load_mesh(*ext)
    full = [o2.get_metric_lap(k) for k in range(13)]
    # the extended grid has more points in y: h_eta = 1/(ny+15) instead of
    # 1/(ny-1), so its cofactors with an eta derivative (Ak*, As*) and its
    # 1/Jacobian scale by the ratio
    r = (ny + 15) / (ny - 1)
    scale = {3: 1 / r, 4: 1 / r, 5: 1 / r, 6: 1.0, 7: 1.0, 8: 1.0, 9: 1 / r, 10: 1 / r, 11: 1 / r, 12: r}
    # interior x and z (what the rule covers), y rows -4..ny+3 of the half grid
    # = rows 4..ny+11 of the extended one (LAP-box rows 0..ny+7 and 8..ny+15)
    xs, zs = slice(4, nx + 4), slice(4, nz + 4)
    for k in range(3, 13):
        a = half[k][zs, 0:ny + 8, xs]
        b = full[k][zs, 8:ny + 16, xs] * scale[k]
        assert np.allclose(a, b, rtol=1e-9, atol=1e-12 * np.abs(b).max()), k


def test_sutherland_at_reference_temperature():
    c = cases.tgv(8, nsteps=1)
    n = 8
    c.prims = [np.ones((n, n, n)), np.zeros((n, n, n)), np.zeros((n, n, n)), np.zeros((n, n, n)),
               np.ones((n, n, n))]
    o = c.load(O.Oracle(c.cfg))
    assert np.allclose(o.get_amu(), 1.0 / c.cfg.re, rtol=1e-15)


def test_jameson_sensor():
    assert O.jameson_dp(2.0, 1.0, 3.0, 1.5, 2.5, 1.9, 2.1) == 0.0
    # a jump of 1 -> 2 in x only: |-1+4-2|/(1+4+2) = 1/7
    assert abs(O.jameson_dp(2.0, 1.0, 2.0, 2.0, 2.0, 2.0, 2.0) - 1.0 / 7.0) < 1e-16


def test_filter_with_zero_sigma_leaves_interior_unchanged():
    r = cases.ramp(40, 24, 8, nsteps=1, filter_every=1)
    a = r.load(O.Oracle(r.cfg))
    r.cfg.filter_rpara[0][1] = 1e30  # rth huge -> sigma = 0
    b = r.load(O.Oracle(r.cfg))
    r2 = cases.ramp(40, 24, 8, nsteps=1, filt=False)
    c = r2.load(O.Oracle(r2.cfg))
    b.step(1)
    c.step(1)
    fb, fc = b.get_state(), c.get_state()
    # faces in the periodic z direction are refreshed from primitives after
    # f_lap is taken (src/OCFD_NS_Solver.cu:83-95), and the filter writes the
    # pre-refresh value back there, so after one step only the points away
    # from those faces are bitwise equal
    assert np.array_equal(fb[:, 1:-1], fc[:, 1:-1])
    assert np.allclose(fb, fc, rtol=1e-13, atol=1e-13)
    a.step(1)
    assert not np.array_equal(a.get_state(), fc)  # the real filter does act


# ---------------------------------------------------------------------------
# Boundary conditions.  The wall values are one-sided extrapolations in the
# wall-normal index with zero slope at the wall: IBC 108 uses
# p_w = (18 p_1 - 9 p_2 + 2 p_3)/11 (exact for a + c j^2 + e j^3; a + b j
# gives a + 6b/11), IBC 124 uses (4 p_1 - p_2)/3 (exact for a + c k^2; a + b k
# gives a + 2b/3).  Inlet, top and upper faces take the tables given; ghost
# layers mirror the interior with the velocity odd at a no-slip wall and the
# normal velocity odd at a symmetry plane.
# ---------------------------------------------------------------------------
def _wall_rows(shape_xz, rng):
    a = 1.0 + 0.2 * rng.random(shape_xz)
    return a, 0.05 * rng.random(shape_xz), 0.01 * rng.random(shape_xz)


@pytest.mark.parametrize("form", ["cubic", "linear"])
def test_ibc108_isothermal_wall_pressure_extrapolation(form):
    c = cases.ramp(24, 16, 6, nsteps=1, filt=False)
    c.cfg.bc_rpara[1] = 0.0  # no wall blowing
    tw = c.cfg.bc_rpara[0]
    d, u, v, w, T = [p.copy() for p in c.prims]
    a, cq, e = _wall_rows((6, 24), np.random.default_rng(3))
    for j in (1, 2, 3):
        T[:, j, :] = 2.0 + 0.1 * j
        p = a + cq * j ** 2 + e * j ** 3 if form == "cubic" else a + cq * j
        d[:, j, :] = p / T[:, j, :]
    c.prims = [d, u, v, w, T]
    o = c.load(O.Oracle(c.cfg))
    (d2, u2, v2, w2, T2), _, _ = o.get_flow()
    # i = 0 takes the inlet profile before the wall is set: the wall there
    # extrapolates inlet values
    pw = d2[:, 0, 1:] * T2[:, 0, 1:]
    want = a[:, 1:] if form == "cubic" else a[:, 1:] + 6.0 * cq[:, 1:] / 11.0
    assert np.array_equal(T2[:, 0, :], np.full((6, 24), tw))
    assert np.abs(pw - want).max() <= 1e-14 * np.abs(want).max()
    assert not u2[:, 0, :].any() and not v2[:, 0, :].any() and not w2[:, 0, :].any()


def test_ibc108_adiabatic_wall_extrapolation():
    c = cases.ramp(24, 16, 6, nsteps=1, filt=False)
    c.cfg.bc_rpara[1] = 0.0
    c.cfg.bc_rpara[0] = -1.0  # TW <= 0: adiabatic
    d, u, v, w, T = [p.copy() for p in c.prims]
    rng = np.random.default_rng(4)
    a, cq, e = _wall_rows((6, 24), rng)
    t0, tq, te = 2.0 + rng.random((6, 24)), 0.02 * rng.random((6, 24)), 0.01 * rng.random((6, 24))
    for j in (1, 2, 3):
        T[:, j, :] = t0 + tq * j ** 2 + te * j ** 3
        d[:, j, :] = (a + cq * j ** 2 + e * j ** 3) / T[:, j, :]
    c.prims = [d, u, v, w, T]
    o = c.load(O.Oracle(c.cfg))
    (d2, u2, v2, w2, T2), _, _ = o.get_flow()
    assert np.abs(T2[:, 0, 1:] - t0[:, 1:]).max() <= 1e-14 * t0.max()
    assert np.abs(d2[:, 0, 1:] * T2[:, 0, 1:] - a[:, 1:]).max() <= 1e-14 * a.max()


def test_ibc108_inlet_top_and_wall_ghosts():
    c = cases.ramp(24, 16, 6, nsteps=1, filt=False)
    o = c.load(O.Oracle(c.cfg))
    (d2, u2, v2, w2, T2), _, _ = o.get_flow()
    prof = c.inlet  # rho, u, v, T over j
    # inlet column (between the wall and the top row)
    for got, want in ((d2, prof[0]), (u2, prof[1]), (v2, prof[2]), (T2, prof[3])):
        assert np.array_equal(got[:, 1:-1, 0], np.broadcast_to(want[1:-1], (6, 14)))
    assert not w2[:, 1:-1, 0].any()
    # top row from x_begin (0 here: no top x below XUB) takes the profile's top
    for got, want in ((d2, prof[0]), (u2, prof[1]), (v2, prof[2]), (T2, prof[3])):
        assert np.array_equal(got[:, -1, :], np.full((6, 24), want[-1]))
    # ghosts below the wall: rho, T even, velocity odd about j = 0
    L = 4
    for k, sgn in ((0, 1.0), (1, -1.0), (2, -1.0), (3, -1.0), (4, 1.0)):
        a = o.get_prim_lap(k)
        for m in range(1, 5):
            assert np.array_equal(a[L:-L, L - m, L:-L], sgn * a[L:-L, L + m, L:-L])


@pytest.mark.parametrize("form", ["quadratic", "linear"])
def test_ibc124_isothermal_wall_pressure_extrapolation(form):
    c = cases.liftbody(20, 12, 10, nsteps=1, tw=1.5)
    c.cfg.bc_rpara[2] = 0.0  # EPSL_WALL: no wall disturbance
    c.cfg.bc_rpara[3] = 0.0  # EPSL_UPPER
    d, u, v, w, T = [p.copy() for p in c.prims]
    a, cq, _ = _wall_rows((12, 20), np.random.default_rng(5))
    for k in (1, 2):
        T[k] = 1.2 + 0.1 * k
        p = a + cq * k ** 2 if form == "quadratic" else a + cq * k
        d[k] = p / T[k]
    c.prims = [d, u, v, w, T]
    o = c.load(O.Oracle(c.cfg))
    (d2, u2, v2, w2, T2), _, _ = o.get_flow()
    want = a if form == "quadratic" else a + 2.0 * cq / 3.0
    # i = 0 is the inlet, whose values are set before the wall
    assert np.array_equal(T2[0, :, 1:], np.full((12, 19), 1.5))
    assert np.abs(d2[0, :, 1:] * 1.5 - want[:, 1:]).max() <= 1e-14 * want.max()
    assert not u2[0, :, 1:].any() and not v2[0, :, 1:].any() and not w2[0, :, 1:].any()


def test_ibc124_adiabatic_wall_upper_and_symmetry_ghosts():
    c = cases.liftbody(20, 12, 10, nsteps=1, tw=-1.0)
    c.cfg.bc_rpara[2] = 0.0
    c.cfg.bc_rpara[3] = 0.0
    d, u, v, w, T = [p.copy() for p in c.prims]
    rng = np.random.default_rng(6)
    t0, tq = 2.0 + rng.random((12, 20)), 0.05 * rng.random((12, 20))
    d0, dq = 0.5 + rng.random((12, 20)), 0.05 * rng.random((12, 20))
    for k in (1, 2):
        T[k] = t0 + tq * k ** 2
        d[k] = d0 + dq * k ** 2
    c.prims = [d, u, v, w, T]
    o = c.load(O.Oracle(c.cfg))
    (d2, u2, v2, w2, T2), _, _ = o.get_flow()
    assert np.abs(T2[0, :, 1:] - t0[:, 1:]).max() <= 1e-14 * t0.max()
    assert np.abs(d2[0, :, 1:] - d0[:, 1:]).max() <= 1e-14 * d0.max()
    # upper face without a table or disturbance: the free stream at AOA
    aoa = math.radians(c.cfg.bc_rpara[0])
    for got, want in ((d2, 1.0), (u2, math.cos(aoa)), (v2, 0.0), (w2, math.sin(aoa)), (T2, 1.0)):
        assert np.array_equal(got[-1, :, 1:], np.full((12, 19), want))
    # symmetry planes j = 0 and j = ny-1 (the plane y = 0): v odd, the rest even
    L = 4
    for k, sgn in ((0, 1.0), (1, 1.0), (2, -1.0), (3, 1.0), (4, 1.0)):
        a = o.get_prim_lap(k)
        for m in range(1, 5):
            assert np.array_equal(a[L:-L, L - m, L:-L], sgn * a[L:-L, L + m, L:-L])
            assert np.array_equal(a[L:-L, L + 11 + m, L:-L], sgn * a[L:-L, L + 11 - m, L:-L])


def test_ibc124_inlet_and_upper_tables():
    c = cases.liftbody(20, 12, 10, nsteps=1, withleading=1, upper=1)
    c.cfg.bc_rpara[3] = 0.0  # no upper disturbance: the upper face is the table
    o = c.load(O.Oracle(c.cfg))
    prims, _, _ = o.get_flow()
    inl, up = c.extra["inlet5"], c.extra["upper5"]
    for m in range(5):
        # inlet i = 0 between the wall (k = 0) and the upper face
        assert np.array_equal(prims[m][1:-1, :, 0], inl[m][1:-1])
        assert np.array_equal(prims[m][-1, :, 1:], up[m][:, 1:])


# ---------------------------------------------------------------------------
# Filter: with RTH = 0 the shock sensor is 1 everywhere and the filter is the
# fixed stencil f_i - s0 [c1 (f_{i+1} - 2 f_i + f_{i-1})
# + c2 (f_{i+2} - f_{i+1} - f_{i-1} + f_{i-2})].  On f = i^2 the bracket is
# 2 c1 + 6 c2; on f = (i - i0)^4 at i0 it is 2 c1 + 30 c2: the two pin c1 and
# c2 (c1 = -0.2103830, c2 = 0.0396170); a linear f is left unchanged.
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("form", ["quadratic", "quartic", "linear"])
def test_filter_stencil_with_unit_sensor(form):
    nx = 40
    c = cases.ramp(nx, 24, 8, nsteps=1, filter_every=1)
    c.cfg.filter_rpara[0][1] = 0.0  # RTH = 0: sigma = 1
    c.cfg.filter_npara[0][2] = 0    # x only
    c.cfg.filter_npara[0][3] = 0
    d, u, v, w, T = [p.copy() for p in c.prims]
    i = np.arange(nx, dtype=float)
    i0 = 20
    beta = {"quadratic": 1e-4, "quartic": 1e-6, "linear": 1e-3}[form]
    prof = {"quadratic": (i - i0) ** 2, "quartic": (i - i0) ** 4, "linear": i - i0}[form]
    d[:] = 1.0 + beta * prof
    u[:] = 0.7
    v[:] = 0.1
    w[:] = 0.0
    T[:] = 2.0
    c.prims = [d, u, v, w, T]
    o = c.load(O.Oracle(c.cfg))
    f0 = o.get_state()
    o.filter_now()
    f1 = o.get_state()
    s0 = c.cfg.filter_rpara[0][0]
    c1, c2 = -0.2103830, 0.0396170
    bracket = {"quadratic": 2 * c1 + 6 * c2, "quartic": 2 * c1 + 30 * c2, "linear": 0.0}[form]
    # away from the wall, inlet and top rows; each conserved variable is a
    # constant multiple of rho here
    for m in range(5):
        scale = f0[m][3, 8, i0] / d[3, 8, i0]
        got = f1[m][2:-2, 6:-6, i0] - f0[m][2:-2, 6:-6, i0]
        want = -s0 * bracket * beta * scale
        assert np.abs(got - want).max() <= 1e-12 * max(abs(scale), 1e-300), (m, got.max(), want)
    if form == "linear":
        assert np.abs(f1[0][2:-2, 6:-6, 8:-8] - f0[0][2:-2, 6:-6, 8:-8]).max() <= 1e-15
