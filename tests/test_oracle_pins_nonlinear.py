"""Pins of the oracle's viscous terms, filter sensor and nonlinear
reconstructions against what the mathematics fixes.

Viscous operator.  The right-hand side is the inviscid part plus a viscous
part that is linear in the viscosity coefficient mu0 = (1 + Tsb)/Re
(src/OCFD_init.cu:381-410; src/OCFD_NS_Jacobian3d.cu:80-281): the difference
of two right-hand sides of the same state at two Reynolds numbers is the
viscous part alone, times mu0(Re1) - mu0(Re2).  On an affine (sheared,
non-orthogonal) mesh the metric cofactors are exact constants, and for a
velocity field that is a quadratic polynomial with constant T the stress is
linear and the energy flux u.tau cubic, so every CD6/CD8 derivative on the
path is exact and the discrete viscous terms equal the continuous ones:

  momentum  mu (lap u_i + 1/3 d_i div u)
  energy    mu Phi + u . mu (lap u + 1/3 grad div u),
  Phi = 2 (ux^2 + vy^2 + wz^2) + (uy + vx)^2 + (uz + wx)^2 + (vz + wy)^2
        - 2/3 (div u)^2

with mu = mu0 T^1.5 / (Tsb + T) (Sutherland).  A linear temperature at rest
gives the heat flux alone: energy = (Cp/Pr) c^2 |n|^2 mu'(T) for
T = T0 + c n.x.  The stress tensor (its -2/3 dilatation term included), the
heat-flux factor Cp/Pr, the Ev assembly with the cofactors and the
Jacobian factor of the divergence all enter these closed forms.

Filter sensor (src/OCFD_filtering.cu:384-396).  For a cubic pressure
p = P0 + A (i - i0)^3 along x the sensor's differences are known exactly:
dp0 - dp1 = 6A/4, dp0 - dp2 = -6A/4, so r = 2.25 A^2 / p^2 + 1e-16 at every
node.  With rth set between r(i*) and r(i* + 1) the sensor
sigma = max(0, 1 - rth/r) is positive at i <= i* and zero beyond, so the
filter changes the state exactly at the nodes i <= i* + 1 and leaves the
rest bitwise unchanged; the change at the active nodes follows from the
closed-form sigma and the filter stencil (its constants are pinned
separately).

Nonlinear reconstructions at a step.  Every candidate stencil of WENO7,
WENO5 and WENO7-SYMBO contains the upwind node of the face, so for step data
the stencils that do not cross the jump all reproduce the upwind node's
value; the nonlinear weights must send the crossing stencils' weight to
zero (ENO behaviour).  The OMP6 schemes' MP limiter must not create a new
extremum (the value stays within the data range).  The linear UP7 on the
same data misses by up to 0.43, which shows the test is sharp.

WENO7, WENO5 and WENO7-SYMBO face values from first principles (at the end
of this file): candidate values, smoothness indicators as integrals of the
stencil polynomials' derivatives, optimal weights, formed in exact rational
arithmetic, and the SYMBO TV limiter's switch between linear and nonlinear
weights.
"""
import math

import numpy as np
import pytest

import oracle as O
from pyocfd import capi

P, M = 0, 1


def affine_mesh(nx, ny, nz, B, x0):
    """x = x0 + B (xi, eta, zeta) with xi = i/(nx-1) etc."""
    xi = np.arange(nx) / (nx - 1)
    eta = np.arange(ny) / (ny - 1)
    zeta = np.arange(nz) / (nz - 1)
    Z, Y, X = np.meshgrid(zeta, eta, xi, indexing="ij")
    return [x0[r] + B[r, 0] * X + B[r, 1] * Y + B[r, 2] * Z for r in range(3)]


def visc_config(n, re, svis, ma=0.8):
    return capi.make_config(
        nx=n, ny=n, nz=n, iperiodic=(0, 0, 0), jac_bound=(1, 1, 1), dif_bound=(1,) * 6, non_ref=(0,) * 6,
        scheme_invis="WENO7", scheme_vis=svis, re=re, ama=ma, gamma=1.4, pr=0.72, t_ref=288.15, epsl_sw=0.0,
        dt=1e-4, end_time=1e-4, ibc=0)


def viscous_part(mesh, prims, n, svis, re1=2.0, re2=5.0):
    """du(Re1) - du(Re2) divided by mu0(Re1) - mu0(Re2): the viscous terms
    per unit mu0."""
    out = []
    for re in (re1, re2):
        o = O.Oracle(visc_config(n, re, svis))
        o.load_mesh(*mesh)
        o.load_flow(*prims, istep=0, tt=0.0)
        o.init()
        o.rhs(1)
        out.append(o.get_du())
        o.close()
    tsb = 110.4 / 288.15
    dmu = (1 + tsb) / re1 - (1 + tsb) / re2
    return (out[0] - out[1]) / dmu


def sutherland(T, tsb=110.4 / 288.15):
    return T ** 1.5 / (tsb + T)


def sutherland_prime(T, tsb=110.4 / 288.15):
    return (1.5 * T ** 0.5 * (tsb + T) - T ** 1.5) / (tsb + T) ** 2


B = np.array([[1.0, 0.35, -0.2], [0.15, 0.9, 0.25], [-0.1, 0.2, 1.1]])
X0 = np.array([0.3, -0.2, 0.1])


@pytest.mark.parametrize("svis", ["CD6", "CD8"])
def test_viscous_terms_of_a_quadratic_velocity_on_an_affine_mesh(svis):
    n = 24
    x, y, z = affine_mesh(n, n, n, B, X0)
    rng = np.random.default_rng(11)
    # u_i = a_i + G_ij x_j + 1/2 H_ijk x_j x_k (H symmetric in j, k), small
    # enough that the flow stays subsonic and positive
    a = rng.uniform(-0.1, 0.1, 3)
    G = rng.uniform(-0.3, 0.3, (3, 3))
    H = rng.uniform(-0.3, 0.3, (3, 3, 3))
    H = 0.5 * (H + H.transpose(0, 2, 1))
    X = np.stack([x, y, z])
    vel = [a[i] + np.einsum("j,j...->...", G[i], X) + 0.5 * np.einsum("jk,j...,k...->...", H[i], X, X)
           for i in range(3)]
    T0 = 1.3
    d = 1.0 + 0.0 * x
    T = T0 + 0.0 * x
    W = viscous_part((x, y, z), [d, vel[0], vel[1], vel[2], T], n, svis)
    mu = sutherland(T0)
    # gradient grad[i][j] = d u_i / d x_j (field), laplacian and grad div (constants)
    grad = [[G[i, j] + np.einsum("k,k...->...", H[i, j], X) for j in range(3)] for i in range(3)]
    lap = np.array([np.trace(H[i]) for i in range(3)])
    graddiv = np.array([sum(H[k, k, i] for k in range(3)) for i in range(3)])
    div = grad[0][0] + grad[1][1] + grad[2][2]
    Phi = 2 * (grad[0][0] ** 2 + grad[1][1] ** 2 + grad[2][2] ** 2) + (grad[0][1] + grad[1][0]) ** 2 + \
        (grad[0][2] + grad[2][0]) ** 2 + (grad[1][2] + grad[2][1]) ** 2 - 2.0 / 3.0 * div ** 2
    mom = mu * (lap + graddiv / 3.0)
    energy = mu * Phi + sum(vel[i] * mom[i] for i in range(3))
    scale = max(np.abs(mom).max(), np.abs(energy).max())
    # nodes whose two derivative stencils (4 + 4 wide) stay clear of the D0
    # closures, which are exact only to lower degree (first order at a face)
    s = (slice(8, -8),) * 3
    for i in range(3):
        err = np.abs(W[1 + i][s] - mom[i]).max() / scale
        assert err <= 1e-9, (svis, i, err)
    err = np.abs(W[4][s] - energy[s]).max() / scale
    assert err <= 1e-9, (svis, err)
    assert np.abs(W[0]).max() <= 1e-9 * scale  # no viscous term in the mass equation


@pytest.mark.parametrize("svis", ["CD6", "CD8"])
def test_heat_flux_of_a_linear_temperature_at_rest(svis):
    n = 28
    x, y, z = affine_mesh(n, n, n, B, X0)
    nvec = np.array([0.4, -0.7, 0.3])
    c = 0.25
    T = 1.2 + c * (nvec[0] * x + nvec[1] * y + nvec[2] * z)
    zero = 0.0 * x
    W = viscous_part((x, y, z), [1.0 + zero, zero, zero, zero, T], n, svis)
    gamma, ma, pr = 1.4, 0.8, 0.72
    vis_c = 1.0 / ((gamma - 1.0) * ma * ma) / pr  # Cp / Pr
    want = vis_c * c * c * (nvec @ nvec) * sutherland_prime(T)
    # interior nodes: the D0 closures of the faces are lower order on this
    # non-polynomial viscosity
    s = (slice(8, -8),) * 3
    err = np.abs(W[4][s] - want[s]).max() / np.abs(want).max()
    print(f"[pin] heat flux {svis}: {err:.2e}")
    tol = 1e-8 if svis == "CD8" else 1e-6
    assert err <= tol, (svis, err)
    for i in range(1, 4):
        assert np.abs(W[i]).max() <= 1e-10 * np.abs(want).max()


def filter_case(nx, rth, T0, beta, i0):
    """Cartesian box (identity metrics), periodic in y and z, rho = 1, a
    cubic temperature along x, the x filter every step with RTH = rth."""
    ny = nz = 8
    gamma, ma = 1.4, 0.5
    cfg = capi.make_config(
        nx=nx, ny=ny, nz=nz, iperiodic=(1, 1, 1), jac_bound=(1, 1, 1), dif_bound=(0,) * 6, non_ref=(0,) * 6,
        scheme_invis="WENO7", scheme_vis="CD6", re=100.0, ama=ma, gamma=gamma, pr=0.72, t_ref=288.15,
        epsl_sw=0.0, dt=1e-4, end_time=1e-4, ibc=0,
        filters=[((1, 1, 0, 0, 0, nx, 0, ny, 0, nz, 2), (1.0, rth, 1e9))])
    i = np.arange(nx, dtype=float)
    jac = np.zeros((13, nz, ny, nx))
    jac[0] = i[None, None, :] / (nx - 1)
    jac[1] = np.arange(ny)[None, :, None] / (ny - 1)
    jac[2] = np.arange(nz)[:, None, None] / (nz - 1)
    jac[3] = jac[7] = jac[11] = jac[12] = 1.0
    T = np.broadcast_to(T0 * (1.0 + beta * (i - i0) ** 3), (nz, ny, nx)).copy()
    ones = np.ones((nz, ny, nx))
    prims = [ones, 0.3 * ones, 0.1 * ones, -0.05 * ones, T]
    o = O.Oracle(cfg)
    o.load_jacobi(jac)
    o.load_flow(*prims, istep=0, tt=0.0)
    o.init()
    return o, cfg


def test_filter_shock_sensor_closed_form():
    nx, T0, beta, i0 = 40, 1.0, 2e-3, -4.0  # (i - i0) >= 4 on the interior: T rises from 1.13 to 119
    istar = 14
    p00 = 1.0 / (1.4 * 0.5 * 0.5)
    A = p00 * T0 * beta
    p = p00 * T0 * (1.0 + beta * (np.arange(nx) - i0) ** 3)
    r = 2.25 * A * A / p ** 2 + 1e-16
    assert r[istar] / r[istar + 1] > 1.3  # the threshold separates the two nodes clearly
    rth = math.sqrt(r[istar] * r[istar + 1])
    o, cfg = filter_case(nx, rth, T0, beta, i0)
    f0 = o.get_state()
    o.filter_now()
    f1 = o.get_state()
    assert np.allclose(o.get_pressure()[2, 2], p, rtol=1e-13)  # the p the sensor read
    changed = np.abs(f1[4] - f0[4]).max(axis=(0, 1)) > 0
    # away from the periodic seam (the cubic wraps around there)
    sl = slice(6, nx - 6)
    want_changed = np.arange(nx) <= istar + 1
    assert np.array_equal(changed[sl], want_changed[sl]), (np.nonzero(changed)[0], istar)
    # mass and momentum are uniform: the filter leaves them alone
    for m in range(4):
        assert np.array_equal(f1[m][:, :, sl], f0[m][:, :, sl])
    # the energy change at the active nodes from the closed-form sigma
    sig = np.maximum(0.0, 1.0 - rth / r)
    c1, c2, s0 = -0.2103830, 0.0396170, 1.0
    E = f0[4][2, 2]
    for i in range(6, istar + 3):
        Sc1 = 0.5 * (sig[i] + sig[i + 1]) * s0
        Sc2 = 0.5 * (sig[i] + sig[i - 1]) * s0
        Dp = c1 * (E[i + 1] - E[i]) + c2 * (E[i + 2] - E[i - 1])
        Dm = c1 * (E[i] - E[i - 1]) + c2 * (E[i + 1] - E[i - 2])
        want = E[i] - (Sc1 * Dp - Sc2 * Dm)
        assert abs(f1[4][2, 2, i] - want) <= 1e-13 * abs(E[i]), i
    # rth just above every r: nothing changes away from the seam
    o2, _ = filter_case(nx, r[6:nx - 6].max() * 1.01, T0, beta, i0)
    g0 = o2.get_state()
    o2.filter_now()
    assert np.array_equal(o2.get_state()[:, :, :, sl.start + 1:sl.stop - 1], g0[:, :, :, sl.start + 1:sl.stop - 1])


UPWIND = {  # the upwind node of the face in the 8-value stencil, P and M side
    "WENO7": (3, 3), "WENO5": (2, 2), "WENO7_SYMBO": (3, 4), "WENO7_SYMBO_LIM": (3, 4),
    "UP7": (3, 4), "OMP6_HR": (3, 4), "OMP6_LD": (3, 4), "OMP6_CD8": (3, 4)}
SPAN = {"WENO7": 7, "WENO5": 5, "WENO7_SYMBO": 8, "WENO7_SYMBO_LIM": 8, "UP7": 7,
        "OMP6_HR": 8, "OMP6_LD": 8, "OMP6_CD8": 8}


def step_values(scheme, side):
    """(face value, upwind node value, lo, hi) for every jump position inside
    the stencil and both jump directions."""
    out = []
    for m in range(SPAN[scheme] - 1):
        for lo, hi in ((1.0, 2.0), (2.0, 1.0), (0.5, 3.0)):
            s = np.full(8, lo)
            s[m + 1:] = hi
            h = O.face_flux(capi.SCHEME_IDS[scheme], side, s)
            out.append((h, s[UPWIND[scheme][side]], min(lo, hi), max(lo, hi)))
    return out


@pytest.mark.parametrize("scheme,tol", [("WENO7", 1e-12), ("WENO5", 1e-9), ("WENO7_SYMBO", 1e-6),
                                        ("WENO7_SYMBO_LIM", 1e-6)])
@pytest.mark.parametrize("side", [P, M])
def test_weno_is_essentially_non_oscillatory_at_a_step(scheme, tol, side):
    for h, up, lo, hi in step_values(scheme, side):
        assert abs(h - up) <= tol * (hi - lo), (scheme, side, h, up)


@pytest.mark.parametrize("side", [P, M])
def test_the_linear_scheme_oscillates_at_the_same_steps(side):
    worst = max(abs(h - up) / (hi - lo) for h, up, lo, hi in step_values("UP7", side))
    assert worst > 0.3


@pytest.mark.parametrize("scheme", ["OMP6_HR", "OMP6_LD", "OMP6_CD8"])
@pytest.mark.parametrize("side", [P, M])
def test_mp_limiter_creates_no_new_extremum(scheme, side):
    for h, up, lo, hi in step_values(scheme, side):
        assert lo - 1e-14 <= h <= hi + 1e-14, (scheme, side, h, lo, hi)


@pytest.mark.parametrize("scheme", ["OMP6_HR", "OMP6_LD", "OMP6_CD8"])
def test_mp_limiter_is_inactive_on_smooth_monotone_data(scheme):
    """On smooth data the limiter must leave the high-order value alone (no
    clipping to first order).  The face value of a flux-difference scheme is
    the function whose cell mean is the data: for exp(x) with spacing D it is
    exp(x_face) (D/2) / sinh(D/2)."""
    dx = 0.05
    x = (np.arange(8) - 3.5) * dx
    want = (dx / 2) / math.sinh(dx / 2)
    for side in (P, M):
        h = O.face_flux(capi.SCHEME_IDS[scheme], side, np.exp(x))
        assert abs(h - want) <= 1e-9, (side, h, want)
        assert abs(h - np.exp(x[UPWIND[scheme][side]])) > 1e-3


# ---------------------------------------------------------------------------
# WENO7 smoothness indicators and weights from first principles
# (src/OCFD_Schemes.cu:616-663).  On the P side the face is x_{i+1/2} and
# s[0..6] are the nodes i-3..i+3; stencil k (k = 0..3) is nodes i-3+k..i+k.
# The face value is sum_k w_k q_k with
#   q_k  the value at the face of the cubic whose means over the cells
#        [j-1/2, j+1/2] of the stencil are the data (finite-difference WENO),
#   w_k  = a_k / sum a, a_k = d_k / (eps + beta_k)^2, d = (1, 12, 18, 4) / 35,
#   beta_k = 34560 * sum_{l=1..3} int_{-1/2}^{1/2} (p_k^(l))^2 dxi for the
#        cubic p_k through the stencil's point values (the Balsara-Shu
#        integral with the point-value polynomial; 34560 * 1e-8 = eps =
#        3.456e-4).
# The reference's stencil 2 takes its second-derivative term at the node
# i+1 (its second difference is s[3] - 2 s[4] + s[5] where p_2''(i) is
# s[2] - 2 s[3] + s[4]); the expected value below does the same, and
# test_weno7_textbook_stencil_2_differs shows the test tells the two apart.
# Everything is exact rational arithmetic on the (exact) binary inputs:
# the polynomials are solved for and integrated here, not transcribed.
from fractions import Fraction as Fr


def _solve(A, b):
    n = len(A)
    M = [list(r) + [b[i]] for i, r in enumerate(A)]
    for c in range(n):
        p = next(r for r in range(c, n) if M[r][c] != 0)
        M[c], M[p] = M[p], M[c]
        for r in range(n):
            if r != c and M[r][c] != 0:
                f = M[r][c] / M[c][c]
                M[r] = [x - f * y for x, y in zip(M[r], M[c])]
    return [M[i][n] / M[i][i] for i in range(n)]


def _cubic(nodes, vals, cell_mean):
    """Coefficients c0..c(n-1) (in xi, node i at xi = 0, unit spacing) of the
    polynomial of degree n-1 (n nodes) with the given point values, or cell
    means, at the nodes."""
    def basis(j, e):
        if not cell_mean:
            return Fr(j) ** e
        a, b = Fr(2 * j - 1, 2), Fr(2 * j + 1, 2)
        return (b ** (e + 1) - a ** (e + 1)) / (e + 1)
    return _solve([[basis(j, e) for e in range(len(nodes))] for j in nodes], vals)


def _pmul(a, b):
    r = [Fr(0)] * (len(a) + len(b) - 1)
    for i, x in enumerate(a):
        for j, y in enumerate(b):
            r[i + j] += x * y
    return r


def _smoothness(c):
    """sum_{l=1..deg} int_{-1/2}^{1/2} (d^l p / dxi^l)^2 dxi."""
    tot, d = Fr(0), list(c)
    for _ in range(len(c) - 1):
        d = [i * d[i] for i in range(1, len(d))]
        sq = _pmul(d, d)
        tot += sum(ci * (Fr(1, 2) ** (i + 1) - Fr(-1, 2) ** (i + 1)) / (i + 1) for i, ci in enumerate(sq))
    return tot


def _weno7_face(s, node_of_second=(0, 0, 1, 0)):
    d = [Fr(1, 35), Fr(12, 35), Fr(18, 35), Fr(4, 35)]
    eps = Fr(3456, 10 ** 7)
    num = den = Fr(0)
    for k in range(4):
        nodes = [k - 3 + m for m in range(4)]
        vals = [Fr(float(x)) for x in s[k:k + 4]]
        cm = _cubic(nodes, vals, True)
        q = sum(ci * Fr(1, 2) ** e for e, ci in enumerate(cm))
        c = _cubic(nodes, vals, False)
        e2 = node_of_second[k]
        # the same cubic's derivatives, its second derivative taken at xi = e2
        c_eff = [c[0], c[1], c[2] + 3 * c[3] * e2, c[3]]
        beta = 34560 * _smoothness(c_eff)
        a = d[k] / (eps + beta) ** 2
        num += a * q
        den += a
    return num / den


def _weno7_inputs():
    rng = np.random.default_rng(7)
    out = [rng.uniform(-1.0, 1.0, 8) for _ in range(12)]
    out += [rng.uniform(0.5, 2.0, 8) * 1e-3 for _ in range(4)]          # small values: eps matters
    out += [np.sin(0.3 * np.arange(8) + ph) for ph in (0.0, 0.7, 2.1)]  # smooth: near-linear weights
    out += [np.where(np.arange(8) > m, 2.0, 1.0) + 0.01 * np.arange(8) for m in range(1, 6)]
    return out


@pytest.mark.parametrize("side", [P, M])
def test_weno7_face_value_from_first_principles(side):
    """The oracle's WENO7 face value against the candidate values, the
    integrated smoothness indicators and the optimal weights formed here; M
    is the P scheme on the mirrored stencil (src/OCFD_Schemes.cu:692-741)."""
    for s in _weno7_inputs():
        want = float(_weno7_face(s if side == P else s[6::-1]))
        got = O.face_flux(capi.SCHEME_IDS["WENO7"], side, s)
        assert abs(got - want) <= 1e-13 * max(1.0, np.abs(s).max()) + 1e-15, (side, s, got, want)


def test_weno7_textbook_stencil_2_differs():
    """With stencil 2's second derivative at node i (the textbook
    Balsara-Shu indicator) the expected values move by far more than the
    tolerance above on rough data: the pin sees where the second derivative
    of each stencil is taken."""
    worst = 0.0
    for s in _weno7_inputs()[:12]:
        got = O.face_flux(capi.SCHEME_IDS["WENO7"], P, s)
        worst = max(worst, abs(got - float(_weno7_face(s, (0, 0, 0, 0)))))
    assert worst > 1e-6, worst


# WENO5 (Jiang-Shu, src/OCFD_Schemes.cu:1118-1210), the same construction:
# node i = s[2], face i+1/2, stencils i-2..i, i-1..i+1, i..i+2 with optimal
# weights 1:6:3 / 10, candidate values of the quadratics whose cell means are
# the data, indicators 12 * sum_{l=1,2} int (p_k^(l))^2 over the cell (for
# quadratics the cell-mean and point-value polynomials differ by a constant,
# so the indicator is the same), eps = 12 * 1e-6.  M is P on s[4::-1].
def _weno5_face(s):
    d = [Fr(1, 10), Fr(6, 10), Fr(3, 10)]
    eps = Fr(12, 10 ** 6)
    num = den = Fr(0)
    for k in range(3):
        nodes = [k - 2 + m for m in range(3)]
        vals = [Fr(float(x)) for x in s[k:k + 3]]
        cm = _cubic(nodes, vals, True)
        q = sum(ci * Fr(1, 2) ** e for e, ci in enumerate(cm))
        a = d[k] / (eps + 12 * _smoothness(cm)) ** 2
        num += a * q
        den += a
    return num / den


@pytest.mark.parametrize("side", [P, M])
def test_weno5_face_value_from_first_principles(side):
    for s in _weno7_inputs():
        want = float(_weno5_face(s if side == P else s[4::-1]))
        got = O.face_flux(capi.SCHEME_IDS["WENO5"], side, s)
        assert abs(got - want) <= 1e-13 * max(1.0, np.abs(s).max()) + 1e-15, (side, s, got, want)


# WENO7-SYMBO (src/OCFD_Schemes.cu:391-478, the nonlinear branch): five
# stencils i-3+k..i+k (k = 0..4, the last one downwind of the face), the
# candidate values of the cubics whose cell means are the data, indicators
# 25920 * sum_{l=1..3} int (p_k^(l))^2 over the cell for those same
# cell-mean cubics (the reference's 720a^2 + 780b^2 + 781c^2 is exactly
# that), stencil 2's second derivative again at node i+1, the downwind
# stencil's indicator replaced by the largest of the five, eps = 25920 *
# 1e-10 = 2.592e-6, and the weights d_k / (eps + beta_k) to the first power
# (WENO7 and WENO5 square the denominator).  The five bandwidth-optimised
# weights d_k are the source's
# numbers (their sum is 1; the scheme's order is pinned by the convergence
# test).  M is P on the mirrored stencil s[::-1].
SYMBO_D = (0.0401954833730, 0.2493800006710, 0.4802686256260, 0.2009775476730, 0.0291783426580)


def _symbo_face(s):
    eps = Fr(2592, 10 ** 9)
    q, beta = [], []
    for k in range(5):
        nodes = [k - 3 + m for m in range(4)]
        vals = [Fr(float(x)) for x in s[k:k + 4]]
        c = _cubic(nodes, vals, True)
        q.append(sum(ci * Fr(1, 2) ** e for e, ci in enumerate(c)))
        e2 = 1 if k == 2 else 0
        beta.append(25920 * _smoothness([c[0], c[1], c[2] + 3 * c[3] * e2, c[3]]))
    beta[4] = max(beta)
    a = [Fr(dk) / (eps + b) for dk, b in zip(SYMBO_D, beta)]
    return sum(x * y for x, y in zip(a, q)) / sum(a)


def test_symbo_weights_sum_to_one():
    assert abs(sum(SYMBO_D) - 1.0) < 1e-11  # 13-digit decimals: 1 + 1e-12


@pytest.mark.parametrize("side", [P, M])
def test_weno7_symbo_face_value_from_first_principles(side):
    for s in _weno7_inputs():
        want = float(_symbo_face(s if side == P else s[::-1]))
        got = O.face_flux(capi.SCHEME_IDS["WENO7_SYMBO"], side, s)
        assert abs(got - want) <= 1e-13 * max(1.0, np.abs(s).max()) + 1e-15, (side, s, got, want)


def _symbo_linear(s):
    """sum_k d_k q_k / sum_k d_k: the bandwidth-optimised linear scheme (the
    weights normalised like the nonlinear ones; the source's d_k sum to
    1 + 1e-12)."""
    out = Fr(0)
    for k in range(5):
        c = _cubic([k - 3 + m for m in range(4)], [Fr(float(x)) for x in s[k:k + 4]], True)
        out += Fr(SYMBO_D[k]) * sum(ci * Fr(1, 2) ** e for e, ci in enumerate(c))
    return out / sum(Fr(d) for d in SYMBO_D)


def _stencil_tv(s):
    return [sum(abs(s[k + m + 1] - s[k + m]) for m in range(3)) for k in range(5)]


def test_symbo_tv_limiter_switch():
    """WENO7-SYMBO with the TV limiter (src/OCFD_Schemes.cu:391-478, lmt = 1):
    when the largest total variation of the five stencils is below tv_max
    and below tv_limiter times the smallest, the face value is the linear
    scheme's; otherwise the nonlinear one.  Smooth data of small amplitude
    take the linear branch, the same data scaled up (above tv_max) or with a
    small kink (one flat stencil: ratio unbounded) the nonlinear one; the
    M side never takes the linear branch (its smallest variation is never
    formed, src/OCFD_Schemes.cu:513-524)."""
    lim, tvm = 5.0, 0.2
    j = np.arange(8)
    cases = [0.01 * np.sin(0.3 * j + ph) for ph in (0.0, 0.9, 2.0)]
    cases += [0.01 * np.exp(0.25 * j)]
    sid = capi.SCHEME_IDS["WENO7_SYMBO_LIM"]
    for s in cases:
        tv = _stencil_tv(s)
        assert max(tv) < tvm and max(tv) < lim * min(tv)
        got = O.face_flux(sid, P, s, tv_limiter=lim, tv_max=tvm)
        lin, nl = float(_symbo_linear(s)), float(_symbo_face(s))
        assert abs(lin - nl) > 1e-9  # the two branches are told apart
        assert abs(got - lin) <= 1e-15, (s, got, lin, nl)
        big = 100.0 * s
        assert max(_stencil_tv(big)) > tvm
        assert abs(O.face_flux(sid, P, big, tv_limiter=lim, tv_max=tvm) - float(_symbo_face(big))) <= 1e-13
        # either side of tv_max: the same shape with its largest variation
        # at 0.95 and 1.25 tv_max
        below, above = s * (0.95 * tvm / max(tv)), s * (1.25 * tvm / max(tv))
        got = O.face_flux(sid, P, below, tv_limiter=lim, tv_max=tvm)
        assert abs(got - float(_symbo_linear(below))) <= 1e-14, (s, got)
        got = O.face_flux(sid, P, above, tv_limiter=lim, tv_max=tvm)
        assert abs(got - float(_symbo_face(above))) <= 1e-14, (s, got)
        kink = np.where(j > 4, s, s[4])
        assert min(_stencil_tv(kink)) * lim <= max(_stencil_tv(kink))
        assert abs(O.face_flux(sid, P, kink, tv_limiter=lim, tv_max=tvm) - float(_symbo_face(kink))) <= 1e-15
        got_m = O.face_flux(sid, M, s, tv_limiter=lim, tv_max=tvm)
        assert abs(got_m - float(_symbo_face(s[::-1]))) <= 1e-15, (s, got_m)


def test_symbo_limiter_is_active_on_the_gpu_parity_case():
    """The wavy box of the per-scheme GPU parity test (20x16x12, CD6) takes
    the limiter's linear branch: its right-hand side with WENO7_SYMBO_LIM
    differs from plain WENO7_SYMBO's by far more than the 1e-12 parity bar,
    so that test checks the HIP path's switch as well."""
    from pyocfd import cases
    du = {}
    for sch in ("WENO7_SYMBO", "WENO7_SYMBO_LIM"):
        case = cases.wavy_box(20, 16, 12, scheme=sch, scheme_vis="CD6")
        o = case.load(O.Oracle(case.cfg))
        o.rhs(1)
        du[sch] = np.asarray(o.get_du()).reshape(5, -1)
    a, b = du["WENO7_SYMBO"], du["WENO7_SYMBO_LIM"]
    rel = np.abs(a - b).max(1) / np.abs(a).max(1)
    assert (rel > 1e-4).all(), rel
