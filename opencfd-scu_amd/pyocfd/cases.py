"""Case generators for the benchmark configurations (synthetic, seeded).

The reference ships its own mesh/initial-field generators as separate C
programs (benchmark_comp/swept-compression-ramp.c, init3d-inlet1d.c); these are
this project's own generators for the workloads it measures:

* ``tgv``   - 3-D Taylor-Green vortex in a periodic box with identity metrics.
* ``wavy_box`` - curvilinear periodic box (every metric term non-zero).
* ``ramp``  - 24-degree compression ramp (optionally swept in z) on a
              tanh-stretched wall-normal grid, IBC 108 inflow/wall, hybrid
              WENO7/UP7 + CD8, the shock-capturing filter.
* ``shock_turbulence`` - Mach 1.5 normal shock in a Cartesian box with a
              seeded solenoidal von Karman upstream field, WENO7.

Each generator returns a :class:`Case` with the config, the grid (mesh
coordinates or the 13 Jacobi arrays), the primitive fields and, for the ramp,
the inlet profile.  ``Case.write(dir)`` writes the reference's input files
(opencfd-scu.in, OCFD3d-Mesh.dat / OCFD3d-Jacobi.dat, opencfd.dat,
flow1d-inlet.dat).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import capi

SEED = 20261202


@dataclass
class Case:
    name: str
    cfg: capi.OcfdConfig
    prims: list                       # d,u,v,w,T each (nz,ny,nx)
    mesh: tuple | None = None         # x,y,z each (nz,ny,nx)
    jacobi: np.ndarray | None = None  # (13,nz,ny,nx)
    inlet: np.ndarray | None = None   # (4,ny): d,u,v,T
    extra: dict = field(default_factory=dict)

    @property
    def npoints(self):
        return self.cfg.nx * self.cfg.ny * self.cfg.nz

    def load(self, solver, init=True):
        """Push grid, flow and inlet profile into a Solver/Oracle and init it."""
        if self.mesh is not None:
            solver.load_mesh(*self.mesh)
        else:
            solver.load_jacobi(self.jacobi)
        solver.load_flow(*self.prims, istep=0, tt=0.0)
        if self.inlet is not None:
            solver.set_inlet_profile(self.inlet)
        if "inlet5" in self.extra or "upper5" in self.extra:
            solver.set_bc_tables(self.extra.get("inlet5"), self.extra.get("upper5"))
        if init:
            solver.init()
        return solver

    def input_text(self) -> str:
        return config_text(self.cfg)

    def write(self, d: Path):
        """Write the reference's input files for this case into directory d."""
        d = Path(d)
        d.mkdir(parents=True, exist_ok=True)
        (d / "opencfd-scu.in").write_text(self.input_text())
        if self.mesh is not None:
            capi.write_records(d / "OCFD3d-Mesh.dat", list(self.mesh))
        else:
            capi.write_records(d / "OCFD3d-Jacobi.dat", list(self.jacobi))
        capi.write_flow_file(d / "opencfd.dat", self.prims, 0, 0.0)
        if "inlet5" in self.extra:
            # one record of nz*ny values per variable (y fastest)
            capi.write_records(d / "flow-inlet-section.dat", [a_[None] for a_ in self.extra["inlet5"]])
        if "upper5" in self.extra:
            capi.write_records(d / "flow-outboundary.dat", [a_[None] for a_ in self.extra["upper5"]])
        if self.inlet is not None:
            ny = self.cfg.ny
            y = self.extra.get("inlet_y", np.arange(ny, dtype=float))
            with open(d / "flow1d-inlet.dat", "w") as fh:
                fh.write(" variables=y,d,u,v,T\n")
                for j in range(ny):
                    fh.write("%24.16e %24.16e %24.16e %24.16e %24.16e\n" %
                             (y[j], self.inlet[0, j], self.inlet[1, j], self.inlet[2, j], self.inlet[3, j]))


def config_text(c: capi.OcfdConfig) -> str:
    """opencfd-scu.in text for a config (keys of src/parameters.c:166-194)."""
    inv = {v: k for k, v in capi.SCHEME_IDS.items()}
    L = [
        "# OpenCFD-SCU input file written by pyocfd.cases",
        f"GRID_3D = {c.nx} {c.ny} {c.nz}",
        f"PARALLEL_3D = {c.npx0 or 1} {c.npy0 or 1} {c.npz0 or 1}",
        f"STREAM = {c.stream_mode}",
        f"CHARTERIC = {c.charteric}",
        f"TEST = {c.test}",
        "IPERIODIC = %d %d %d" % tuple(c.iperiodic),
        "JAC_BOUND = %d %d %d" % tuple(c.jac_bound),
        "DIF_BOUND = %d %d %d %d %d %d" % tuple(c.dif_bound),
        "NON_REFLETION = %d %d %d %d %d %d" % tuple(c.non_ref),
        f"SCHEME_INVIS = {inv[c.scheme_invis]}",
        f"SCHEME_VIS = {inv[c.scheme_vis]}",
        "HY_DP_INTV = %.17g %.17g" % tuple(c.hy_dp_intv),
        f"HY_SMOOTH_DP = {c.hy_smooth_dp}",
        f"HY_STYLE = {c.hy_style}",
        f"HY_PATCH_ZONE = {c.hy_patch_zone}",
        f"RE = {c.re!r}",
        f"AMA = {c.ama!r}",
        f"GAMMA = {c.gamma!r}",
        f"PR = {c.pr!r}",
        f"T_REF = {c.t_ref!r}",
        f"EPSL_SW = {c.epsl_sw!r}",
        f"DT = {c.dt!r}",
        f"END_TIME = {c.end_time!r}",
        f"KSTEP_SHOW = {c.kstep_show}",
        f"KSTEP_SAVE = {c.kstep_save}",
        f"INIT_STAT = {c.init_stat}",
        f"IBC = {c.ibc}",
    ]
    if c.n_bc_npara:
        L.append("BC_NPARA = " + " ".join(str(c.bc_npara[i]) for i in range(c.n_bc_npara)))
    if c.n_bc_rpara:
        L.append("BC_RPARA = " + " ".join(repr(c.bc_rpara[i]) for i in range(c.n_bc_rpara)))
    for m in range(c.nfilter):
        L.append(f"FILTER_NPARA{m} = " + " ".join(str(c.filter_npara[m][i]) for i in range(11)))
        L.append(f"FILTER_RPARA{m} = " + " ".join(repr(c.filter_rpara[m][i]) for i in range(3)))
    for m in range(c.nana):
        L.append(f"ANA_EVENT{m} = {c.ana_event[m][0]} {c.ana_event[m][1]}")
        if c.n_ana_npara[m]:
            L.append(f"ANA_NPARA{m} = " + " ".join(str(c.ana_npara[m][i]) for i in range(c.n_ana_npara[m])))
        if c.n_ana_rpara[m]:
            L.append(f"ANA_RPARA{m} = " + " ".join(repr(c.ana_rpara[m][i]) for i in range(c.n_ana_rpara[m])))
    return "\n".join(L) + "\n"


def end_time_for(nsteps: int, dt: float) -> float:
    """END_TIME giving exactly nsteps with the reference's while(tt<end_time)."""
    return (nsteps - 0.5) * dt


# --------------------------------------------------------------------------
# Taylor-Green vortex
# --------------------------------------------------------------------------
def tgv(n=64, ma=0.1, re=1600.0, noise=1e-4, seed=SEED, cfl=0.5, scheme="WENO7",
        scheme_vis="CD6", nsteps=200, precision=capi.PREC_FP64) -> Case:
    """Periodic TGV box, identity Jacobian (Akx=Aiy=Asz=1, Ajac=1).

    With h = 1/(N-1) (src/OCFD_init.cu:29-31) the periodic length of the box
    in computational units is L = N*h; the vortex has wavenumber 2*pi/L.
    """
    nx = ny = nz = n
    gamma = 1.4
    h = 1.0 / (n - 1.0)
    L = n * h
    k = 2.0 * math.pi / L
    xs = np.arange(n) * h
    z3, y3, x3 = np.meshgrid(xs, xs, xs, indexing="ij")
    rng = np.random.default_rng(seed)
    u = np.sin(k * x3) * np.cos(k * y3) * np.cos(k * z3)
    v = -np.cos(k * x3) * np.sin(k * y3) * np.cos(k * z3)
    w = np.zeros_like(u)
    if noise:
        u = u + noise * rng.standard_normal(u.shape)
        v = v + noise * rng.standard_normal(v.shape)
        w = w + noise * rng.standard_normal(w.shape)
    d = np.ones_like(u)
    p = 1.0 / (gamma * ma * ma) + (np.cos(2 * k * x3) + np.cos(2 * k * y3)) * (np.cos(2 * k * z3) + 2.0) / 16.0
    T = gamma * ma * ma * p / d
    # CFL 0.5 on the inviscid spectral radius (computational units)
    c = np.sqrt(T) / ma
    lam = (np.abs(u) + c) / h + (np.abs(v) + c) / h + (np.abs(w) + c) / h
    dt = float(cfl / lam.max())
    jac = np.zeros((13, nz, ny, nx))
    jac[0], jac[1], jac[2] = x3, y3, z3
    jac[3] = 1.0   # Akx
    jac[7] = 1.0   # Aiy
    jac[11] = 1.0  # Asz
    jac[12] = 1.0  # Ajac
    cfg = capi.make_config(
        nx=nx, ny=ny, nz=nz, iperiodic=(1, 1, 1), jac_bound=(1, 1, 1),
        dif_bound=(0,) * 6, non_ref=(0,) * 6, scheme_invis=scheme, scheme_vis=scheme_vis,
        re=re, ama=ma, gamma=gamma, pr=0.72, t_ref=288.15, epsl_sw=0.0, dt=dt,
        end_time=end_time_for(nsteps, dt), ibc=0, precision=precision,
        hy_dp_intv=(0.01, 0.2), hy_style=2)
    return Case("tgv%d" % n, cfg, [d, u, v, w, T], jacobi=jac, extra={"nsteps": nsteps})


# --------------------------------------------------------------------------
# Curvilinear periodic box (parity workload: every metric term non-zero)
# --------------------------------------------------------------------------
def wavy_box(nx=20, ny=16, nz=12, ma=0.5, re=500.0, amp=0.04, seed=SEED, cfl=0.3, scheme="WENO7",
             scheme_vis="CD6", periodic=(1, 1, 1), precision=capi.PREC_FP64) -> Case:
    """Periodic box whose grid lines are smooth periodic waves (so all nine
    cofactors and the Jacobian vary) with a smooth seeded flow.  Metrics come
    from the mesh (JAC_BOUND = 1 1 1), as in OCFD3d-Mesh.dat runs."""
    gamma = 1.4
    rng = np.random.default_rng(seed)
    i = np.arange(nx) / nx
    j = np.arange(ny) / ny
    k = np.arange(nz) / nz
    K, J, I = np.meshgrid(k, j, i, indexing="ij")
    tp = 2 * np.pi
    x = I + amp * np.sin(tp * J) * np.cos(tp * K)
    y = 0.8 * J + amp * np.sin(tp * I + 0.3) * np.cos(tp * K)
    z = 0.6 * K + amp * np.sin(tp * I) * np.sin(tp * J + 0.7)
    ph = rng.uniform(0, tp, 12)
    d = 1.0 + 0.1 * np.sin(tp * I + ph[0]) * np.cos(tp * J + ph[1]) * np.cos(tp * K + ph[2])
    u = 0.3 + 0.2 * np.sin(tp * I + ph[3]) * np.cos(tp * J + ph[4]) * np.sin(tp * K + ph[5])
    v = -0.1 + 0.2 * np.cos(tp * I + ph[6]) * np.sin(tp * J + ph[7])
    w = 0.05 + 0.15 * np.sin(tp * J + ph[8]) * np.cos(tp * K + ph[9])
    T = 1.0 + 0.1 * np.cos(tp * I + ph[10]) * np.sin(tp * K + ph[11])
    h = min(1.0 / (nx - 1), 1.0 / (ny - 1), 1.0 / (nz - 1))
    c = np.sqrt(T) / ma
    dt = float(cfl * h / (np.abs(u) + np.abs(v) + np.abs(w) + 3 * c).max() * 0.5)
    cfg = capi.make_config(
        nx=nx, ny=ny, nz=nz, iperiodic=periodic, jac_bound=(1, 1, 1), dif_bound=(0,) * 6, non_ref=(0,) * 6,
        scheme_invis=scheme, scheme_vis=scheme_vis, hy_style=2, hy_dp_intv=(0.01, 0.2), re=re, ama=ma,
        gamma=gamma, pr=0.72, t_ref=288.15, epsl_sw=0.0, dt=dt, end_time=end_time_for(10, dt), ibc=0,
        precision=precision)
    return Case("wavy%dx%dx%d" % (nx, ny, nz), cfg, [d, u, v, w, T], mesh=(x, y, z))


# --------------------------------------------------------------------------
# Compression ramp (2-D extruded or swept)
# --------------------------------------------------------------------------
def _tanh_stretch(ny, first):
    """s_j in [0,1], s_0=0, s_1=first, tanh clustering at the wall."""
    j = np.arange(ny) / (ny - 1.0)

    def first_cell(b):
        return 1.0 - np.tanh(b * (1.0 - j[1])) / np.tanh(b)

    lo, hi = 1e-6, 50.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if first_cell(mid) > first:
            lo = mid
        else:
            hi = mid
    b = 0.5 * (lo + hi)
    return 1.0 - np.tanh(b * (1.0 - j)) / np.tanh(b)


def ramp_mesh(nx, ny, nz, angle_deg=24.0, sweep_deg=0.0, height=0.5, span=0.1, first=1e-4,
              corner=0.5, arc=0.02):
    """Wall follows y_w(x): flat to x=corner, then angle_deg, joined by a
    quadratic blend of half-width `arc`; wall-normal lines vertical; top flat at
    y=height; z uniform over `span` (periodic); sweep shears x by z*tan(sweep).
    """
    t = math.tan(math.radians(angle_deg))
    xs = np.linspace(0.0, 1.0, nx)
    yw = np.where(xs < corner - arc, 0.0,
                  np.where(xs > corner + arc, (xs - corner) * t,
                           t * (xs - corner + arc) ** 2 / (4 * arc)))
    s = _tanh_stretch(ny, first)
    zs = np.arange(nz) * (span / nz)
    X2 = np.broadcast_to(xs[None, :], (ny, nx))
    Y2 = yw[None, :] + (height - yw[None, :]) * s[:, None]
    x = np.empty((nz, ny, nx)); y = np.empty((nz, ny, nx)); z = np.empty((nz, ny, nx))
    shear = math.tan(math.radians(sweep_deg))
    for k in range(nz):
        x[k] = X2 + zs[k] * shear
        y[k] = Y2
        z[k] = zs[k]
    return x, y, z, s * height


def blasius_like(yn, delta, ma, gamma, tw, pr=0.72):
    """Pohlhausen quartic velocity profile of thickness delta with a
    Crocco-Busemann temperature, uniform pressure p = 1/(gamma Ma^2)."""
    eta = np.clip(yn / delta, 0.0, 1.0)
    U = 2 * eta - 2 * eta ** 3 + eta ** 4
    r = math.sqrt(pr)
    tr = 1.0 + r * 0.5 * (gamma - 1.0) * ma * ma
    T = tw + (tr - tw) * U + (1.0 - tr) * U * U
    d = 1.0 / T
    return d, U, T


def oblique_shock(ma, theta_deg, gamma=1.4):
    """Weak oblique-shock solution for a wedge of theta: (beta, rho2/rho1,
    p2/p1, u2, v2) with the upstream velocity (1, 0)."""
    th = math.radians(theta_deg)

    def f(b):
        return 2 / math.tan(b) * (ma * ma * math.sin(b) ** 2 - 1) / (ma * ma * (gamma + math.cos(2 * b)) + 2) - \
            math.tan(th)

    lo, hi = math.asin(1.0 / ma) + 1e-9, math.radians(65.0)  # weak branch below the max-deflection angle
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if f(mid) < 0:
            lo = mid
        else:
            hi = mid
    b = 0.5 * (lo + hi)
    mn2 = (ma * math.sin(b)) ** 2
    r21 = (gamma + 1) * mn2 / ((gamma - 1) * mn2 + 2)
    p21 = 1 + 2 * gamma / (gamma + 1) * (mn2 - 1)
    ut, un = math.cos(b), math.sin(b) / r21
    u2 = ut * math.cos(b) + un * math.sin(b)
    v2 = ut * math.sin(b) - un * math.cos(b)
    return b, r21, p21, u2, v2


def ramp(nx=512, ny=128, nz=64, ma=2.9, re=5000.0, angle=24.0, sweep=0.0, delta=0.02, tw=2.5,
         pert=0.01, seed=SEED, cfl=0.5, scheme="SCHEME_HYBRIDAUTO", filt=True, filter_every=10,
         nsteps=1000, precision=capi.PREC_FP64, height=0.5, span=0.1, t_ref=108.1, shock=True,
         corner=0.5, progress=None) -> Case:
    """Compression ramp.  With shock=True the initial inviscid field already
    holds the weak oblique shock of the ramp (from the corner at the
    theta-beta-Mach angle, smoothed over two cells) with the post-shock state
    turned parallel to the ramp, under the boundary-layer profile; upstream
    and at the inflow it is the freestream.  progress(msg), if given, is
    called after each stage of the generation (large grids take minutes)."""
    say = progress or (lambda msg: None)
    gamma = 1.4
    x, y, z, ycol = ramp_mesh(nx, ny, nz, angle, sweep, height=height, span=span, corner=corner)
    say("mesh")
    # wall distance along each grid line (vertical lines)
    yn = y - y[:, :1, :]
    d1, U1, T1 = blasius_like(ycol, delta, ma, gamma, tw)
    d, U, T = blasius_like(yn, delta, ma, gamma, tw)
    ui, vi = 1.0, 0.0
    if shock:
        beta, r21, p21, u2, v2 = oblique_shock(ma, angle, gamma)
        xs = x - z * math.tan(math.radians(sweep))  # streamwise coordinate of the unswept section
        dist = (y - (xs - corner) * math.tan(beta)) * math.cos(beta)
        ws = 0.5 * (1.0 - np.tanh(dist / (2.0 / (nx - 1))))
        rho_i = 1.0 + (r21 - 1.0) * ws
        p_i = (1.0 + (p21 - 1.0) * ws) / (gamma * ma * ma)
        ui = 1.0 + (u2 - 1.0) * ws
        vi = v2 * ws
        T_i = gamma * ma * ma * p_i / rho_i
        # the boundary-layer temperature profile runs from the wall value to
        # the local inviscid temperature
        T = tw + (T - tw) * (T_i - tw) / (1.0 - tw)
        d = gamma * ma * ma * p_i / T
    # seeded perturbations: white noise smoothed over a few cells (periodic in
    # z), scaled to a peak of `pert`, with an envelope that vanishes at the
    # wall and peaks at y = delta
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed)
    eta = yn / delta
    env = eta * np.exp(1.0 - eta)
    comps = []
    for _ in range(3):
        # FP32 noise: the same statistics at half the memory and time on
        # large grids
        r = gaussian_filter(rng.standard_normal(U.shape, dtype=np.float32), sigma=(2.0, 2.0, 3.0),
                            mode=("wrap", "nearest", "nearest"))
        comps.append(pert * (r / max(float(np.abs(r).max()), 1e-30)).astype(np.float64) * env)
        say(f"perturbation component {len(comps)} of 3")
    u = U * ui + comps[0]
    v = U * vi + comps[1]
    w = comps[2]
    inlet = np.stack([d1, U1, np.zeros(ny), T1])
    # CFL from the inviscid spectral radius in computational space (estimate
    # with CD2 metrics of the generated grid)
    hx, hy, hz = 1.0 / (nx - 1), 1.0 / (ny - 1), 1.0 / (nz - 1)
    xx = np.gradient(x, axis=2) / hx; yx = np.gradient(y, axis=2) / hx
    xy = np.gradient(x, axis=1) / hy; yy = np.gradient(y, axis=1) / hy
    zz = span / nz / hz
    det = (xx * yy - xy * yx) * zz
    c = np.sqrt(T) / ma
    # contravariant velocities per unit computational length
    kx, ky = yy * zz / det, -xy * zz / det
    ex, ey = -yx * zz / det, xx * zz / det
    sz = 1.0 / zz
    lam = (np.abs(kx * u + ky * v) + c * np.hypot(kx, ky)) / hx + \
          (np.abs(ex * u + ey * v) + c * np.hypot(ex, ey)) / hy + (np.abs(w * sz) + c * sz) / hz
    # viscous spectral radius: 2*gamma/Pr * mu/rho * sum |grad xi|^2 / h^2
    tsb = 110.4 / t_ref
    mu = (1.0 + tsb) / re * T ** 1.5 / (tsb + T)
    lam_v = 2.0 * gamma / 0.70 * mu / d * ((kx * kx + ky * ky) / hx ** 2 + (ex * ex + ey * ey) / hy ** 2
                                          + (sz / hz) ** 2)
    dt = float(cfl / (lam + lam_v).max())
    filters = []
    if filt:
        filters = [((filter_every, 1, 1, 1, 0, nx, 0, ny, 0, nz, 2), (1.0, 1e-5, 1e9))]
    cfg = capi.make_config(
        nx=nx, ny=ny, nz=nz, iperiodic=(0, 0, 1), jac_bound=(1, 1, 1),
        dif_bound=(0, 1, 0, 1, 0, 0), non_ref=(0, 1, 0, 1, 0, 0), scheme_invis=scheme, scheme_vis="CD8",
        hy_style=2, hy_dp_intv=(0.01, 0.2), re=re, ama=ma, gamma=gamma, pr=0.70, t_ref=t_ref, epsl_sw=0.0,
        dt=dt, end_time=end_time_for(nsteps, dt), ibc=108,
        bc_npara=(10, 5, 1, 0),
        bc_rpara=(tw, 0.02, 0.05, 0.1, 0.1, -1.0, -1.0, span),
        filters=filters, precision=precision)
    return Case("ramp%dx%dx%d" % (nx, ny, nz), cfg, [d, u, v, w, T], mesh=(x, y, z), inlet=inlet,
                extra={"inlet_y": ycol, "nsteps": nsteps})


# --------------------------------------------------------------------------
# The reference's benchmark deck (benchmark_comp/opencfd-scu.in)
# --------------------------------------------------------------------------
def _exp_stretch(n, first):
    """s_j in [0,1] with s_0 = 0, s_1 = first: s = (e^{b j/(n-1)} - 1)/(e^b - 1),
    b by bisection.  This is the wall-normal distribution the deck's inlet
    profile (flow1d-inlet.dat, first cell 0.01 of a height of 35) is given on."""
    j = np.arange(n) / (n - 1.0)
    lo, hi = 1e-6, 60.0
    for _ in range(200):
        b = 0.5 * (lo + hi)
        if math.expm1(b / (n - 1.0)) / math.expm1(b) > first:
            lo = b
        else:
            hi = b
    b = 0.5 * (lo + hi)
    return np.expm1(b * j) / math.expm1(b)


def deck_ramp(deck, inlet_file, nx_inlet=12, nx_corner=12, l_inlet=300.0, l_corner=50.0, buff_ratio=1.01,
              angle=24.0, height=35.0, first=0.01, span=14.0) -> Case:
    """Mesh and initial field for the reference's benchmark deck (its
    opencfd-scu.in, read by this library's own parser, and its
    flow1d-inlet.dat).  The reference builds them with separate generators
    (benchmark_comp/swept-compression-ramp.c with swept-compression-grid.in,
    benchmark_comp/init3d-inlet1d.c); this is this project's generator for the
    same kind of grid, from the grid file's parameters (the defaults):

    * wall line: a flat plate from x = -(l_inlet + l_corner) (nx_inlet
      uniform points), then nx_corner points of spacing 2 l_corner /
      (nx_corner - 1) through the corner at x = 0, where the wall turns up by
      `angle`, then the remaining points continuing the ramp with the spacing
      growing by buff_ratio per cell;
    * grid lines leave the wall along a normal that turns from vertical to the
      ramp normal across the corner points, `height` long, with the
      exponential distribution of first cell `first` (at the flat plate it
      reproduces the y column of flow1d-inlet.dat);
    * z uniform over [0, span] (z is periodic in the deck).

    The initial field is the inlet profile at every x and z, w = 0 (what
    init3d-inlet1d.c writes to its flow3d0.dat)."""
    cfg = capi.read_config(deck)
    nx, ny, nz = cfg.nx, cfg.ny, cfg.nz
    prof = np.loadtxt(inlet_file, skiprows=1)[:ny]
    th = math.radians(angle)
    h0 = 2.0 * l_corner / (nx_corner - 1)
    xw, yw, phi = [], [], []
    for i in range(nx_inlet):
        xw.append(-(l_inlet + l_corner) + i * (l_inlet - h0) / (nx_inlet - 1))
        yw.append(0.0)
        phi.append(0.0)
    for i in range(nx_corner):
        s = -l_corner + i * h0
        xw.append(s if s < 0 else s * math.cos(th))
        yw.append(0.0 if s < 0 else s * math.sin(th))
        phi.append(th * i / (nx_corner - 1))
    ds = h0
    while len(xw) < nx:
        ds *= buff_ratio
        xw.append(xw[-1] + ds * math.cos(th))
        yw.append(yw[-1] + ds * math.sin(th))
        phi.append(th)
    xw, yw, phi = np.array(xw[:nx]), np.array(yw[:nx]), np.array(phi[:nx])
    sj = _exp_stretch(ny, first / height)
    X2 = xw[None, :] - height * np.sin(phi)[None, :] * sj[:, None]
    Y2 = yw[None, :] + height * np.cos(phi)[None, :] * sj[:, None]
    zs = np.arange(nz) * (span / (nz - 1))
    x = np.broadcast_to(X2, (nz, ny, nx)).copy()
    y = np.broadcast_to(Y2, (nz, ny, nx)).copy()
    z = np.broadcast_to(zs[:, None, None], (nz, ny, nx)).copy()
    d, u, v, T = (np.broadcast_to(prof[:, c][None, :, None], (nz, ny, nx)).copy() for c in (1, 2, 3, 4))
    w = np.zeros((nz, ny, nx))
    inlet = np.stack([prof[:, 1], prof[:, 2], prof[:, 3], prof[:, 4]])
    return Case("deck%dx%dx%d" % (nx, ny, nz), cfg, [d, u, v, w, T], mesh=(x, y, z), inlet=inlet,
                extra={"inlet_y": prof[:, 0]})


# --------------------------------------------------------------------------
# Lifting body / body of revolution (IBC 124)
# --------------------------------------------------------------------------
def liftbody(nx=48, ny=24, nz=20, ma=2.0, re=2000.0, aoa=2.0, tw=1.5, delta=0.15, r_out=2.0, cfl=0.4,
             scheme="WENO7_SYMBO", mzmax=0, mtmax=5, beta=0.3, upper=0, withleading=0, nsteps=10,
             precision=capi.PREC_FP64, seed=SEED) -> Case:
    """Half of a body of revolution, the geometry IBC 124 is written for: x
    along the axis, (y, z) = r (cos t, sin t) with t from pi/2 to -pi/2, so
    both j = 0 and j = ny-1 lie on the symmetry plane y = 0; the wall is r = 1
    (k = 0, with a gentle bump along x so that all metrics vary) and the upper
    face r = r_out (k = nz-1, tanh clustering at the wall).  Free stream at
    angle of attack aoa (u = cos, w = sin) with a wall layer of thickness
    delta, seeded noise.  BC_NPARA = IF_SYMMETRY 1, IF_WITHLEADING,
    IFLAG_UPPERBOUNDARY, MZMAX, MTMAX; BC_RPARA = AOA, TW (<= 0: adiabatic),
    EPSL_WALL, EPSL_UPPER, BETA, WALL_DIS_BEGIN, WALL_DIS_END.  With
    withleading / upper set, case.extra holds the inlet / upper tables."""
    gamma = 1.4
    rng = np.random.default_rng(seed)
    xi = np.linspace(0.0, 1.0, nx)
    th = 0.5 * np.pi - np.pi * np.arange(ny) / (ny - 1)  # decreasing: a right-handed grid
    s = _tanh_stretch(nz, 0.01)
    K, J, I = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    r0 = 1.0 + 0.05 * np.sin(np.pi * xi[I]) ** 2
    r = r0 + (r_out - r0) * s[K]
    x = xi[I] + 0.02 * s[K] * np.sin(2 * np.pi * xi[I])
    y = r * np.cos(th[J])
    z = r * np.sin(th[J])
    y[:, 0, :] = 0.0
    y[:, -1, :] = 0.0
    a = math.radians(aoa)
    eta = np.clip((r - r0) / delta, 0.0, None)
    f = np.tanh(2.0 * eta)
    T_wall = tw if tw > 0 else 1.0 + 0.5 * (gamma - 1) * ma * ma * 0.85
    T = T_wall + (1.0 - T_wall) * f
    d = 1.0 / T
    noise = [1e-3 * rng.standard_normal(x.shape) * f * (1 - f + 0.1) for _ in range(3)]
    u = math.cos(a) * f + noise[0]
    v = noise[1]
    w = math.sin(a) * f + noise[2]
    h = min(1.0 / (nx - 1), 1.0 / (ny - 1), 1.0 / (nz - 1))
    dr = np.diff(r, axis=0).min()
    dt = float(cfl * min(dr, 1.0 / (nx - 1), np.pi / (ny - 1)) / (1.0 + math.sqrt(T.max()) / ma) * 0.5)
    cfg = capi.make_config(
        nx=nx, ny=ny, nz=nz, iperiodic=(0, 0, 0), jac_bound=(1, 0, 1),
        dif_bound=(0, 1, 0, 0, 1, 1), non_ref=(0, 1, 0, 0, 0, 1), scheme_invis=scheme, scheme_vis="CD6",
        hy_style=2, hy_dp_intv=(0.01, 0.2), re=re, ama=ma, gamma=gamma, pr=0.72, t_ref=288.15, epsl_sw=0.0,
        dt=dt, end_time=end_time_for(nsteps, dt), ibc=124,
        bc_npara=(1, withleading, upper, mzmax, mtmax),
        bc_rpara=(aoa, tw, 1e-3, 1e-4, beta, 0.2, 0.6),
        precision=precision)
    extra = {"nsteps": nsteps}
    if withleading:
        extra["inlet5"] = np.stack([a_[:, :, 0] for a_ in (d, u, v, w, T)]) * (1.0 + 1e-3)
    if upper:
        extra["upper5"] = np.stack([np.full((ny, nx), 1.0), np.full((ny, nx), math.cos(a)), np.zeros((ny, nx)),
                                    np.full((ny, nx), math.sin(a)), np.full((ny, nx), 1.0)]) * (1.0 + 2e-3)
    return Case("liftbody%dx%dx%d" % (nx, ny, nz), cfg, [d, u, v, w, T], mesh=(x, y, z), extra=extra)


# --------------------------------------------------------------------------
# Shock-turbulence interaction (Cartesian)
# --------------------------------------------------------------------------
def solenoidal_field(shape, lengths, k0, u_rms, seed=SEED):
    """Seeded divergence-free random velocity (periodic box, `shape` =
    (nz, ny, nx), `lengths` = (Lz, Ly, Lx)) with the von Karman energy
    spectrum E(k) ~ (k/k0)^4 / (1 + (k/k0)^2)^(17/6), k0 in units of the
    fundamental wavenumber of the y period, scaled to the given u_rms (rms of
    each component)."""
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape
    Lz, Ly, Lx = lengths
    kz = np.fft.fftfreq(nz, Lz / nz) * 2 * np.pi
    ky = np.fft.fftfreq(ny, Ly / ny) * 2 * np.pi
    kx = np.fft.rfftfreq(nx, Lx / nx) * 2 * np.pi
    KZ, KY, KX = np.meshgrid(kz, ky, kx, indexing="ij")
    k2 = KX * KX + KY * KY + KZ * KZ
    kk = np.sqrt(k2)
    kref = k0 * 2 * np.pi / Ly
    with np.errstate(divide="ignore", invalid="ignore"):
        E = (kk / kref) ** 4 / (1.0 + (kk / kref) ** 2) ** (17.0 / 6.0)
        amp = np.where(kk > 0, np.sqrt(E / (4 * np.pi * np.maximum(k2, 1e-300))), 0.0)
    comps = []
    for _ in range(3):
        c = rng.standard_normal(KX.shape) + 1j * rng.standard_normal(KX.shape)
        comps.append(c * amp)
    # projection onto divergence-free fields: u - k (k.u) / |k|^2
    with np.errstate(divide="ignore", invalid="ignore"):
        dot = (KX * comps[0] + KY * comps[1] + KZ * comps[2]) / np.where(k2 > 0, k2, 1.0)
    comps = [comps[0] - KX * dot, comps[1] - KY * dot, comps[2] - KZ * dot]
    u = [np.fft.irfftn(c, s=shape, axes=(0, 1, 2)) for c in comps]
    rms = math.sqrt(sum(float((a * a).mean()) for a in u) / 3.0)
    return [a * (u_rms / rms) for a in u]


def shock_turbulence(nx=512, ny=256, nz=256, ma=1.5, re=2000.0, x_shock=0.25, k0=4.0, u_rms=0.1,
                     seed=SEED, cfl=0.5, scheme="WENO7", nsteps=500, precision=capi.PREC_FP64) -> Case:
    """Normal shock at Mach `ma` at x = x_shock*Lx in a Cartesian box
    (Lx = 4*pi, Ly = Lz = 2*pi, y and z periodic); upstream a seeded
    solenoidal von Karman velocity field with constant density and pressure
    (rho = 1, T = 1, mean u = 1); downstream the Rankine-Hugoniot state.  The
    jump is smoothed over two cells and the fluctuations are tapered off over
    a few cells ahead of the shock."""
    gamma = 1.4
    Lx, Ly, Lz = 4 * math.pi, 2 * math.pi, 2 * math.pi
    xs = np.arange(nx) * (Lx / (nx - 1))
    ys = np.arange(ny) * (Ly / ny)
    zs = np.arange(nz) * (Lz / nz)
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    dx = Lx / (nx - 1)
    m2 = ma * ma
    r21 = (gamma + 1) * m2 / ((gamma - 1) * m2 + 2)          # rho2/rho1
    p21 = 1 + 2 * gamma / (gamma + 1) * (m2 - 1)              # p2/p1
    xs0 = x_shock * Lx
    s = 0.5 * (1 + np.tanh((xs - xs0) / (2 * dx)))            # 0 upstream, 1 downstream
    d1 = 1.0 + (r21 - 1.0) * s
    u1 = 1.0 + (1.0 / r21 - 1.0) * s
    p1 = (1.0 + (p21 - 1.0) * s) / (gamma * m2)
    T1 = gamma * m2 * p1 / d1
    fl = solenoidal_field((nz, ny, nx), (Lz, Ly, Lx), k0, u_rms, seed)
    taper = 0.5 * (1 - np.tanh((xs - (xs0 - 6 * dx)) / (2 * dx)))
    d = np.broadcast_to(d1, X.shape).copy()
    u = u1[None, None, :] + fl[0] * taper
    v = fl[1] * taper
    w = fl[2] * taper
    T = np.broadcast_to(T1, X.shape).copy()
    c = np.sqrt(T) / ma
    # inviscid spectral radius; in computational units xi = i*h, so
    # (|u|+c) * (dxi/dx) / h = (|u|+c) / dx per direction
    lam = (np.abs(u) + c) / dx + (np.abs(v) + c) / (Ly / ny) + (np.abs(w) + c) / (Lz / nz)
    dt = float(cfl / lam.max())
    cfg = capi.make_config(
        nx=nx, ny=ny, nz=nz, iperiodic=(0, 1, 1), jac_bound=(1, 1, 1),
        dif_bound=(1, 1, 0, 0, 0, 0), non_ref=(0, 0, 0, 0, 0, 0), scheme_invis=scheme, scheme_vis="CD6",
        hy_style=2, hy_dp_intv=(0.01, 0.2), re=re, ama=ma, gamma=gamma, pr=0.72, t_ref=288.15, epsl_sw=0.0,
        dt=dt, end_time=end_time_for(nsteps, dt), ibc=0, precision=precision)
    return Case("shockturb%dx%dx%d" % (nx, ny, nz), cfg, [d, u, v, w, T], mesh=(X, Y, Z),
                extra={"nsteps": nsteps})
