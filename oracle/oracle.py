"""ctypes wrapper of the CPU oracle (oracle/ocfd_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker, never as the product.
"""
from __future__ import annotations

import ctypes as C
import subprocess
import sys
from pathlib import Path

import numpy as np

ORACLE_DIR = Path(__file__).resolve().parent
REPO = ORACLE_DIR.parent
LIB = ORACLE_DIR / "liboracle.so"
sys.path.insert(0, str(REPO / "opencfd-scu_amd"))
from pyocfd.capi import OcfdConfig  # noqa: E402  (interface type only)


LIB_FMA = ORACLE_DIR / "liboracle_fma.so"


def build(force: bool = False, path: Path = LIB) -> Path:
    """Compile liboracle.so (or liboracle_fma.so) with the committed Makefile."""
    srcs = [ORACLE_DIR / n for n in ("ocfd_oracle.c", "ocfd_oracle.h", "Makefile")]
    if force or not path.exists() or path.stat().st_mtime < max(p.stat().st_mtime for p in srcs):
        subprocess.run(["make", "-C", str(ORACLE_DIR), path.name], check=True,
                       stdout=subprocess.DEVNULL)
    return path


_libs = {}


def lib(fma: bool = False) -> C.CDLL:
    """The oracle library; fma=True: the build with contracted multiply-adds."""
    path = LIB_FMA if fma else LIB
    if path not in _libs:
        build(path=path)
        L = C.CDLL(str(path))
        dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
        i8 = np.ctypeslib.ndpointer(dtype=np.int8, flags="C_CONTIGUOUS")
        vp = C.c_void_p
        sig = {
            "orc_create": (vp, [C.POINTER(OcfdConfig)]),
            "orc_destroy": (None, [vp]),
            "orc_load_mesh": (C.c_int, [vp, dp, dp, dp]),
            "orc_load_jacobi": (C.c_int, [vp, dp]),
            "orc_load_flow": (C.c_int, [vp, dp, dp, dp, dp, dp, C.c_int, C.c_double]),
            "orc_set_inlet_profile": (C.c_int, [vp, dp]),
            "orc_set_bc_tables": (C.c_int, [vp, C.c_void_p, C.c_void_p]),
            "orc_init": (C.c_int, [vp]),
            "orc_rhs": (C.c_int, [vp, C.c_int]),
            "orc_step": (C.c_int, [vp, C.c_int]),
            "orc_num_threads": (C.c_int, []),
            "orc_get_flow": (C.c_int, [vp, dp, dp, dp, dp, dp, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
            "orc_get_state": (C.c_int, [vp, dp]),
            "orc_get_du": (C.c_int, [vp, dp]),
            "orc_get_pressure": (C.c_int, [vp, dp]),
            "orc_get_pressure_lap": (C.c_int, [vp, dp]),
            "orc_get_prim_lap": (C.c_int, [vp, C.c_int, dp]),
            "orc_filter_now": (C.c_int, [vp]),
            "orc_get_metric_lap": (C.c_int, [vp, C.c_int, dp]),
            "orc_get_amu": (C.c_int, [vp, dp]),
            "orc_get_metrics": (C.c_int, [vp, dp]),
            "orc_get_flags": (C.c_int, [vp, i8, i8, i8]),
            "orc_get_q": (C.c_int, [vp, dp]),
            "orc_char_matrices": (None, [dp, dp, C.c_double, dp, dp]),
            "orc_line_flux_char": (None, [C.c_int, C.c_int, C.c_int, C.c_int, dp, dp, dp, C.c_double, C.c_int,
                                          C.c_int, C.c_int, C.c_void_p, C.c_double, dp]),
            "orc_face_flux": (C.c_double, [C.c_int, C.c_int, C.c_int, dp, C.c_double, C.c_double]),
            "orc_split_point": (None, [C.c_double] * 10 + [dp]),
            "orc_line_flux_derivative": (None, [C.c_int, C.c_int, C.c_int, dp, dp, C.c_double, C.c_int,
                                                C.c_int, C.c_int, C.c_void_p, dp]),
            "orc_line_central": (None, [C.c_int, C.c_int, dp, C.c_double, C.c_int, C.c_int, dp]),
            "orc_jameson_dp": (C.c_double, [C.c_double] * 7),
            "orc_rk_stage": (None, [C.c_int, C.c_int, C.c_double, dp, dp, dp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _ip = ip  # noqa: F841
        _libs[path] = L
    return _libs[path]


class Oracle:
    """Mirror of pyocfd.capi.Solver backed by the CPU restatement."""

    def __init__(self, cfg: OcfdConfig, fma: bool = False):
        self.L = lib(fma)
        self.cfg = cfg
        self.nx, self.ny, self.nz = cfg.nx, cfg.ny, cfg.nz
        self.h = self.L.orc_create(C.byref(cfg))

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def shape(self):
        return (self.nz, self.ny, self.nx)

    @staticmethod
    def _flat(a):
        return np.ascontiguousarray(a, dtype=np.float64).ravel()

    def load_mesh(self, x, y, z):
        self.L.orc_load_mesh(self.h, self._flat(x), self._flat(y), self._flat(z))

    def load_jacobi(self, a13):
        self.L.orc_load_jacobi(self.h, self._flat(a13))

    def load_flow(self, d, u, v, w, T, istep=0, tt=0.0):
        self.L.orc_load_flow(self.h, *(self._flat(a) for a in (d, u, v, w, T)), int(istep), float(tt))

    def set_inlet_profile(self, prof):
        self.L.orc_set_inlet_profile(self.h, self._flat(prof))

    def set_bc_tables(self, inlet5=None, upper5=None):
        a = None if inlet5 is None else self._flat(inlet5)
        b = None if upper5 is None else self._flat(upper5)
        self.L.orc_set_bc_tables(self.h, None if a is None else a.ctypes.data_as(C.c_void_p),
                                 None if b is None else b.ctypes.data_as(C.c_void_p))

    def init(self):
        self.L.orc_init(self.h)

    def rhs(self, krk=1):
        self.L.orc_rhs(self.h, krk)

    def step(self, n=1):
        self.L.orc_step(self.h, n)

    def get_flow(self):
        n = self.nx * self.ny * self.nz
        arrs = [np.empty(n) for _ in range(5)]
        istep, tt = C.c_int(), C.c_double()
        self.L.orc_get_flow(self.h, *arrs, C.byref(istep), C.byref(tt))
        return [a.reshape(self.shape) for a in arrs], istep.value, tt.value

    def get_state(self):
        out = np.empty(5 * self.nx * self.ny * self.nz)
        self.L.orc_get_state(self.h, out)
        return out.reshape((5,) + self.shape)

    def get_du(self):
        out = np.empty(5 * self.nx * self.ny * self.nz)
        self.L.orc_get_du(self.h, out)
        return out.reshape((5,) + self.shape)

    def get_pressure(self):
        out = np.empty(self.nx * self.ny * self.nz)
        self.L.orc_get_pressure(self.h, out)
        return out.reshape(self.shape)

    def get_pressure_lap(self, ghosts=1):
        """Pressure with `ghosts` ghost layers on each side (at most 4)."""
        n2 = (self.nx + 8) * (self.ny + 8) * (self.nz + 8)
        out = np.empty(n2)
        self.L.orc_get_pressure_lap(self.h, out)
        out = out.reshape(self.nz + 8, self.ny + 8, self.nx + 8)
        c = 4 - ghosts
        return out[c:out.shape[0] - c, c:out.shape[1] - c, c:out.shape[2] - c]

    def get_prim_lap(self, k):
        """Primitive k (0..4 = d,u,v,w,T) with all 4 ghost layers, [z][y][x]."""
        out = np.empty((self.nx + 8) * (self.ny + 8) * (self.nz + 8))
        self.L.orc_get_prim_lap(self.h, int(k), out)
        return out.reshape(self.nz + 8, self.ny + 8, self.nx + 8)

    def filter_now(self):
        """The filter alone on the current state (every filter due)."""
        self.L.orc_filter_now(self.h)

    def get_metric_lap(self, k):
        """Metric array k (0..12) over the whole LAP box (nz+8, ny+8, nx+8)."""
        out = np.empty((self.nx + 8) * (self.ny + 8) * (self.nz + 8))
        self.L.orc_get_metric_lap(self.h, int(k), out)
        return out.reshape(self.nz + 8, self.ny + 8, self.nx + 8)

    def get_amu(self):
        out = np.empty(self.nx * self.ny * self.nz)
        self.L.orc_get_amu(self.h, out)
        return out.reshape(self.shape)

    def get_metrics(self):
        out = np.empty(13 * self.nx * self.ny * self.nz)
        self.L.orc_get_metrics(self.h, out)
        return out.reshape((13,) + self.shape)

    def get_flags(self):
        nx, ny, nz = self.nx, self.ny, self.nz
        fx = np.empty((nx + 1) * ny * nz, np.int8)
        fy = np.empty(nx * (ny + 1) * nz, np.int8)
        fz = np.empty(nx * ny * (nz + 1), np.int8)
        self.L.orc_get_flags(self.h, fx, fy, fz)
        return fx.reshape(nz, ny, nx + 1), fy.reshape(nz, ny + 1, nx), fz.reshape(nz + 1, ny, nx)

    def get_q(self):
        out = np.empty(self.nx * self.ny * self.nz)
        self.L.orc_get_q(self.h, out)
        return out.reshape(self.shape)


# ---- unit-level helpers -------------------------------------------------
def face_flux(scheme, side, s8, hyflag=0, tv_limiter=5.0, tv_max=0.2):
    s = np.zeros(8)
    s[: len(s8)] = s8
    return lib().orc_face_flux(scheme, side, hyflag, s, tv_limiter, tv_max)


def split_point(d, u, v, w, cc, A1, A2, A3, gamma=1.4, eps=0.0):
    out = np.empty(10)
    lib().orc_split_point(d, u, v, w, cc, A1, A2, A3, gamma, eps, out)
    return out[:5], out[5:]


def line_flux_derivative(scheme, side, f, ajac, h, bound_l=0, bound_r=0, nonref=0, flags=None):
    """f has n+8 values (nodes -4..n+3); returns du (n) = -ajac*(h_r-h_l)/h."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    n = f.size - 8
    aj = np.ascontiguousarray(ajac, dtype=np.float64)
    du = np.zeros(n)
    fl = None
    if flags is not None:
        fl_arr = np.ascontiguousarray(flags, dtype=np.int32)
        fl = fl_arr.ctypes.data_as(C.c_void_p)
    lib().orc_line_flux_derivative(scheme, side, n, f, aj, h, bound_l, bound_r, nonref, fl, du)
    return du


def char_matrices(s0, s1, gamma=1.4):
    """(L, R) of a face from the u v w c a1 a2 a3 of its two nodes."""
    L, R = np.empty(25), np.empty(25)
    lib().orc_char_matrices(np.ascontiguousarray(s0, dtype=np.float64), np.ascontiguousarray(s1, dtype=np.float64),
                            gamma, L, R)
    return L.reshape(5, 5), R.reshape(5, 5)


def line_flux_char(scheme, side, dir, F, prim, ajac, h, bound_l=0, bound_r=0, nonref=0, flags=None, gamma=1.4):
    """F (5, n+8) split fluxes at nodes -4..n+3, prim (7, n+8) = u v w c a1 a2
    a3; returns du (5, n) of the CHARTERIC = 1 flux derivative."""
    F = np.ascontiguousarray(F, dtype=np.float64)
    n = F.shape[1] - 8
    du = np.zeros(5 * n)
    fl = None
    if flags is not None:
        fl_arr = np.ascontiguousarray(flags, dtype=np.int32)
        fl = fl_arr.ctypes.data_as(C.c_void_p)
    lib().orc_line_flux_char(scheme, side, dir, n, F, np.ascontiguousarray(prim, dtype=np.float64),
                             np.ascontiguousarray(ajac, dtype=np.float64), h, bound_l, bound_r, nonref, fl, gamma, du)
    return du.reshape(5, n)


def line_central(scheme_vis, f, h, bound_l=0, bound_r=0):
    f = np.ascontiguousarray(f, dtype=np.float64)
    n = f.size - 8
    out = np.zeros(n)
    lib().orc_line_central(scheme_vis, n, f, h, bound_l, bound_r, out)
    return out


def rk_stage(krk, dt, f, fn, du):
    f = np.ascontiguousarray(f, dtype=np.float64).copy()
    lib().orc_rk_stage(f.size, krk, dt, f, np.ascontiguousarray(fn, dtype=np.float64),
                       np.ascontiguousarray(du, dtype=np.float64))
    return f


def jameson_dp(pc, pxm, pxp, pym, pyp, pzm, pzp):
    return lib().orc_jameson_dp(pc, pxm, pxp, pym, pyp, pzm, pzp)
