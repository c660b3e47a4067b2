"""ctypes bindings of libocfd_hip.so (declared in include/ocfd.h).

This module is the Python host-side mirror of the C ABI: the ``OcfdConfig``
structure mirrors ``ocfd_config`` and :class:`Solver` wraps one ``ocfd_ctx``.
All compute runs in the HIP library; loading fails loudly when it is missing.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent.parent          # opencfd-scu_amd/
LIB_PATH = PKG_DIR / "libocfd_hip.so"

LAP = 4
NVARS = 5
MAX_FILTERS = 10
MAX_ANA = 16
ANA_NPARA = 100

CD6, CD8 = 203, 204
UP7, WENO5, WENO7, WENO7_SYMBO, WENO7_SYMBO_LIM, NND2 = 301, 302, 303, 304, 305, 306
OMP6_HR, OMP6_LD, OMP6_CD8, HYBRIDAUTO = 307, 308, 309, 310
PREC_FP64, PREC_MIXED = 0, 1

SCHEME_IDS = {
    "UP7": UP7, "WENO5": WENO5, "WENO7": WENO7, "WENO7_SYMBO": WENO7_SYMBO,
    "WENO7_SYMBO_LIM": WENO7_SYMBO_LIM, "NND2": NND2, "OMP6_HR": OMP6_HR,
    "OMP6_LD": OMP6_LD, "OMP6_CD8": OMP6_CD8, "SCHEME_HYBRIDAUTO": HYBRIDAUTO,
    "CD6": CD6, "CD8": CD8,
}


class OcfdConfig(C.Structure):
    """Mirror of ``ocfd_config`` (include/ocfd.h)."""

    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("npx0", C.c_int), ("npy0", C.c_int), ("npz0", C.c_int),
        ("stream_mode", C.c_int), ("charteric", C.c_int), ("test", C.c_int),
        ("iperiodic", C.c_int * 3), ("jac_bound", C.c_int * 3),
        ("dif_bound", C.c_int * 6), ("non_ref", C.c_int * 6),
        ("scheme_invis", C.c_int), ("scheme_vis", C.c_int),
        ("hy_style", C.c_int), ("hy_smooth_dp", C.c_int), ("hy_patch_zone", C.c_int),
        ("hy_dp_intv", C.c_double * 2),
        ("re", C.c_double), ("ama", C.c_double), ("gamma", C.c_double), ("pr", C.c_double),
        ("t_ref", C.c_double), ("epsl_sw", C.c_double),
        ("dt", C.c_double), ("end_time", C.c_double),
        ("kstep_show", C.c_int), ("kstep_save", C.c_int), ("init_stat", C.c_int),
        ("ibc", C.c_int),
        ("n_bc_npara", C.c_int), ("n_bc_rpara", C.c_int),
        ("bc_npara", C.c_int * 32), ("bc_rpara", C.c_double * 32),
        ("nfilter", C.c_int),
        ("filter_npara", (C.c_int * 11) * MAX_FILTERS),
        ("filter_rpara", (C.c_double * 3) * MAX_FILTERS),
        ("precision", C.c_int),
        ("nana", C.c_int),
        ("ana_event", (C.c_int * 2) * MAX_ANA),
        ("n_ana_npara", C.c_int * MAX_ANA), ("n_ana_rpara", C.c_int * MAX_ANA),
        ("ana_npara", (C.c_int * ANA_NPARA) * MAX_ANA),
        ("ana_rpara", (C.c_double * ANA_NPARA) * MAX_ANA),
    ]


def make_config(**kw) -> OcfdConfig:
    """Build an ``OcfdConfig`` from keyword arguments (field names above).

    Array fields accept sequences; ``filters`` is a list of
    ``(npara11, rpara3)`` pairs; ``analysis`` is a list of
    ``(style, kstep, npara, rpara)`` tuples (ANA_EVENT& / ANA_NPARA& /
    ANA_RPARA&); ``bc_npara``/``bc_rpara`` set their counts.
    """
    c = OcfdConfig()
    c.npx0 = c.npy0 = c.npz0 = 1
    c.gamma = 1.4
    c.pr = 0.72
    c.t_ref = 288.15
    c.scheme_vis = CD6
    c.scheme_invis = WENO7
    c.hy_style = 2
    c.kstep_show = 1
    c.kstep_save = 1 << 30
    c.init_stat = 1
    filters = kw.pop("filters", [])
    analysis = kw.pop("analysis", [])
    for k, v in kw.items():
        if k in ("iperiodic", "jac_bound", "dif_bound", "non_ref", "hy_dp_intv"):
            arr = getattr(c, k)
            for i, x in enumerate(v):
                arr[i] = x
        elif k == "bc_npara":
            for i, x in enumerate(v):
                c.bc_npara[i] = int(x)
            c.n_bc_npara = len(v)
        elif k == "bc_rpara":
            for i, x in enumerate(v):
                c.bc_rpara[i] = float(x)
            c.n_bc_rpara = len(v)
        elif k in ("scheme_invis", "scheme_vis") and isinstance(v, str):
            setattr(c, k, SCHEME_IDS[v])
        else:
            setattr(c, k, v)
    c.nfilter = len(filters)
    for m, (npara, rpara) in enumerate(filters):
        for i in range(11):
            c.filter_npara[m][i] = int(npara[i])
        for i in range(3):
            c.filter_rpara[m][i] = float(rpara[i])
    set_analysis(c, analysis)
    return c


def set_analysis(c: OcfdConfig, events):
    """Replace the analysis events of a config: (style, kstep, npara, rpara)."""
    assert len(events) <= MAX_ANA
    c.nana = len(events)
    for m, ev in enumerate(events):
        style, kstep = ev[0], ev[1]
        npara = list(ev[2]) if len(ev) > 2 else []
        rpara = list(ev[3]) if len(ev) > 3 else []
        c.ana_event[m][0], c.ana_event[m][1] = int(style), int(kstep)
        c.n_ana_npara[m], c.n_ana_rpara[m] = len(npara), len(rpara)
        for i, v in enumerate(npara):
            c.ana_npara[m][i] = int(v)
        for i, v in enumerate(rpara):
            c.ana_rpara[m][i] = float(v)


def config_to_dict(c: OcfdConfig) -> dict:
    out = {}
    for name, _ in OcfdConfig._fields_:
        v = getattr(c, name)
        if hasattr(v, "__len__"):
            v = [list(x) if hasattr(x, "__len__") else x for x in v]
        out[name] = v
    return out


_lib = None


def load_library(path: os.PathLike | None = None) -> C.CDLL:
    """Load libocfd_hip.so; raise if it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(
            f"{p} is missing: build it with `python __graft_entry__.py build` "
            "(hipcc --offload-arch=gfx950); there is no CPU fallback")
    lib = C.CDLL(str(p))
    dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
    i8p = np.ctypeslib.ndpointer(dtype=np.int8, flags="C_CONTIGUOUS")
    vp = C.c_void_p
    sig = {
        "ocfd_read_config": (C.c_int, [C.c_char_p, C.POINTER(OcfdConfig)]),
        "ocfd_read_flow_file": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, dp, dp, dp, dp, dp,
                                          C.POINTER(C.c_int), C.POINTER(C.c_double)]),
        "ocfd_write_flow_file": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, dp, dp, dp, dp, dp,
                                           C.c_int, C.c_double]),
        "ocfd_read_records": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, dp]),
        "ocfd_write_records": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, dp]),
        "ocfd_create": (C.c_int, [C.POINTER(OcfdConfig), C.c_int, C.POINTER(vp)]),
        "ocfd_destroy": (C.c_int, [vp]),
        "ocfd_last_error": (C.c_char_p, []),
        "ocfd_load_mesh": (C.c_int, [vp, dp, dp, dp]),
        "ocfd_load_jacobi": (C.c_int, [vp, dp]),
        "ocfd_load_flow": (C.c_int, [vp, dp, dp, dp, dp, dp, C.c_int, C.c_double]),
        "ocfd_set_inlet_profile": (C.c_int, [vp, dp]),
        "ocfd_set_bc_tables": (C.c_int, [vp, C.c_void_p, C.c_void_p]),
        "ocfd_init": (C.c_int, [vp]),
        "ocfd_rhs": (C.c_int, [vp, C.c_int]),
        "ocfd_step": (C.c_int, [vp, C.c_int]),
        "ocfd_sync": (C.c_int, [vp]),
        "ocfd_get_flow": (C.c_int, [vp, dp, dp, dp, dp, dp, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
        "ocfd_get_state": (C.c_int, [vp, dp]),
        "ocfd_get_du": (C.c_int, [vp, dp]),
        "ocfd_get_pressure": (C.c_int, [vp, dp]),
        "ocfd_get_amu": (C.c_int, [vp, dp]),
        "ocfd_get_metrics": (C.c_int, [vp, dp]),
        "ocfd_write_jacobi": (C.c_int, [vp, C.c_char_p]),
        "ocfd_set_sensor": (C.c_int, [vp, C.c_void_p, C.c_float, C.c_float]),
        "ocfd_set_sensor_fused": (C.c_int, [vp, C.c_int]),
        "ocfd_get_sensor_prob": (C.c_int, [vp, fp]),
        "ocfd_mlp_create": (C.c_int, [C.c_int, C.c_longlong, C.POINTER(C.c_void_p)]),
        "ocfd_mlp_destroy": (C.c_int, [vp]),
        "ocfd_mlp_set_params": (C.c_int, [vp, fp]),
        "ocfd_mlp_get_params": (C.c_int, [vp, fp]),
        "ocfd_mlp_sample_count": (C.c_longlong, [vp]),
        "ocfd_mlp_clear": (C.c_int, [vp]),
        "ocfd_mlp_collect": (C.c_int, [vp, vp, C.c_int]),
        "ocfd_mlp_put_samples": (C.c_int, [vp, fp, fp, C.c_longlong]),
        "ocfd_mlp_get_samples": (C.c_int, [vp, C.c_longlong, C.c_longlong, fp, fp]),
        "ocfd_mlp_grad": (C.c_int, [vp, C.c_longlong, C.c_int, C.c_float, fp, C.POINTER(C.c_float)]),
        "ocfd_mlp_predict": (C.c_int, [vp, C.c_longlong, C.c_int, fp]),
        "ocfd_mlp_train": (C.c_int, [vp, C.c_int, C.c_int, C.c_float, C.c_float, C.c_ulonglong, fp]),
        "ocfd_set_sensor_from": (C.c_int, [vp, vp, C.c_float, C.c_float]),
        "ocfd_diagnostics": (C.c_int, [vp, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                       C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
        "ocfd_diagnostics_noalloc": (C.c_int, [vp, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                               C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
        "ocfd_get_hybrid_flags": (C.c_int, [vp, i8p, i8p, i8p]),
        "ocfd_analysis": (C.c_int, [vp, C.c_int, C.c_int]),
        "ocfd_get_q": (C.c_int, [vp, dp]),
        "ocfd_get_time_average": (C.c_int, [vp, dp, dp, dp, dp, dp, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
        "ocfd_hybrid_proportions": (C.c_int, [vp, dp]),
        "ocfd_profile_enable": (C.c_int, [vp, C.c_int]),
        "ocfd_profile_count": (C.c_int, [vp]),
        "ocfd_profile_entry": (C.c_int, [vp, C.c_int, C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
        "ocfd_profile_reset": (C.c_int, [vp]),
        "ocfd_last_step_ms": (C.c_double, [vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def _check(rc: int, what: str):
    if rc != 0:
        msg = load_library().ocfd_last_error()
        raise RuntimeError(f"{what} failed (status {rc}): {msg.decode() if msg else ''}")


def read_config(path: os.PathLike) -> OcfdConfig:
    c = OcfdConfig()
    _check(load_library().ocfd_read_config(str(path).encode(), C.byref(c)), "ocfd_read_config")
    return c


def read_flow_file(path, nx, ny, nz):
    arrs = [np.empty(nx * ny * nz) for _ in range(5)]
    istep, tt = C.c_int(), C.c_double()
    _check(load_library().ocfd_read_flow_file(str(path).encode(), nx, ny, nz, *arrs,
                                              C.byref(istep), C.byref(tt)), "read_flow_file")
    return [a.reshape(nz, ny, nx) for a in arrs], istep.value, tt.value


def write_flow_file(path, prims, istep, tt):
    nz, ny, nx = prims[0].shape
    arrs = [np.ascontiguousarray(a, dtype=np.float64).ravel() for a in prims]
    _check(load_library().ocfd_write_flow_file(str(path).encode(), nx, ny, nz, *arrs, int(istep),
                                               float(tt)), "write_flow_file")


def write_records(path, arrays):
    nz, ny, nx = arrays[0].shape
    buf = np.ascontiguousarray(np.stack(arrays), dtype=np.float64).ravel()
    _check(load_library().ocfd_write_records(str(path).encode(), nx, ny, nz, len(arrays), buf),
           "write_records")


def read_records(path, nx, ny, nz, nfield):
    buf = np.empty(nfield * nx * ny * nz)
    _check(load_library().ocfd_read_records(str(path).encode(), nx, ny, nz, nfield, buf), "read_records")
    return buf.reshape(nfield, nz, ny, nx)


MLP_NPARAM = 1889


# Learned-sensor weights shipped with the package (trained by
# pyocfd.train_sensor).
SENSOR_PARAMS = Path(__file__).resolve().parent / "sensor_params.npy"


def mlp_pack(p) -> np.ndarray:
    """Dict W1 b1 W2 b2 W3 b3 -> the 1889-float parameter vector."""
    return np.concatenate([np.ravel(p[k]) for k in ("W1", "b1", "W2", "b2", "W3", "b3")]).astype(np.float32)


def mlp_unpack(v) -> dict:
    v = np.asarray(v)
    return {"W1": v[0:768].reshape(32, 24), "b1": v[768:800], "W2": v[800:1824].reshape(32, 32),
            "b2": v[1824:1856], "W3": v[1856:1888], "b3": v[1888:1889]}


class Trainer:
    """The learned-sensor trainer (ocfd_mlp) on one device."""

    def __init__(self, capacity, device=0):
        self.lib = load_library()
        h = C.c_void_p()
        _check(self.lib.ocfd_mlp_create(device, int(capacity), C.byref(h)), "mlp_create")
        self.h = h

    def close(self):
        if self.h:
            self.lib.ocfd_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, v):
        _check(self.lib.ocfd_mlp_set_params(self.h, np.ascontiguousarray(v, np.float32)), "mlp_set_params")

    def get_params(self):
        out = np.empty(MLP_NPARAM, np.float32)
        _check(self.lib.ocfd_mlp_get_params(self.h, out), "mlp_get_params")
        return out

    @property
    def n(self):
        return self.lib.ocfd_mlp_sample_count(self.h)

    def clear(self):
        _check(self.lib.ocfd_mlp_clear(self.h), "mlp_clear")

    def collect(self, solver, stride):
        _check(self.lib.ocfd_mlp_collect(self.h, solver.h, int(stride)), "mlp_collect")

    def put(self, X, y):
        X = np.ascontiguousarray(X, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        _check(self.lib.ocfd_mlp_put_samples(self.h, X, y, len(y)), "mlp_put_samples")

    def get(self, first=0, n=None):
        n = self.n - first if n is None else n
        X = np.empty((n, 24), np.float32)
        y = np.empty(n, np.float32)
        _check(self.lib.ocfd_mlp_get_samples(self.h, first, n, X, y), "mlp_get_samples")
        return X, y

    def grad(self, first, n, wpos=1.0):
        g = np.empty(MLP_NPARAM, np.float32)
        loss = C.c_float()
        _check(self.lib.ocfd_mlp_grad(self.h, first, n, wpos, g, C.byref(loss)), "mlp_grad")
        return loss.value, g

    def predict(self, first=0, n=None):
        n = self.n - first if n is None else n
        out = np.empty(n, np.float32)
        _check(self.lib.ocfd_mlp_predict(self.h, first, n, out), "mlp_predict")
        return out

    def train(self, epochs, batch, lr, wpos=1.0, seed=1):
        out = np.empty(epochs, np.float32)
        _check(self.lib.ocfd_mlp_train(self.h, epochs, batch, lr, wpos, seed, out), "mlp_train")
        return out


class Solver:
    """One ``ocfd_ctx`` on a HIP device.  Arrays are (nz, ny, nx) float64."""

    def __init__(self, cfg: OcfdConfig, device: int = 0):
        self.lib = load_library()
        self.cfg = cfg
        self.nx, self.ny, self.nz = cfg.nx, cfg.ny, cfg.nz
        h = C.c_void_p()
        _check(self.lib.ocfd_create(C.byref(cfg), device, C.byref(h)), "ocfd_create")
        self.h = h

    def close(self):
        if self.h:
            self.lib.ocfd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def shape(self):
        return (self.nz, self.ny, self.nx)

    def _flat(self, a):
        return np.ascontiguousarray(a, dtype=np.float64).ravel()

    def load_mesh(self, x, y, z):
        _check(self.lib.ocfd_load_mesh(self.h, self._flat(x), self._flat(y), self._flat(z)), "load_mesh")

    def load_jacobi(self, a13):
        _check(self.lib.ocfd_load_jacobi(self.h, self._flat(a13)), "load_jacobi")

    def load_flow(self, d, u, v, w, T, istep=0, tt=0.0):
        _check(self.lib.ocfd_load_flow(self.h, *(self._flat(a) for a in (d, u, v, w, T)), int(istep),
                                       float(tt)), "load_flow")

    def set_inlet_profile(self, prof):
        _check(self.lib.ocfd_set_inlet_profile(self.h, self._flat(prof)), "set_inlet_profile")

    def set_bc_tables(self, inlet5=None, upper5=None):
        """IBC 124 tables: inlet (5, nz, ny), upper (5, ny, nx); either None."""
        a = None if inlet5 is None else self._flat(inlet5)
        b = None if upper5 is None else self._flat(upper5)
        _check(self.lib.ocfd_set_bc_tables(self.h, None if a is None else a.ctypes.data_as(C.c_void_p),
                                           None if b is None else b.ctypes.data_as(C.c_void_p)), "set_bc_tables")
        self._keep = (a, b)

    def init(self):
        _check(self.lib.ocfd_init(self.h), "ocfd_init")

    def rhs(self, krk=1):
        _check(self.lib.ocfd_rhs(self.h, krk), "ocfd_rhs")

    def step(self, n=1):
        _check(self.lib.ocfd_step(self.h, n), "ocfd_step")

    def sync(self):
        _check(self.lib.ocfd_sync(self.h), "ocfd_sync")

    def get_flow(self):
        n = self.nx * self.ny * self.nz
        arrs = [np.empty(n) for _ in range(5)]
        istep, tt = C.c_int(), C.c_double()
        _check(self.lib.ocfd_get_flow(self.h, *arrs, C.byref(istep), C.byref(tt)), "get_flow")
        return [a.reshape(self.shape) for a in arrs], istep.value, tt.value

    def get_state(self):
        out = np.empty(5 * self.nx * self.ny * self.nz)
        _check(self.lib.ocfd_get_state(self.h, out), "get_state")
        return out.reshape((5,) + self.shape)

    def get_du(self):
        out = np.empty(5 * self.nx * self.ny * self.nz)
        _check(self.lib.ocfd_get_du(self.h, out), "get_du")
        return out.reshape((5,) + self.shape)

    def get_pressure(self):
        out = np.empty(self.nx * self.ny * self.nz)
        _check(self.lib.ocfd_get_pressure(self.h, out), "get_pressure")
        return out.reshape(self.shape)

    def get_amu(self):
        out = np.empty(self.nx * self.ny * self.nz)
        _check(self.lib.ocfd_get_amu(self.h, out), "get_amu")
        return out.reshape(self.shape)

    def get_metrics(self):
        out = np.empty(13 * self.nx * self.ny * self.nz)
        _check(self.lib.ocfd_get_metrics(self.h, out), "get_metrics")
        return out.reshape((13,) + self.shape)

    def diagnostics(self):
        """(mean total energy, mean T, NaN count of T, negative-T count)."""
        e, t, nn, ng = C.c_double(), C.c_double(), C.c_longlong(), C.c_longlong()
        _check(self.lib.ocfd_diagnostics(self.h, C.byref(e), C.byref(t), C.byref(nn), C.byref(ng)), "diagnostics")
        return e.value, t.value, nn.value, ng.value

    def diagnostics_noalloc(self):
        """The same through ocfd_diagnostics_noalloc (no allocation: a
        file-scope device scratch array)."""
        e, t, nn, ng = C.c_double(), C.c_double(), C.c_longlong(), C.c_longlong()
        _check(self.lib.ocfd_diagnostics_noalloc(self.h, C.byref(e), C.byref(t), C.byref(nn), C.byref(ng)),
               "diagnostics_noalloc")
        return e.value, t.value, nn.value, ng.value

    def analysis(self, style, event=0):
        """OCFD_ana: run analysis `style` with the parameters of event `event`
        (files go to the working directory)."""
        _check(self.lib.ocfd_analysis(self.h, int(style), int(event)), f"analysis {style}")

    def get_q(self):
        out = np.empty(self.nx * self.ny * self.nz)
        _check(self.lib.ocfd_get_q(self.h, out), "get_q")
        return out.reshape(self.shape)

    def get_time_average(self):
        """(d, u, v, w, T averages, samples, accumulated time)."""
        n = self.nx * self.ny * self.nz
        a = [np.empty(n) for _ in range(5)]
        ns, tt = C.c_int(), C.c_double()
        _check(self.lib.ocfd_get_time_average(self.h, *a, C.byref(ns), C.byref(tt)), "get_time_average")
        return [x.reshape(self.shape) for x in a], ns.value, tt.value

    def hybrid_proportions(self):
        """(3, 3): fractions of flag types 1, 2, 3 for x, y, z."""
        out = np.empty(9)
        _check(self.lib.ocfd_hybrid_proportions(self.h, out), "hybrid_proportions")
        return out.reshape(3, 3)

    def set_sensor(self, params=None, lo=0.2, hi=0.8):
        """Learned hybrid sensor with these 1889 parameters (None: plain)."""
        if params is None:
            _check(self.lib.ocfd_set_sensor(self.h, None, 0.0, 1.0), "set_sensor")
            return
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.size == MLP_NPARAM
        _check(self.lib.ocfd_set_sensor(self.h, p.ctypes.data_as(C.c_void_p), lo, hi), "set_sensor")

    def set_sensor_fused(self, on=True):
        """Run the learned sensor inside the x sweep of the first stage (True,
        the default) or in its own pass (False); the same flags either way."""
        _check(self.lib.ocfd_set_sensor_fused(self.h, 1 if on else 0), "set_sensor_fused")

    def get_sensor_prob(self):
        out = np.empty(self.nx * self.ny * self.nz, np.float32)
        _check(self.lib.ocfd_get_sensor_prob(self.h, out), "get_sensor_prob")
        return out.reshape(self.shape)

    def write_jacobi(self, path):
        """OCFD3d-Jacobi.dat with the coordinates and metrics of this context."""
        _check(self.lib.ocfd_write_jacobi(self.h, str(path).encode()), "write_jacobi")

    def get_flags(self):
        nx, ny, nz = self.nx, self.ny, self.nz
        fx = np.empty((nx + 1) * ny * nz, np.int8)
        fy = np.empty(nx * (ny + 1) * nz, np.int8)
        fz = np.empty(nx * ny * (nz + 1), np.int8)
        _check(self.lib.ocfd_get_hybrid_flags(self.h, fx, fy, fz), "get_hybrid_flags")
        return fx.reshape(nz, ny, nx + 1), fy.reshape(nz, ny + 1, nx), fz.reshape(nz + 1, ny, nx)

    # -- kernel timing -----------------------------------------------------
    def profile(self, on=True):
        _check(self.lib.ocfd_profile_enable(self.h, 1 if on else 0), "profile_enable")

    def profile_reset(self):
        _check(self.lib.ocfd_profile_reset(self.h), "profile_reset")

    def profile_entries(self):
        out = {}
        buf = C.create_string_buffer(64)
        for i in range(self.lib.ocfd_profile_count(self.h)):
            n, ms = C.c_int(), C.c_double()
            _check(self.lib.ocfd_profile_entry(self.h, i, buf, C.byref(n), C.byref(ms)), "profile_entry")
            out[buf.value.decode()] = (n.value, ms.value)
        return out

    def last_step_ms(self):
        return self.lib.ocfd_last_step_ms(self.h)
