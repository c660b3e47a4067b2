"""C ABI of libocfd_hip.so on the CPU: the library loads, exports every
function declared in include/ocfd.h, and its host-side format code (the
opencfd-scu.in parser, Fortran-record and flow files) follows the
reference's rules.  No GPU is touched here."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from pyocfd import capi, cases

REPO = Path(__file__).resolve().parent.parent


def declared_functions():
    text = (REPO / "include" / "ocfd.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*[A-Za-z_][\w\s\*]*?\b(\w+)\s*\([^;{]*\)\s*;", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "while")))


def test_every_declared_symbol_is_exported(hiplib):
    names = declared_functions()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(hiplib, n)]
    assert not missing, missing


def test_config_roundtrip_through_parser(tmp_path):
    for case in (cases.tgv(16, nsteps=3), cases.ramp(64, 32, 8, nsteps=3)):
        capi.set_analysis(case.cfg, [(100, 5), (104, 10, [2, 1, 3, 7], [0.5]), (107, 1)])
        p = tmp_path / "opencfd-scu.in"
        p.write_text(case.input_text())
        c = capi.read_config(p)
        a, b = capi.config_to_dict(case.cfg), capi.config_to_dict(c)
        for k in a:
            if k in ("filter_npara", "filter_rpara"):
                assert a[k][: case.cfg.nfilter] == b[k][: case.cfg.nfilter], k
            elif k in ("ana_event", "n_ana_npara", "n_ana_rpara", "ana_npara", "ana_rpara"):
                assert a[k][: case.cfg.nana] == b[k][: case.cfg.nana], k
            elif k in ("bc_npara", "bc_rpara"):
                n = case.cfg.n_bc_npara if k == "bc_npara" else case.cfg.n_bc_rpara
                assert a[k][:n] == b[k][:n], k
            elif k in ("npx0", "npy0", "npz0", "precision", "kstep_save"):
                continue
            elif k.startswith("hy_") and case.cfg.scheme_invis != capi.HYBRIDAUTO:
                continue
            else:
                assert a[k] == b[k], (k, a[k], b[k])


def test_parser_follows_reference_key_rules(tmp_path):
    """src/parameters.c:667-772: spaces in keys ignored, last occurrence
    wins, '#' keys never match, unknown keys ignored, numbered filter
    families discovered by stripping digits, T_REF <= 0 -> 288.15."""
    text = "\n".join([
        "GRID_3D = 10 11 12",
        "G RID_3D = 20 21 22",           # spaces removed from the key: wins (last)
        "#GRID_3D = 1 1 1",              # never matches
        "SCHEME_INVIS = WENO7",
        "SCHEME_VIS = CD8",
        "UNKNOWN_KEY = 7",
        "T_REF = -1",
        "FILTER_NPARA3 = 5 1 0 1 0 20 0 21 0 22 2",
        "FILTER_RPARA3 = 0.5 1e-5 100",
        "#FILTER_NPARA4 = 9 9 9 9 9 9 9 9 9 9 9",
        "IBC = 108",
        "BC_NPARA = 10 5 1 0",
        "BC_RPARA = 2.5 0.2 -1 1 0.1 -1 -1 14",
        "RE = 100",                       # last line without newline is ignored
    ])
    p = tmp_path / "opencfd-scu.in"
    p.write_text(text)
    c = capi.read_config(p)
    assert (c.nx, c.ny, c.nz) == (20, 21, 22)
    assert c.scheme_invis == capi.WENO7 and c.scheme_vis == capi.CD8
    assert c.t_ref == 288.15
    assert c.nfilter == 1 and list(c.filter_npara[0]) == [5, 1, 0, 1, 0, 20, 0, 21, 0, 22, 2]
    assert list(c.filter_rpara[0]) == [0.5, 1e-5, 100.0]
    assert c.n_bc_npara == 4 and c.n_bc_rpara == 8 and c.bc_rpara[7] == 14.0
    assert c.re == 0.0


def test_parser_reads_analysis_events(tmp_path):
    """ANA_EVENT& = style Kstep with ANA_NPARA& / ANA_RPARA& of the same
    number, in file order (src/parameters.c:328-358)."""
    text = "\n".join([
        "GRID_3D = 8 8 8",
        "SCHEME_INVIS = WENO7",
        "SCHEME_VIS = CD6",
        "ANA_EVENT7 = 104 20",
        "ANA_NPARA7 = 2 1 3 5",
        "ANA_RPARA7 = 0.25",
        "ANA_EVENT2 = 100 1",
        "#ANA_EVENT3 = 101 1",
        "ANA_EVENT = 101 1",              # no number: not an event here
        "",
    ])
    p = tmp_path / "opencfd-scu.in"
    p.write_text(text)
    c = capi.read_config(p)
    assert c.nana == 2
    assert list(c.ana_event[0]) == [104, 20] and list(c.ana_event[1]) == [100, 1]
    assert c.n_ana_npara[0] == 4 and list(c.ana_npara[0])[:4] == [2, 1, 3, 5]
    assert c.n_ana_rpara[0] == 1 and c.ana_rpara[0][0] == 0.25
    assert c.n_ana_npara[1] == 0 and c.n_ana_rpara[1] == 0


def test_parser_rejects_bad_scheme(tmp_path):
    p = tmp_path / "opencfd-scu.in"
    p.write_text("GRID_3D = 8 8 8\nSCHEME_INVIS = WENO9\nSCHEME_VIS = CD6\n")
    with pytest.raises(RuntimeError):
        capi.read_config(p)


def test_flow_and_record_files_roundtrip(tmp_path):
    rng = np.random.default_rng(0)
    nx, ny, nz = 7, 5, 3
    prims = [rng.standard_normal((nz, ny, nx)) for _ in range(5)]
    capi.write_flow_file(tmp_path / "f.dat", prims, 42, 1.25)
    raw = (tmp_path / "f.dat").read_bytes()
    # header record {int32 12 | int32 Istep | f64 tt | int32 12}, src/OCFD_IO.c:184-193
    hdr = np.frombuffer(raw[:20], dtype=np.uint8)
    assert np.frombuffer(hdr[:4].tobytes(), np.int32)[0] == 12
    assert np.frombuffer(hdr[4:8].tobytes(), np.int32)[0] == 42
    assert np.frombuffer(hdr[8:16].tobytes(), np.float64)[0] == 1.25
    assert len(raw) == 20 + 5 * nz * (8 + 8 * nx * ny)
    back, istep, tt = capi.read_flow_file(tmp_path / "f.dat", nx, ny, nz)
    assert istep == 42 and tt == 1.25
    for a, b in zip(prims, back):
        assert np.array_equal(a, b)
    arrs = [rng.standard_normal((nz, ny, nx)) for _ in range(3)]
    capi.write_records(tmp_path / "m.dat", arrs)
    assert np.array_equal(capi.read_records(tmp_path / "m.dat", nx, ny, nz, 3), np.stack(arrs))


def test_create_without_gpu_fails_loudly(hiplib):
    """No silent CPU path: creating a context without a usable device is an error."""
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("a GPU is visible")
    cfg = cases.tgv(8, nsteps=1).cfg
    h = C.c_void_p()
    rc = hiplib.ocfd_create(C.byref(cfg), 0, C.byref(h))
    assert rc != 0


def test_create_refuses_more_wall_modes_than_the_tables_hold(hiplib):
    """MTMAX above OCFD_MTMAX (64) is a configuration error, for IBC 108
    (BC_NPARA[1]) and IBC 124 (BC_NPARA[4]); the check comes before any
    device call, so it holds on a host without a GPU too."""
    c = cases.ramp(64, 32, 8, nsteps=3).cfg
    c.bc_npara[1] = 65
    h = C.c_void_p()
    assert hiplib.ocfd_create(C.byref(c), 0, C.byref(h)) == 2  # OCFD_ERR_CONFIG
    assert b"MTMAX" in hiplib.ocfd_last_error()
    c = cases.liftbody(nsteps=1).cfg
    c.bc_npara[4] = 65
    assert hiplib.ocfd_create(C.byref(c), 0, C.byref(h)) == 2
    assert b"MTMAX" in hiplib.ocfd_last_error()
