"""ocfd_diagnostics_noalloc, the allocation-free variant of
ocfd_diagnostics (its partial sums and counts in a file-scope device
scratch array instead of buffers allocated on first use), against
ocfd_diagnostics on the same contexts, through the C ABI: calls in a row on
grids of decreasing size (FP64 and mixed path, NaN and negative T in the
field), the first context again after the smaller ones, and a grid whose
partials do not fit the scratch, refused.

Bars: the NaN and negative counts exactly; the means to 1e-13 relative
(the partials are the same; only the order of the final sum differs).
"""
import numpy as np
import pytest

from pyocfd import capi, cases

pytestmark = pytest.mark.gpu


def assert_same(a, b):
    (e1, t1, n1, g1), (e2, t2, n2, g2) = a, b
    assert n1 == n2 and g1 == g2, (a, b)
    for x, y in ((e1, e2), (t1, t2)):
        if np.isnan(y):
            assert np.isnan(x)
        else:
            assert abs(x - y) <= 1e-13 * abs(y), (x, y)


def test_noalloc_diagnostics_equal_the_allocating_entry_point():
    big = cases.wavy_box(200, 40, 24, scheme="WENO7")  # 4 x 10 x 24 partial blocks
    g = big.load(capi.Solver(big.cfg))
    g.step(2)
    first = g.diagnostics_noalloc()
    want = g.diagnostics()
    print(f"[no_alloc] 200x40x24 after 2 steps: {first}", flush=True)
    assert_same(first, want)
    f = g.get_state()
    T = g.get_flow()[0][4]
    assert abs(first[0] - f[4].mean()) <= 1e-13 * abs(f[4].mean())
    assert abs(first[1] - T.mean()) <= 1e-13 * abs(T.mean())
    assert first[2] == 0 and first[3] == 0

    # a smaller grid next, with three negative temperatures and one NaN
    small = cases.wavy_box(20, 16, 12, scheme="WENO7")
    d, u, v, w, T0 = [a.copy() for a in small.prims]
    T0[5, 6, 7] = T0[6, 7, 8] = T0[5, 8, 9] = -1.0
    T0[6, 6, 6] = np.nan
    for precision in (capi.PREC_FP64, capi.PREC_MIXED):
        cfg = capi.OcfdConfig.from_buffer_copy(small.cfg)
        cfg.precision = precision
        h = capi.Solver(cfg)
        h.load_mesh(*small.mesh)
        h.load_flow(d, u, v, w, T0, istep=0, tt=0.0)
        h.init()
        got = h.diagnostics_noalloc()
        print(f"[no_alloc] 20x16x12 precision {precision}: {got}", flush=True)
        assert_same(got, h.diagnostics())
        assert got[2] == 1 and got[3] == 3 and np.isnan(got[0])
        h.close()

    # the first context again: nothing of the smaller calls is left in use
    assert_same(g.diagnostics_noalloc(), want)
    g.close()


def test_noalloc_diagnostics_refuse_grids_beyond_the_scratch():
    """8 x 8 x 262000: 2 x 262000 partial blocks of 64 x 4 points, 524000 >
    523774, the most that OCFD_NOALLOC_SCRATCH_BYTES (8 MiB) holds; the
    allocating entry point is the one for such grids."""
    cfg = capi.OcfdConfig.from_buffer_copy(cases.wavy_box(8, 8, 8).cfg)
    cfg.nz = 262000
    h = capi.Solver(cfg)
    with pytest.raises(RuntimeError, match="scratch"):
        h.diagnostics_noalloc()
    h.close()
