"""C-ABI checks that need no GPU: the library builds and loads, exports every
entry point the public headers declare, and reports argument errors (and a
missing device) through the OSKAR status codes without touching outputs."""
import ctypes as ct
import os
import re

import numpy as np
import pytest

import fpga_gridding_amd as fg
from fpga_gridding_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _ensure_built():
    if not os.path.exists(fg.LIB_PATH):
        from fpga_gridding_amd import build
        build.build()


def _declared_symbols():
    names = set()
    for h in sorted(f for f in os.listdir(os.path.join(ROOT, "include"))
                    if f.endswith(".h")):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = "\n".join(l for l in text.splitlines()
                         if not l.lstrip().startswith("#"))
        # GW_EXPORT <ret> <name>(   possibly across lines
        for m in re.finditer(r"GW_(?:IO_)?EXPORT\s+[\w\s\*]*?\b(\w+)\s*\(",
                             text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    _ensure_built()
    declared = _declared_symbols()
    assert {"oskar_grid_wproj_f", "oskar_grid_wproj_d",
            "oskar_grid_wproj_f_diagnostics", "oskar_grid_weights_write_f",
            "oskar_grid_weights_read_f", "oskar_grid_weights_write_d",
            "oskar_grid_weights_read_d"} <= declared
    lib = ct.CDLL(fg.LIB_PATH)
    for name in sorted(declared):
        assert hasattr(lib, name), name
    from fpga_gridding_amd import io as gio
    assert declared == set(fg.EXPORTED_SYMBOLS) | set(gio.IO_SYMBOLS)


def test_error_strings_and_codes():
    _ensure_built()
    L = fg.lib()
    assert L.gw_error_string(0) == b"no error"
    assert L.gw_error_string(-3) == b"invalid function argument"
    assert L.gw_error_string(-19) == b"compute device error"


def _wl():
    return synth.random_small(n=50, grid_size=32, num_planes=3, seed=1)


@pytest.mark.parametrize("change,fragment", [
    (dict(conv_size_half=3), "conv_size_half too small"),
    (dict(num_w_planes=0), "num_w_planes"),
    (dict(oversample=0), "oversample"),
    (dict(grid_size=0), "grid_size"),
    (dict(support_neg=True), "support values"),
])
def test_invalid_arguments_rejected_before_device_use(change, fragment):
    """Argument errors come back as OSKAR_ERR_INVALID_ARGUMENT (-3) and leave
    grid, norm and num_skipped untouched (checked before any GPU call)."""
    _ensure_built()
    wl = _wl()
    support = wl.support.copy()
    if change.get("support_neg"):
        support[1] = -1
    G = change.get("grid_size", wl.grid_size)
    grid = np.full(2 * wl.grid_size ** 2, 7.0, np.float32)
    skipped = ct.c_size_t(12345)
    norm = ct.c_double(3.5)
    st = ct.c_int(0)
    fg.lib().oskar_grid_wproj_f_status(
        change.get("num_w_planes", wl.num_w_planes), support,
        change.get("oversample", wl.oversample),
        change.get("conv_size_half", wl.conv_size_half), wl.conv_func,
        wl.num_points, wl.uu, wl.vv, wl.ww, wl.vis, None, wl.cell_size_rad,
        wl.w_scale, G, ct.byref(skipped), ct.byref(norm), grid,
        ct.byref(st))
    assert st.value == -3
    assert fragment in fg.lib().gw_last_error_detail().decode()
    assert skipped.value == 12345 and norm.value == 3.5
    assert np.all(grid == 7.0)


def test_status_in_is_honoured():
    """OSKAR convention: a call with *status != 0 on entry does nothing."""
    _ensure_built()
    wl = _wl()
    grid = np.zeros(2 * wl.grid_size ** 2, np.float32)
    skipped = ct.c_size_t(99)
    norm = ct.c_double(1.0)
    st = ct.c_int(-7)
    fg.lib().oskar_grid_wproj_f_status(
        wl.num_w_planes, wl.support, wl.oversample, wl.conv_size_half,
        wl.conv_func, wl.num_points, wl.uu, wl.vv, wl.ww, wl.vis, None,
        wl.cell_size_rad, wl.w_scale, wl.grid_size, ct.byref(skipped),
        ct.byref(norm), grid, ct.byref(st))
    assert st.value == -7 and skipped.value == 99 and norm.value == 1.0


def test_void_entry_point_records_last_error():
    _ensure_built()
    wl = _wl()
    grid = np.zeros(2 * wl.grid_size ** 2, np.float32)
    skipped = ct.c_size_t(0)
    norm = ct.c_double(0.0)
    fg.lib().oskar_grid_wproj_f(
        0, wl.support, wl.oversample, wl.conv_size_half, wl.conv_func,
        wl.num_points, wl.uu, wl.vv, wl.ww, wl.vis, None, wl.cell_size_rad,
        wl.w_scale, wl.grid_size, ct.byref(skipped), ct.byref(norm), grid)
    assert fg.lib().oskar_grid_wproj_last_error() == -3


def test_no_device_reports_compute_device_error():
    """Without a GPU the product path fails loudly (no CPU fallback)."""
    _ensure_built()
    if fg.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(fg.GridderError) as e:
        fg.grid_workload_f(_wl())
    assert e.value.status == -19


@pytest.mark.parametrize("precision", [0, 6, -1])
def test_plan_rejects_unknown_precision(precision):
    """gw_plan_create checks the precision code before any GPU call."""
    _ensure_built()
    wl = _wl()
    with pytest.raises(fg.GridderError) as e:
        fg.Plan.for_workload(wl, precision=precision)
    assert e.value.status == -3
    assert "precision" in fg.lib().gw_last_error_detail().decode()


def test_storage_rounding_is_nearest_even():
    """The 16-bit storage conversions the reduced-precision plans use."""
    a = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -2.5, 65504.0,
                  1e-8], np.float32)
    bf = fg.from_storage(fg.to_storage(a, fg.GW_BFLOAT16), fg.GW_BFLOAT16)
    assert list(bf[:4]) == [1.0, 1.0, 1.0 + 2.0 ** -6, -2.5]
    h = fg.from_storage(fg.to_storage(a, fg.GW_HALF), fg.GW_HALF)
    assert list(h[:5]) == [1.0, 1.00390625, 1.01171875, -2.5, 65504.0]
    assert h[5] == 0.0
