"""CPU oracle for the W-projection gridder -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this package, and only as the checker or the timed CPU
baseline.  The product library never imports, loads or calls it.

Wraps two shared libraries with ctypes:

* ``oracle/libgridref.so`` -- the C restatement in ``oracle/gridref.c`` of
  ``oskar_grid_wproj_f`` / ``_d`` / ``_f_diagnostics``
  (oskar_grid_wproj.c:39-370) and of the uniform-weighting functions
  (utils/oskar_grid_weights.c:38-172), looping over the full ``num_points``.
* ``oracle/_ref/libref_gridder.so`` -- a plain gcc build of the reference's
  own ``oskar_grid_wproj.c`` and ``utils/oskar_grid_weights.c`` (made by
  ``make -C oracle ref`` where the reference tree is mounted).  Its
  ``oskar_grid_wproj_f`` grids only the first 10000 points and prints a
  progress line every 1000 points (oskar_grid_wproj.c:154-157).

Parity pins of the restatement (see tests/test_oracle.py): bit-exact against
the reference build (grid bits, norm bits, num_skipped) where that library is
present; against the committed golden fixtures made from it
(tests/golden/); and against closed forms and invariants (all-ones taps give
footprint counts, linearity in vis, conjugation under w -> -w, norm
independent of vis, S = 0 nearest-cell histogram, oversample 1, edge skips,
round-half-away ties).
"""

from __future__ import annotations

import ctypes as ct
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgridref.so")
REF_PATH = os.path.join(HERE, "_ref", "libref_gridder.so")

_lib = None
_ref = None

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_sz = ct.c_size_t
_szp = ct.POINTER(ct.c_size_t)
_dp = ct.POINTER(ct.c_double)


def _wproj_argtypes(fp, real):
    return [_sz, _i32p, ct.c_int, ct.c_int, fp, _sz, fp, fp, fp, fp,
            ct.c_void_p, real, real, ct.c_int, _szp, _dp, fp]


def _bind_gridder(lib, prefix, wprefix):
    f = getattr(lib, prefix + "oskar_grid_wproj_f")
    f.argtypes = _wproj_argtypes(_f32p, ct.c_float)
    f.restype = None
    d = getattr(lib, prefix + "oskar_grid_wproj_d")
    d.argtypes = _wproj_argtypes(_f64p, ct.c_double)
    d.restype = None
    g = getattr(lib, prefix + "oskar_grid_wproj_f_diagnostics")
    g.argtypes = _wproj_argtypes(_f32p, ct.c_float) + [
        _i64p, _i64p, _i64p, _f32p, ct.c_void_p]
    g.restype = None
    for suffix, fp, real in (("f", _f32p, ct.c_float),
                             ("d", _f64p, ct.c_double)):
        w = getattr(lib, wprefix + "grid_weights_write_" + suffix)
        w.argtypes = [_sz, fp, fp, fp, real, ct.c_int, _szp, fp]
        w.restype = None
        r = getattr(lib, wprefix + "grid_weights_read_" + suffix)
        r.argtypes = [_sz, fp, fp, fp, fp, real, ct.c_int, _szp, fp]
        r.restype = None


def build():
    """Compile libgridref.so (and the reference build if the tree exists)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if os.path.isdir("/root/reference"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.run(["make", "-s", "-C", HERE], check=True)
        _lib = ct.CDLL(LIB_PATH, mode=ct.RTLD_LOCAL)
        _bind_gridder(_lib, "gridref_", "gridref_")
        _lib.gridref_locate_f.argtypes = [
            _sz, _i32p, ct.c_int, _sz, _f32p, _f32p, _f32p, ct.c_float,
            ct.c_float, ct.c_int, ct.c_void_p, ct.c_void_p]
        _lib.gridref_locate_f.restype = ct.c_ulonglong
        _lib.gridref_locate_d.argtypes = [
            _sz, _i32p, ct.c_int, _sz, _f64p, _f64p, _f64p, ct.c_double,
            ct.c_double, ct.c_int, ct.c_void_p, ct.c_void_p]
        _lib.gridref_locate_d.restype = ct.c_ulonglong
        _lib.gridref_cells_f.argtypes = [
            _sz, _i32p, ct.c_int, ct.c_int, _f32p, _sz, _f32p, _f32p, _f32p,
            _f32p, ct.c_float, ct.c_float, ct.c_int, _sz, _i64p, _f64p]
        _lib.gridref_cells_f.restype = None
        _lib.gridref_cells_d.argtypes = [
            _sz, _i32p, ct.c_int, ct.c_int, _f64p, _sz, _f64p, _f64p, _f64p,
            _f64p, ct.c_double, ct.c_double, ct.c_int, _sz, _i64p, _f64p]
        _lib.gridref_cells_d.restype = None
        _lib.gridref_oskar_grid_wproj_f_exact.argtypes = [
            _sz, _i32p, ct.c_int, ct.c_int, _f32p, _sz, _f32p, _f32p, _f32p,
            _f32p, ct.c_float, ct.c_float, ct.c_int, _szp, _dp, _f64p]
        _lib.gridref_oskar_grid_wproj_f_exact.restype = None
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    """The gcc build of the reference's own sources (None if absent)."""
    global _ref
    if _ref is None and ref_available():
        _ref = ct.CDLL(REF_PATH, mode=ct.RTLD_LOCAL)
        _bind_gridder(_ref, "", "oskar_")
    return _ref


# --------------------------------------------------------------------------
# numpy-level wrappers (same argument meaning as the reference C API)
# --------------------------------------------------------------------------

def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def _call_wproj(fn, dt, real, num_w_planes, support, oversample,
                conv_size_half, conv_func, uu, vv, ww, vis, cell_size_rad,
                w_scale, grid_size, grid, norm, num_points=None, extra=()):
    uu, vv, ww = _c(uu, dt), _c(vv, dt), _c(ww, dt)
    vis = _c(vis, dt)
    n = uu.shape[0] if num_points is None else num_points
    skipped = ct.c_size_t(0)
    nrm = ct.c_double(norm)
    fn(num_w_planes, _c(support, np.int32), oversample, conv_size_half,
       _c(conv_func, dt), n, uu, vv, ww, vis, None, real(cell_size_rad),
       real(w_scale), grid_size, ct.byref(skipped), ct.byref(nrm), grid,
       *extra)
    return int(skipped.value), float(nrm.value)


def grid_wproj_f(wl, grid=None, norm=0.0, lib_=None, num_points=None):
    """Run the FP32 gridder on a synth.Workload; returns (grid, skipped, norm).

    ``grid`` (float32 [2*G*G]) is accumulated in place when given."""
    g = wl.grid_size
    if grid is None:
        grid = np.zeros(2 * g * g, dtype=np.float32)
    L = lib_ if lib_ is not None else lib()
    fn = L.gridref_oskar_grid_wproj_f if L is lib() else L.oskar_grid_wproj_f
    sk, nrm = _call_wproj(fn, np.float32, ct.c_float, wl.num_w_planes,
                          wl.support, wl.oversample, wl.conv_size_half,
                          wl.conv_func, wl.uu, wl.vv, wl.ww, wl.vis,
                          wl.cell_size_rad, wl.w_scale, g, grid, norm,
                          num_points)
    return grid, sk, nrm


def grid_wproj_d(wl, grid=None, norm=0.0, lib_=None, uvw64=None):
    """FP64 oracle on a workload (inputs widened from float32 unless
    ``uvw64=(uu, vv, ww)`` gives float64 coordinates)."""
    g = wl.grid_size
    if grid is None:
        grid = np.zeros(2 * g * g, dtype=np.float64)
    L = lib_ if lib_ is not None else lib()
    fn = L.gridref_oskar_grid_wproj_d if L is lib() else L.oskar_grid_wproj_d
    uu, vv, ww = uvw64 if uvw64 is not None else (wl.uu, wl.vv, wl.ww)
    sk, nrm = _call_wproj(fn, np.float64, ct.c_double, wl.num_w_planes,
                          wl.support, wl.oversample, wl.conv_size_half,
                          wl.conv_func.astype(np.float64), uu, vv, ww,
                          wl.vis.astype(np.float64),
                          float(np.float32(wl.cell_size_rad)),
                          float(np.float32(wl.w_scale)), g, grid, norm)
    return grid, sk, nrm


def grid_wproj_f_exact(wl):
    """FP32 index math with float64 products and sums: the exact value of
    the FP32 gridder's result.  Returns (grid float64, skipped, norm)."""
    g = wl.grid_size
    grid = np.zeros(2 * g * g, dtype=np.float64)
    skipped = ct.c_size_t(0)
    nrm = ct.c_double(0.0)
    lib().gridref_oskar_grid_wproj_f_exact(
        wl.num_w_planes, _c(wl.support, np.int32), wl.oversample,
        wl.conv_size_half, _c(wl.conv_func, np.float32), wl.num_points,
        _c(wl.uu, np.float32), _c(wl.vv, np.float32), _c(wl.ww, np.float32),
        _c(wl.vis, np.float32), wl.cell_size_rad, wl.w_scale, g,
        ct.byref(skipped), ct.byref(nrm), grid)
    return grid, int(skipped.value), float(nrm.value)


def grid_wproj_f_diagnostics(wl, max_support=None, max_plane=None,
                             lib_=None):
    """Diagnostics variant; returns (grid, skipped, norm, hists...)."""
    g = wl.grid_size
    grid = np.zeros(2 * g * g, dtype=np.float32)
    ms = int(wl.support.max()) + 1 if max_support is None else max_support
    mp = wl.num_w_planes if max_plane is None else max_plane
    per_s = np.zeros(ms, np.int64)
    per_w = np.zeros(mp, np.int64)
    per_cell = np.zeros(g * g, np.int64)
    avg_s = np.zeros(g * g, np.float32)
    L = lib_ if lib_ is not None else lib()
    fn = (L.gridref_oskar_grid_wproj_f_diagnostics if L is lib()
          else L.oskar_grid_wproj_f_diagnostics)
    sk, nrm = _call_wproj(fn, np.float32, ct.c_float, wl.num_w_planes,
                          wl.support, wl.oversample, wl.conv_size_half,
                          wl.conv_func, wl.uu, wl.vv, wl.ww, wl.vis,
                          wl.cell_size_rad, wl.w_scale, g, grid, 0.0,
                          extra=(per_s, per_w, per_cell, avg_s, None))
    return grid, sk, nrm, per_s, per_w, per_cell, avg_s


def locate_f(wl):
    """Per-point index stage: dict of arrays + U (the update count)."""
    n = wl.num_points
    out = np.zeros((n, 8), np.int32)
    gw = np.zeros(n, np.uint64)
    u = lib().gridref_locate_f(
        wl.num_w_planes, _c(wl.support, np.int32), wl.oversample, n,
        _c(wl.uu, np.float32), _c(wl.vv, np.float32), _c(wl.ww, np.float32),
        wl.cell_size_rad, wl.w_scale, wl.grid_size,
        out.ctypes.data_as(ct.c_void_p), gw.ctypes.data_as(ct.c_void_p))
    names = ["grid_u", "grid_v", "off_u", "off_v", "plane", "support",
             "skip", "conj"]
    rec = {k: out[:, i].copy() for i, k in enumerate(names)}
    rec["grid_w"] = gw
    return rec, int(u)


def locate_d(wl):
    """Per-point FP64 index stage (oskar_grid_wproj.c:71-101) on the
    float32 coordinates widened to double, with the cell size and w scale
    as float32 values: dict of arrays + U."""
    n = wl.num_points
    out = np.zeros((n, 8), np.int32)
    gw = np.zeros(n, np.uint64)
    u = lib().gridref_locate_d(
        wl.num_w_planes, _c(wl.support, np.int32), wl.oversample, n,
        _c(wl.uu, np.float64), _c(wl.vv, np.float64), _c(wl.ww, np.float64),
        float(np.float32(wl.cell_size_rad)), float(np.float32(wl.w_scale)),
        wl.grid_size, out.ctypes.data_as(ct.c_void_p),
        gw.ctypes.data_as(ct.c_void_p))
    names = ["grid_u", "grid_v", "off_u", "off_v", "plane", "support",
             "skip", "conj"]
    rec = {k: out[:, i].copy() for i, k in enumerate(names)}
    rec["grid_w"] = gw
    return rec, int(u)


def cells_d(wl, cells):
    """float64 values of the given grid cells (v*G + u) from all points with
    the FP64 oracle's index math and arithmetic (inputs widened from the
    workload's float32 arrays, cell size and w scale as float32 values)."""
    cells = _c(cells, np.int64)
    out = np.zeros(2 * cells.shape[0], np.float64)
    lib().gridref_cells_d(
        wl.num_w_planes, _c(wl.support, np.int32), wl.oversample,
        wl.conv_size_half, _c(wl.conv_func, np.float64), wl.num_points,
        _c(wl.uu, np.float64), _c(wl.vv, np.float64), _c(wl.ww, np.float64),
        _c(wl.vis, np.float64), float(np.float32(wl.cell_size_rad)),
        float(np.float32(wl.w_scale)), wl.grid_size, cells.shape[0], cells,
        out)
    return out


def cells_f(wl, cells):
    """float64 values of the given grid cells (v*G + u) from all points."""
    cells = _c(cells, np.int64)
    out = np.zeros(2 * cells.shape[0], np.float64)
    lib().gridref_cells_f(
        wl.num_w_planes, _c(wl.support, np.int32), wl.oversample,
        wl.conv_size_half, _c(wl.conv_func, np.float32), wl.num_points,
        _c(wl.uu, np.float32), _c(wl.vv, np.float32), _c(wl.ww, np.float32),
        _c(wl.vis, np.float32), wl.cell_size_rad, wl.w_scale, wl.grid_size,
        cells.shape[0], cells, out)
    return out


def grid_weights_write_f(uu, vv, weight, cell_size_rad, grid_size, grid,
                         lib_=None):
    L = lib_ if lib_ is not None else lib()
    fn = (L.gridref_grid_weights_write_f if L is lib()
          else L.oskar_grid_weights_write_f)
    sk = ct.c_size_t(0)
    uu, vv, w = _c(uu, np.float32), _c(vv, np.float32), _c(weight, np.float32)
    fn(uu.shape[0], uu, vv, w, cell_size_rad, grid_size, ct.byref(sk), grid)
    return int(sk.value)


def grid_weights_write_d(uu, vv, weight, cell_size_rad, grid_size, grid):
    """gridref_grid_weights_write_d (utils/oskar_grid_weights.c:38-70):
    adds into ``grid`` (float64); returns num_skipped."""
    sk = ct.c_size_t(0)
    uu, vv, w = _c(uu, np.float64), _c(vv, np.float64), _c(weight, np.float64)
    lib().gridref_grid_weights_write_d(uu.shape[0], uu, vv, w, cell_size_rad,
                                       grid_size, ct.byref(sk), grid)
    return int(sk.value)


def grid_weights_read_d(uu, vv, weight_in, weight_out, cell_size_rad,
                        grid_size, grid):
    """gridref_grid_weights_read_d (utils/oskar_grid_weights.c:72-104)."""
    sk = ct.c_size_t(0)
    uu, vv = _c(uu, np.float64), _c(vv, np.float64)
    lib().gridref_grid_weights_read_d(uu.shape[0], uu, vv,
                                      _c(weight_in, np.float64), weight_out,
                                      cell_size_rad, grid_size, ct.byref(sk),
                                      grid)
    return int(sk.value)


def grid_weights_read_f(uu, vv, weight_in, weight_out, cell_size_rad,
                        grid_size, grid, lib_=None):
    L = lib_ if lib_ is not None else lib()
    fn = (L.gridref_grid_weights_read_f if L is lib()
          else L.oskar_grid_weights_read_f)
    sk = ct.c_size_t(0)
    uu, vv = _c(uu, np.float32), _c(vv, np.float32)
    fn(uu.shape[0], uu, vv, _c(weight_in, np.float32), weight_out,
       cell_size_rad, grid_size, ct.byref(sk), grid)
    return int(sk.value)
