"""MI355X-native W-projection gridder (drop-in for oskar_grid_wproj_f/_d).

The product is the C-ABI shared library ``libfpga_gridding_amd.so`` (HIP
kernels for gfx950 behind the reference's C signatures, see
``include/oskar_grid_wproj.h`` and ``include/fpga_gridding_amd.h``).  This
module binds it with ctypes so that Python tests and benchmarks call the
very same entry points a C caller links against.  There is no CPU fallback:
if the library is missing, or no GPU is present, the calls raise.
"""

from __future__ import annotations

import ctypes as ct
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfpga_gridding_amd.so")

GW_SINGLE = 0x01
GW_DOUBLE = 0x02
GW_HALF = 0x03          # FP16 storage, FP32 products, compensated grid
GW_BFLOAT16 = 0x04      # BF16 storage, FP32 products, compensated grid
GW_SINGLE_INDEX64 = 0x05  # GW_SINGLE data with FP64 index math

_lib = None

# Every symbol the public headers declare (checked by the CPU tests).
EXPORTED_SYMBOLS = (
    "oskar_grid_wproj_f", "oskar_grid_wproj_d",
    "oskar_grid_wproj_f_diagnostics",
    "oskar_grid_wproj_f_status", "oskar_grid_wproj_d_status",
    "oskar_grid_wproj_last_error", "gw_error_string", "gw_last_error_detail",
    "gw_release_cached",
    "gw_plan_create", "gw_plan_free", "gw_plan_set_stream", "gw_grid_device",
    "gw_grid_device_noalloc",
    "gw_plan_results", "gw_plan_records", "gw_plan_timings",
    "gw_plan_set_tuning",
    "gw_plan_get_tuning", "gw_plan_autotune",
    "gw_device_count", "gw_device_malloc", "gw_device_free",
    "gw_host_malloc", "gw_host_free",
    "gw_memcpy_h2d", "gw_memcpy_d2h", "gw_memset_device",
    "gw_device_synchronize", "gw_device_copy", "gw_compensated_resolve",
    "gw_batch_create", "gw_batch_free", "gw_batch_add", "gw_batch_clear",
    "gw_batch_set_stream", "gw_batch_run", "gw_batch_capture",
    "gw_batch_launch", "gw_batch_results", "gw_batch_last_ms",
    "gw_server_create", "gw_server_free", "gw_server_add_kernel_set",
    "gw_server_create_grid", "gw_server_read_grid", "gw_server_free_grid",
    "gw_server_submit", "gw_server_wait", "gw_server_get_stats",
    "oskar_grid_weights_write_f", "oskar_grid_weights_read_f",
    "oskar_grid_weights_write_d", "oskar_grid_weights_read_d",
    "gw_grid_weights_write_device", "gw_grid_weights_read_device",
    "gw_plan_create_device_cube", "gw_wkernel_generate",
    "gw_server_add_wkernel_set",
)


class GridderError(RuntimeError):
    """A non-zero status from the library (OSKAR error codes)."""

    def __init__(self, status: int, where: str):
        L = lib()
        msg = L.gw_error_string(status).decode()
        detail = L.gw_last_error_detail().decode()
        super().__init__(f"{where}: status {status} ({msg}) {detail}".strip())
        self.status = status


_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_sz = ct.c_size_t
_szp = ct.POINTER(ct.c_size_t)
_dp = ct.POINTER(ct.c_double)
_ip = ct.POINTER(ct.c_int)
_vp = ct.c_void_p


class RequestResult(ct.Structure):
    """gw_request_result."""
    _fields_ = [("status", ct.c_int), ("num_skipped", ct.c_size_t),
                ("norm_delta", ct.c_double),
                ("num_updates", ct.c_ulonglong),
                ("t_submit", ct.c_double), ("t_admit", ct.c_double),
                ("t_start", ct.c_double), ("t_done", ct.c_double)]


class ServerStats(ct.Structure):
    """gw_server_stats."""
    _fields_ = [(name, ct.c_size_t) for name in (
        "completed", "batches", "max_batch", "plans_evicted", "points",
        "device_bytes", "resident_plans", "pending")]


def _wproj_args(fp, real):
    return [_sz, _i32p, ct.c_int, ct.c_int, fp, _sz, fp, fp, fp, fp, _vp,
            real, real, ct.c_int, _szp, _dp, fp]


def lib():
    """Load the library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m "
            "fpga_gridding_amd.build` (there is no CPU fallback)")
    L = ct.CDLL(LIB_PATH, mode=ct.RTLD_LOCAL)
    L.oskar_grid_wproj_f_status.argtypes = _wproj_args(_f32p, ct.c_float) + [_ip]
    L.oskar_grid_wproj_f_status.restype = None
    L.oskar_grid_wproj_d_status.argtypes = _wproj_args(_f64p, ct.c_double) + [_ip]
    L.oskar_grid_wproj_d_status.restype = None
    L.oskar_grid_wproj_f.argtypes = _wproj_args(_f32p, ct.c_float)
    L.oskar_grid_wproj_f.restype = None
    L.oskar_grid_wproj_d.argtypes = _wproj_args(_f64p, ct.c_double)
    L.oskar_grid_wproj_d.restype = None
    L.oskar_grid_wproj_f_diagnostics.argtypes = _wproj_args(
        _f32p, ct.c_float) + [_i64p, _i64p, _i64p, _f32p, _vp]
    L.oskar_grid_wproj_f_diagnostics.restype = None
    L.oskar_grid_wproj_last_error.argtypes = []
    L.oskar_grid_wproj_last_error.restype = ct.c_int
    L.gw_error_string.argtypes = [ct.c_int]
    L.gw_error_string.restype = ct.c_char_p
    L.gw_last_error_detail.argtypes = []
    L.gw_last_error_detail.restype = ct.c_char_p
    L.gw_release_cached.argtypes = []
    L.gw_release_cached.restype = None
    L.gw_plan_create.argtypes = [ct.c_int, _sz, _i32p, ct.c_int, ct.c_int,
                                 _vp, ct.c_double, ct.c_double, ct.c_int, _ip]
    L.gw_plan_create.restype = _vp
    L.gw_plan_free.argtypes = [_vp]
    L.gw_plan_free.restype = None
    L.gw_plan_set_stream.argtypes = [_vp, _vp]
    L.gw_plan_set_stream.restype = None
    L.gw_grid_device.argtypes = [_vp, _sz, _vp, _vp, _vp, _vp, _vp, _ip]
    L.gw_grid_device.restype = None
    L.gw_grid_device_noalloc.argtypes = [_vp, _sz, _vp, _vp, _vp, _vp, _vp,
                                         _ip]
    L.gw_grid_device_noalloc.restype = None
    L.gw_plan_results.argtypes = [_vp, _szp, _dp,
                                  ct.POINTER(ct.c_ulonglong), _ip]
    L.gw_plan_results.restype = None
    L.gw_plan_records.argtypes = [_vp, ct.POINTER(ct.c_ulonglong),
                                  ct.POINTER(ct.c_ulonglong), _ip]
    L.gw_plan_records.restype = None
    L.gw_plan_timings.argtypes = [_vp, _dp, ct.c_int]
    L.gw_plan_timings.restype = ct.c_int
    L.gw_plan_set_tuning.argtypes = [_vp, ct.c_int, ct.c_int, ct.c_int, _ip]
    L.gw_plan_set_tuning.restype = None
    L.gw_plan_get_tuning.argtypes = [_vp, ct.POINTER(ct.c_int), ct.c_int]
    L.gw_plan_get_tuning.restype = ct.c_int
    L.gw_plan_autotune.argtypes = [_vp, _sz, _vp, _vp, _vp, _vp, _vp,
                                   ct.c_int, _dp, _ip]
    L.gw_plan_autotune.restype = None
    L.gw_device_count.argtypes = []
    L.gw_device_count.restype = ct.c_int
    L.gw_device_malloc.argtypes = [_sz, _ip]
    L.gw_device_malloc.restype = _vp
    L.gw_device_free.argtypes = [_vp]
    L.gw_device_free.restype = None
    L.gw_host_malloc.argtypes = [_sz, _ip]
    L.gw_host_malloc.restype = _vp
    L.gw_host_free.argtypes = [_vp]
    L.gw_host_free.restype = None
    for name in ("gw_memcpy_h2d", "gw_memcpy_d2h"):
        getattr(L, name).argtypes = [_vp, _vp, _sz, _ip]
        getattr(L, name).restype = None
    L.gw_memset_device.argtypes = [_vp, ct.c_int, _sz, _ip]
    L.gw_memset_device.restype = None
    L.gw_device_synchronize.argtypes = [_ip]
    L.gw_device_synchronize.restype = None
    L.gw_device_copy.argtypes = [_vp, _vp, _sz, _ip]
    L.gw_device_copy.restype = None
    L.gw_compensated_resolve.argtypes = [_vp, _sz, _vp, _vp, _ip]
    L.gw_compensated_resolve.restype = None
    L.gw_batch_create.argtypes = [_ip]
    L.gw_batch_create.restype = _vp
    L.gw_batch_free.argtypes = [_vp]
    L.gw_batch_free.restype = None
    L.gw_batch_add.argtypes = [_vp, _vp, _sz, _vp, _vp, _vp, _vp, _vp, _ip]
    L.gw_batch_add.restype = ct.c_int
    L.gw_batch_clear.argtypes = [_vp]
    L.gw_batch_clear.restype = None
    L.gw_batch_set_stream.argtypes = [_vp, _vp]
    L.gw_batch_set_stream.restype = None
    for name in ("gw_batch_run", "gw_batch_capture", "gw_batch_launch"):
        getattr(L, name).argtypes = [_vp, _ip]
        getattr(L, name).restype = None
    L.gw_batch_results.argtypes = [_vp, ct.c_int, _szp, _dp,
                                   ct.POINTER(ct.c_ulonglong), _ip]
    L.gw_batch_results.restype = None
    L.gw_batch_last_ms.argtypes = [_vp]
    L.gw_batch_last_ms.restype = ct.c_double
    L.gw_server_create.argtypes = [_sz, ct.c_int, ct.c_int, ct.c_int, _ip]
    L.gw_server_create.restype = _vp
    L.gw_server_free.argtypes = [_vp]
    L.gw_server_free.restype = None
    L.gw_server_add_kernel_set.argtypes = [_vp, _sz, _i32p, ct.c_int,
                                           ct.c_int, _f32p, _ip]
    L.gw_server_add_kernel_set.restype = ct.c_int
    L.gw_server_create_grid.argtypes = [_vp, ct.c_int, _ip]
    L.gw_server_create_grid.restype = ct.c_int
    L.gw_server_read_grid.argtypes = [_vp, ct.c_int, _f32p, _ip]
    L.gw_server_read_grid.restype = None
    L.gw_server_free_grid.argtypes = [_vp, ct.c_int]
    L.gw_server_free_grid.restype = None
    L.gw_server_submit.argtypes = [_vp, ct.c_int, ct.c_int, ct.c_double,
                                   ct.c_double, _sz, _vp, _vp, _vp, _vp,
                                   ct.c_int, _vp, _ip]
    L.gw_server_submit.restype = ct.c_longlong
    L.gw_server_wait.argtypes = [_vp, ct.c_longlong, ct.c_double,
                                 ct.POINTER(RequestResult), _ip]
    L.gw_server_wait.restype = ct.c_int
    L.gw_server_get_stats.argtypes = [_vp, ct.POINTER(ServerStats)]
    L.gw_server_get_stats.restype = None
    for sfx, fp, real in (("f", _f32p, ct.c_float), ("d", _f64p, ct.c_double)):
        w = getattr(L, "oskar_grid_weights_write_" + sfx)
        w.argtypes = [_sz, fp, fp, fp, real, ct.c_int, _szp, fp]
        w.restype = None
        r = getattr(L, "oskar_grid_weights_read_" + sfx)
        r.argtypes = [_sz, fp, fp, fp, fp, real, ct.c_int, _szp, fp]
        r.restype = None
    L.gw_grid_weights_write_device.argtypes = [
        ct.c_int, _sz, _vp, _vp, _vp, ct.c_double, ct.c_int, _szp, _vp, _vp,
        _ip]
    L.gw_grid_weights_write_device.restype = None
    L.gw_grid_weights_read_device.argtypes = [
        ct.c_int, _sz, _vp, _vp, _vp, _vp, ct.c_double, ct.c_int, _szp, _vp,
        _vp, _ip]
    L.gw_grid_weights_read_device.restype = None
    L.gw_plan_create_device_cube.argtypes = [ct.c_int, _sz, _i32p, ct.c_int,
                                             ct.c_int, _vp, ct.c_double,
                                             ct.c_double, ct.c_int, _ip]
    L.gw_plan_create_device_cube.restype = _vp
    L.gw_wkernel_generate.argtypes = [_sz, _i32p, ct.c_int, ct.c_int,
                                      ct.c_double, _vp, _vp, _ip]
    L.gw_wkernel_generate.restype = None
    L.gw_server_add_wkernel_set.argtypes = [_vp, _sz, _i32p, ct.c_int,
                                            ct.c_int, ct.c_double, _ip]
    L.gw_server_add_wkernel_set.restype = ct.c_int
    _lib = L
    return L


def _check(status: ct.c_int, where: str):
    if status.value != 0:
        raise GridderError(status.value, where)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def to_storage(a, precision):
    """float32 values as the 16-bit storage of a GW_HALF (numpy float16) or
    GW_BFLOAT16 (uint16 holding the top half of the FP32, rounded to
    nearest even) plan."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    if precision == GW_HALF:
        return a.astype(np.float16)
    if precision == GW_BFLOAT16:
        u = a.view(np.uint32).astype(np.uint64)
        nan = np.isnan(a)
        r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
        r[nan] = ((u[nan] >> 16) | 0x40).astype(np.uint16)
        return r
    raise ValueError("precision must be GW_HALF or GW_BFLOAT16")


def from_storage(a, precision):
    """Float32 values of 16-bit storage (inverse of to_storage)."""
    if precision == GW_HALF:
        return np.asarray(a, dtype=np.float16).astype(np.float32)
    if precision == GW_BFLOAT16:
        return (np.asarray(a, dtype=np.uint16).astype(np.uint32)
                << 16).view(np.float32)
    raise ValueError("precision must be GW_HALF or GW_BFLOAT16")


def resolve_compensated(d_comp, num_cells, d_out):
    """A compensated two-float grid (re_hi, re_lo, im_hi, im_lo per cell,
    device) to a float grid (re, im per cell, device); waits."""
    st = ct.c_int(0)
    lib().gw_compensated_resolve(d_comp.ptr, num_cells, d_out.ptr, None,
                                 ct.byref(st))
    _check(st, "gw_compensated_resolve")
    device_synchronize()


def grid_workload_reduced(wl, precision, comp=None):
    """Grid a synth.Workload with a GW_HALF / GW_BFLOAT16 plan: vis stored
    in 16 bits, float coordinates, compensated device grid.  cell_size_rad
    and w_scale are passed rounded to float (as the FP64 oracle receives
    them).  Returns (grid float32 [2G^2] = hi + lo, skipped, norm, the
    compensated grid float32 [4G^2]).  ``comp`` (float32 [4G^2]) is
    accumulated when given."""
    g = wl.grid_size
    plan = Plan(precision, wl.num_w_planes, wl.support, wl.oversample,
                wl.conv_size_half, wl.conv_func,
                float(np.float32(wl.cell_size_rad)),
                float(np.float32(wl.w_scale)), g)
    try:
        d = [DeviceArray.from_host(np.ascontiguousarray(a, np.float32))
             for a in (wl.uu, wl.vv, wl.ww)]
        d_vis = DeviceArray.from_host(to_storage(wl.vis, precision))
        d_comp = DeviceArray(16 * g * g)
        if comp is None:
            d_comp.zero()
        else:
            d_comp.upload(np.ascontiguousarray(comp, np.float32))
        plan.grid(wl.num_points, *d, d_vis, d_comp)
        sk, nd, _ = plan.results()
        d_out = DeviceArray(8 * g * g)
        resolve_compensated(d_comp, g * g, d_out)
        out = d_out.download(np.empty(2 * g * g, np.float32))
        comp_out = d_comp.download(np.empty(4 * g * g, np.float32))
        for x in (*d, d_vis, d_comp, d_out):
            x.close()
        return out, sk, nd, comp_out
    finally:
        plan.close()


def grid_workload_index64(wl):
    """Grid a synth.Workload with a GW_SINGLE_INDEX64 plan (FP32 data, FP64
    index math); cell_size_rad and w_scale passed rounded to float, as the
    FP64 oracle receives them.  Returns (grid float32 [2G^2], skipped,
    norm)."""
    g = wl.grid_size
    plan = Plan(GW_SINGLE_INDEX64, wl.num_w_planes, wl.support,
                wl.oversample, wl.conv_size_half, wl.conv_func,
                float(np.float32(wl.cell_size_rad)),
                float(np.float32(wl.w_scale)), g)
    try:
        d = [DeviceArray.from_host(np.ascontiguousarray(a, np.float32))
             for a in (wl.uu, wl.vv, wl.ww, wl.vis)]
        d_grid = DeviceArray(8 * g * g)
        d_grid.zero()
        plan.grid(wl.num_points, *d, d_grid)
        sk, nd, _ = plan.results()
        out = d_grid.download(np.empty(2 * g * g, np.float32))
        for x in (*d, d_grid):
            x.close()
        return out, sk, nd
    finally:
        plan.close()


# --------------------------------------------------------------------------
# Drop-in calls (host arrays), same argument meaning as the reference
# oskar_grid_wproj_f / _d (oskar_grid_wproj.h:50-66).
# --------------------------------------------------------------------------

def grid_wproj_f(num_w_planes, support, oversample, conv_size_half, conv_func,
                 uu, vv, ww, vis, weight, cell_size_rad, w_scale, grid_size,
                 grid, norm=0.0):
    """FP32 gridding into ``grid`` (float32 [2*G*G], accumulated in place).

    Returns ``(num_skipped, norm)`` with ``norm`` the updated value."""
    if grid.dtype != np.float32 or not grid.flags.c_contiguous:
        raise TypeError("grid must be a C-contiguous float32 array")
    if grid.size < 2 * grid_size * grid_size:
        raise ValueError("grid is smaller than 2*grid_size^2")
    uu, vv, ww = _c(uu, np.float32), _c(vv, np.float32), _c(ww, np.float32)
    vis = _c(vis, np.float32)
    n = uu.shape[0]
    if vv.shape[0] != n or ww.shape[0] != n or vis.shape[0] != 2 * n:
        raise ValueError("uu, vv, ww need N values and vis 2N")
    sk = ct.c_size_t(0)
    nrm = ct.c_double(norm)
    st = ct.c_int(0)
    lib().oskar_grid_wproj_f_status(
        num_w_planes, _c(support, np.int32), oversample, conv_size_half,
        _c(conv_func, np.float32), n, uu, vv, ww, vis, None,
        cell_size_rad, w_scale, grid_size, ct.byref(sk), ct.byref(nrm), grid,
        ct.byref(st))
    _check(st, "oskar_grid_wproj_f")
    return int(sk.value), float(nrm.value)


def grid_wproj_d(num_w_planes, support, oversample, conv_size_half, conv_func,
                 uu, vv, ww, vis, weight, cell_size_rad, w_scale, grid_size,
                 grid, norm=0.0):
    """FP64 gridding into ``grid`` (float64 [2*G*G], accumulated)."""
    if grid.dtype != np.float64 or not grid.flags.c_contiguous:
        raise TypeError("grid must be a C-contiguous float64 array")
    if grid.size < 2 * grid_size * grid_size:
        raise ValueError("grid is smaller than 2*grid_size^2")
    uu, vv, ww = _c(uu, np.float64), _c(vv, np.float64), _c(ww, np.float64)
    vis = _c(vis, np.float64)
    n = uu.shape[0]
    if vv.shape[0] != n or ww.shape[0] != n or vis.shape[0] != 2 * n:
        raise ValueError("uu, vv, ww need N values and vis 2N")
    sk = ct.c_size_t(0)
    nrm = ct.c_double(norm)
    st = ct.c_int(0)
    lib().oskar_grid_wproj_d_status(
        num_w_planes, _c(support, np.int32), oversample, conv_size_half,
        _c(conv_func, np.float64), n, uu, vv, ww, vis, None,
        cell_size_rad, w_scale, grid_size, ct.byref(sk), ct.byref(nrm), grid,
        ct.byref(st))
    _check(st, "oskar_grid_wproj_d")
    return int(sk.value), float(nrm.value)


def grid_wproj_f_diagnostics(num_w_planes, support, oversample,
                             conv_size_half, conv_func, uu, vv, ww, vis,
                             weight, cell_size_rad, w_scale, grid_size, grid,
                             per_support, per_plane, per_cell, avg_support,
                             norm=0.0):
    """oskar_grid_wproj_f_diagnostics (oskar_grid_wproj.c:245-370)."""
    uu, vv, ww = _c(uu, np.float32), _c(vv, np.float32), _c(ww, np.float32)
    vis = _c(vis, np.float32)
    n = uu.shape[0]
    sk = ct.c_size_t(0)
    nrm = ct.c_double(norm)
    L = lib()
    L.oskar_grid_wproj_f_diagnostics(
        num_w_planes, _c(support, np.int32), oversample, conv_size_half,
        _c(conv_func, np.float32), n, uu, vv, ww, vis, None,
        cell_size_rad, w_scale, grid_size, ct.byref(sk), ct.byref(nrm), grid,
        per_support, per_plane, per_cell, avg_support, None)
    st = L.oskar_grid_wproj_last_error()
    if st:
        raise GridderError(st, "oskar_grid_wproj_f_diagnostics")
    return int(sk.value), float(nrm.value)


def wkernel_generate_device(support, oversample, conv_size_half, w_max,
                            d_out=None):
    """gw_wkernel_generate into a DeviceArray (allocated if not given);
    returns it."""
    sup = _c(support, np.int32)
    n = 2 * sup.size * conv_size_half * conv_size_half
    if d_out is None:
        d_out = DeviceArray(4 * n)
    st = ct.c_int(0)
    lib().gw_wkernel_generate(sup.size, sup, oversample, conv_size_half,
                              float(w_max), d_out.ptr, None, ct.byref(st))
    _check(st, "gw_wkernel_generate")
    return d_out


def wkernel_generate(support, oversample, conv_size_half, w_max):
    """The device-generated W-kernel cube as a host float32 array
    [2 * P * C * C] (reference layout)."""
    sup = _c(support, np.int32)
    d = wkernel_generate_device(sup, oversample, conv_size_half, w_max)
    try:
        return d.download(np.empty(2 * sup.size * conv_size_half ** 2,
                                   np.float32))
    finally:
        d.close()


def grid_weights_write(uu, vv, weight, cell_size_rad, grid_size, grid,
                       double=False):
    """oskar_grid_weights_write_f / _d on host arrays: adds into ``grid``
    (float32 or float64 [grid_size^2], modified in place); returns
    num_skipped."""
    dt = np.float64 if double else np.float32
    if grid.dtype != dt or not grid.flags["C_CONTIGUOUS"]:
        raise ValueError("grid must be a contiguous array of the call's type")
    if grid.size != grid_size * grid_size:
        raise ValueError("grid must hold grid_size^2 values")
    L = lib()
    fn = L.oskar_grid_weights_write_d if double else \
        L.oskar_grid_weights_write_f
    sk = ct.c_size_t(0)
    uu, vv, w = _c(uu, dt), _c(vv, dt), _c(weight, dt)
    if not (uu.size == vv.size == w.size):
        raise ValueError("uu, vv and weight need the same length")
    L.oskar_grid_wproj_last_error()
    fn(uu.size, uu, vv, w, cell_size_rad, grid_size, ct.byref(sk), grid)
    st = L.oskar_grid_wproj_last_error()
    if st:
        raise GridderError(st, fn.__name__)
    return int(sk.value)


def grid_weights_read(uu, vv, weight_in, weight_out, cell_size_rad,
                      grid_size, grid, double=False):
    """oskar_grid_weights_read_f / _d on host arrays: fills ``weight_out``
    (modified in place, skipped points untouched); returns num_skipped."""
    dt = np.float64 if double else np.float32
    if weight_out.dtype != dt or not weight_out.flags["C_CONTIGUOUS"]:
        raise ValueError("weight_out must be a contiguous array of the "
                         "call's type")
    L = lib()
    fn = L.oskar_grid_weights_read_d if double else \
        L.oskar_grid_weights_read_f
    sk = ct.c_size_t(0)
    uu, vv, w = _c(uu, dt), _c(vv, dt), _c(weight_in, dt)
    if not (uu.size == vv.size == w.size == weight_out.size):
        raise ValueError("uu, vv, weight_in and weight_out need the same "
                         "length")
    g = _c(grid, dt)
    if g.size != grid_size * grid_size:
        raise ValueError("grid must hold grid_size^2 values")
    fn(uu.size, uu, vv, w, weight_out, cell_size_rad, grid_size,
       ct.byref(sk), g)
    st = L.oskar_grid_wproj_last_error()
    if st:
        raise GridderError(st, fn.__name__)
    return int(sk.value)


def grid_workload_f(wl, grid=None, norm=0.0):
    """Drop-in FP32 call on a synth.Workload; returns (grid, skipped, norm)."""
    g = wl.grid_size
    if grid is None:
        grid = np.zeros(2 * g * g, dtype=np.float32)
    sk, nrm = grid_wproj_f(wl.num_w_planes, wl.support, wl.oversample,
                           wl.conv_size_half, wl.conv_func, wl.uu, wl.vv,
                           wl.ww, wl.vis, wl.weight, wl.cell_size_rad,
                           wl.w_scale, g, grid, norm)
    return grid, sk, nrm


def grid_workload_d(wl, grid=None, norm=0.0):
    """Drop-in FP64 call on a synth.Workload (inputs widened)."""
    g = wl.grid_size
    if grid is None:
        grid = np.zeros(2 * g * g, dtype=np.float64)
    sk, nrm = grid_wproj_d(wl.num_w_planes, wl.support, wl.oversample,
                           wl.conv_size_half, wl.conv_func.astype(np.float64),
                           wl.uu, wl.vv, wl.ww, wl.vis.astype(np.float64),
                           None, float(np.float32(wl.cell_size_rad)),
                           float(np.float32(wl.w_scale)), g, grid, norm)
    return grid, sk, nrm


# --------------------------------------------------------------------------
# Device buffers and the resident plan
# --------------------------------------------------------------------------

class PinnedArray:
    """Page-locked host memory (gw_host_malloc) seen as a numpy array
    (``.array``); freed on close/GC.  Copies from it to the device run as
    DMA, e.g. the points of server requests."""

    def __init__(self, shape, dtype):
        dtype = np.dtype(dtype)
        count = int(np.prod(shape))
        st = ct.c_int(0)
        self.ptr = lib().gw_host_malloc(count * dtype.itemsize, ct.byref(st))
        _check(st, "gw_host_malloc")
        buf = (ct.c_char * max(1, count * dtype.itemsize)).from_address(
            self.ptr)
        self.array = np.frombuffer(buf, dtype=dtype, count=count).reshape(
            shape)

    @classmethod
    def copy_of(cls, a: np.ndarray):
        p = cls(a.shape, a.dtype)
        p.array[...] = a
        return p

    def close(self):
        if getattr(self, "ptr", None):
            self.array = None
            lib().gw_host_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceArray:
    """A device allocation owned by Python (freed on close/GC)."""

    def __init__(self, nbytes: int):
        st = ct.c_int(0)
        self.ptr = lib().gw_device_malloc(nbytes, ct.byref(st))
        _check(st, "gw_device_malloc")
        self.nbytes = nbytes

    @classmethod
    def from_host(cls, a: np.ndarray):
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        if a.nbytes > self.nbytes:
            raise ValueError("host array larger than the device buffer")
        st = ct.c_int(0)
        lib().gw_memcpy_h2d(self.ptr, a.ctypes.data, a.nbytes, ct.byref(st))
        _check(st, "gw_memcpy_h2d")

    def download(self, out: np.ndarray):
        if out.nbytes > self.nbytes or not out.flags.c_contiguous:
            raise ValueError("bad output array")
        st = ct.c_int(0)
        lib().gw_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes,
                            ct.byref(st))
        _check(st, "gw_memcpy_d2h")
        return out

    def copy_from(self, src: "DeviceArray", nbytes: int = None):
        """Device-to-device copy of nbytes (default: all of src)."""
        n = src.nbytes if nbytes is None else nbytes
        if n > self.nbytes or n > src.nbytes:
            raise ValueError("copy larger than a buffer")
        st = ct.c_int(0)
        lib().gw_device_copy(self.ptr, src.ptr, n, ct.byref(st))
        _check(st, "gw_device_copy")

    def zero(self):
        st = ct.c_int(0)
        lib().gw_memset_device(self.ptr, 0, self.nbytes, ct.byref(st))
        _check(st, "gw_memset_device")

    def close(self):
        if getattr(self, "ptr", None):
            lib().gw_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_synchronize():
    st = ct.c_int(0)
    lib().gw_device_synchronize(ct.byref(st))
    _check(st, "gw_device_synchronize")


def device_count() -> int:
    return lib().gw_device_count()


class Plan:
    """Resident plan: W-kernel cube and tables on the device; grids
    device-resident points into a device-resident grid."""

    def __init__(self, precision, num_w_planes, support, oversample,
                 conv_size_half, conv_func, cell_size_rad, w_scale,
                 grid_size):
        dt = np.float64 if precision == GW_DOUBLE else np.float32
        cube = _c(conv_func, dt)
        st = ct.c_int(0)
        self._cube = cube
        self.precision = precision
        self.grid_size = grid_size
        self.ptr = lib().gw_plan_create(
            precision, num_w_planes, _c(support, np.int32), oversample,
            conv_size_half, cube.ctypes.data, cell_size_rad, w_scale,
            grid_size, ct.byref(st))
        _check(st, "gw_plan_create")

    @classmethod
    def from_device_cube(cls, precision, num_w_planes, support, oversample,
                         conv_size_half, d_conv_func, cell_size_rad, w_scale,
                         grid_size):
        """gw_plan_create_device_cube: tables built from a cube already in
        device memory (a DeviceArray)."""
        self = cls.__new__(cls)
        st = ct.c_int(0)
        self.ptr = lib().gw_plan_create_device_cube(
            precision, num_w_planes, _c(support, np.int32), oversample,
            conv_size_half, _ptr(d_conv_func), cell_size_rad, w_scale,
            grid_size, ct.byref(st))
        _check(st, "gw_plan_create_device_cube")
        self.precision = precision
        self.grid_size = grid_size
        return self

    @classmethod
    def for_workload(cls, wl, precision=GW_SINGLE):
        return cls(precision, wl.num_w_planes, wl.support, wl.oversample,
                   wl.conv_size_half, wl.conv_func, wl.cell_size_rad,
                   wl.w_scale, wl.grid_size)

    def grid(self, n, d_uu, d_vv, d_ww, d_vis, d_grid):
        st = ct.c_int(0)
        lib().gw_grid_device(self.ptr, n, _ptr(d_uu), _ptr(d_vv), _ptr(d_ww),
                             _ptr(d_vis), _ptr(d_grid), ct.byref(st))
        _check(st, "gw_grid_device")

    def grid_noalloc(self, n, d_uu, d_vv, d_ww, d_vis, d_grid):
        """gw_grid_device_noalloc: as grid(), with the per-call workspace in
        the library's static device scratch (nothing allocated)."""
        st = ct.c_int(0)
        lib().gw_grid_device_noalloc(self.ptr, n, _ptr(d_uu), _ptr(d_vv),
                                     _ptr(d_ww), _ptr(d_vis), _ptr(d_grid),
                                     ct.byref(st))
        _check(st, "gw_grid_device_noalloc")

    def results(self):
        st = ct.c_int(0)
        sk = ct.c_size_t(0)
        nd = ct.c_double(0)
        nu = ct.c_ulonglong(0)
        lib().gw_plan_results(self.ptr, ct.byref(sk), ct.byref(nd),
                              ct.byref(nu), ct.byref(st))
        _check(st, "gw_plan_results")
        return int(sk.value), float(nd.value), int(nu.value)

    def records(self):
        """(records, record_updates) the last grid() call's scatter took:
        one record per run of points with equal origin, plane and sub-cell
        table, and the taps it applied."""
        st = ct.c_int(0)
        nr = ct.c_ulonglong(0)
        nu = ct.c_ulonglong(0)
        lib().gw_plan_records(self.ptr, ct.byref(nr), ct.byref(nu),
                              ct.byref(st))
        _check(st, "gw_plan_records")
        return int(nr.value), int(nu.value)

    def timings(self):
        buf = (ct.c_double * 8)()
        n = lib().gw_plan_timings(self.ptr, buf, 8)
        return list(buf[:n])

    def set_tuning(self, tile_log2=-1, slot_cap=0, min_share=0):
        st = ct.c_int(0)
        lib().gw_plan_set_tuning(self.ptr, tile_log2, slot_cap, min_share,
                                 ct.byref(st))
        _check(st, "gw_plan_set_tuning")

    def tuning(self):
        """dict(tile_log2 (-1 for a side that is not a power of two),
        slot_cap, min_share, region, max_tile_log2, tile, max_tile)."""
        buf = (ct.c_int * 7)()
        n = lib().gw_plan_get_tuning(self.ptr, buf, 7)
        keys = ("tile_log2", "slot_cap", "min_share", "region",
                "max_tile_log2", "tile", "max_tile")
        return dict(zip(keys[:n], list(buf[:n])))

    def autotune(self, n, d_uu, d_vv, d_ww, d_vis, d_scratch, reps=2):
        """Keeps the fastest scatter shape for these points; returns its
        device ms per call."""
        st = ct.c_int(0)
        ms = ct.c_double(0)
        lib().gw_plan_autotune(self.ptr, n, _ptr(d_uu), _ptr(d_vv),
                               _ptr(d_ww), _ptr(d_vis), _ptr(d_scratch),
                               reps, ct.byref(ms), ct.byref(st))
        _check(st, "gw_plan_autotune")
        return ms.value

    def close(self):
        if getattr(self, "ptr", None):
            lib().gw_plan_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many FP32 jobs gridded by one pipeline with one persistent scatter
    launch; capture() records it as a HIP graph, launch() replays it."""

    def __init__(self):
        st = ct.c_int(0)
        self.ptr = lib().gw_batch_create(ct.byref(st))
        _check(st, "gw_batch_create")
        self._keep = []

    def add(self, plan, n, d_uu, d_vv, d_ww, d_vis, d_grid):
        st = ct.c_int(0)
        j = lib().gw_batch_add(self.ptr, plan.ptr, n, _ptr(d_uu), _ptr(d_vv),
                               _ptr(d_ww), _ptr(d_vis), _ptr(d_grid),
                               ct.byref(st))
        _check(st, "gw_batch_add")
        self._keep.append((plan, d_uu, d_vv, d_ww, d_vis, d_grid))
        return j

    def clear(self):
        lib().gw_batch_clear(self.ptr)
        self._keep = []

    def set_stream(self, hip_stream):
        """Run on a caller's HIP stream (an integer handle; work already
        queued on the previous stream is waited for first)."""
        lib().gw_batch_set_stream(self.ptr, ct.c_void_p(hip_stream))

    def _call(self, name):
        st = ct.c_int(0)
        getattr(lib(), name)(self.ptr, ct.byref(st))
        _check(st, name)

    def run(self):
        self._call("gw_batch_run")

    def capture(self):
        self._call("gw_batch_capture")

    def launch(self):
        self._call("gw_batch_launch")

    def results(self, job):
        st = ct.c_int(0)
        sk = ct.c_size_t(0)
        nd = ct.c_double(0)
        nu = ct.c_ulonglong(0)
        lib().gw_batch_results(self.ptr, job, ct.byref(sk), ct.byref(nd),
                               ct.byref(nu), ct.byref(st))
        _check(st, "gw_batch_results")
        return int(sk.value), float(nd.value), int(nu.value)

    def last_ms(self):
        return lib().gw_batch_last_ms(self.ptr)

    def close(self):
        if getattr(self, "ptr", None):
            lib().gw_batch_free(self.ptr)
            self.ptr = None
            self._keep = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Request loop on the current device (gw_server_*): requests are the
    arguments of oskar_grid_wproj_f on host arrays; admitted by the device
    memory budget and gridded in batches on several streams."""

    def __init__(self, hbm_budget, max_batch_jobs=64, num_streams=2,
                 precision=None):
        st = ct.c_int(0)
        prec = GW_SINGLE if precision is None else precision
        self.ptr = lib().gw_server_create(int(hbm_budget), max_batch_jobs,
                                          num_streams, prec, ct.byref(st))
        _check(st, "gw_server_create")
        self._keep = {}

    def add_kernel_set(self, num_w_planes, support, oversample,
                       conv_size_half, conv_func):
        st = ct.c_int(0)
        k = lib().gw_server_add_kernel_set(
            self.ptr, num_w_planes, _c(support, np.int32), oversample,
            conv_size_half, _c(conv_func, np.float32), ct.byref(st))
        _check(st, "gw_server_add_kernel_set")
        return k

    def add_wkernel_set(self, support, oversample, conv_size_half, w_max):
        """A kernel set generated on the device by its parameters."""
        st = ct.c_int(0)
        sup = _c(support, np.int32)
        k = lib().gw_server_add_wkernel_set(self.ptr, sup.size, sup,
                                            oversample, conv_size_half,
                                            float(w_max), ct.byref(st))
        _check(st, "gw_server_add_wkernel_set")
        return k

    def create_grid(self, grid_size):
        st = ct.c_int(0)
        g = lib().gw_server_create_grid(self.ptr, grid_size, ct.byref(st))
        _check(st, "gw_server_create_grid")
        return g

    def read_grid(self, grid_id, grid_size):
        out = np.empty(2 * grid_size * grid_size, np.float32)
        st = ct.c_int(0)
        lib().gw_server_read_grid(self.ptr, grid_id, out, ct.byref(st))
        _check(st, "gw_server_read_grid")
        return out

    def free_grid(self, grid_id):
        lib().gw_server_free_grid(self.ptr, grid_id)

    def submit(self, kernel_set, grid_size, cell_size_rad, w_scale, uu, vv,
               ww, vis, grid_id=-1):
        """Queues a request; returns its id.  Without a resident grid the
        result grid is returned by wait()."""
        arrs = [_c(uu, np.float32), _c(vv, np.float32), _c(ww, np.float32),
                _c(vis, np.float32)]
        n = arrs[0].size
        if not (arrs[1].size == arrs[2].size == n and arrs[3].size == 2 * n):
            raise ValueError("uu, vv, ww need n values and vis 2n")
        out = None
        if grid_id < 0:
            out = np.empty(2 * grid_size * grid_size, np.float32)
        st = ct.c_int(0)
        rid = lib().gw_server_submit(
            self.ptr, kernel_set, grid_size, float(cell_size_rad),
            float(w_scale), n, *(a.ctypes.data for a in arrs), grid_id,
            out.ctypes.data if out is not None else None, ct.byref(st))
        _check(st, "gw_server_submit")
        self._keep[rid] = (arrs, out)
        return rid

    def wait(self, rid, timeout=-1.0):
        """Returns (RequestResult, grid or None); None on timeout."""
        st = ct.c_int(0)
        res = RequestResult()
        done = lib().gw_server_wait(self.ptr, rid, float(timeout),
                                    ct.byref(res), ct.byref(st))
        _check(st, "gw_server_wait")
        if not done:
            return None
        _, out = self._keep.pop(rid)
        if res.status != 0:
            raise GridderError(res.status, f"request {rid}")
        return res, out

    def stats(self):
        s = ServerStats()
        lib().gw_server_get_stats(self.ptr, ct.byref(s))
        return {name: int(getattr(s, name)) for name, _ in s._fields_}

    def close(self):
        if getattr(self, "ptr", None):
            lib().gw_server_free(self.ptr)
            self.ptr = None
            self._keep = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _ptr(x):
    if isinstance(x, DeviceArray):
        return x.ptr
    return x
