"""Python side of the library's host I/O (include/fpga_gridding_amd_io.h):
OSKAR binary visibility files, FITS kernel cubes in, FITS grid cubes out."""
from __future__ import annotations

import ctypes as ct

import numpy as np

from . import lib as _lib

OSKAR_INT, OSKAR_SINGLE, OSKAR_DOUBLE = 0x02, 0x04, 0x08
OSKAR_COMPLEX, OSKAR_CHAR = 0x20, 0x01
VIS_GROUP_HEADER, VIS_GROUP_BLOCK = 11, 12

IO_SYMBOLS = (
    "gw_binary_open", "gw_binary_free", "gw_binary_num_tags",
    "gw_binary_set_query_search_start",
    "gw_binary_query", "gw_binary_read", "gw_binary_read_ext",
    "gw_binary_read_int", "gw_binary_read_double", "gw_vis_read_coords",
    "gw_vis_scale_coords", "gw_read_kernel", "gw_free",
    "gw_write_fits_cube",
)

_vp, _sz, _ip = ct.c_void_p, ct.c_size_t, ct.POINTER(ct.c_int)
_done = False


class IOError_(RuntimeError):
    """A non-zero OSKAR status from an I/O call."""

    def __init__(self, status, where):
        super().__init__(f"{where}: status {status}")
        self.status = status


def io_lib():
    global _done
    L = _lib()
    if not _done:
        L.gw_binary_open.argtypes = [ct.c_char_p, _ip]
        L.gw_binary_open.restype = _vp
        L.gw_binary_free.argtypes = [_vp]
        L.gw_binary_free.restype = None
        L.gw_binary_num_tags.argtypes = [_vp]
        L.gw_binary_num_tags.restype = ct.c_int
        L.gw_binary_set_query_search_start.argtypes = [_vp, ct.c_int, _ip]
        L.gw_binary_set_query_search_start.restype = None
        L.gw_binary_query.argtypes = [_vp, ct.c_ubyte, ct.c_ubyte, ct.c_ubyte,
                                      ct.c_int, ct.POINTER(ct.c_size_t), _ip]
        L.gw_binary_query.restype = ct.c_int
        L.gw_binary_read.argtypes = [_vp, ct.c_ubyte, ct.c_ubyte, ct.c_ubyte,
                                     ct.c_int, _sz, _vp, _ip]
        L.gw_binary_read.restype = None
        L.gw_binary_read_ext.argtypes = [_vp, ct.c_ubyte, ct.c_char_p,
                                         ct.c_char_p, ct.c_int, _sz, _vp, _ip]
        L.gw_binary_read_ext.restype = None
        L.gw_binary_read_int.argtypes = [_vp, ct.c_ubyte, ct.c_ubyte,
                                         ct.c_int, _ip, _ip]
        L.gw_binary_read_int.restype = None
        L.gw_binary_read_double.argtypes = [_vp, ct.c_ubyte, ct.c_ubyte,
                                            ct.c_int,
                                            ct.POINTER(ct.c_double), _ip]
        L.gw_binary_read_double.restype = None
        L.gw_vis_read_coords.argtypes = [_vp, ct.c_int, ct.c_int, _sz, _vp,
                                         _vp, _vp, _ip]
        L.gw_vis_read_coords.restype = None
        L.gw_vis_scale_coords.argtypes = [ct.c_double, ct.c_int, _sz, _vp,
                                          _vp, _vp]
        L.gw_vis_scale_coords.restype = None
        L.gw_read_kernel.argtypes = [ct.c_char_p, ct.c_int, _ip, _ip,
                                     ct.POINTER(ct.POINTER(ct.c_int)), _ip,
                                     _ip, ct.POINTER(ct.c_double),
                                     ct.POINTER(ct.c_double), _ip]
        L.gw_read_kernel.restype = _vp
        L.gw_free.argtypes = [_vp]
        L.gw_free.restype = None
        L.gw_write_fits_cube.argtypes = [ct.c_int, _vp, ct.c_char_p, ct.c_int,
                                         ct.c_int, ct.c_int, ct.c_int, _ip]
        L.gw_write_fits_cube.restype = None
        _done = True
    return L


def _check(st, where):
    if st.value != 0:
        raise IOError_(st.value, where)


class BinaryFile:
    """An OSKAR binary file opened for reading (gw_binary_*)."""

    def __init__(self, path):
        st = ct.c_int(0)
        self.h = io_lib().gw_binary_open(str(path).encode(), ct.byref(st))
        _check(st, "gw_binary_open")

    def num_tags(self):
        return io_lib().gw_binary_num_tags(self.h)

    def set_query_search_start(self, start):
        st = ct.c_int(0)
        io_lib().gw_binary_set_query_search_start(self.h, start,
                                                  ct.byref(st))
        _check(st, "gw_binary_set_query_search_start")

    def payload_size(self, data_type, group, tag, index=0):
        st = ct.c_int(0)
        n = ct.c_size_t(0)
        io_lib().gw_binary_query(self.h, data_type, group, tag, index,
                                 ct.byref(n), ct.byref(st))
        _check(st, "gw_binary_query")
        return int(n.value)

    def read(self, data_type, group, tag, index, dtype, count=None):
        """A tag's payload as a numpy array of dtype."""
        nbytes = self.payload_size(data_type, group, tag, index)
        out = np.empty(nbytes // np.dtype(dtype).itemsize if count is None
                       else count, dtype)
        st = ct.c_int(0)
        io_lib().gw_binary_read(self.h, data_type, group, tag, index,
                                out.nbytes, out.ctypes.data, ct.byref(st))
        _check(st, "gw_binary_read")
        return out

    def read_ext(self, data_type, group, tag, index, nbytes):
        out = np.empty(nbytes, np.uint8)
        st = ct.c_int(0)
        io_lib().gw_binary_read_ext(self.h, data_type, group.encode(),
                                    tag.encode(), index, out.nbytes,
                                    out.ctypes.data, ct.byref(st))
        _check(st, "gw_binary_read_ext")
        return out

    def read_int(self, group, tag, index=0):
        st = ct.c_int(0)
        v = ct.c_int(0)
        io_lib().gw_binary_read_int(self.h, group, tag, index, ct.byref(v),
                                    ct.byref(st))
        _check(st, "gw_binary_read_int")
        return int(v.value)

    def read_double(self, group, tag, index=0):
        st = ct.c_int(0)
        v = ct.c_double(0)
        io_lib().gw_binary_read_double(self.h, group, tag, index,
                                       ct.byref(v), ct.byref(st))
        _check(st, "gw_binary_read_double")
        return float(v.value)

    def read_coords(self, block, precision, n):
        dt = np.float64 if precision == OSKAR_DOUBLE else np.float32
        uu, vv, ww = (np.empty(n, dt) for _ in range(3))
        st = ct.c_int(0)
        io_lib().gw_vis_read_coords(self.h, block, precision, n,
                                    uu.ctypes.data, vv.ctypes.data,
                                    ww.ctypes.data, ct.byref(st))
        _check(st, "gw_vis_read_coords")
        return uu, vv, ww

    def close(self):
        if getattr(self, "h", None):
            io_lib().gw_binary_free(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def scale_coords(freq_hz, uu, vv, ww):
    """gw_vis_scale_coords in place on float32 or float64 arrays."""
    prec = OSKAR_DOUBLE if uu.dtype == np.float64 else OSKAR_SINGLE
    io_lib().gw_vis_scale_coords(float(freq_hz), prec, uu.size,
                                 uu.ctypes.data, vv.ctypes.data,
                                 ww.ctypes.data)


def read_kernel(path, precision=OSKAR_SINGLE, support=True):
    """gw_read_kernel: (cube [P, C, C] float32/float64, dict of the header
    values, support int32 [P] or None)."""
    L = io_lib()
    st = ct.c_int(0)
    C, P, ov, G = (ct.c_int(0) for _ in range(4))
    cell, ws = ct.c_double(0), ct.c_double(0)
    sup = ct.POINTER(ct.c_int)()
    p = L.gw_read_kernel(str(path).encode(), precision, ct.byref(C),
                         ct.byref(P), ct.byref(sup) if support else None,
                         ct.byref(ov), ct.byref(G), ct.byref(cell),
                         ct.byref(ws), ct.byref(st))
    _check(st, "gw_read_kernel")
    dt = np.float64 if precision == OSKAR_DOUBLE else np.float32
    n = P.value * C.value * C.value
    cube = np.ctypeslib.as_array(ct.cast(p, ct.POINTER(
        ct.c_double if dt == np.float64 else ct.c_float)), (n,)).copy()
    L.gw_free(p)
    s = None
    if support:
        s = np.ctypeslib.as_array(sup, (P.value,)).astype(np.int32)
        L.gw_free(ct.cast(sup, ct.c_void_p))
    return (cube.reshape(P.value, C.value, C.value),
            dict(conv_size_half=C.value, num_w_planes=P.value,
                 oversample=ov.value, grid_size=G.value,
                 cellsize_rad=cell.value, w_scale=ws.value), s)


def write_fits_cube(data, root_name, width, height, num_planes, i_plane=-1,
                    complex_=False):
    """gw_write_fits_cube on a float32 / float64 array (interleaved re, im
    when complex_)."""
    a = np.ascontiguousarray(data)
    dt = (OSKAR_DOUBLE if a.dtype == np.float64 else OSKAR_SINGLE) | \
        (OSKAR_COMPLEX if complex_ else 0)
    st = ct.c_int(0)
    io_lib().gw_write_fits_cube(dt, a.ctypes.data, str(root_name).encode(),
                                width, height, num_planes, i_plane,
                                ct.byref(st))
    _check(st, "gw_write_fits_cube")
