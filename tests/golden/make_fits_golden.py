"""Writes small FITS fixtures with the reference's OWN FITS code: its
utils/write_fits_cube.c and utils/read_kernel.c on its vendored cfitsio
3.41 (a gcc build in oracle/_ref, `make -C oracle ref`), as pins for the
library's from-scratch FITS reader and writer.  Only the written files and
the values the reference read back are committed.

Fixtures:
  ref_grid_REAL.fits / ref_grid_IMAG.fits  write_fits_cube of a complex
      FP32 cube (7 x 5 x 2, all planes);
  ref_cube_d.fits  write_fits_cube of a real FP64 cube (7 x 5 x 3), then
      plane 1 rewritten in place;
  ref_kernel_e.fits  a W-kernel file written through cfitsio (FP32 image,
      OVERSAMP / GRIDSIZE / CELLSIZE / W_SCALE keywords, W_KERNELS table
      with a non-integer float support column);
  ref_kernel_j.fits  the same with a 16-bit scaled image (BSCALE, BZERO)
      and an integer support column;
  ref_fits.npz  the arrays written, and everything read_kernel returned for
      each kernel file (FP32 and FP64 cubes, supports, header values).

Usage (where the reference tree is mounted):
  make -C oracle ref && python tests/golden/make_fits_golden.py
"""
import ctypes as ct
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
LIB = os.path.join(ROOT, "oracle", "_ref", "libref_fits.so")

OSKAR_SINGLE, OSKAR_DOUBLE, OSKAR_COMPLEX = 0x04, 0x08, 0x20
# cfitsio constants (fitsio.h)
TFLOAT, TDOUBLE, TINT, TSTRING = 42, 82, 31, 16
FLOAT_IMG, SHORT_IMG = -32, 16
BINARY_TBL = 2


def _lib():
    L = ct.CDLL(LIB)
    L.write_fits_cube.argtypes = [ct.c_int, ct.c_void_p, ct.c_char_p, ct.c_int,
                                  ct.c_int, ct.c_int, ct.c_int,
                                  ct.POINTER(ct.c_int)]
    L.read_kernel.restype = ct.c_void_p
    L.read_kernel.argtypes = [ct.c_char_p, ct.c_int] + \
        [ct.POINTER(ct.c_int)] * 2 + [ct.POINTER(ct.POINTER(ct.c_int))] + \
        [ct.POINTER(ct.c_int)] * 2 + [ct.POINTER(ct.c_double)] * 2 + \
        [ct.POINTER(ct.c_int)]
    return L


def _check(st):
    assert st.value == 0, st.value


def write_kernel_file(L, path, cube, support, tform, bitpix, bscale=None,
                      bzero=None):
    """A W-kernel FITS file through cfitsio's own calls."""
    f = ct.c_void_p()
    st = ct.c_int(0)
    if os.path.exists(path):
        os.remove(path)
    L.ffinit(ct.byref(f), path.encode(), ct.byref(st))
    _check(st)
    P, C, _ = cube.shape
    naxes = (ct.c_long * 3)(C, C, P)
    L.ffcrim(f, bitpix, 3, naxes, ct.byref(st))
    _check(st)

    def key(name, dtype, value):
        if dtype == TINT:
            v = ct.c_int(value)
        elif dtype == TDOUBLE:
            v = ct.c_double(value)
        L.ffpky(f, dtype, name.encode(), ct.byref(v), None, ct.byref(st))
        _check(st)

    key("OVERSAMP", TINT, 4)
    key("GRIDSIZE", TINT, 2048)
    key("CELLSIZE", TDOUBLE, 1.7044230976507e-05)
    key("W_SCALE", TDOUBLE, 2.5)
    if bscale is not None:
        key("BSCALE", TDOUBLE, bscale)
        key("BZERO", TDOUBLE, bzero)
        # Scaling applies from the next write on (cfitsio: ffrdef).
        L.ffrdef(f, ct.byref(st))
        _check(st)
    first = (ct.c_long * 3)(1, 1, 1)
    data = np.ascontiguousarray(cube, np.float64)
    L.ffppx(f, TDOUBLE, first, ct.c_longlong(data.size),
            data.ctypes.data_as(ct.c_void_p), ct.byref(st))
    _check(st)
    ttype = (ct.c_char_p * 1)(b"SUPPORT")
    tf = (ct.c_char_p * 1)(tform.encode())
    L.ffcrtb(f, BINARY_TBL, ct.c_longlong(0), 1, ttype, tf, None,
             b"W_KERNELS", ct.byref(st))
    _check(st)
    sup = np.ascontiguousarray(support, np.float64)
    L.ffpcl(f, TDOUBLE, 1, ct.c_longlong(1), ct.c_longlong(1),
            ct.c_longlong(sup.size), sup.ctypes.data_as(ct.c_void_p),
            ct.byref(st))
    _check(st)
    L.ffclos(f, ct.byref(st))
    _check(st)


def read_kernel_ref(L, path, precision):
    C, P, ov, G, st = (ct.c_int(0) for _ in range(5))
    cell, ws = ct.c_double(0), ct.c_double(0)
    sup = ct.POINTER(ct.c_int)()
    p = L.read_kernel(path.encode(), precision, ct.byref(C), ct.byref(P),
                      ct.byref(sup), ct.byref(ov), ct.byref(G),
                      ct.byref(cell), ct.byref(ws), ct.byref(st))
    _check(st)
    n = P.value * C.value * C.value
    ty = ct.c_double if precision == OSKAR_DOUBLE else ct.c_float
    cube = np.ctypeslib.as_array(ct.cast(p, ct.POINTER(ty)), (n,)).copy()
    s = np.ctypeslib.as_array(sup, (P.value,)).astype(np.int32).copy()
    hdr = np.array([C.value, P.value, ov.value, G.value], np.int64)
    return cube.reshape(P.value, C.value, C.value), s, hdr, \
        np.array([cell.value, ws.value])


def main():
    L = _lib()
    rng = np.random.default_rng(2024)
    out = {}
    W, H = 7, 5
    st = ct.c_int(0)
    # Complex FP32 grid, all planes.
    grid = rng.normal(size=2 * W * H * 2).astype(np.float32)
    root = os.path.join(HERE, "ref_grid")
    for suffix in ("_REAL.fits", "_IMAG.fits"):
        if os.path.exists(root + suffix):
            os.remove(root + suffix)
    L.write_fits_cube(OSKAR_SINGLE | OSKAR_COMPLEX,
                      grid.ctypes.data_as(ct.c_void_p), root.encode(), W, H,
                      2, -1, ct.byref(st))
    _check(st)
    out["grid_c"] = grid
    # Real FP64 cube, then plane 1 rewritten in place.
    cube = rng.normal(size=(3, H, W))
    plane = rng.normal(size=(H, W))
    path = os.path.join(HERE, "ref_cube_d.fits")
    if os.path.exists(path):
        os.remove(path)
    L.write_fits_cube(OSKAR_DOUBLE, cube.ctypes.data_as(ct.c_void_p),
                      path.encode(), W, H, 3, -1, ct.byref(st))
    _check(st)
    L.write_fits_cube(OSKAR_DOUBLE, plane.ctypes.data_as(ct.c_void_p),
                      path.encode(), W, H, 3, 1, ct.byref(st))
    _check(st)
    out["cube_d"] = cube
    out["plane_d"] = plane
    # Kernel files, and what the reference's read_kernel returns for them.
    P, C = 3, 9
    k_e = rng.normal(size=(P, C, C)).astype(np.float32)
    sup_e = np.array([1.7, 2.99, -1.6])
    k_j = rng.integers(-30000, 30000, size=(P, C, C)).astype(np.float64)
    sup_j = np.array([1, 3, 4])
    files = {"e": ("ref_kernel_e.fits", k_e, sup_e, "1E", FLOAT_IMG, None,
                   None),
             "j": ("ref_kernel_j.fits", k_j * 0.5 + 3.0, sup_j, "1J",
                   SHORT_IMG, 0.5, 3.0)}
    for tag, (name, cube_k, sup, tform, bp, bs, bz) in files.items():
        path = os.path.join(HERE, name)
        write_kernel_file(L, path, cube_k, sup, tform, bp, bs, bz)
        for prec, pn in ((OSKAR_SINGLE, "f"), (OSKAR_DOUBLE, "d")):
            c, s, hdr, dbl = read_kernel_ref(L, path, prec)
            out[f"kernel_{tag}_cube_{pn}"] = c
            out[f"kernel_{tag}_support_{pn}"] = s
            out[f"kernel_{tag}_header_{pn}"] = hdr
            out[f"kernel_{tag}_cell_ws_{pn}"] = dbl
    np.savez(os.path.join(HERE, "ref_fits.npz"), **out)
    print("wrote", sorted(out))


if __name__ == "__main__":
    main()
