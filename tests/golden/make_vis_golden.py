"""Writes small OSKAR binary visibility files with the reference's OWN
binary-file library (a gcc build of utils/binary/src/*.c in oracle/_ref,
`make -C oracle ref`), as fixtures for the library's from-scratch reader.
Only the written files and the arrays written into them are committed.

Fixtures:
  ref_vis_f.vis / ref_vis_d.vis  header tags (group 11) as main.cpp reads
      them, two visibility blocks (group 12) with DIM_START_AND_SIZE,
      CROSS_CORRELATIONS (complex float) and BASELINE_UU/VV/WW (float or
      double), plus an extended tag and a double array tag;
  ref_vis.npz  every array and value written.

Usage (where the reference tree is mounted):
  make -C oracle ref && python tests/golden/make_vis_golden.py
"""
import ctypes as ct
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
LIB = os.path.join(ROOT, "oracle", "_ref", "libref_binary.so")

INT, SINGLE, DOUBLE, COMPLEX, CHAR = 0x02, 0x04, 0x08, 0x20, 0x01
HDR, BLK = 11, 12


def write(path, coord_type, seed):
    L = ct.CDLL(LIB)
    L.oskar_binary_create.restype = ct.c_void_p
    L.oskar_binary_create.argtypes = [ct.c_char_p, ct.c_char,
                                      ct.POINTER(ct.c_int)]
    L.oskar_binary_write.argtypes = [ct.c_void_p, ct.c_ubyte, ct.c_ubyte,
                                     ct.c_ubyte, ct.c_int, ct.c_size_t,
                                     ct.c_void_p, ct.POINTER(ct.c_int)]
    L.oskar_binary_write_int.argtypes = [ct.c_void_p, ct.c_ubyte, ct.c_ubyte,
                                         ct.c_int, ct.c_int,
                                         ct.POINTER(ct.c_int)]
    L.oskar_binary_write_double.argtypes = [ct.c_void_p, ct.c_ubyte,
                                            ct.c_ubyte, ct.c_int, ct.c_double,
                                            ct.POINTER(ct.c_int)]
    L.oskar_binary_write_ext.argtypes = [ct.c_void_p, ct.c_ubyte, ct.c_char_p,
                                         ct.c_char_p, ct.c_int, ct.c_size_t,
                                         ct.c_void_p, ct.POINTER(ct.c_int)]
    L.oskar_binary_free.argtypes = [ct.c_void_p]
    st = ct.c_int(0)
    h = L.oskar_binary_create(path.encode(), b"w", ct.byref(st))
    assert st.value == 0
    rng = np.random.default_rng(seed)
    n_st, n_t = 5, 3
    n_bl = n_st * (n_st - 1) // 2
    out = {}
    header = dict(num_tags_per_block=5, amp_type=SINGLE | COMPLEX,
                  coord_precision=coord_type, max_times_per_block=n_t,
                  num_times_total=2 * n_t, num_channels_total=1,
                  num_stations=n_st)
    ids = dict(num_tags_per_block=2, amp_type=5, coord_precision=6,
               max_times_per_block=7, num_times_total=8,
               num_channels_total=10, num_stations=11)
    for k, v in header.items():
        L.oskar_binary_write_int(h, HDR, ids[k], 0, v, ct.byref(st))
        out["hdr_" + k] = np.int64(v)
    L.oskar_binary_write_double(h, HDR, 23, 0, 150.25e6, ct.byref(st))
    out["hdr_freq_start_hz"] = np.float64(150.25e6)
    cdt = np.float64 if coord_type == DOUBLE else np.float32
    for b in range(2):
        dims = np.array([b * n_t, 0, n_t, 1, n_bl, n_st], np.int32)
        L.oskar_binary_write(h, INT, BLK, 1, b, dims.nbytes,
                             dims.ctypes.data, ct.byref(st))
        vis = (rng.normal(size=2 * n_t * n_bl)).astype(np.float32)
        L.oskar_binary_write(h, SINGLE | COMPLEX, BLK, 3, b, vis.nbytes,
                             vis.ctypes.data, ct.byref(st))
        for t, name in ((4, "uu"), (5, "vv"), (6, "ww")):
            a = (rng.normal(size=n_t * n_bl) * 1000.0).astype(cdt)
            L.oskar_binary_write(h, coord_type, BLK, t, b, a.nbytes,
                                 a.ctypes.data, ct.byref(st))
            out[f"b{b}_{name}"] = a
        out[f"b{b}_dims"] = dims
        out[f"b{b}_vis"] = vis
    note = b"written by the reference binary library\0"
    buf = ct.create_string_buffer(note, len(note))
    L.oskar_binary_write_ext(h, CHAR, b"fixture", b"note", 7, len(note),
                             buf, ct.byref(st))
    out["ext_note"] = np.frombuffer(note, np.uint8)
    extra = rng.normal(size=9)
    L.oskar_binary_write(h, DOUBLE, 1, 99, 3, extra.nbytes,
                         extra.ctypes.data, ct.byref(st))
    out["extra_double"] = extra
    assert st.value == 0, st.value
    L.oskar_binary_free(h)
    return out


def main():
    f = write(os.path.join(HERE, "ref_vis_f.vis"), SINGLE, seed=1)
    d = write(os.path.join(HERE, "ref_vis_d.vis"), DOUBLE, seed=2)
    np.savez(os.path.join(HERE, "ref_vis.npz"),
             **{"f_" + k: v for k, v in f.items()},
             **{"d_" + k: v for k, v in d.items()})


if __name__ == "__main__":
    main()
