"""FITS files written and parsed from the standard (FITS 4.0: 2880-byte
blocks of 80-character cards, big-endian data), independently of the
library's reader and writer (test infrastructure).

kernel_fits writes a W-kernel cube in the layout read_kernel.c:36-118
expects: a primary image (NAXIS1 = NAXIS2 = conv_size_half, NAXIS3 =
planes) with keywords OVERSAMP, GRIDSIZE, CELLSIZE and W_SCALE, then a
binary table whose first column holds the per-plane support sizes.
"""
import numpy as np


def cards(pairs):
    out = b""
    for k, v in pairs:
        if k == "END":
            out += b"END".ljust(80)
            continue
        out += (k.ljust(8) + "= " + str(v).rjust(20)).ljust(80).encode()
    return out.ljust(((len(out) + 2879) // 2880) * 2880, b" ")


def pad(b):
    return b.ljust(((len(b) + 2879) // 2880) * 2880, b"\0")


def kernel_fits(path, cube, support, bitpix=-32, extra=(), tform="1J",
                oversample=4, grid_size=2048,
                cell_size="1.7044230976507E-05", w_scale="0.25D+01"):
    P, C, _ = cube.shape
    hdr = [("SIMPLE", "T"), ("BITPIX", bitpix), ("NAXIS", 3),
           ("NAXIS1", C), ("NAXIS2", C), ("NAXIS3", P), ("EXTEND", "T"),
           ("OVERSAMP", oversample), ("GRIDSIZE", grid_size),
           ("CELLSIZE", cell_size), ("W_SCALE", w_scale),
           *extra, ("END", None)]
    fmt = {-32: ">f4", -64: ">f8", 16: ">i2", 32: ">i4"}[bitpix]
    data = pad(np.ascontiguousarray(cube).astype(fmt).tobytes())
    code, width = {"1J": (">i4", 4), "1E": (">f4", 4), "1I": (">i2", 2)}[tform]
    rows = np.asarray(support).astype(code).tobytes()
    tab = [("XTENSION", "'BINTABLE'"), ("BITPIX", 8), ("NAXIS", 2),
           ("NAXIS1", width), ("NAXIS2", len(support)), ("PCOUNT", 0),
           ("GCOUNT", 1), ("TFIELDS", 1), ("TTYPE1", "'SUPPORT '"),
           ("TFORM1", f"'{tform}      '"), ("EXTNAME", "'W_KERNELS'"),
           ("END", None)]
    open(path, "wb").write(cards(hdr) + data + cards(tab) + pad(rows))


def parse_image(path):
    """Primary header cards and the big-endian image."""
    raw = open(path, "rb").read()
    found, at = {}, 0
    while True:
        c = raw[at:at + 80].decode()
        at += 80
        if c.startswith("END"):
            break
        if c[8:10] == "= ":
            found[c[:8].strip()] = c[10:].split("/")[0].strip()
    data_at = ((at + 2879) // 2880) * 2880
    bp = int(found["BITPIX"])
    shape = tuple(int(found[f"NAXIS{k}"]) for k in (3, 2, 1))
    n = int(np.prod(shape))
    a = np.frombuffer(raw, ">f4" if bp == -32 else ">f8", n, data_at)
    assert len(raw) % 2880 == 0
    return found, a.reshape(shape)
