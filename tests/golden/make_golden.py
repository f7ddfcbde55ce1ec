"""Write the golden fixtures from the reference's own gridder.

Runs ``oracle/_ref/libref_gridder.so`` -- a plain gcc build
(``make -C oracle ref``) of the reference's unmodified
``oskar_grid_wproj.c`` and ``utils/oskar_grid_weights.c`` -- on small seeded
inputs and stores inputs and outputs as ``.npz`` files next to this script.
Only the vectors are committed; the reference sources are not.

Fixtures:
  ref_f_diag.npz   oskar_grid_wproj_f_diagnostics (full-length FP32 loop,
                   arithmetic identical to oskar_grid_wproj_f) on a random
                   64x64 workload: grid, norm, num_skipped, histograms.
  ref_f_10000.npz  oskar_grid_wproj_f itself (its loop is fixed to the first
                   10000 points, oskar_grid_wproj.c:155) on the edge-case
                   points (ties, grid edges, w = +-0, NaN, +-inf, huge
                   values) repeated to 10000 points.
  ref_d.npz        oskar_grid_wproj_d on the random workload, float64 input.
  ref_weights.npz  oskar_grid_weights_write_f/_read_f and _d.
  ref_asym.npz     oskar_grid_wproj_f_diagnostics and oskar_grid_wproj_d on
                   a workload whose kernel cube is random (no symmetry
                   between rows and columns, or within either), so that a
                   transposed or mirrored tap index shows.

Usage (where the reference tree is mounted):
  make -C oracle ref && python tests/golden/make_golden.py
"""

from __future__ import annotations

import ctypes as ct
import dataclasses
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from fpga_gridding_amd import synth  # noqa: E402

_libc = ct.CDLL(None)


def _quiet(fn, *a, **k):
    """Run fn with C-level stdout sent to /dev/null (the reference prints
    a progress line every 1000 points)."""
    sys.stdout.flush()
    saved = os.dup(1)
    null = os.open(os.devnull, os.O_WRONLY)
    os.dup2(null, 1)
    try:
        return fn(*a, **k)
    finally:
        _libc.fflush(None)
        os.dup2(saved, 1)
        os.close(saved)
        os.close(null)


def padded_edges(n=10000):
    wl = synth.edge_cases()
    reps = -(-n // wl.num_points)
    vis = np.tile(wl.vis.reshape(-1, 2), (reps, 1))[:n].reshape(-1)
    return dataclasses.replace(
        wl, uu=np.tile(wl.uu, reps)[:n].copy(),
        vv=np.tile(wl.vv, reps)[:n].copy(), ww=np.tile(wl.ww, reps)[:n].copy(),
        vis=vis.copy(), weight=np.ones(n, np.float32))


def random_workload():
    return synth.random_small(n=3000, grid_size=64, num_planes=6, s_min=1,
                              s_max=6, ov=4, seed=11, edge_fraction=0.15)


def asymmetric_workload():
    wl = synth.random_small(n=2000, grid_size=48, num_planes=4, s_min=1,
                            s_max=5, ov=4, seed=12, edge_fraction=0.1)
    rng = np.random.default_rng(13)
    cube = rng.standard_normal(wl.conv_func.shape).astype(np.float32)
    return dataclasses.replace(wl, conv_func=cube, name="asym")


def main():
    ref = oracle.ref()
    if ref is None:
        raise SystemExit("oracle/_ref/libref_gridder.so missing: "
                         "run `make -C oracle ref` first")
    out = {}

    wl = random_workload()
    g, sk, nrm, per_s, per_w, per_cell, avg_s = oracle.grid_wproj_f_diagnostics(
        wl, lib_=ref)
    out["ref_f_diag"] = dict(**synth.workload_arrays(wl), grid=g,
                             num_skipped=sk, norm=nrm, per_support=per_s,
                             per_plane=per_w, per_cell=per_cell,
                             avg_support=avg_s)

    ed = padded_edges()
    g, sk, nrm = _quiet(oracle.grid_wproj_f, ed, lib_=ref)
    out["ref_f_10000"] = dict(**synth.workload_arrays(ed), grid=g,
                              num_skipped=sk, norm=nrm)

    gd, skd, nrmd = oracle.grid_wproj_d(wl, lib_=ref)
    out["ref_d"] = dict(**synth.workload_arrays(wl), grid=gd,
                        num_skipped=skd, norm=nrmd)

    wa = asymmetric_workload()
    g, sk, nrm, per_s, per_w, per_cell, avg_s = oracle.grid_wproj_f_diagnostics(
        wa, lib_=ref)
    gd, skd, nrmd = oracle.grid_wproj_d(wa, lib_=ref)
    out["ref_asym"] = dict(**synth.workload_arrays(wa), grid=g,
                           num_skipped=sk, norm=nrm, grid_d=gd,
                           num_skipped_d=skd, norm_d=nrmd)

    G = wl.grid_size
    wgrid = np.zeros(G * G, np.float32)
    sk_w = oracle.grid_weights_write_f(wl.uu, wl.vv, wl.weight,
                                       wl.cell_size_rad, G, wgrid, lib_=ref)
    wout = np.zeros(wl.num_points, np.float32)
    sk_r = oracle.grid_weights_read_f(wl.uu, wl.vv, wl.weight, wout,
                                      wl.cell_size_rad, G, wgrid, lib_=ref)
    wgrid_d = np.zeros(G * G, np.float64)
    sk_wd = ct.c_size_t(0)
    uu_d = wl.uu.astype(np.float64)
    vv_d = wl.vv.astype(np.float64)
    w_d = wl.weight.astype(np.float64)
    ref.oskar_grid_weights_write_d(wl.num_points, uu_d, vv_d, w_d,
                                   float(np.float32(wl.cell_size_rad)), G,
                                   ct.byref(sk_wd), wgrid_d)
    wout_d = np.zeros(wl.num_points, np.float64)
    sk_rd = ct.c_size_t(0)
    ref.oskar_grid_weights_read_d(wl.num_points, uu_d, vv_d, w_d, wout_d,
                                  float(np.float32(wl.cell_size_rad)), G,
                                  ct.byref(sk_rd), wgrid_d)
    out["ref_weights"] = dict(uu=wl.uu, vv=wl.vv, weight=wl.weight,
                              cell_size_rad=np.float32(wl.cell_size_rad),
                              grid_size=G, grid_f=wgrid, weight_out_f=wout,
                              skipped_write_f=sk_w, skipped_read_f=sk_r,
                              grid_d=wgrid_d, weight_out_d=wout_d,
                              skipped_write_d=sk_wd.value,
                              skipped_read_d=sk_rd.value)

    for name, d in out.items():
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **d)
        print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
