"""Helpers for the GPU parity tests (test infrastructure).

Tolerances used by the tests (relative to max|grid| of the reference side):

* num_skipped: bit-exact against the CPU oracle (FP32 index math of
  oskar_grid_wproj.c:169-212, reproduced on the device).
* grid, FP32 path: against the *exact* value of the FP32 gridder's result
  (same FP32 index math and taps, float64 products and sums;
  oracle.grid_wproj_f_exact) the GPU must be within 1e-5 * max|grid| AND no
  less accurate than the reference's own sequential FP32 loop (small slack
  1e-6 * max|grid|).  Against the sequential FP32 oracle itself the
  difference must be within 1e-5 * max|grid| plus that oracle's own error
  against the exact value: at 10^5 contributions per cell the sequential
  float sum is itself ~2e-5 off, so no reordered FP32 sum can match it
  closer than that.
* grid, FP64 path: within 1e-12 * max|grid| of the FP64 oracle.
* norm: relative 1e-12 (the per-point tap sums are identical; only the order
  of their float64 summation differs).
"""
from __future__ import annotations

import numpy as np

import oracle


def max_rel(a, b, scale=None):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    s = np.abs(b).max() if scale is None else scale
    return float(np.abs(a - b).max() / s) if s > 0 else float(np.abs(a - b).max())


def rms_rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = np.sqrt(np.mean((a - b) ** 2))
    n = np.sqrt(np.mean(b ** 2))
    return float(d / n) if n > 0 else float(d)


def fp32_report(g_gpu, wl):
    """Errors of an FP32 GPU grid against the sequential FP32 oracle, the
    exact FP32-index result and the FP64 oracle; returns a dict."""
    g_f, sk_f, n_f = oracle.grid_wproj_f(wl)
    g_e, sk_e, n_e = oracle.grid_wproj_f_exact(wl)
    g_d, sk_d, n_d = oracle.grid_wproj_d(wl)
    scale = np.abs(g_e).max()
    return dict(
        skipped_f=sk_f, norm_f=n_f, norm_exact=n_e,
        gpu_vs_f=max_rel(g_gpu, g_f, scale),
        gpu_vs_exact=max_rel(g_gpu, g_e, scale),
        f_vs_exact=max_rel(g_f, g_e, scale),
        gpu_vs_d=max_rel(g_gpu, g_d), f_vs_d=max_rel(g_f, g_d),
        gpu_rms_vs_d=rms_rel(g_gpu, g_d), f_rms_vs_d=rms_rel(g_f, g_d),
        scale=scale)


def assert_fp32_parity(r, tol=1e-5, slack=1e-6):
    assert r["gpu_vs_exact"] <= tol, r
    assert r["gpu_vs_exact"] <= r["f_vs_exact"] + slack, r
    assert r["gpu_vs_f"] <= tol + r["f_vs_exact"], r


def tap_tables(wl):
    """Per (plane, off_v + h, off_u + h): real tap sum (the reference's
    per-point norm term, oskar_grid_wproj.c:238) and complex tap sum."""
    ov = wl.oversample
    h = (ov + 1) // 2
    K = 2 * h + 1
    C = wl.conv_size_half
    cube = wl.conv_func.reshape(wl.num_w_planes, C, C, 2).astype(np.float64)
    re = np.zeros((wl.num_w_planes, K, K))
    cx = np.zeros((wl.num_w_planes, K, K), np.complex128)
    for p in range(wl.num_w_planes):
        s = int(wl.support[p])
        j = np.arange(-s, s + 1)
        for iv in range(K):
            rows = np.abs(iv - h + j * ov)
            for iu in range(K):
                cols = np.abs(iu - h + j * ov)
                t = cube[p][np.ix_(rows, cols)]
                re[p, iv, iu] = t[..., 0].sum()
                cx[p, iv, iu] = t[..., 0].sum() + 1j * t[..., 1].sum()
    return re, cx, h


def expected_sums(wl, rec, tabs=None):
    """Closed-form norm increment and complex grid checksum (sum over all
    cells) from the per-point index results; also the scale
    sum_i |v_i| * sum_footprint |taps| that bounds the checksum's error.
    ``tabs``: tap_tables(wl) when already computed (workloads sharing a
    kernel set)."""
    re, cx, h = tap_tables(wl) if tabs is None else tabs
    ok = rec["skip"] == 0
    p = rec["plane"][ok]
    iv = rec["off_v"][ok] + h
    iu = rec["off_u"][ok] + h
    conj = rec["conj"][ok].astype(bool)
    t = cx[p, iv, iu]
    t = np.where(conj, np.conj(t), t)
    v = wl.vis.reshape(-1, 2).astype(np.float64)[ok]
    vc = v[:, 0] + 1j * v[:, 1]
    norm = float(np.sum(re[p, iv, iu]))
    checksum = complex(np.sum(vc * t))
    # bound for the absolute error of the checksum
    C = wl.conv_size_half
    cube = wl.conv_func.reshape(wl.num_w_planes, C, C, 2).astype(np.float64)
    amax = np.abs(cube).max()
    w = (2 * rec["support"][ok].astype(np.float64) + 1) ** 2
    scale = float(np.sum(np.abs(vc) * w * amax))
    return norm, checksum, scale
