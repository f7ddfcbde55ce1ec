"""Uniform weighting on the device (oskar_grid_weights_write/_read drop-ins
and the device-pointer API), through the C ABI.

The bar is bit-exact, for FP32 and FP64: the device sums each cell's
weights in input order (stable sort by cell, then one sequential add per
point), which is the reference's loop order (utils/oskar_grid_weights.c:
58-69, :126-137), and the read back is one IEEE division per point.

* against the reference's own outputs (tests/golden/ref_weights.npz, made by
  a gcc build of utils/oskar_grid_weights.c);
* against the oracle (gridref_grid_weights_*) on workloads with thousands
  of points per cell, non-zero starting grids, edge and NaN/inf points;
* unit weights give the nearest-cell histogram (closed form).
"""
import os

import numpy as np
import pytest

import oracle
import fpga_gridding_amd as fg
from fpga_gridding_amd import synth

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module", autouse=True)
def _gpu_present():
    if fg.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    yield
    fg.lib().gw_release_cached()


def _bits(a):
    a = np.asarray(a)
    return a.view(np.uint32 if a.dtype == np.float32 else np.uint64)


def _same(a, b):
    return np.array_equal(_bits(a), _bits(b))


def test_weights_golden_fp32_and_fp64():
    d = np.load(os.path.join(HERE, "golden", "ref_weights.npz"))
    G = int(d["grid_size"])
    cell = float(d["cell_size_rad"])
    for double, sfx in ((False, "f"), (True, "d")):
        dt = np.float64 if double else np.float32
        g = np.zeros(G * G, dt)
        sk = fg.grid_weights_write(d["uu"].astype(dt), d["vv"].astype(dt),
                                   d["weight"].astype(dt), cell, G, g,
                                   double=double)
        assert sk == int(d["skipped_write_" + sfx])
        assert _same(g, d["grid_" + sfx])
        wo = np.zeros(d["uu"].size, dt)
        sk = fg.grid_weights_read(d["uu"].astype(dt), d["vv"].astype(dt),
                                  d["weight"].astype(dt), wo, cell, G, g,
                                  double=double)
        assert sk == int(d["skipped_read_" + sfx])
        assert _same(wo, d["weight_out_" + sfx])


def _dense(n, G, seed):
    """Points crowded into a few central cells (thousands per cell), some
    off the grid, random weights in [0.5, 1]."""
    rng = np.random.default_rng(seed)
    cell = float(np.float32(1.0 / (G * 64.0)))
    scale = G * cell
    uu = (rng.normal(0.0, 3.0, n) / scale).astype(np.float32)
    vv = (rng.normal(0.0, 3.0, n) / scale).astype(np.float32)
    far = rng.random(n) < 0.02
    uu[far] = (rng.choice([-1, 1], far.sum()) * (G / 2 + 5) / scale)
    w = rng.uniform(0.5, 1.0, n).astype(np.float32)
    return uu, vv, w, cell


@pytest.mark.parametrize("double", [False, True], ids=["f", "d"])
def test_weights_dense_cells_bit_exact(double):
    G = 64
    uu, vv, w, cell = _dense(300_000, G, seed=3)
    dt = np.float64 if double else np.float32
    uu, vv, w = uu.astype(dt), vv.astype(dt), w.astype(dt)
    rng = np.random.default_rng(4)
    start = rng.uniform(0, 3, G * G).astype(dt)       # accumulating grid
    g_ref, g_dev = start.copy(), start.copy()
    wr = oracle.grid_weights_write_d if double else \
        oracle.grid_weights_write_f
    rd = oracle.grid_weights_read_d if double else oracle.grid_weights_read_f
    sk_ref = wr(uu, vv, w, cell, G, g_ref)
    sk_dev = fg.grid_weights_write(uu, vv, w, cell, G, g_dev, double=double)
    assert sk_dev == sk_ref > 0
    assert _same(g_dev, g_ref)
    wo_ref = np.full(uu.size, -7.0, dt)
    wo_dev = wo_ref.copy()
    assert rd(uu, vv, w, wo_ref, cell, G, g_ref) == \
        fg.grid_weights_read(uu, vv, w, wo_dev, cell, G, g_dev, double=double)
    assert _same(wo_dev, wo_ref)
    assert np.sum(wo_dev == -7.0) == sk_ref     # skipped left untouched


def test_weights_edge_points_and_zero_cells():
    """The gridder's edge-case points (ties, edges, NaN, inf, huge values):
    same cells and skips; cells with zero weight read back 0."""
    wl = synth.edge_cases()
    G = wl.grid_size
    rng = np.random.default_rng(5)
    w = rng.uniform(0.5, 1.0, wl.num_points).astype(np.float32)
    w[::7] = 0.0
    g_ref = np.zeros(G * G, np.float32)
    g_dev = g_ref.copy()
    sk = oracle.grid_weights_write_f(wl.uu, wl.vv, w, wl.cell_size_rad, G,
                                     g_ref)
    assert fg.grid_weights_write(wl.uu, wl.vv, w, wl.cell_size_rad, G,
                                 g_dev) == sk
    assert _same(g_dev, g_ref)
    wo_ref = np.zeros(wl.num_points, np.float32)
    wo_dev = wo_ref.copy()
    oracle.grid_weights_read_f(wl.uu, wl.vv, w, wo_ref, wl.cell_size_rad, G,
                               g_ref)
    fg.grid_weights_read(wl.uu, wl.vv, w, wo_dev, wl.cell_size_rad, G, g_dev)
    assert _same(wo_dev, wo_ref)


def test_weights_unit_weights_are_counts():
    wl = synth.random_small(n=50_000, grid_size=128, num_planes=3, seed=6)
    G = wl.grid_size
    g = np.zeros(G * G, np.float32)
    sk = fg.grid_weights_write(wl.uu, wl.vv, np.ones(wl.num_points), 
                               wl.cell_size_rad, G, g)
    scale = np.float32(G) * np.float32(wl.cell_size_rad)
    u = np.round(-wl.uu * scale).astype(np.int64) + G // 2
    v = np.round(wl.vv * scale).astype(np.int64) + G // 2
    ok = (u >= 0) & (u < G) & (v >= 0) & (v < G)
    # np.round ties to even; the reference rounds half away from zero, and
    # random float positions hit no exact tie here.
    hist = np.bincount((v * G + u)[ok], minlength=G * G)
    assert sk == int((~ok).sum())
    assert np.array_equal(g, hist.astype(np.float32))


def test_weights_device_api_and_empty():
    wl = synth.random_small(n=20_000, grid_size=96, num_planes=3, seed=7)
    G = wl.grid_size
    rng = np.random.default_rng(8)
    w = rng.uniform(0.5, 1.0, wl.num_points).astype(np.float32)
    g_host = np.zeros(G * G, np.float32)
    sk_host = fg.grid_weights_write(wl.uu, wl.vv, w, wl.cell_size_rad, G,
                                    g_host)
    d = [fg.DeviceArray.from_host(a) for a in (wl.uu, wl.vv, w)]
    dg = fg.DeviceArray(4 * G * G)
    dg.zero()
    L = fg.lib()
    st = fg.ct.c_int(0)
    sk = fg.ct.c_size_t(0)
    L.gw_grid_weights_write_device(fg.GW_SINGLE, wl.num_points, d[0].ptr,
                                   d[1].ptr, d[2].ptr,
                                   float(wl.cell_size_rad), G,
                                   fg.ct.byref(sk), dg.ptr, None,
                                   fg.ct.byref(st))
    assert st.value == 0 and sk.value == sk_host
    assert _same(dg.download(np.empty(G * G, np.float32)), g_host)
    # Empty input: nothing skipped, grid unchanged.
    g0 = g_host.copy()
    e = np.zeros(0, np.float32)
    assert fg.grid_weights_write(e, e, e, wl.cell_size_rad, G, g0) == 0
    assert _same(g0, g_host)
    for x in (*d, dg):
        x.close()


def test_weights_full_scale_positions():
    """The 512-station workload's 13,081,600 positions on its 8192^2 grid
    with its weights (uniform in [0.5, 1]): bit-exact write and read."""
    wl = synth.full_scale()
    G = wl.grid_size
    g_ref = np.zeros(G * G, np.float32)
    g_dev = np.zeros(G * G, np.float32)
    sk = oracle.grid_weights_write_f(wl.uu, wl.vv, wl.weight,
                                     wl.cell_size_rad, G, g_ref)
    assert fg.grid_weights_write(wl.uu, wl.vv, wl.weight, wl.cell_size_rad,
                                 G, g_dev) == sk
    assert _same(g_dev, g_ref)
    wo_ref = np.zeros(wl.num_points, np.float32)
    wo_dev = np.zeros(wl.num_points, np.float32)
    oracle.grid_weights_read_f(wl.uu, wl.vv, wl.weight, wo_ref,
                               wl.cell_size_rad, G, g_ref)
    fg.grid_weights_read(wl.uu, wl.vv, wl.weight, wo_dev, wl.cell_size_rad,
                         G, g_dev)
    assert _same(wo_dev, wo_ref)
