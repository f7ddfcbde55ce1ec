"""GW_SINGLE_INDEX64 plans: FP32 data, taps, products and grid with the
index math of oskar_grid_wproj_d (oskar_grid_wproj.c:52-112), through the C
ABI, alone, in batches and through the server.

Tolerances against the FP64 oracle (gridref_oskar_grid_wproj_d on the same
float inputs widened):

* num_skipped: bit-exact (the same FP64 index decisions);
* norm: relative 1e-12 (FP64 sums of the same taps, another order);
* grid: within 1e-6 * max|grid| (only FP32 products and FP32 sums of
  position runs differ from the oracle's FP64 arithmetic).
"""
import dataclasses

import numpy as np
import pytest

import oracle
import fpga_gridding_amd as fg
from fpga_gridding_amd import synth

from parity import max_rel

pytestmark = pytest.mark.gpu

GiB = 1 << 30


@pytest.fixture(scope="module", autouse=True)
def _gpu_present():
    if fg.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    yield
    fg.lib().gw_release_cached()


CASES = [
    ("random", dict(n=3000, grid_size=96, num_planes=6, s_min=1, s_max=6,
                    seed=31)),
    ("random-edges", dict(n=20000, grid_size=256, num_planes=16, s_min=2,
                          s_max=14, seed=32, edge_fraction=0.2)),
]


def _check(wl, g, sk, nrm):
    g_d, sk_d, n_d = oracle.grid_wproj_d(wl)
    assert sk == sk_d
    assert abs(nrm - n_d) <= 1e-12 * max(abs(n_d), 1.0)
    if np.any(g_d):
        assert max_rel(g, g_d) <= 1e-6
    else:
        assert not np.any(g)
    return max_rel(g, g_d) if np.any(g_d) else 0.0


@pytest.mark.parametrize("name,case", CASES, ids=[c[0] for c in CASES])
def test_index64_matches_fp64_oracle(name, case):
    wl = synth.random_small(**case)
    _check(wl, *fg.grid_workload_index64(wl))


def test_index64_snapshot64():
    """The 64-antenna snapshot workload (725,760 points, grid 2048^2)."""
    wl = synth.snapshot_small()
    err = _check(wl, *fg.grid_workload_index64(wl))
    print(f"index64 snapshot: max_rel vs FP64 oracle {err:.3e}")


def test_index64_edge_cases():
    """Ties, grid edges, w = +-0, NaN, +-inf and huge values take the FP64
    oracle's plane and skip decisions."""
    wl = synth.edge_cases()
    _check(wl, *fg.grid_workload_index64(wl))


def test_index64_follows_fp64_where_fp32_indices_flip():
    """Coordinates a few ulps from rounding ties of the position: for about
    one point in ten the FP32 product rounds to the other cell.  The FP64
    form follows the FP64 oracle there, the FP32 form the FP32 one, and the
    two results differ."""
    wl = synth.random_small(n=20000, grid_size=100, num_planes=8, s_min=1,
                            s_max=5, seed=33)
    cell = np.float32(wl.cell_size_rad)
    scale32 = np.float32(wl.grid_size) * cell
    scale64 = np.float64(wl.grid_size) * np.float64(cell)
    rng = np.random.default_rng(36)
    k = rng.integers(-30, 30, wl.num_points)
    uu = (-(k + 0.5) / scale64).astype(np.float32)
    uu = uu + (rng.integers(-3, 4, uu.size) *
               np.spacing(uu)).astype(np.float32)
    wl = dataclasses.replace(wl, uu=uu)

    def rnd(x):
        return np.sign(x) * np.floor(np.abs(x) + 0.5)
    flips = rnd((-uu * scale32).astype(np.float64)) != \
        rnd(-uu.astype(np.float64) * scale64)
    assert flips.sum() > wl.num_points // 20
    _check(wl, *fg.grid_workload_index64(wl))
    g_f, sk_f, _ = fg.grid_workload_f(wl)
    g_x, sk_x, _ = oracle.grid_wproj_f_exact(wl)
    assert sk_f == sk_x
    assert max_rel(g_f, g_x) <= 1e-5
    g_d, _, _ = oracle.grid_wproj_d(wl)
    assert max_rel(g_x, g_d) > 1e-3


def test_index64_batch_matches_plans():
    base = synth.snapshot_small()
    jobs, _, _ = synth.batch_mix(n_jobs=5, seed=34, n_range=(1500, 40000),
                                 grids=(256, 512), n_kernel_sets=2,
                                 base=base,
                                 fov_deg_at_8192=2.0 * 2048 / 8192)
    plans, dev, ref = {}, [], []
    try:
        for wl in jobs:
            key = (wl.params["kernel_set"], wl.grid_size)
            if key not in plans:
                plans[key] = fg.Plan.for_workload(wl, fg.GW_SINGLE_INDEX64)
            d = [fg.DeviceArray.from_host(a)
                 for a in (wl.uu, wl.vv, wl.ww, wl.vis)]
            g = fg.DeviceArray(8 * wl.grid_size ** 2)
            dev.append((d, g))
        for j, wl in enumerate(jobs):
            d, g = dev[j]
            g.zero()
            pl = plans[(wl.params["kernel_set"], wl.grid_size)]
            pl.grid(wl.num_points, *d, g)
            r = pl.results()
            ref.append((g.download(np.empty(2 * wl.grid_size ** 2,
                                            np.float32)), r))
            g_d, sk_d, _ = oracle.grid_wproj_d(wl)
            assert r[0] == sk_d
            assert max_rel(ref[-1][0], g_d) <= 1e-6
        b = fg.Batch()
        for j, wl in enumerate(jobs):
            d, g = dev[j]
            g.zero()
            b.add(plans[(wl.params["kernel_set"], wl.grid_size)],
                  wl.num_points, *d, g)
        b.run()
        for j, wl in enumerate(jobs):
            sk, nd, nu = b.results(j)
            assert (sk, nu) == (ref[j][1][0], ref[j][1][2])
            gj = dev[j][1].download(np.empty(2 * wl.grid_size ** 2,
                                             np.float32))
            assert max_rel(gj, ref[j][0]) <= 2e-6
        b.close()
    finally:
        for d, g in dev:
            for x in (*d, g):
                x.close()
        for p in plans.values():
            p.close()


def test_index64_server_matches_fp64_oracle():
    base = synth.snapshot_small()
    jobs, _, sets = synth.batch_mix(n_jobs=6, seed=35, n_range=(1500, 40000),
                                    grids=(256, 512), n_kernel_sets=2,
                                    base=base,
                                    fov_deg_at_8192=2.0 * 2048 / 8192)
    sv = fg.Server(8 * GiB, max_batch_jobs=4, num_streams=2,
                   precision=fg.GW_SINGLE_INDEX64)
    try:
        ids = [sv.add_kernel_set(k.num_w_planes, k.support, k.oversample,
                                 k.conv_size_half, k.conv_func) for k in sets]
        rids = [sv.submit(ids[wl.params["kernel_set"]], wl.grid_size,
                          wl.cell_size_rad, wl.w_scale, wl.uu, wl.vv, wl.ww,
                          wl.vis) for wl in jobs]
        for wl, rid in zip(jobs, rids):
            res, g = sv.wait(rid, timeout=60.0)
            _check(wl, g, res.num_skipped, res.norm_delta)
    finally:
        sv.close()
