"""gw_grid_device_noalloc, the allocation-free gridding call, through the C
ABI: its results equal gw_grid_device's on the same plan and inputs (skip
count, update count and norm identical; grid within 2e-6 * max|grid|, the
same FP32 products and FP64 per-item sums added in another order), on two
calls in a row, the second with fewer points (the static scratch keeps the
first call's contents); a call beyond the scratch's capacity fails with
OSKAR_ERR_OUT_OF_RANGE and leaves the grid untouched."""
import numpy as np
import pytest

import fpga_gridding_amd as fg
from fpga_gridding_amd import synth

from parity import max_rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu_present():
    if fg.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    yield
    fg.lib().gw_release_cached()


def _call(plan, fn, wl, n, G):
    d = [fg.DeviceArray.from_host(np.ascontiguousarray(a[:k]))
         for a, k in ((wl.uu, n), (wl.vv, n), (wl.ww, n), (wl.vis, 2 * n))]
    dg = fg.DeviceArray(8 * G * G)
    try:
        dg.zero()
        fn(n, *d, dg)
        res = plan.results()
        return dg.download(np.empty(2 * G * G, np.float32)), res
    finally:
        for x in (*d, dg):
            x.close()


def test_noalloc_matches_grid_device_twice():
    wl = synth.snapshot_small()
    G = wl.grid_size
    plan = fg.Plan.for_workload(wl)
    try:
        for n in (500_000, 120_000):           # second call smaller
            g_ref, r_ref = _call(plan, plan.grid, wl, n, G)
            g, r = _call(plan, plan.grid_noalloc, wl, n, G)
            assert r[0] == r_ref[0] and r[2] == r_ref[2], n
            assert abs(r[1] - r_ref[1]) <= 1e-12 * abs(r_ref[1]), n
            assert max_rel(g, g_ref) <= 2e-6, n
    finally:
        plan.close()


def test_noalloc_rejects_more_points_than_its_scratch():
    wl = synth.random_small(n=10, grid_size=64, num_planes=3, seed=3)
    plan = fg.Plan.for_workload(wl)
    n = 4_000_000                              # > 64 MiB of workspace
    d = [fg.DeviceArray(4 * n) for _ in range(3)] + [fg.DeviceArray(8 * n)]
    dg = fg.DeviceArray(8 * 64 * 64)
    try:
        dg.zero()
        with pytest.raises(fg.GridderError) as e:
            plan.grid_noalloc(n, *d, dg)
        assert e.value.status == -5            # OSKAR_ERR_OUT_OF_RANGE
        assert "points per call" in str(e.value)
        assert not np.any(dg.download(np.empty(2 * 64 * 64, np.float32)))
    finally:
        for x in (*d, dg):
            x.close()
        plan.close()
