"""W-kernel cubes generated on the device (gw_wkernel_generate), plans built
from a cube in device memory (gw_plan_create_device_cube) and server kernel
sets given by their parameters (gw_server_add_wkernel_set).

Checks:
* the device cube against the host construction the workloads use
  (synth.wkernel_cube, numpy FP64, rounded to float): every value within
  two float ulps (the FP64 values differ only in the last bits: summation
  order of the normalising sum, libm vs device cos/sin);
* closed form: each plane sums to 1 + 0i over its footprint taps at offset
  zero (the normalisation), to float rounding;
* a plan built on the device from a cube grids exactly as a plan built on
  the host from the same cube (skips and norm identical, grid within
  2e-6 * max|grid|: the same tables, FP32 sums in another order);
* a server kernel set generated on the device gives the results of the
  same set uploaded from the host.
"""
import dataclasses

import numpy as np
import pytest

import fpga_gridding_amd as fg
from fpga_gridding_amd import synth

from parity import assert_fp32_parity, fp32_report, max_rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu_present():
    if fg.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    yield
    fg.lib().gw_release_cached()


def _params(wl):
    return wl.support, wl.oversample, wl.conv_size_half, wl.params["w_max"]


@pytest.mark.parametrize("make", [synth.snapshot_small, synth.full_scale],
                         ids=["snapshot64", "ska512"])
def test_generated_cube_matches_host_construction(make):
    wl = make()
    sup, ov, C, w_max = _params(wl)
    dev = fg.wkernel_generate(sup, ov, C, w_max)
    ref = synth.cube_to_conv_func(synth.wkernel_cube(sup, ov, C, w_max))
    assert np.array_equal(ref, wl.conv_func)
    d = np.abs(dev.astype(np.float64) - ref.astype(np.float64))
    ulp = np.spacing(np.abs(ref)).astype(np.float64)
    assert np.all(d <= 2 * ulp + 1e-40)
    print(f"{wl.name}: {np.count_nonzero(d)} of {d.size} values differ "
          f"(max {np.max(d / (ulp + 1e-40)):.0f} ulp)")


def test_generated_planes_have_unit_sum():
    sup = synth.support_table(24, 3, 20)
    ov, C = 4, synth.cube_half_size(4, 20)
    cube = fg.wkernel_generate(sup, ov, C, 1234.5).astype(np.float64)
    cube = cube.reshape(len(sup), C, C, 2)
    for p, s in enumerate(sup):
        t = np.abs(np.arange(-s, s + 1)) * ov
        tot = cube[p][np.ix_(t, t)].sum(axis=(0, 1))
        assert abs(tot[0] - 1.0) <= 1e-5 and abs(tot[1]) <= 1e-5, p


def test_plan_from_device_cube_grids_as_host_plan():
    wl = synth.random_small(n=30000, grid_size=256, num_planes=12, s_min=2,
                            s_max=14, seed=41, edge_fraction=0.05)
    sup, ov, C = wl.support, wl.oversample, wl.conv_size_half
    d_cube = fg.wkernel_generate_device(sup, ov, C, wl.params["w_max"])
    host_cube = d_cube.download(np.empty(2 * len(sup) * C * C, np.float32))
    d = [fg.DeviceArray.from_host(a) for a in (wl.uu, wl.vv, wl.ww, wl.vis)]
    G = wl.grid_size
    out = []
    plans = [fg.Plan.from_device_cube(fg.GW_SINGLE, len(sup), sup, ov, C,
                                      d_cube, wl.cell_size_rad, wl.w_scale,
                                      G),
             fg.Plan(fg.GW_SINGLE, len(sup), sup, ov, C, host_cube,
                     wl.cell_size_rad, wl.w_scale, G)]
    try:
        for pl in plans:
            g = fg.DeviceArray(8 * G * G)
            g.zero()
            pl.grid(wl.num_points, *d, g)
            out.append((g.download(np.empty(2 * G * G, np.float32)),
                        pl.results()))
            g.close()
        (g0, r0), (g1, r1) = out
        assert r0 == r1
        assert max_rel(g0, g1) <= 2e-6
    finally:
        for p in plans:
            p.close()
        for x in (*d, d_cube):
            x.close()


def test_plan_from_device_cube_asymmetric():
    """Tables built on the device from a random cube (no symmetry) grid as
    the host-built tables of the same cube, and the grid meets the FP32
    parity bar against the oracle."""
    wl = synth.random_small(n=8000, grid_size=96, num_planes=5, s_min=1,
                            s_max=7, ov=4, seed=43, edge_fraction=0.05)
    rng = np.random.default_rng(44)
    cube = rng.standard_normal(wl.conv_func.shape).astype(np.float32)
    wl = dataclasses.replace(wl, conv_func=cube)
    sup, ov, C = wl.support, wl.oversample, wl.conv_size_half
    d_cube = fg.DeviceArray.from_host(cube)
    d = [fg.DeviceArray.from_host(a) for a in (wl.uu, wl.vv, wl.ww, wl.vis)]
    G = wl.grid_size
    out = []
    plans = [fg.Plan.from_device_cube(fg.GW_SINGLE, len(sup), sup, ov, C,
                                      d_cube, wl.cell_size_rad, wl.w_scale,
                                      G),
             fg.Plan(fg.GW_SINGLE, len(sup), sup, ov, C, cube,
                     wl.cell_size_rad, wl.w_scale, G)]
    try:
        for pl in plans:
            g = fg.DeviceArray(8 * G * G)
            g.zero()
            pl.grid(wl.num_points, *d, g)
            out.append((g.download(np.empty(2 * G * G, np.float32)),
                        pl.results()))
            g.close()
        (g0, r0), (g1, r1) = out
        assert r0 == r1
        assert max_rel(g0, g1) <= 2e-6
        assert_fp32_parity(fp32_report(g0, wl))
    finally:
        for p in plans:
            p.close()
        for x in (*d, d_cube):
            x.close()


def test_server_generated_kernel_set():
    wl = synth.random_small(n=20000, grid_size=256, num_planes=10, s_min=2,
                            s_max=12, seed=42, edge_fraction=0.05)
    sup, ov, C, w_max = _params(wl)
    host_cube = fg.wkernel_generate(sup, ov, C, w_max)
    sv = fg.Server(4 << 30, max_batch_jobs=4, num_streams=2)
    try:
        k_dev = sv.add_wkernel_set(sup, ov, C, w_max)
        k_host = sv.add_kernel_set(len(sup), sup, ov, C, host_cube)
        res = []
        for k in (k_dev, k_host):
            rid = sv.submit(k, wl.grid_size, wl.cell_size_rad, wl.w_scale,
                            wl.uu, wl.vv, wl.ww, wl.vis)
            res.append(sv.wait(rid, timeout=60.0))
        (a, ga), (b, gb) = res
        assert (a.num_skipped, a.num_updates) == (b.num_skipped, b.num_updates)
        assert a.norm_delta == b.norm_delta
        assert max_rel(ga, gb) <= 2e-6
        assert sv.stats()["resident_plans"] == 2
    finally:
        sv.close()
