"""The precision, batched and served forms at their own workload sizes,
through the C ABI, against the oracle.

* Reduced precision (GW_HALF, GW_BFLOAT16) on the 512-station workload
  (13,081,600 points, grid 8192^2, 128 planes): skip count against the FP64
  index math (oskar_grid_wproj.c:71-101), exact update count, norm against
  the closed form on the stored taps (relative 1e-12), grid checksum against
  the closed form on the stored inputs (1e-6 of sum |v| |taps|), and 41
  spot cells against the FP64 brute force on the stored inputs
  (oracle.cells_d): within 1e-5 * max|grid|.
* Batched: the 64-job mix (10k-5M points per job, grids 1024^2-8192^2, 8
  kernel sets of 16-128 planes, support up to 36) in one pipeline, eager and
  as a replayed HIP graph: every job's skip count and update count equal the
  oracle's FP32 index stage, norm and checksum match the closed forms, spot
  cells match the FP32-index brute force (oracle.cells_f, float64 sums)
  within 1e-5 * max|grid|; replays agree with the eager run to 2e-6.
* Served: the same 64 jobs as requests arriving as a Poisson stream at 200
  requests/s: each result checked as above (counts, norm, checksum), a sample
  also cell by cell against the oracle and against the request served alone.

Every job keeps the semantics of oskar_grid_wproj.c:128-243 (full-length
loop, weights ignored).
"""
import dataclasses
import threading
import time

import numpy as np
import pytest

import oracle
import fpga_gridding_amd as fg
from fpga_gridding_amd import synth

from parity import expected_sums, max_rel, tap_tables

pytestmark = pytest.mark.gpu

GiB = 1 << 30


@pytest.fixture(scope="module", autouse=True)
def _gpu_present():
    if fg.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    yield
    fg.lib().gw_release_cached()


def _spot_cells(g, n, seed):
    """The largest cell and n - 1 random covered cells (indices v*G + u)."""
    mag = np.hypot(g[0::2], g[1::2])
    covered = np.flatnonzero(mag > 0)
    rng = np.random.default_rng(seed)
    pick = rng.choice(covered, min(n - 1, covered.size), replace=False)
    return np.unique(np.concatenate([[int(np.argmax(mag))], pick])), mag.max()


def _cells_of(g, cells):
    return np.stack([g[2 * cells], g[2 * cells + 1]], 1).reshape(-1)


def _checksum(g):
    return complex(np.sum(g[0::2], dtype=np.float64),
                   np.sum(g[1::2], dtype=np.float64))


# --------------------------------------------------------------------------
# Reduced precision at full scale
# --------------------------------------------------------------------------

@pytest.fixture(scope="module")
def full_scale():
    wl = synth.full_scale()
    rec, U = oracle.locate_d(wl)
    return wl, rec, U


@pytest.mark.parametrize("prec,name", [(fg.GW_HALF, "fp16"),
                                       (fg.GW_BFLOAT16, "bf16")],
                         ids=["fp16", "bf16"])
def test_reduced_precision_full_scale(full_scale, prec, name):
    wl, rec, U = full_scale
    G = wl.grid_size
    stored = dataclasses.replace(
        wl, vis=fg.from_storage(fg.to_storage(wl.vis, prec), prec),
        conv_func=fg.from_storage(fg.to_storage(wl.conv_func, prec), prec))
    plan = fg.Plan(prec, wl.num_w_planes, wl.support, wl.oversample,
                   wl.conv_size_half, wl.conv_func,
                   float(np.float32(wl.cell_size_rad)),
                   float(np.float32(wl.w_scale)), G)
    d = [fg.DeviceArray.from_host(a) for a in (wl.uu, wl.vv, wl.ww)]
    d_vis = fg.DeviceArray.from_host(fg.to_storage(wl.vis, prec))
    d_comp = fg.DeviceArray(16 * G * G)
    d_out = fg.DeviceArray(8 * G * G)
    try:
        d_comp.zero()
        plan.grid(wl.num_points, *d, d_vis, d_comp)
        sk, nd, nu = plan.results()
        fg.resolve_compensated(d_comp, G * G, d_out)
        g = d_out.download(np.empty(2 * G * G, np.float32))
    finally:
        for x in (*d, d_vis, d_comp, d_out):
            x.close()
        plan.close()
    assert sk == int(np.sum(rec["skip"] != 0))
    assert nu == U
    norm, checksum, scale = expected_sums(stored, rec)
    assert abs(nd - norm) <= 1e-12 * abs(norm)
    assert abs(_checksum(g) - checksum) <= 1e-6 * scale
    cells, gmax = _spot_cells(g, 41, seed=7)
    ref = oracle.cells_d(stored, cells)
    err_stored = float(np.abs(_cells_of(g, cells) - ref).max() / gmax)
    assert err_stored <= 1e-5
    ref_orig = oracle.cells_d(wl, cells)
    err_orig = float(np.abs(_cells_of(g, cells) - ref_orig).max() / gmax)
    print(f"{name} full scale: spot cells vs FP64 oracle on stored inputs "
          f"{err_stored:.2e}, on the original inputs {err_orig:.2e}")


@pytest.mark.parametrize("prec,name", [(fg.GW_HALF, "fp16"),
                                       (fg.GW_BFLOAT16, "bf16")],
                         ids=["fp16", "bf16"])
def test_compensated_grid_beats_fp32_grid(full_scale, prec, name):
    """What the compensated two-float grid of a 16-bit plan buys: on the same
    stored (16-bit) inputs, an FP32 grid from GW_SINGLE_INDEX64 (the same
    FP64 index math, FP32 products and FP32 position-run sums, FP64 region)
    and the compensated grid (hi + lo read in float64) against the FP64
    brute force on those inputs (oracle.cells_d), over 256 covered cells:
    the compensated grid's RMS error is below the FP32 grid's, because it
    drops the rounding of every region add into a float cell."""
    wl, _, _ = full_scale
    G = wl.grid_size
    stored = dataclasses.replace(
        wl, vis=fg.from_storage(fg.to_storage(wl.vis, prec), prec),
        conv_func=fg.from_storage(fg.to_storage(wl.conv_func, prec), prec))
    args = (wl.num_w_planes, wl.support, wl.oversample, wl.conv_size_half)
    cell = float(np.float32(wl.cell_size_rad))
    wsc = float(np.float32(wl.w_scale))
    d = [fg.DeviceArray.from_host(a) for a in (wl.uu, wl.vv, wl.ww)]
    p16 = fg.Plan(prec, *args, wl.conv_func, cell, wsc, G)
    p32 = fg.Plan(fg.GW_SINGLE_INDEX64, *args,
                  np.ascontiguousarray(stored.conv_func, np.float32), cell,
                  wsc, G)
    d_v16 = fg.DeviceArray.from_host(fg.to_storage(wl.vis, prec))
    d_v32 = fg.DeviceArray.from_host(
        np.ascontiguousarray(stored.vis, np.float32))
    d_comp = fg.DeviceArray(16 * G * G)
    d_g32 = fg.DeviceArray(8 * G * G)
    try:
        d_comp.zero()
        d_g32.zero()
        p16.grid(wl.num_points, *d, d_v16, d_comp)
        r16 = p16.results()
        p32.grid(wl.num_points, *d, d_v32, d_g32)
        r32 = p32.results()
        comp = d_comp.download(np.empty(4 * G * G, np.float32))
        g32 = d_g32.download(np.empty(2 * G * G, np.float32))
    finally:
        for x in (*d, d_v16, d_v32, d_comp, d_g32):
            x.close()
        p16.close()
        p32.close()
    assert r16[0] == r32[0] and r16[2] == r32[2]
    cells, gmax = _spot_cells(g32, 256, seed=3)
    ref = oracle.cells_d(stored, cells)
    q = comp.reshape(-1, 4)[cells].astype(np.float64)
    v16 = np.stack([q[:, 0] + q[:, 1], q[:, 2] + q[:, 3]], 1).reshape(-1)
    e16 = float(np.sqrt(np.mean((v16 - ref) ** 2)) / gmax)
    e32 = float(np.sqrt(np.mean((_cells_of(g32, cells) - ref) ** 2)) / gmax)
    print(f"{name} stored inputs, RMS error over 256 cells / max|grid|: "
          f"compensated {e16:.3e}, FP32 grid {e32:.3e}")
    assert e16 < e32


# --------------------------------------------------------------------------
# The 64-job mix: batched and served
# --------------------------------------------------------------------------

@pytest.fixture(scope="module")
def mix():
    """The 64 jobs, their kernel sets, and the oracle's per-job counts and
    closed forms (tap tables shared per kernel set)."""
    base = synth.full_scale()
    jobs, _, sets = synth.batch_mix(64, seed=7, base=base)
    tabs = {}
    expect = []
    for wl in jobs:
        k = wl.params["kernel_set"]
        if k not in tabs:
            tabs[k] = tap_tables(wl)
        rec, U = oracle.locate_f(wl)
        norm, checksum, scale = expected_sums(wl, rec, tabs[k])
        expect.append(dict(skipped=int(np.sum(rec["skip"] != 0)), U=U,
                           norm=norm, checksum=checksum, scale=scale))
    return jobs, sets, expect


def _check_job(wl, ex, sk, nd, nu, g, spot=0, seed=0):
    assert sk == ex["skipped"] and nu == ex["U"], wl.name
    assert abs(nd - ex["norm"]) <= 1e-12 * max(abs(ex["norm"]), 1.0), \
        wl.name
    assert abs(_checksum(g) - ex["checksum"]) <= 1e-6 * ex["scale"], wl.name
    if spot and np.any(g):
        cells, gmax = _spot_cells(g, spot, seed)
        ref = oracle.cells_f(wl, cells)
        assert np.abs(_cells_of(g, cells) - ref).max() <= 1e-5 * gmax, \
            wl.name


def test_batched_mix64_against_oracle(mix):
    jobs, sets, expect = mix
    plans = {}
    for wl in jobs:
        key = (wl.params["kernel_set"], wl.grid_size)
        if key not in plans:
            plans[key] = fg.Plan.for_workload(wl)
    dev = []
    b = fg.Batch()
    try:
        for j, wl in enumerate(jobs):
            d = [fg.DeviceArray.from_host(a)
                 for a in (wl.uu, wl.vv, wl.ww, wl.vis)]
            g = fg.DeviceArray(8 * wl.grid_size ** 2)
            g.zero()
            dev.append((d, g))
            b.add(plans[(wl.params["kernel_set"], wl.grid_size)],
                  wl.num_points, *d, g)

        def grid(j):
            G = jobs[j].grid_size
            return dev[j][1].download(np.empty(2 * G * G, np.float32))

        b.run()
        eager = []
        for j, wl in enumerate(jobs):
            sk, nd, nu = b.results(j)
            g = grid(j)
            _check_job(wl, expect[j], sk, nd, nu, g, spot=4, seed=j)
            eager.append((sk, nd, nu, _checksum(g),
                          g[::4097].astype(np.float64)))
        # The same pipeline as a graph, replayed twice from cleared grids.
        b.capture()
        for rep in range(2):
            for _, g in dev:
                g.zero()
            fg.device_synchronize()
            b.launch()
            for j, wl in enumerate(jobs):
                sk, nd, nu = b.results(j)
                e = eager[j]
                assert (sk, nu) == (e[0], e[2]), (rep, j)
                assert abs(nd - e[1]) <= 1e-12 * max(abs(e[1]), 1.0)
                g = grid(j)
                assert abs(_checksum(g) - e[3]) <= 1e-6 * expect[j]["scale"]
                s = g[::4097].astype(np.float64)
                sc = np.abs(e[4]).max()
                if sc > 0:
                    assert np.abs(s - e[4]).max() <= 2e-6 * max(
                        sc, np.abs(g).max()), (rep, j)
    finally:
        b.close()
        for d, g in dev:
            for x in (*d, g):
                x.close()
        for p in plans.values():
            p.close()


def test_served_mix64_poisson_200(mix):
    jobs, sets, expect = mix
    rng = np.random.default_rng(11)
    order = rng.permutation(len(jobs))
    arrive = np.cumsum(rng.exponential(1.0 / 200.0, len(jobs)))
    sample = set(sorted(range(len(jobs)),
                        key=lambda j: jobs[j].num_points)[:3]) | {
        int(np.argsort([w.num_points for w in jobs])[len(jobs) // 2])}
    sv = fg.Server(64 * GiB, max_batch_jobs=16, num_streams=2)
    kept = {}
    try:
        ids = [sv.add_kernel_set(k.num_w_planes, k.support, k.oversample,
                                 k.conv_size_half, k.conv_func) for k in sets]

        def submit(j):
            wl = jobs[j]
            return sv.submit(ids[wl.params["kernel_set"]], wl.grid_size,
                             wl.cell_size_rad, wl.w_scale, wl.uu, wl.vv,
                             wl.ww, wl.vis)

        rids = [None] * len(jobs)

        def producer():
            t0 = time.perf_counter()
            for r, j in enumerate(order):
                wait = t0 + arrive[r] - time.perf_counter()
                if wait > 0:
                    time.sleep(wait)
                rids[r] = submit(int(j))

        th = threading.Thread(target=producer)
        th.start()
        for r, j in enumerate(order):
            j = int(j)
            while rids[r] is None:
                time.sleep(1e-4)
            res, g = sv.wait(rids[r], timeout=120.0)
            _check_job(jobs[j], expect[j], res.num_skipped, res.norm_delta,
                       res.num_updates, g, spot=3 if j in sample else 0,
                       seed=100 + j)
            if j in sample:
                kept[j] = (res.num_skipped, res.norm_delta, res.num_updates,
                           g)
        th.join()
        st = sv.stats()
        assert st["completed"] >= len(jobs) and st["pending"] == 0
        # The sampled requests alone: the same result as in the mix.
        for j, (sk, nd, nu, g) in kept.items():
            res, ga = sv.wait(submit(j), timeout=120.0)
            assert (res.num_skipped, res.num_updates) == (sk, nu)
            assert abs(res.norm_delta - nd) <= 1e-12 * max(abs(nd), 1.0)
            assert max_rel(g, ga) <= 2e-6
        # The smallest requests against the exact FP32 result, every cell.
        for j in sorted(kept, key=lambda i: jobs[i].num_points)[:2]:
            g_x, sk_x, _ = oracle.grid_wproj_f_exact(jobs[j])
            assert kept[j][0] == sk_x
            assert max_rel(kept[j][3], g_x) <= 1e-5
    finally:
        sv.close()
