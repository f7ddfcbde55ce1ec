"""The serving loop (gw_server_*), through the C ABI.

Requests are the arguments of oskar_grid_wproj_f on host arrays; the server
admits them by its device-memory budget and grids admitted groups as
batches on several streams.  A request's result must not depend on what it
was gridded with:

* num_skipped and the update count: identical alone and in the mix;
* norm increment: relative 1e-12 (FP64 sums of the same taps);
* grid: within 2e-6 * max|grid| (FP32 products, FP64 per-job sums, added in
  another order), and within 1e-5 * max|grid| of the exact value of the FP32
  gridder's result (oracle.grid_wproj_f_exact).
"""
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


def _mix(n_jobs=9, seed=21):
    base = synth.snapshot_small()
    return synth.batch_mix(n_jobs=n_jobs, seed=seed, n_range=(1500, 60000),
                           grids=(256, 512, 1024), n_kernel_sets=3,
                           base=base, fov_deg_at_8192=2.0 * 2048 / 8192)


def _register(server, sets):
    return [server.add_kernel_set(k.num_w_planes, k.support, k.oversample,
                                  k.conv_size_half, k.conv_func)
            for k in sets]


def _submit(server, ids, wl, grid_id=-1):
    return server.submit(ids[wl.params["kernel_set"]], wl.grid_size,
                         wl.cell_size_rad, wl.w_scale, wl.uu, wl.vv, wl.ww,
                         wl.vis, grid_id=grid_id)


def _alone(jobs, sets, budget=8 * GiB):
    """Each request submitted and waited for on its own."""
    sv = fg.Server(budget, max_batch_jobs=8, num_streams=2)
    try:
        ids = _register(sv, sets)
        out = []
        for wl in jobs:
            res, g = sv.wait(_submit(sv, ids, wl), timeout=60.0)
            out.append((res.num_skipped, res.norm_delta, res.num_updates, g))
        return out
    finally:
        sv.close()


def _check(got, ref):
    for j, ((sk, nd, nu, g), (sk_r, nd_r, nu_r, g_r)) in enumerate(
            zip(got, ref)):
        assert (sk, nu) == (sk_r, nu_r), j
        assert abs(nd - nd_r) <= 1e-12 * max(abs(nd_r), 1.0), j
        if np.any(g_r):
            assert max_rel(g, g_r) <= 2e-6, j
        else:
            assert not np.any(g), j


def test_server_alone_matches_exact_fp32():
    """Requests served one at a time agree with the oracle's exact FP32
    result (the reference's index math and taps, float64 arithmetic)."""
    jobs, _, sets = _mix(n_jobs=3, seed=22)
    alone = _alone(jobs, sets)
    for wl, (sk, nd, nu, g) in zip(jobs, alone):
        g_x, sk_x, n_x = oracle.grid_wproj_f_exact(wl)
        assert sk == sk_x
        assert abs(nd - n_x) <= 1e-12 * abs(n_x)
        assert max_rel(g, g_x) <= 1e-5


def test_server_same_alone_and_in_mix():
    """All requests submitted at once (continuous batching over two
    streams, at most four per batch) give each request's result alone."""
    jobs, _, sets = _mix()
    ref = _alone(jobs, sets)
    sv = fg.Server(8 * GiB, max_batch_jobs=4, num_streams=2)
    try:
        ids = _register(sv, sets)
        rids = [_submit(sv, ids, wl) for wl in jobs]
        got = []
        for rid in rids:
            res, g = sv.wait(rid, timeout=60.0)
            assert res.t_submit <= res.t_admit <= res.t_start <= res.t_done
            got.append((res.num_skipped, res.norm_delta, res.num_updates, g))
        _check(got, ref)
        st = sv.stats()
        assert st["completed"] == len(jobs) and st["pending"] == 0
        assert st["max_batch"] <= 4
        assert st["resident_plans"] == len({(wl.params["kernel_set"],
                                             wl.grid_size) for wl in jobs})
    finally:
        sv.close()


def test_server_pinned_requests_match():
    """Request points in page-locked host memory (gw_host_malloc), all
    submitted at once, give each request's result alone from pageable
    arrays."""
    jobs, _, sets = _mix(n_jobs=5, seed=24)
    ref = _alone(jobs, sets)
    pinned = [[fg.PinnedArray.copy_of(a) for a in
               (wl.uu, wl.vv, wl.ww, wl.vis)] for wl in jobs]
    sv = fg.Server(8 * GiB, max_batch_jobs=4, num_streams=2)
    try:
        ids = _register(sv, sets)
        rids = [sv.submit(ids[wl.params["kernel_set"]], wl.grid_size,
                          wl.cell_size_rad, wl.w_scale,
                          *(p.array for p in ps))
                for wl, ps in zip(jobs, pinned)]
        got = []
        for rid in rids:
            res, g = sv.wait(rid, timeout=60.0)
            got.append((res.num_skipped, res.norm_delta, res.num_updates, g))
        _check(got, ref)
    finally:
        sv.close()
        for ps in pinned:
            for p in ps:
                p.close()


def test_server_tight_budget_evicts_and_matches():
    """A budget that holds about one request at a time: plans are evicted
    and re-created, batches shrink, results do not change."""
    jobs, _, sets = _mix(n_jobs=6, seed=23)
    ref = _alone(jobs, sets)
    # The footprint of the largest request alone on a fresh server.
    need = 0
    for wl in jobs:
        sv = fg.Server(8 * GiB, max_batch_jobs=1, num_streams=1)
        ids = _register(sv, sets)
        sv.wait(_submit(sv, ids, wl), timeout=60.0)
        need = max(need, sv.stats()["device_bytes"])
        sv.close()
    sv = fg.Server(int(need * 1.25), max_batch_jobs=8, num_streams=2)
    try:
        ids = _register(sv, sets)
        rids = [_submit(sv, ids, wl) for wl in jobs]
        got = []
        for rid in rids:
            res, g = sv.wait(rid, timeout=60.0)
            got.append((res.num_skipped, res.norm_delta, res.num_updates, g))
        _check(got, ref)
        st = sv.stats()
        assert st["device_bytes"] <= int(need * 1.25)
        assert st["plans_evicted"] > 0
    finally:
        sv.close()


def test_server_resident_grid_accumulates():
    """Requests into one resident grid add up to the sum of their grids
    alone; the grid's kernel-set plan stays resident."""
    jobs, _, sets = _mix(n_jobs=12, seed=24)
    G = 512
    same = [wl for wl in jobs if wl.grid_size == G and
            wl.params["kernel_set"] == jobs[0].params["kernel_set"]][:1]
    import dataclasses
    wl0 = dataclasses.replace(jobs[0], grid_size=G) if not same else same[0]
    rng = np.random.default_rng(5)
    parts = []
    for _ in range(3):
        idx = np.sort(rng.choice(wl0.num_points, wl0.num_points // 2,
                                 replace=False))
        v2 = wl0.vis.reshape(-1, 2)[idx].reshape(-1)
        parts.append(dataclasses.replace(
            wl0, uu=wl0.uu[idx].copy(), vv=wl0.vv[idx].copy(),
            ww=wl0.ww[idx].copy(), vis=np.ascontiguousarray(v2),
            weight=np.ones(idx.size, np.float32)))
    ref = _alone(parts, sets)
    sv = fg.Server(8 * GiB, max_batch_jobs=2, num_streams=2)
    try:
        ids = _register(sv, sets)
        gid = sv.create_grid(G)
        rids = [_submit(sv, ids, wl, grid_id=gid) for wl in parts]
        norm = 0.0
        for rid, r in zip(rids, ref):
            res, g = sv.wait(rid, timeout=60.0)
            assert g is None
            assert (res.num_skipped, res.num_updates) == (r[0], r[2])
            norm += res.norm_delta
        acc = sv.read_grid(gid, G)
        total = sum(r[3].astype(np.float64) for r in ref)
        assert max_rel(acc, total) <= 4e-6
        assert abs(norm - sum(r[1] for r in ref)) <= 1e-12 * abs(norm)
        assert sv.stats()["resident_plans"] == 1
        sv.free_grid(gid)
    finally:
        sv.close()


def test_server_errors():
    jobs, _, sets = _mix(n_jobs=1, seed=25)
    wl = jobs[0]
    sv = fg.Server(1 << 20, max_batch_jobs=4, num_streams=1)
    try:
        ids = _register(sv, sets)
        with pytest.raises(fg.GridderError):
            sv.submit(99, wl.grid_size, wl.cell_size_rad, wl.w_scale,
                      wl.uu, wl.vv, wl.ww, wl.vis)
        with pytest.raises(fg.GridderError):
            sv.submit(ids[0], wl.grid_size, wl.cell_size_rad, wl.w_scale,
                      wl.uu, wl.vv, wl.ww, wl.vis, grid_id=7)
        # A request that can never fit a 1 MiB budget fails; the queue
        # goes on.
        rid = _submit(sv, ids, wl)
        with pytest.raises(fg.GridderError):
            sv.wait(rid, timeout=60.0)
        assert sv.stats()["pending"] == 0
    finally:
        sv.close()
