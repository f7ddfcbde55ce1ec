"""Batched jobs (gw_batch_*) and the plan autotuner, through the C ABI.

A batch grids many independent FP32 jobs (own plan, points and grid each)
with one pipeline and one persistent scatter launch; captured as a HIP
graph it is replayed with one launch.  Every job's result must equal the
same job gridded alone with the plain per-plan call:

* num_skipped, the norm increment and the update count: identical;
* grid: within 2e-6 * max|grid| (the same FP32 products and FP64 per-job
  sums, added in another order).

Graph replays are each checked against the eager result, with the
program's own work (grid clearing, result reads) between replays.
"""
import numpy as np
import pytest

import oracle
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


def _mix(n_jobs=7, seed=3, n_range=(1500, 60000), grids=(256, 512, 1024)):
    base = synth.snapshot_small()
    return synth.batch_mix(n_jobs=n_jobs, seed=seed, n_range=n_range,
                           grids=grids, n_kernel_sets=3, base=base,
                           fov_deg_at_8192=2.0 * 2048 / 8192)


class _Jobs:
    """Plans (one per kernel set and grid side), device points and grids."""

    def __init__(self, jobs):
        self.jobs = jobs
        self.plans = {}
        self.dev = []
        for wl in jobs:
            key = (wl.params["kernel_set"], wl.grid_size)
            if key not in self.plans:
                self.plans[key] = fg.Plan.for_workload(wl)
            d = [fg.DeviceArray.from_host(a)
                 for a in (wl.uu, wl.vv, wl.ww, wl.vis)]
            g = fg.DeviceArray(8 * wl.grid_size ** 2)
            g.zero()
            self.dev.append((d, g))

    def plan(self, j):
        wl = self.jobs[j]
        return self.plans[(wl.params["kernel_set"], wl.grid_size)]

    def zero(self):
        for _, g in self.dev:
            g.zero()

    def grid(self, j):
        G = self.jobs[j].grid_size
        return self.dev[j][1].download(np.empty(2 * G * G, np.float32))

    def close(self):
        for d, g in self.dev:
            for x in (*d, g):
                x.close()
        for p in self.plans.values():
            p.close()


def _sequential(jobs):
    """Each job alone through its plan (the plain per-job launch)."""
    out = []
    for j, wl in enumerate(jobs.jobs):
        d, g = jobs.dev[j]
        g.zero()
        pl = jobs.plan(j)
        pl.grid(wl.num_points, *d, g)
        res = pl.results()          # waits for the plan's stream
        out.append((jobs.grid(j), res))
    jobs.zero()
    return out


def _check(batch, jobs, ref):
    for j in range(len(jobs.jobs)):
        sk, nd, nu = batch.results(j)
        g_ref, (sk_r, nd_r, nu_r) = ref[j]
        assert (sk, nu) == (sk_r, nu_r), j
        assert abs(nd - nd_r) <= 1e-12 * max(abs(nd_r), 1.0), j
        g = jobs.grid(j)
        if np.any(g_ref):
            assert max_rel(g, g_ref) <= 2e-6, j
        else:
            assert not np.any(g), j
    # The smallest job with points against the oracle directly (skip count,
    # norm and grid to the FP32 parity bar), not only against the plain
    # path the batch shares kernels with.
    live = [j for j in range(len(jobs.jobs)) if jobs.jobs[j].num_points]
    if live:
        j = min(live, key=lambda i: jobs.jobs[i].num_points)
        g_e, sk_e, n_e = oracle.grid_wproj_f_exact(jobs.jobs[j])
        sk, nd, _ = batch.results(j)
        assert sk == sk_e
        assert abs(nd - n_e) <= 1e-12 * max(abs(n_e), 1.0)
        if np.any(g_e):
            assert max_rel(jobs.grid(j), g_e) <= 1e-5


def test_batch_matches_sequential_jobs():
    jobs = _Jobs(_mix()[0])
    try:
        ref = _sequential(jobs)
        b = fg.Batch()
        for j, wl in enumerate(jobs.jobs):
            d, g = jobs.dev[j]
            assert b.add(jobs.plan(j), wl.num_points, *d, g) == j
        b.run()
        _check(b, jobs, ref)
        b.close()
    finally:
        jobs.close()


def test_batch_graph_replays():
    """Capture once, replay three times; every replay matches the plain
    per-job results (grids cleared between replays)."""
    jobs = _Jobs(_mix(seed=5)[0])
    try:
        ref = _sequential(jobs)
        b = fg.Batch()
        for j, wl in enumerate(jobs.jobs):
            d, g = jobs.dev[j]
            b.add(jobs.plan(j), wl.num_points, *d, g)
        b.capture()
        for _ in range(3):
            jobs.zero()
            fg.device_synchronize()
            b.launch()
            _check(b, jobs, ref)
            assert b.last_ms() > 0.0
        b.close()
    finally:
        jobs.close()


def test_batch_empty_and_skipped_jobs():
    """Jobs with no points and with every point off the grid, between
    ordinary jobs."""
    mix = _mix(n_jobs=3, seed=9)[0]
    empty = synth.random_small(n=10, grid_size=64, num_planes=3, seed=2)
    import dataclasses
    empty = dataclasses.replace(empty, uu=empty.uu[:0], vv=empty.vv[:0],
                                ww=empty.ww[:0], vis=empty.vis[:0],
                                params=dict(kernel_set=100))
    far = synth.random_small(n=300, grid_size=64, num_planes=3, seed=4)
    far = dataclasses.replace(far, uu=np.full_like(far.uu, 1e9),
                              params=dict(kernel_set=101))
    jobs = _Jobs([mix[0], empty, mix[1], far, mix[2]])
    try:
        ref = _sequential(jobs)
        b = fg.Batch()
        for j, wl in enumerate(jobs.jobs):
            d, g = jobs.dev[j]
            b.add(jobs.plan(j), wl.num_points, *d, g)
        b.run()
        _check(b, jobs, ref)
        assert b.results(1)[0] == 0 and b.results(3)[0] == 300
        b.close()
    finally:
        jobs.close()


def test_empty_batch():
    b = fg.Batch()
    b.run()
    b.capture()
    b.launch()
    fg.device_synchronize()
    b.close()


def test_autotune_keeps_results():
    """The autotuner only changes the scatter shape: a tuned plan grids the
    same values, and each shape it may pick does too."""
    wl = synth.random_small(n=40000, grid_size=512, num_planes=12, s_min=3,
                            s_max=20, seed=8, edge_fraction=0.05)
    pl = fg.Plan.for_workload(wl)
    d = [fg.DeviceArray.from_host(a) for a in (wl.uu, wl.vv, wl.ww, wl.vis)]
    g = fg.DeviceArray(8 * wl.grid_size ** 2)
    scratch = fg.DeviceArray(8 * wl.grid_size ** 2)

    def grid_once():
        g.zero()
        pl.grid(wl.num_points, *d, g)
        r = pl.results()
        return g.download(np.empty(2 * wl.grid_size ** 2, np.float32)), r

    try:
        g0, r0 = grid_once()
        t = pl.tuning()
        # The default side is the largest whose region fits (here 89 - 41 =
        # 48, not a power of two); the power-of-two sides below agree with it.
        assert t["tile"] == t["max_tile"] == 48 and t["tile_log2"] == -1
        assert t["max_tile_log2"] == 5 and t["region"] == 48 + 40
        for tile in range(max(0, t["max_tile_log2"] - 2),
                          t["max_tile_log2"] + 1):
            for cap in (4, 6, 8):
                for ms in (1, 8, 64):
                    pl.set_tuning(tile, cap, ms)
                    g1, r1 = grid_once()
                    assert r1[0] == r0[0] and r1[2] == r0[2]
                    assert abs(r1[1] - r0[1]) <= 1e-12 * abs(r0[1])
                    assert max_rel(g1, g0) <= 2e-6, (tile, cap, ms)
        pl.set_tuning()
        ms = pl.autotune(wl.num_points, *d, scratch, reps=1)
        assert ms > 0.0
        tuned = pl.tuning()
        assert tuned["tile"] in (48, 32, 16, 8)
        assert tuned["tile_log2"] == (-1 if tuned["tile"] == 48 else
                                      tuned["tile"].bit_length() - 1)
        assert tuned["slot_cap"] in (4, 6, 8)
        g2, r2 = grid_once()
        assert r2[0] == r0[0] and max_rel(g2, g0) <= 2e-6
    finally:
        for x in (*d, g, scratch):
            x.close()
        pl.close()


class _HipStreams:
    """Caller-owned HIP streams (test infrastructure: the HIP runtime the
    library uses, through ctypes)."""

    def __init__(self, n):
        import ctypes as ct
        self.hip = ct.CDLL("libamdhip64.so")
        self.streams = []
        for _ in range(n):
            s = ct.c_void_p()
            assert self.hip.hipStreamCreate(ct.byref(s)) == 0
            self.streams.append(s.value)

    def close(self):
        import ctypes as ct
        for s in self.streams:
            self.hip.hipStreamSynchronize(ct.c_void_p(s))
            self.hip.hipStreamDestroy(ct.c_void_p(s))
        self.streams = []


def test_batch_reuse_back_to_back():
    """One Batch, job lists changed between runs with nothing waited for in
    between: a run within capacity, a run that grows the buffers, a graph
    capture and replay, then a job change and a replay (recapture while the
    previous replay may still run) on another caller stream.  Every round
    grids into its own grids; all are checked at the end against the jobs
    gridded alone and against the oracle."""
    rounds = [_mix(n_jobs=3, seed=21),
              _mix(n_jobs=3, seed=22),
              _mix(n_jobs=5, seed=23, n_range=(60000, 150000)),
              _mix(n_jobs=4, seed=24),
              _mix(n_jobs=3, seed=25)]
    sets = [_Jobs(m[0]) for m in rounds]
    streams = _HipStreams(2)
    try:
        refs = [_sequential(js) for js in sets]
        b = fg.Batch()

        def add_all(js):
            b.clear()
            for j, wl in enumerate(js.jobs):
                d, g = js.dev[j]
                b.add(js.plan(j), wl.num_points, *d, g)

        add_all(sets[0])
        b.run()                       # round 0: first allocation
        add_all(sets[1])
        b.run()                       # round 1: same capacity
        add_all(sets[2])
        b.run()                       # round 2: buffers grow
        b.set_stream(streams.streams[0])
        add_all(sets[3])
        b.capture()
        b.launch()                    # round 3: graph on a caller stream
        b.set_stream(streams.streams[1])
        add_all(sets[4])
        b.launch()                    # round 4: jobs changed, recaptured
        for j in range(len(sets[4].jobs)):
            sk, nd, nu = b.results(j)
            assert (sk, nu) == (refs[4][j][1][0], refs[4][j][1][2])
        fg.device_synchronize()
        for r, js in enumerate(sets):
            for j, wl in enumerate(js.jobs):
                g_ref = refs[r][j][0]
                g = js.grid(j)
                if np.any(g_ref):
                    assert max_rel(g, g_ref) <= 2e-6, (r, j)
                else:
                    assert not np.any(g), (r, j)
        # The oracle on the smallest job of each round (FP32 index math,
        # FP64 sums: the batch's FP32 grid to the FP32 parity bar).
        for r, js in enumerate(sets):
            j = min(range(len(js.jobs)), key=lambda i: js.jobs[i].num_points)
            g_e, sk_e, _ = oracle.grid_wproj_f_exact(js.jobs[j])
            assert refs[r][j][1][0] == sk_e
            assert max_rel(js.grid(j), g_e) <= 1e-5, (r, j)
        b.close()
    finally:
        streams.close()
        for js in sets:
            js.close()
