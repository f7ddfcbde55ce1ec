"""Parity of the HIP gridder (through its C ABI) with the CPU oracle.

Small and medium workloads are compared cell by cell with the oracle and
with the golden fixtures written from the reference build; the 13M-point
full-scale workload is checked through size-independent properties (skip
count, update count, norm, grid checksum, spot cells, linearity).
Tolerances are stated in tests/parity.py.
"""
import dataclasses
import os

import numpy as np
import pytest

import oracle
import fpga_gridding_amd as fg
from fpga_gridding_amd import synth

from parity import (assert_fp32_parity, expected_sums, fp32_report, max_rel)

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _load(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)


@pytest.fixture(scope="module", autouse=True)
def _gpu_present():
    if fg.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    yield
    fg.lib().gw_release_cached()


# --------------------------------------------------------------------------
# Golden fixtures (outputs of the reference's own gridder)
# --------------------------------------------------------------------------

def test_golden_fp32_fixture():
    d = _load("ref_f_diag")
    wl = synth.workload_from_arrays(d)
    g, sk, nrm = fg.grid_workload_f(wl)
    assert sk == int(d["num_skipped"])
    assert abs(nrm - float(d["norm"])) <= 1e-12 * abs(float(d["norm"]))
    assert max_rel(g, d["grid"]) <= 1e-5
    assert_fp32_parity(fp32_report(g, wl))


def test_golden_fp32_edge_cases_10000():
    """Ties, grid edges, w = +-0, NaN, +-inf and huge coordinates, against
    the reference oskar_grid_wproj_f itself (its fixed 10000-point loop)."""
    d = _load("ref_f_10000")
    wl = synth.workload_from_arrays(d)
    g, sk, nrm = fg.grid_workload_f(wl)
    assert sk == int(d["num_skipped"])
    assert abs(nrm - float(d["norm"])) <= 1e-12 * abs(float(d["norm"]))
    assert max_rel(g, d["grid"]) <= 1e-5


def test_golden_fp64_fixture():
    d = _load("ref_d")
    wl = synth.workload_from_arrays(d)
    g, sk, nrm = fg.grid_workload_d(wl)
    assert sk == int(d["num_skipped"])
    assert abs(nrm - float(d["norm"])) <= 1e-12 * abs(float(d["norm"]))
    assert max_rel(g, d["grid"]) <= 1e-12


def test_golden_asymmetric_cube():
    """A random kernel cube (no symmetry between or within rows and
    columns): the scatter's tap tables, including the taps of negative
    sub-cell offsets read at mirrored cells, against the reference's
    outputs."""
    d = _load("ref_asym")
    wl = synth.workload_from_arrays(d)
    g, sk, nrm = fg.grid_workload_f(wl)
    assert sk == int(d["num_skipped"])
    assert abs(nrm - float(d["norm"])) <= 1e-12 * abs(float(d["norm"]))
    assert max_rel(g, d["grid"]) <= 1e-5
    assert_fp32_parity(fp32_report(g, wl))
    gd, skd, nrmd = fg.grid_workload_d(wl)
    assert skd == int(d["num_skipped_d"])
    assert abs(nrmd - float(d["norm_d"])) <= 1e-12 * abs(float(d["norm_d"]))
    assert max_rel(gd, d["grid_d"]) <= 1e-12


# --------------------------------------------------------------------------
# Shapes against the oracle
# --------------------------------------------------------------------------

SHAPES = [
    dict(n=4000, grid_size=64, num_planes=5, s_min=1, s_max=5, ov=4, seed=1),
    dict(n=3000, grid_size=48, num_planes=4, s_min=0, s_max=4, ov=1, seed=2),
    dict(n=3000, grid_size=50, num_planes=3, s_min=2, s_max=6, ov=3, seed=3),
    dict(n=2000, grid_size=96, num_planes=7, s_min=1, s_max=7, ov=8, seed=4),
    dict(n=2000, grid_size=33, num_planes=1, s_min=3, s_max=3, ov=4, seed=5),
    dict(n=20000, grid_size=256, num_planes=16, s_min=4, s_max=40, ov=4,
         seed=6),
    dict(n=500, grid_size=300, num_planes=3, s_min=44, s_max=50, ov=2,
         seed=7),
]


@pytest.mark.parametrize("case", SHAPES, ids=lambda c: f"G{c['grid_size']}-"
                         f"ov{c['ov']}-S{c['s_max']}-s{c['seed']}")
def test_shapes_fp32(case):
    wl = synth.random_small(**case)
    g, sk, nrm = fg.grid_workload_f(wl)
    r = fp32_report(g, wl)
    assert sk == r["skipped_f"]
    assert abs(nrm - r["norm_f"]) <= 1e-12 * abs(r["norm_f"])
    assert_fp32_parity(r)


@pytest.mark.parametrize("n", [1, 4095, 4096, 4097, 8193])
def test_sort_chunk_boundaries(n):
    """Point counts around the key sort's 4096-key chunks (a lone key, a
    chunk short of full, exactly full, one over, two chunks and one key)."""
    wl = synth.random_small(n=n, grid_size=128, num_planes=6, s_min=2,
                            s_max=6, seed=20 + n)
    g, sk, nrm = fg.grid_workload_f(wl)
    r = fp32_report(g, wl)
    assert sk == r["skipped_f"]
    assert abs(nrm - r["norm_f"]) <= 1e-12 * max(1.0, abs(r["norm_f"]))
    assert_fp32_parity(r)


def test_all_points_one_cell():
    """9000 points with the same uvw: every sort key and every digit is the
    same (one run, each chunk's keys in one digit), then two cells taking
    turns (two runs, interleaved input)."""
    base = synth.random_small(n=9000, grid_size=128, num_planes=6, s_min=2,
                              s_max=6, seed=31, edge_fraction=0.0)
    kept = np.flatnonzero(oracle.locate_f(base)[0]["skip"] == 0)[:2]
    for cells in (1, 2):
        pick = kept[np.arange(base.num_points) % cells]
        wl = dataclasses.replace(
            base, uu=base.uu[pick].copy(), vv=base.vv[pick].copy(),
            ww=base.ww[pick].copy())
        g, sk, nrm = fg.grid_workload_f(wl)
        r = fp32_report(g, wl)
        assert sk == r["skipped_f"] == 0
        assert abs(nrm - r["norm_f"]) <= 1e-12 * abs(r["norm_f"])
        assert_fp32_parity(r)
        pl = fg.Plan.for_workload(wl)
        try:
            d = [fg.DeviceArray.from_host(a)
                 for a in (wl.uu, wl.vv, wl.ww, wl.vis)]
            gd = fg.DeviceArray(8 * wl.grid_size ** 2)
            gd.zero()
            pl.grid(wl.num_points, *d, gd)
            pl.results()
            assert pl.records()[0] == cells
            for x in (*d, gd):
                x.close()
        finally:
            pl.close()


@pytest.mark.parametrize("case", SHAPES[:4], ids=lambda c: f"ov{c['ov']}")
def test_shapes_fp64(case):
    wl = synth.random_small(**case)
    g, sk, nrm = fg.grid_workload_d(wl)
    gd, skd, nd = oracle.grid_wproj_d(wl)
    assert sk == skd
    assert abs(nrm - nd) <= 1e-12 * abs(nd)
    assert max_rel(g, gd) <= 1e-12


def _runs_workload(seed=11):
    """Points repeated in runs that share cell, plane and sub-cell table:
    each base point appears as itself, an exact copy, with w negated (same
    plane, conjugated kernel) and as a copy of that, in shuffled order and
    with its own visibilities, so merged runs mix conjugation signs."""
    base = synth.random_small(n=400, grid_size=128, num_planes=8, s_min=2,
                              s_max=12, ov=4, seed=seed, edge_fraction=0.05)
    rng = np.random.default_rng(seed)
    reps = 4
    uu = np.tile(base.uu, reps)
    vv = np.tile(base.vv, reps)
    ww = np.concatenate([base.ww, base.ww, -base.ww, -base.ww])
    order = rng.permutation(uu.size)
    n = uu.size
    return dataclasses.replace(
        base, uu=uu[order], vv=vv[order], ww=ww[order].astype(np.float32),
        vis=rng.standard_normal(2 * n).astype(np.float32),
        weight=np.ones(n, np.float32), name="runs")


def test_merged_runs_mixed_conjugation():
    """Runs of points with equal cell, plane and sub-cell table become one
    scatter record whose visibility terms are summed with each point's own
    conjugation sign: FP32 to the FP32 bar, FP64 to 1e-12 of the FP64
    oracle, skip counts exact, the batch path to the same bars."""
    wl = _runs_workload()
    g, sk, nrm = fg.grid_workload_f(wl)
    r = fp32_report(g, wl)
    assert sk == r["skipped_f"]
    assert abs(nrm - r["norm_f"]) <= 1e-12 * abs(r["norm_f"])
    assert_fp32_parity(r)
    gd, skd, nd = fg.grid_workload_d(wl)
    g_o, sk_o, n_o = oracle.grid_wproj_d(wl)
    assert skd == sk_o
    assert abs(nd - n_o) <= 1e-12 * abs(n_o)
    assert max_rel(gd, g_o) <= 1e-12
    # The plan API reports one record per run: the distinct (cell, plane,
    # sub-cell) of the gridded points, and their footprint sizes.
    plan = fg.Plan.for_workload(wl, fg.GW_SINGLE)
    G = wl.grid_size
    d = [fg.DeviceArray.from_host(a) for a in (wl.uu, wl.vv, wl.ww, wl.vis)]
    dg = fg.DeviceArray(8 * G * G)
    dg.zero()
    plan.grid(wl.num_points, *d, dg)
    skp, _, upd = plan.results()
    nrec, nrec_upd = plan.records()
    rec, u_alg = oracle.locate_f(wl)
    ok = rec["skip"] == 0
    keys = np.stack([rec[k][ok] for k in ("grid_u", "grid_v", "plane",
                                          "off_u", "off_v")], axis=1)
    uniq, first = np.unique(keys, axis=0, return_index=True)
    w = 2 * rec["support"][ok][first].astype(np.int64) + 1
    assert skp == sk and upd == u_alg
    assert nrec == len(uniq) and nrec_upd == int((w * w).sum())
    assert 4 * nrec == ok.sum()
    for x in (*d, dg):
        x.close()
    # The same points through a two-job batch (FP32 index math).
    jobs = []
    for part in (slice(0, 700), slice(700, None)):
        sub = dataclasses.replace(wl, uu=wl.uu[part], vv=wl.vv[part],
                                  ww=wl.ww[part],
                                  vis=wl.vis.reshape(-1, 2)[part].ravel())
        d = [fg.DeviceArray.from_host(a) for a in (sub.uu, sub.vv, sub.ww,
                                                    sub.vis)]
        dg = fg.DeviceArray(8 * G * G)
        dg.zero()
        jobs.append((sub, d, dg))
    batch = fg.Batch()
    for sub, d, dg in jobs:
        batch.add(plan, sub.num_points, *d, dg)
    batch.run()
    res = [batch.results(j) for j in range(len(jobs))]
    for (sub, d, dg), (skb, nb, _) in zip(jobs, res):
        gb = np.empty(2 * G * G, np.float32)
        dg.download(gb)
        rb = fp32_report(gb, sub)
        assert skb == rb["skipped_f"]
        assert abs(nb - rb["norm_f"]) <= 1e-12 * abs(rb["norm_f"])
        assert_fp32_parity(rb)
    batch.close()
    plan.close()


def test_plane_index_ties():
    """w placed on and one ulp beside the plane rounding ties
    sqrt(|w| w_scale) = m + 1/2: any device/host difference in sqrt or
    round would pick another plane's taps and show in the grid."""
    rng = np.random.default_rng(11)
    wl = synth.random_small(n=4000, grid_size=128, num_planes=64, s_min=1,
                            s_max=6, seed=12)
    m = rng.integers(0, 63, wl.num_points)
    w = ((m + 0.5) ** 2).astype(np.float32)
    step = rng.integers(-1, 2, wl.num_points)
    w = np.where(step < 0, np.nextafter(w, np.float32(0)), w)
    w = np.where(step > 0, np.nextafter(w, np.float32(np.inf)), w)
    w = w * rng.choice([-1, 1], wl.num_points).astype(np.float32)
    wl = dataclasses.replace(wl, ww=w.astype(np.float32), w_scale=1.0)
    g, sk, _ = fg.grid_workload_f(wl)
    r = fp32_report(g, wl)
    assert sk == r["skipped_f"]
    assert_fp32_parity(r)


def test_snapshot64_full_workload():
    """The 64-antenna 2 km snapshot (725,760 points, 2048^2, 32 planes)."""
    wl = synth.snapshot_small()
    g, sk, nrm = fg.grid_workload_f(wl)
    r = fp32_report(g, wl)
    print({k: v for k, v in r.items() if isinstance(v, float)})
    assert sk == r["skipped_f"] == 0
    assert abs(nrm - r["norm_f"]) <= 1e-12 * abs(r["norm_f"])
    assert_fp32_parity(r)


# --------------------------------------------------------------------------
# Call semantics and edge inputs
# --------------------------------------------------------------------------

def test_accumulates_and_resets_skipped():
    wl = synth.random_small(n=3000, grid_size=64, num_planes=5, seed=21)
    g1, sk1, n1 = fg.grid_workload_f(wl)
    g2, sk2, n2 = fg.grid_workload_f(wl, grid=g1.copy(), norm=n1)
    assert sk2 == sk1
    assert abs(n2 - 2 * n1) <= 1e-12 * abs(n1)
    assert max_rel(g2, 2 * g1) <= 1e-6
    # Untouched cells keep the caller's values.
    base = np.full_like(g1, 0.25)
    g3, _, _ = fg.grid_workload_f(wl, grid=base.copy())
    touched = g1 != 0
    assert np.all(g3[~touched] == 0.25)


def test_empty_input():
    wl = synth.random_small(n=100, grid_size=32, seed=22)
    wl = dataclasses.replace(wl, uu=wl.uu[:0], vv=wl.vv[:0], ww=wl.ww[:0],
                             vis=wl.vis[:0])
    grid = np.full(2 * 32 * 32, 1.5, np.float32)
    g, sk, nrm = fg.grid_workload_f(wl, grid=grid, norm=2.0)
    assert sk == 0 and nrm == 2.0 and np.all(g == 1.5)


def test_all_points_skipped():
    wl = synth.random_small(n=300, grid_size=32, seed=23)
    far = np.float32(1e9)
    wl = dataclasses.replace(wl, uu=np.full_like(wl.uu, far))
    grid = np.full(2 * 32 * 32, -3.0, np.float32)
    g, sk, nrm = fg.grid_workload_f(wl, grid=grid, norm=5.0)
    assert sk == wl.num_points and nrm == 5.0 and np.all(g == -3.0)


def test_diagnostics_variant():
    """oskar_grid_wproj_f_diagnostics: histograms counted before the skip
    test, unclamped plane index, per-cell footprint counts and support+plane
    sums (oskar_grid_wproj.c:317-360)."""
    d = _load("ref_f_diag")
    wl = synth.workload_from_arrays(d)
    G = wl.grid_size
    grid = np.zeros(2 * G * G, np.float32)
    per_s = np.zeros(d["per_support"].size, np.int64)
    per_w = np.zeros(d["per_plane"].size, np.int64)
    per_cell = np.zeros(G * G, np.int64)
    avg_s = np.zeros(G * G, np.float32)
    sk, nrm = fg.grid_wproj_f_diagnostics(
        wl.num_w_planes, wl.support, wl.oversample, wl.conv_size_half,
        wl.conv_func, wl.uu, wl.vv, wl.ww, wl.vis, None, wl.cell_size_rad,
        wl.w_scale, G, grid, per_s, per_w, per_cell, avg_s)
    assert sk == int(d["num_skipped"])
    assert np.array_equal(per_s, d["per_support"])
    assert np.array_equal(per_w, d["per_plane"])
    assert np.array_equal(per_cell, d["per_cell"])
    assert np.array_equal(avg_s, d["avg_support"])
    assert max_rel(grid, d["grid"]) <= 1e-5


def test_diagnostics_variant_asymmetric_cube():
    """The diagnostics variant (which adds straight into the grid, not
    through the LDS region) on the random cube: its grid against the
    reference's."""
    d = _load("ref_asym")
    wl = synth.workload_from_arrays(d)
    G = wl.grid_size
    grid = np.zeros(2 * G * G, np.float32)
    per_s = np.zeros(int(wl.support.max()) + 1, np.int64)
    per_w = np.zeros(4 * wl.num_w_planes, np.int64)
    per_cell = np.zeros(G * G, np.int64)
    avg_s = np.zeros(G * G, np.float32)
    sk, nrm = fg.grid_wproj_f_diagnostics(
        wl.num_w_planes, wl.support, wl.oversample, wl.conv_size_half,
        wl.conv_func, wl.uu, wl.vv, wl.ww, wl.vis, None, wl.cell_size_rad,
        wl.w_scale, G, grid, per_s, per_w, per_cell, avg_s)
    assert sk == int(d["num_skipped"])
    assert abs(nrm - float(d["norm"])) <= 1e-12 * abs(float(d["norm"]))
    assert max_rel(grid, d["grid"]) <= 1e-5


def test_resident_plan_matches_dropin():
    """Device-pointer plan API and drop-in call give the same grid."""
    wl = synth.snapshot_small()
    g_drop, sk, nrm = fg.grid_workload_f(wl)
    pl = fg.Plan.for_workload(wl)
    d = [fg.DeviceArray.from_host(a) for a in (wl.uu, wl.vv, wl.ww, wl.vis)]
    dg = fg.DeviceArray(8 * wl.grid_size ** 2)
    dg.zero()
    pl.grid(wl.num_points, *d, dg)
    sk2, nd, nu = pl.results()
    g = dg.download(np.empty(2 * wl.grid_size ** 2, np.float32))
    assert sk2 == sk
    assert abs(nd - nrm) <= 1e-12 * abs(nrm)
    assert max_rel(g, g_drop) <= 2e-6
    _, U = oracle.locate_f(wl)
    assert nu == U


def test_resident_plan_empty_first_call():
    """A fresh plan's first call with zero points grids nothing and leaves
    the plan usable (its workspace exists from the first call)."""
    wl = synth.random_small(n=500, grid_size=64, num_planes=4, s_min=1,
                            s_max=4, seed=21)
    pl = fg.Plan.for_workload(wl)
    dg = fg.DeviceArray(8 * wl.grid_size ** 2)
    dg.zero()
    pl.grid(0, None, None, None, None, dg)
    assert pl.results() == (0, 0.0, 0)
    d = [fg.DeviceArray.from_host(a) for a in (wl.uu, wl.vv, wl.ww, wl.vis)]
    pl.grid(wl.num_points, *d, dg)
    sk, nd, _ = pl.results()
    g = dg.download(np.empty(2 * wl.grid_size ** 2, np.float32))
    g_ref, sk_ref, n_ref = fg.grid_workload_f(wl)
    assert sk == sk_ref and abs(nd - n_ref) <= 1e-12 * abs(n_ref)
    assert max_rel(g, g_ref) <= 2e-6
    pl.close()


# --------------------------------------------------------------------------
# Full scale: 512 stations, 13,081,600 points, 8192^2, 128 planes
# --------------------------------------------------------------------------

@pytest.fixture(scope="module")
def full_scale():
    wl = synth.full_scale()
    rec, U = oracle.locate_f(wl)
    pl = fg.Plan.for_workload(wl)
    d = [fg.DeviceArray.from_host(a) for a in (wl.uu, wl.vv, wl.ww, wl.vis)]
    dg = fg.DeviceArray(8 * wl.grid_size ** 2)
    dg.zero()
    pl.grid(wl.num_points, *d, dg)
    res = pl.results()
    g = dg.download(np.empty(2 * wl.grid_size ** 2, np.float32))
    yield wl, rec, U, pl, d, dg, res, g
    for a in d:
        a.close()
    dg.close()
    pl.close()


def test_full_scale_counts_and_norm(full_scale):
    wl, rec, U, pl, d, dg, (sk, nd, nu), g = full_scale
    assert sk == int(np.sum(rec["skip"] != 0))
    assert nu == U
    norm, _, _ = expected_sums(wl, rec)
    assert abs(nd - norm) <= 1e-12 * abs(norm)


def test_full_scale_checksum(full_scale):
    wl, rec, U, pl, d, dg, res, g = full_scale
    _, checksum, scale = expected_sums(wl, rec)
    got = complex(np.sum(g[0::2].astype(np.float64)),
                  np.sum(g[1::2].astype(np.float64)))
    assert abs(got - checksum) <= 1e-6 * scale


def test_full_scale_spot_cells(full_scale):
    """64 cells -- the largest, the densest region and random covered ones
    -- against a float64 brute force over all 13M points."""
    wl, rec, U, pl, d, dg, res, g = full_scale
    mag = np.hypot(g[0::2], g[1::2])
    rng = np.random.default_rng(5)
    covered = np.flatnonzero(mag > 0)
    cells = np.unique(np.concatenate([
        [int(np.argmax(mag))], rng.choice(covered, 63, replace=False)]))
    ref = oracle.cells_f(wl, cells)
    got = np.stack([g[2 * cells], g[2 * cells + 1]], 1).reshape(-1)
    assert np.abs(got - ref).max() <= 1e-5 * mag.max()


def test_full_scale_linearity(full_scale):
    """Gridding 2*vis doubles the grid (up to the reordering of float
    sums)."""
    wl, rec, U, pl, d, dg, res, g = full_scale
    d2 = fg.DeviceArray.from_host((2 * wl.vis).astype(np.float32))
    dg.zero()
    pl.grid(wl.num_points, d[0], d[1], d[2], d2, dg)
    pl.results()
    g2 = dg.download(np.empty_like(g))
    assert max_rel(g2, 2 * g) <= 1e-6


def test_full_scale_dropin(full_scale):
    """The drop-in host-pointer call at full scale (grid rows moved across
    PCIe) agrees with the resident plan."""
    wl, rec, U, pl, d, dg, (sk, nd, nu), g = full_scale
    gh, sk2, n2 = fg.grid_workload_f(wl)
    assert sk2 == sk
    assert abs(n2 - nd) <= 1e-12 * abs(nd)
    assert max_rel(gh, g) <= 2e-6
