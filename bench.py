"""Benchmark of the MI355X W-projection gridder on the 512-station workload.

One step = one pass of the hot path (oskar_grid_wproj_f semantics) over the
whole 13,081,600-point batch: index math, sort, work planning and the scatter
into an 8192^2 grid, with points, visibilities and the grid resident in HBM
(the plan API: gw_grid_device + gw_plan_results).  Prints ONE JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
  python bench.py --copy-bytes B     (profiler calibration: one plain copy)

The scatter kernel's duration comes from HIP events the library records on
its own stream around each stage; the roofline uses the algorithmic FLOPs of
the reference's inner loop (8 per grid update, oskar_grid_wproj.c:236-237)
over that duration.  The CPU baseline times the reference's own
oskar_grid_wproj_f (a gcc build of its sources, oracle/_ref) on one host
core over a random sample of the same points; it is a reported baseline,
not a target.
"""
from __future__ import annotations

import argparse
import ctypes as ct
import dataclasses
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3      # MI355X FP32 peak (AMD published)
HBM_PEAK_GBS = 8000.0
WORKLOAD = ("512 stations (core + log-radial to 40 km), 130816 baselines x "
            "100 x 10 s, 150 MHz, 13,081,600 points, grid 8192^2, 5 deg FOV, "
            "128 w-planes, support 4..36, oversample 4")


def _traffic_from_profile():
    """HBM bytes per scatter launch, from the committed rocprofv3 counter
    summary (FETCH_SIZE / WRITE_SIZE scaled by a copy-kernel calibration),
    or None if no such summary exists."""
    path = os.path.join(ROOT, "profiles", "k_grid_ska512_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("hbm_bytes_per_launch"):
            try:
                return float(line.split("=")[1].split()[0])
            except (IndexError, ValueError):
                return None
    return None


def accuracy_sample(wl, n_sample=200_000, seed=4):
    """Errors of the HIP gridder on a random sample of the workload against
    the FP64 oracle (oskar_grid_wproj_d semantics: FP64 index math and sums)
    and against the exact oracle (FP32 index math as oskar_grid_wproj_f,
    FP64 sums): max|d|/max|g| and RMS(d)/RMS(g)."""
    import oracle
    import fpga_gridding_amd as fg
    rng = np.random.default_rng(seed)
    idx = np.sort(rng.choice(wl.num_points, n_sample, replace=False))
    vis = wl.vis.reshape(-1, 2)[idx].reshape(-1)
    sub = dataclasses.replace(wl, uu=np.ascontiguousarray(wl.uu[idx]),
                              vv=np.ascontiguousarray(wl.vv[idx]),
                              ww=np.ascontiguousarray(wl.ww[idx]),
                              vis=np.ascontiguousarray(vis))
    g, _, _ = fg.grid_workload_f(sub)
    out = {"sample": f"{n_sample} random points of the workload"}

    def err(a, ref):
        d = a.astype(np.float64) - ref
        return {"max_rel": float(np.abs(d).max() / np.abs(ref).max()),
                "rms_rel": float(np.sqrt(np.mean(d * d)) /
                                 np.sqrt(np.mean(ref * ref)))}

    ref_d = oracle.grid_wproj_d(sub)[0]
    out["fp64_oracle"] = err(g, ref_d)
    out["exact_oracle"] = err(g, oracle.grid_wproj_f_exact(sub)[0])
    # FP32 data with FP64 index math (the FP64 oracle's cells).
    out["fp32_index64_vs_fp64_oracle"] = err(
        fg.grid_workload_index64(sub)[0], ref_d)
    # Reduced-precision plans (FP64 index math, so the FP64 oracle's cells).
    for prec, name in ((fg.GW_HALF, "fp16"), (fg.GW_BFLOAT16, "bf16")):
        out[f"{name}_vs_fp64_oracle"] = err(
            fg.grid_workload_reduced(sub, prec)[0], ref_d)
    return out


def cpu_baseline(wl, budget_s=12.0, seed=3):
    """Time the reference's oskar_grid_wproj_f on one core over random
    10000-point slices of the workload (its loop handles exactly 10000
    points per call, oskar_grid_wproj.c:155)."""
    import oracle
    lib = oracle.ref()
    kind = "reference"
    if lib is None:
        lib = oracle.lib()
        kind = "port"
    fn = lib.oskar_grid_wproj_f if kind == "reference" else \
        lib.gridref_oskar_grid_wproj_f
    rng = np.random.default_rng(seed)
    n = wl.num_points
    G = wl.grid_size
    grid = np.zeros(2 * G * G, np.float32)
    slice_n = 10000
    support = np.ascontiguousarray(wl.support, np.int32)
    cube = np.ascontiguousarray(wl.conv_func, np.float32)
    libc = ct.CDLL(None)
    done_pts = 0
    done_upd = 0
    elapsed = 0.0
    slices = 0
    rec_all = None
    while (elapsed < budget_s or slices < 3) and slices < 4000:
        idx = np.sort(rng.choice(n, slice_n, replace=False))
        uu = np.ascontiguousarray(wl.uu[idx])
        vv = np.ascontiguousarray(wl.vv[idx])
        ww = np.ascontiguousarray(wl.ww[idx])
        vis = np.ascontiguousarray(wl.vis.reshape(-1, 2)[idx].reshape(-1))
        sk = ct.c_size_t(0)
        nrm = ct.c_double(0.0)
        # The reference prints a progress line every 1000 points.
        sys.stdout.flush()
        saved = os.dup(1)
        null = os.open(os.devnull, os.O_WRONLY)
        os.dup2(null, 1)
        try:
            t0 = time.perf_counter()
            fn(wl.num_w_planes, support, wl.oversample, wl.conv_size_half,
               cube, slice_n, uu, vv, ww, vis, None, wl.cell_size_rad,
               wl.w_scale, G, ct.byref(sk), ct.byref(nrm), grid)
            elapsed += time.perf_counter() - t0
        finally:
            libc.fflush(None)
            os.dup2(saved, 1)
            os.close(saved)
            os.close(null)
        slices += 1
        done_pts += slice_n
        sub = dataclasses.replace(wl, uu=uu, vv=vv, ww=ww, vis=vis)
        done_upd += oracle.locate_f(sub)[1]
    return dict(value=done_pts / elapsed / 1e6, unit="Mvis/s", cores=1,
                kind=kind,
                sample=f"{slices} random slices of 10000 points "
                       f"({done_pts} points, {done_upd} grid updates) of the "
                       f"same workload, oskar_grid_wproj_f single-threaded",
                updates_per_s=done_upd / elapsed, seconds=elapsed)


def cpu_baseline_all_cores(wl, seconds=1.5, seed=4, threads=None):
    """The same reference call on every host core at once (up to 16
    threads, each gridding its own random 10000-point slices into its own
    grid; the call releases the GIL), for ``seconds`` of wall time: the
    host's aggregate rate for independent jobs."""
    import threading
    import oracle
    lib = oracle.ref()
    kind = "reference"
    if lib is None:
        lib = oracle.lib()
        kind = "port"
    fn = lib.oskar_grid_wproj_f if kind == "reference" else \
        lib.gridref_oskar_grid_wproj_f
    threads = threads or max(1, min(16, os.cpu_count() or 1))
    G = wl.grid_size
    slice_n = 10000
    support = np.ascontiguousarray(wl.support, np.int32)
    cube = np.ascontiguousarray(wl.conv_func, np.float32)
    rng = np.random.default_rng(seed)
    work = []
    for _ in range(threads):
        idx = [np.sort(rng.choice(wl.num_points, slice_n, replace=False))
               for _ in range(4)]
        work.append([tuple(np.ascontiguousarray(a[i]) for a in
                           (wl.uu, wl.vv, wl.ww)) +
                     (np.ascontiguousarray(wl.vis.reshape(-1, 2)[i]
                                           .reshape(-1)),) for i in idx])
    grids = [np.zeros(2 * G * G, np.float32) for _ in range(threads)]
    counts = [0] * threads
    stop = time.perf_counter() + seconds

    def run(t):
        sk = ct.c_size_t(0)
        nrm = ct.c_double(0.0)
        k = 0
        while time.perf_counter() < stop:
            uu, vv, ww, vis = work[t][k % len(work[t])]
            fn(wl.num_w_planes, support, wl.oversample, wl.conv_size_half,
               cube, slice_n, uu, vv, ww, vis, None, wl.cell_size_rad,
               wl.w_scale, G, ct.byref(sk), ct.byref(nrm), grids[t])
            counts[t] += slice_n
            k += 1

    # The reference prints a progress line every 1000 points.
    libc = ct.CDLL(None)
    sys.stdout.flush()
    saved = os.dup(1)
    null = os.open(os.devnull, os.O_WRONLY)
    os.dup2(null, 1)
    try:
        t0 = time.perf_counter()
        th = [threading.Thread(target=run, args=(t,)) for t in range(threads)]
        for x in th:
            x.start()
        for x in th:
            x.join()
        wall = time.perf_counter() - t0
    finally:
        libc.fflush(None)
        os.dup2(saved, 1)
        os.close(saved)
        os.close(null)
    return {"value": sum(counts) / wall / 1e6, "unit": "Mvis/s",
            "cores": threads, "kind": kind,
            "sample": f"{sum(counts)} points in {wall:.1f} s wall, "
                      f"{threads} threads, each its own grid"}


def dropin_rate(wl, calls=2):
    """Mvis/s of the drop-in oskar_grid_wproj_f on host arrays (uploads of
    the points, read-modify-write of the touched grid rows over PCIe)."""
    import fpga_gridding_amd as fg
    grid = np.zeros(2 * wl.grid_size ** 2, np.float32)
    fg.grid_workload_f(wl, grid=grid)           # warm-up (tables uploaded)
    t0 = time.perf_counter()
    for _ in range(calls):
        fg.grid_workload_f(wl, grid=grid)
    ms = (time.perf_counter() - t0) / calls * 1e3
    return {"value": wl.num_points / (ms * 1e-3) / 1e6, "unit": "Mvis/s",
            "ms_per_call": ms,
            "what": "oskar_grid_wproj_f on host arrays, grid accumulated "
                    "in host memory"}


def precision_variants(wl, d_uvw, d_vis32, d_grid32, args):
    """Mvis/s of the FP32 plan with FP64 index math and of the FP16 and BF16
    plans (16-bit visibilities, compensated two-float grid) on the same
    resident points, same steps as the headline."""
    import fpga_gridding_amd as fg
    out = {}
    G = wl.grid_size
    plan = fg.Plan(fg.GW_SINGLE_INDEX64, wl.num_w_planes, wl.support,
                   wl.oversample, wl.conv_size_half, wl.conv_func,
                   float(np.float32(wl.cell_size_rad)),
                   float(np.float32(wl.w_scale)), G)
    for _ in range(args.warmup):
        plan.grid(wl.num_points, *d_uvw, d_vis32, d_grid32)
        plan.results()
    fg.device_synchronize()
    stage = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.grid(wl.num_points, *d_uvw, d_vis32, d_grid32)
        plan.results()
        stage.append(plan.timings())
    fg.device_synchronize()
    ms = (time.perf_counter() - t0) / args.steps * 1e3
    out["fp32_index64"] = {"value": wl.num_points / (ms * 1e-3) / 1e6,
                           "unit": "Mvis/s", "ms_per_step": ms,
                           "scatter_ms": float(np.mean([x[3] for x in stage])),
                           "storage": "fp32 taps, visibilities and grid, FP64 "
                                      "index math"}
    plan.close()
    for prec, name in ((fg.GW_HALF, "fp16"), (fg.GW_BFLOAT16, "bf16")):
        plan = fg.Plan(prec, wl.num_w_planes, wl.support, wl.oversample,
                       wl.conv_size_half, wl.conv_func,
                       float(np.float32(wl.cell_size_rad)),
                       float(np.float32(wl.w_scale)), G)
        d_vis = fg.DeviceArray.from_host(fg.to_storage(wl.vis, prec))
        d_comp = fg.DeviceArray(16 * G * G)
        d_comp.zero()
        for _ in range(args.warmup):
            plan.grid(wl.num_points, *d_uvw, d_vis, d_comp)
            plan.results()
        fg.device_synchronize()
        stage = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            plan.grid(wl.num_points, *d_uvw, d_vis, d_comp)
            plan.results()
            stage.append(plan.timings())
        fg.device_synchronize()
        ms = (time.perf_counter() - t0) / args.steps * 1e3
        out[name] = {"value": wl.num_points / (ms * 1e-3) / 1e6,
                     "unit": "Mvis/s", "ms_per_step": ms,
                     "scatter_ms": float(np.mean([s[3] for s in stage])),
                     "storage": f"{name} taps and visibilities, FP32 products,"
                                " FP64 index math, compensated two-float grid"}
        plan.close()
        d_vis.close()
        d_comp.close()
    return out


MIX_WORKLOAD = ("64 jobs from the 512-station workload: random subsets of "
                "log-uniform size 10k-5M points, grid from {1024, 2048, 4096, "
                "8192}^2 (5 deg FOV at 8192), 8 kernel sets of 16-128 "
                "w-planes, support 4..4-36, oversample 4 (seed 7)")


def _mix_setup(precision, autotune=True):
    """The 64-job mix on the device: plans (one per kernel set and grid
    side, autotuned on their largest job), points and one grid per job."""
    import fpga_gridding_amd as fg
    from fpga_gridding_amd import synth
    base = synth.full_scale()
    jobs, _, sets = synth.batch_mix(64, seed=7, base=base)
    plans = {}
    for wl in jobs:
        key = (wl.params["kernel_set"], wl.grid_size)
        if key not in plans:
            plans[key] = fg.Plan.for_workload(wl, precision)
    dev = []
    for wl in jobs:
        d = [fg.DeviceArray.from_host(a) for a in (wl.uu, wl.vv, wl.ww,
                                                    wl.vis)]
        g = fg.DeviceArray(8 * wl.grid_size ** 2)
        g.zero()
        dev.append((d, g))
    tune_s = 0.0
    if autotune:
        t0 = time.perf_counter()
        largest = {}
        for j, wl in enumerate(jobs):
            key = (wl.params["kernel_set"], wl.grid_size)
            if key not in largest or wl.num_points > \
                    jobs[largest[key]].num_points:
                largest[key] = j
        for key, j in largest.items():
            d, g = dev[j]
            plans[key].autotune(jobs[j].num_points, *d, g, reps=1)
            g.zero()
        tune_s = time.perf_counter() - t0
    return jobs, sets, plans, dev, tune_s


def _vs_fp64_oracle(jobs, grids):
    """max|d|/max|g| and RMS(d)/RMS(g) of each grid against the FP64 oracle
    on its job (worst over the jobs)."""
    import oracle
    mx, rms = 0.0, 0.0
    for wl, g in zip(jobs, grids):
        ref = oracle.grid_wproj_d(wl)[0]
        d = g.astype(np.float64) - ref
        if np.any(ref):
            mx = max(mx, float(np.abs(d).max() / np.abs(ref).max()))
            rms = max(rms, float(np.sqrt(np.mean(d * d)) /
                                 np.sqrt(np.mean(ref * ref))))
    return {"jobs": len(jobs), "fp64_oracle": {"max_rel": mx,
                                                "rms_rel": rms}}


def bench_batched(args):
    """Batched jobs: the 64-job mix in one pipeline with one persistent
    scatter launch, replayed from a HIP graph, against 64 sequential plain
    per-job calls (gw_grid_device + gw_plan_results each) on the same plans
    and resident data."""
    import fpga_gridding_amd as fg
    prec = fg.GW_SINGLE if args.index32 else fg.GW_SINGLE_INDEX64
    jobs, sets, plans, dev, tune_s = _mix_setup(prec)
    key = [(wl.params["kernel_set"], wl.grid_size) for wl in jobs]
    n_total = sum(wl.num_points for wl in jobs)

    def sequential():
        out = []
        for j, wl in enumerate(jobs):
            d, g = dev[j]
            plans[key[j]].grid(wl.num_points, *d, g)
            out.append(plans[key[j]].results())
        return out

    b = fg.Batch()
    for j, wl in enumerate(jobs):
        d, g = dev[j]
        b.add(plans[key[j]], wl.num_points, *d, g)
    b.capture()

    # One pass each from zeroed grids: per-job results must agree.
    small = sorted(range(len(jobs)), key=lambda j: jobs[j].grid_size)[:4]
    for _, g in dev:
        g.zero()
    ref = sequential()
    ref_g = {j: dev[j][1].download(np.empty(2 * jobs[j].grid_size ** 2,
                                            np.float32)) for j in small}
    for _, g in dev:
        g.zero()
    b.launch()
    agree = True
    worst = 0.0
    for j in range(len(jobs)):
        sk, nd, nu = b.results(j)
        agree &= (sk, nu) == (ref[j][0], ref[j][2])
        agree &= abs(nd - ref[j][1]) <= 1e-12 * max(abs(ref[j][1]), 1.0)
    acc = {}
    for j in small:
        gj = dev[j][1].download(np.empty(2 * jobs[j].grid_size ** 2,
                                         np.float32))
        s = np.abs(ref_g[j]).max()
        if s > 0:
            worst = max(worst, float(np.abs(gj.astype(np.float64) -
                                            ref_g[j]).max() / s))
        acc[j] = gj
    accuracy = _vs_fp64_oracle([jobs[j] for j in small],
                               [acc[j] for j in small])

    for _ in range(args.warmup):
        sequential()
    fg.device_synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sequential()
    fg.device_synchronize()
    seq_ms = (time.perf_counter() - t0) / args.steps * 1e3

    for _ in range(args.warmup):
        b.launch()
        b.results(0)
    fg.device_synchronize()
    dev_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.launch()
        b.results(0)
        dev_ms.append(b.last_ms())
    fg.device_synchronize()
    ms = (time.perf_counter() - t0) / args.steps * 1e3
    updates = sum(r[2] for r in ref)
    out = {
        "metric": "Mvis/s, batched jobs (one hipGraph launch) vs sequential "
                  "plain launches",
        "value": n_total / (ms * 1e-3) / 1e6, "unit": "Mvis/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": MIX_WORKLOAD, "num_jobs": len(jobs),
                   "num_points": n_total, "plans": len(plans),
                   "inputs": "resident in HBM"},
        "graph_device_ms": float(np.mean(dev_ms)),
        "grid_updates_per_s": updates / (ms * 1e-3),
        "sequential": {"value": n_total / (seq_ms * 1e-3) / 1e6,
                       "unit": "Mvis/s", "ms_per_step": seq_ms},
        "speedup_vs_sequential": seq_ms / ms,
        "autotune_s": tune_s,
        "check": {"results_identical": bool(agree),
                  "grid_max_rel_4_smallest": worst},
        "precision": "fp32_index64" if prec == fg.GW_SINGLE_INDEX64 else
                     "fp32",
        "accuracy": accuracy,
    }
    b.close()
    return out


def bench_served(args):
    """Served: requests drawn from the 64-job mix arrive as a Poisson stream;
    each adds into a resident grid of its kernel set and grid side.  The
    server (gw_server_*) is compared with the plain per-job loop (upload,
    gw_grid_device, gw_plan_results, one request at a time) on the same
    arrivals.  Request points sit in page-locked host memory (the server's
    receive buffers; --pageable: ordinary numpy arrays), for both forms."""
    import threading
    import fpga_gridding_amd as fg
    prec = fg.GW_SINGLE if args.index32 else fg.GW_SINGLE_INDEX64
    jobs, sets, plans, dev, tune_s = _mix_setup(prec, autotune=False)
    if args.pageable:
        host = [(wl.uu, wl.vv, wl.ww, wl.vis) for wl in jobs]
        pinned = []
    else:
        pinned = [[fg.PinnedArray.copy_of(a) for a in
                   (wl.uu, wl.vv, wl.ww, wl.vis)] for wl in jobs]
        host = [tuple(p.array for p in ps) for ps in pinned]
    rng = np.random.default_rng(11)
    R = args.requests
    order = rng.integers(0, len(jobs), R)
    arrive = np.cumsum(rng.exponential(1.0 / args.rate, R))
    n_req = [jobs[i].num_points for i in order]
    n_total = int(sum(n_req))

    # Plain per-job loop on the plans above (kernel tables resident, a
    # device grid per job), one staging buffer for the uploaded points.
    n_max = max(wl.num_points for wl in jobs)
    stage = [fg.DeviceArray(4 * n_max) for _ in range(3)] + \
        [fg.DeviceArray(8 * n_max)]
    key = [(wl.params["kernel_set"], wl.grid_size) for wl in jobs]
    lat_plain = []
    t0 = time.perf_counter()
    for r in range(R):
        wait = t0 + arrive[r] - time.perf_counter()
        if wait > 0:
            time.sleep(wait)
        wl = jobs[order[r]]
        for d, a in zip(stage, host[order[r]]):
            d.upload(a)
        pl = plans[key[order[r]]]
        pl.grid(wl.num_points, *stage, dev[order[r]][1])
        pl.results()
        lat_plain.append(time.perf_counter() - (t0 + arrive[r]))
    plain_span = time.perf_counter() - t0
    for x in stage:
        x.close()
    for d, g in dev:
        for x in (*d, g):
            x.close()
    for p in plans.values():
        p.close()

    sv = fg.Server(args.hbm_budget_gib << 30, max_batch_jobs=args.max_batch,
                   num_streams=args.streams, precision=prec)
    kid = [sv.add_kernel_set(k.num_w_planes, k.support, k.oversample,
                             k.conv_size_half, k.conv_func) for k in sets]
    gids = {}
    for wl in jobs:
        k = (wl.params["kernel_set"], wl.grid_size)
        if k not in gids:
            gids[k] = sv.create_grid(wl.grid_size)

    def submit(i):
        wl = jobs[i]
        return sv.submit(kid[wl.params["kernel_set"]], wl.grid_size,
                         wl.cell_size_rad, wl.w_scale, *host[i],
                         grid_id=gids[key[i]])

    # Warm-up: every job once (plans built and resident).
    for rid in [submit(i) for i in range(len(jobs))]:
        sv.wait(rid)
    rids = [None] * R

    def producer():
        base = time.perf_counter()
        for r in range(R):
            wait = base + arrive[r] - time.perf_counter()
            if wait > 0:
                time.sleep(wait)
            rids[r] = submit(order[r])

    th = threading.Thread(target=producer)
    t0 = time.perf_counter()
    th.start()
    res = []
    for r in range(R):
        while rids[r] is None:
            time.sleep(1e-4)
        res.append(sv.wait(rids[r])[0])
    th.join()
    served_span = time.perf_counter() - t0
    lat = np.array([x.t_done - x.t_submit for x in res])
    first = min(x.t_submit for x in res)
    last = max(x.t_done for x in res)
    st = sv.stats()
    # Accuracy: the four smallest jobs once more, each into a fresh grid.
    small = sorted(range(len(jobs)), key=lambda j: jobs[j].num_points)[:4]
    outs = []
    for j in small:
        wl = jobs[j]
        rid = sv.submit(kid[wl.params["kernel_set"]], wl.grid_size,
                        wl.cell_size_rad, wl.w_scale, wl.uu, wl.vv, wl.ww,
                        wl.vis)
        outs.append(sv.wait(rid)[1])
    accuracy = _vs_fp64_oracle([jobs[j] for j in small], outs)
    sv.close()
    for ps in pinned:
        for p in ps:
            p.close()
    lp = np.array(lat_plain)
    return {
        "metric": "sustained Mvis/s and p50/p99 job latency, served "
                  "requests vs the plain per-job loop",
        "value": n_total / (last - first) / 1e6, "unit": "Mvis/s",
        "n_gpus": 1, "steps": R, "warmup": len(jobs),
        "ms_per_step": (last - first) / R * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": MIX_WORKLOAD, "requests": R,
                   "arrivals": f"Poisson, {args.rate} requests/s",
                   "num_points": n_total, "grids": "resident per kernel "
                   "set and grid side", "hbm_budget_gib":
                   args.hbm_budget_gib, "max_batch_jobs": args.max_batch,
                   "streams": args.streams, "request_points": "pageable host memory"
                   if args.pageable else "page-locked host memory"},
        "latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3),
                       "p99": float(np.percentile(lat, 99) * 1e3)},
        "offered_mvis_s": n_total / arrive[-1] / 1e6,
        "served_span_s": served_span,
        "server_stats": st,
        "precision": "fp32_index64" if prec == fg.GW_SINGLE_INDEX64 else
                     "fp32",
        "accuracy": accuracy,
        "plain_loop": {"value": n_total / plain_span / 1e6,
                       "unit": "Mvis/s",
                       "latency_ms": {"p50": float(np.percentile(lp, 50)
                                                   * 1e3),
                                      "p99": float(np.percentile(lp, 99)
                                                   * 1e3)}},
    }


def _batched_summary(r):
    """The batched run's figures for the default line."""
    return {"workload": MIX_WORKLOAD, "num_jobs": r["config"]["num_jobs"],
            "num_points": r["config"]["num_points"],
            "value": r["value"], "unit": "Mvis/s",
            "ms_per_batch": r["ms_per_step"],
            "graph_device_ms": r["graph_device_ms"],
            "sequential": r["sequential"],
            "speedup_vs_sequential": r["speedup_vs_sequential"],
            "results_identical_to_sequential": r["check"]["results_identical"],
            "grid_max_rel_vs_sequential": r["check"][
                "grid_max_rel_4_smallest"],
            "fp64_oracle": r["accuracy"]["fp64_oracle"],
            "fp64_oracle_jobs": r["accuracy"]["jobs"],
            "precision": r["precision"], "autotune_s": r["autotune_s"]}


def _served_summary(r):
    """The served run's figures for the default line."""
    return {"workload": MIX_WORKLOAD, "requests": r["config"]["requests"],
            "arrivals": r["config"]["arrivals"],
            "value": r["value"], "unit": "Mvis/s",
            "latency_ms": r["latency_ms"],
            "offered_mvis_s": r["offered_mvis_s"],
            "plain_loop": r["plain_loop"],
            "fp64_oracle": r["accuracy"]["fp64_oracle"],
            "fp64_oracle_jobs": r["accuracy"]["jobs"],
            "precision": r["precision"],
            "request_points": r["config"]["request_points"],
            "hbm_budget_gib": r["config"]["hbm_budget_gib"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true",
                    help="skip the CPU baseline and the oracle error sample")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the FP16 / BF16 plan timings")
    ap.add_argument("--no-mix", action="store_true",
                    help="skip the batched and served runs of the 64-job "
                         "mix in the default (plain) run")
    ap.add_argument("--config", choices=("plain", "batched", "served"),
                    default="plain",
                    help="plain: the 512-station workload (default); "
                         "batched: the 64-job mix in one graph launch; "
                         "served: Poisson requests through the server")
    ap.add_argument("--rate", type=float, default=200.0,
                    help="served: request arrival rate (requests/s)")
    ap.add_argument("--requests", type=int, default=400,
                    help="served: number of requests")
    ap.add_argument("--index32", action="store_true",
                    help="batched / served: FP32 index math (GW_SINGLE) "
                         "instead of FP64 (GW_SINGLE_INDEX64)")
    ap.add_argument("--streams", type=int, default=2,
                    help="served: server streams (slots)")
    ap.add_argument("--max-batch", type=int, default=16,
                    help="served: most requests in one server batch")
    ap.add_argument("--pageable", action="store_true",
                    help="served: request points in pageable host memory "
                         "instead of page-locked")
    ap.add_argument("--hbm-budget-gib", type=int, default=64,
                    help="served: device-memory budget of the server")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="profiler calibration: copy this many bytes once "
                         "with the library's k_copy kernel and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not "
              "supported (the gridder runs on a single MI355X; sharding "
              "over several GPUs is designed but not built)", file=sys.stderr)
        sys.exit(2)

    import fpga_gridding_amd as fg
    from fpga_gridding_amd import synth

    if args.copy_bytes > 0:
        src = fg.DeviceArray(args.copy_bytes)
        dst = fg.DeviceArray(args.copy_bytes)
        src.zero()
        dst.copy_from(src)
        print(json.dumps({"copy_bytes": args.copy_bytes, "kernel": "k_copy"}))
        return
    if args.config == "batched":
        print(json.dumps(bench_batched(args)))
        return
    if args.config == "served":
        print(json.dumps(bench_served(args)))
        return

    wl = synth.full_scale()
    plan = fg.Plan.for_workload(wl)
    d_uvw = [fg.DeviceArray.from_host(a) for a in (wl.uu, wl.vv, wl.ww)]
    d_vis = fg.DeviceArray.from_host(wl.vis)
    d_grid = fg.DeviceArray(8 * wl.grid_size ** 2)
    d_grid.zero()
    fg.device_synchronize()

    def step():
        plan.grid(wl.num_points, *d_uvw, d_vis, d_grid)
        return plan.results()

    for _ in range(args.warmup):
        step()
    fg.device_synchronize()
    stage_ms = []
    step_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        skipped, norm_delta, updates = step()
        step_ms.append((time.perf_counter() - ts) * 1e3)
        stage_ms.append(plan.timings())
    fg.device_synchronize()
    wall = time.perf_counter() - t0
    records, record_updates = plan.records()

    n = wl.num_points
    ms_per_step = wall / args.steps * 1e3
    stage = np.array(stage_ms)          # [steps, 5]
    grid_ms = float(stage[:, 3].mean())
    flops = 8.0 * updates
    achieved = flops / (grid_ms * 1e-3) / 1e12
    traffic = _traffic_from_profile()

    out = {
        "metric": "Mvis/s and grid updates/s vs FP32 roofline; p50/p99 "
                  "job latency at fixed load",
        "value": n / (ms_per_step * 1e-3) / 1e6,
        "unit": "Mvis/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": WORKLOAD, "num_points": n,
                   "grid_size": wl.grid_size,
                   "num_w_planes": wl.num_w_planes,
                   "support": [int(wl.support.min()), int(wl.support.max())],
                   "oversample": wl.oversample,
                   "inputs": "resident in HBM (plan API)"},
        "grid_updates_per_s": updates / (ms_per_step * 1e-3),
        "grid_updates": int(updates),
        "num_skipped": int(skipped),
        "step_latency_ms": {"p50": float(np.percentile(step_ms, 50)),
                            "p99": float(np.percentile(step_ms, 99))},
        "stage_ms_mean": {"locate": float(stage[:, 0].mean()),
                          "sort": float(stage[:, 1].mean()),
                          "gather_plan": float(stage[:, 2].mean()),
                          "scatter": grid_ms,
                          "total": float(stage[:, 4].mean())},
        "hbm_counters": None if traffic is None else {
            "bytes_per_launch": traffic,
            "achieved_gbs": traffic / (grid_ms * 1e-3) / 1e9,
            "peak_gbs": HBM_PEAK_GBS,
            "frac": traffic / (grid_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "source": "rocprofv3 FETCH_SIZE + WRITE_SIZE of k_grid<float>, "
                      "copy-calibrated (profiles/k_grid_ska512_traffic.txt), "
                      "over this run's scatter time"},
        "roofline": {"bound": "valu", "achieved": achieved,
                     "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS,
                     "traffic": traffic,
                     "kernel": "k_grid<float> (scatter)",
                     "compute": "FP32 VALU FMA, no MFMA; peak is MI355X "
                                "FP32 157.3 TFLOP/s",
                     "flops_per_launch": flops,
                     "flops": "algorithmic: 8 per grid update of the "
                              "reference loop, point by point"},
        "executed": {"records": int(records),
                     "grid_updates": int(record_updates),
                     "frac_of_algorithmic": record_updates / updates
                     if updates else None,
                     "tflops": 8.0 * record_updates / (grid_ms * 1e-3) / 1e12,
                     "what": "the scatter takes one record per run of "
                             "points with equal origin, w-plane and "
                             "sub-cell table (visibility terms summed)"},
    }
    if not args.no_variants:
        out["precision_variants"] = precision_variants(wl, d_uvw, d_vis,
                                                       d_grid, args)
        out["dropin_pcie_inclusive"] = dropin_rate(wl)
    if not args.no_cpu:
        cb = cpu_baseline(wl)
        out["cpu_baseline"] = {k: cb[k] for k in
                               ("value", "unit", "cores", "kind", "sample")}
        out["cpu_baseline"]["grid_updates_per_s"] = cb["updates_per_s"]
        out["cpu_baseline"]["all_cores"] = cpu_baseline_all_cores(wl)
        out["accuracy"] = accuracy_sample(wl)
    else:
        out["cpu_baseline"] = None
    if not args.no_mix:
        plan.close()
        for x in (*d_uvw, d_vis, d_grid):
            x.close()
        out["batched"] = _batched_summary(bench_batched(args))
        out["served"] = _served_summary(bench_served(args))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
