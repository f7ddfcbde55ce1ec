"""Reduced-precision plans (GW_HALF, GW_BFLOAT16) through the C ABI.

These plans store taps and visibilities in 16 bits, take their indices from
FP64 index math (those of oskar_grid_wproj_d), form FP32 products, sum each
job in FP64 and add it to a compensated two-float grid.  Tolerances:

* num_skipped: bit-exact against the FP64 oracle on the same float
  coordinates.
* grid against the FP64 oracle run on the *stored* inputs (visibilities and
  kernel cube rounded to 16 bits, then widened): within 1e-5 * max|grid| --
  only the FP32 products and per-run FP32 sums differ from the oracle.
* norm against the same oracle: relative 1e-12 (FP64 sums of the stored
  taps in another order).
* grid against the FP64 oracle on the original FP32 inputs: the storage
  rounding itself, bounded by 2e-3 (FP16, 11-bit significand) and 2e-2
  (BF16, 8-bit significand) of max|grid|; the measured values are printed.
"""
import dataclasses

import numpy as np
import pytest

import oracle
import fpga_gridding_amd as fg
from fpga_gridding_amd import synth

from parity import max_rel, rms_rel

pytestmark = pytest.mark.gpu

PRECISIONS = [(fg.GW_HALF, "fp16", 2e-3), (fg.GW_BFLOAT16, "bf16", 2e-2)]


@pytest.fixture(scope="module", autouse=True)
def _gpu_present():
    if fg.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    yield
    fg.lib().gw_release_cached()


def _stored(wl, prec):
    """The workload with visibilities and kernel cube as the plan stores
    them (rounded to 16 bits, widened back to float32)."""
    return dataclasses.replace(
        wl, vis=fg.from_storage(fg.to_storage(wl.vis, prec), prec),
        conv_func=fg.from_storage(fg.to_storage(wl.conv_func, prec), prec))


CASES = [
    dict(n=3000, grid_size=96, num_planes=6, s_min=1, s_max=6, seed=11),
    dict(n=20000, grid_size=256, num_planes=16, s_min=2, s_max=14, seed=12,
         edge_fraction=0.1),
]


@pytest.mark.parametrize("prec,name,storage_tol", PRECISIONS,
                         ids=[p[1] for p in PRECISIONS])
@pytest.mark.parametrize("case", CASES, ids=["small", "medium"])
def test_reduced_matches_fp64_oracle_on_stored_inputs(prec, name,
                                                      storage_tol, case):
    wl = synth.random_small(**case)
    g, sk, nrm, _ = fg.grid_workload_reduced(wl, prec)
    g_s, sk_s, n_s = oracle.grid_wproj_d(_stored(wl, prec))
    assert sk == sk_s
    assert abs(nrm - n_s) <= 1e-12 * abs(n_s)
    assert max_rel(g, g_s) <= 1e-5
    g_d, sk_d, _ = oracle.grid_wproj_d(wl)
    assert sk == sk_d
    err = max_rel(g, g_d)
    print(f"{name}: max_rel vs FP64 oracle {err:.3e}, "
          f"rms_rel {rms_rel(g, g_d):.3e}")
    assert err <= storage_tol


@pytest.mark.parametrize("prec,name,storage_tol", PRECISIONS,
                         ids=[p[1] for p in PRECISIONS])
def test_reduced_snapshot64(prec, name, storage_tol):
    """The 64-antenna snapshot workload (725,760 points, grid 2048^2)."""
    wl = synth.snapshot_small()
    g, sk, nrm, _ = fg.grid_workload_reduced(wl, prec)
    g_s, sk_s, n_s = oracle.grid_wproj_d(_stored(wl, prec))
    assert sk == sk_s == 0
    assert abs(nrm - n_s) <= 1e-12 * abs(n_s)
    assert max_rel(g, g_s) <= 1e-5
    g_d, _, _ = oracle.grid_wproj_d(wl)
    err = max_rel(g, g_d)
    print(f"{name} snapshot: max_rel vs FP64 oracle {err:.3e}, "
          f"rms_rel {rms_rel(g, g_d):.3e}")
    assert err <= storage_tol


@pytest.mark.parametrize("prec", [fg.GW_HALF, fg.GW_BFLOAT16],
                         ids=["fp16", "bf16"])
def test_compensated_grid_accumulates(prec):
    """Two calls into one compensated grid give twice one call, and lo
    stays a correction: each add leaves its exact rounding error in lo, so
    |lo| is a few ulp of the largest partial sums of its cell (where the
    adds cancel, hi can end below lo), far below max|hi|."""
    wl = synth.random_small(n=5000, grid_size=128, num_planes=8, s_min=2,
                            s_max=10, seed=13)
    g1, sk1, n1, c1 = fg.grid_workload_reduced(wl, prec)
    g2, sk2, n2, c2 = fg.grid_workload_reduced(wl, prec, comp=c1)
    assert sk1 == sk2 and n1 == n2
    assert max_rel(g2, 2.0 * g1.astype(np.float64)) <= 1e-6
    hi = c2.reshape(-1, 2)[:, 0].astype(np.float64)
    lo = c2.reshape(-1, 2)[:, 1].astype(np.float64)
    assert np.abs(lo).max() <= 2.0 ** -18 * np.abs(hi).max()
    assert np.count_nonzero(lo) > 0


@pytest.mark.parametrize("prec", [fg.GW_HALF, fg.GW_BFLOAT16],
                         ids=["fp16", "bf16"])
def test_reduced_empty_and_all_skipped(prec):
    wl = synth.random_small(n=200, grid_size=64, num_planes=4, s_min=1,
                            s_max=4, seed=14)
    empty = dataclasses.replace(wl, uu=wl.uu[:0], vv=wl.vv[:0], ww=wl.ww[:0],
                                vis=wl.vis[:0])
    g, sk, nrm, _ = fg.grid_workload_reduced(empty, prec)
    assert sk == 0 and nrm == 0.0 and not np.any(g)
    far = dataclasses.replace(wl, uu=np.full_like(wl.uu, 1e9))
    g, sk, nrm, _ = fg.grid_workload_reduced(far, prec)
    assert sk == wl.num_points and nrm == 0.0 and not np.any(g)
