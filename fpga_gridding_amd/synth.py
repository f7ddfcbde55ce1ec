"""Synthetic W-projection gridding workloads.

Builds the inputs of ``oskar_grid_wproj_f`` / ``oskar_grid_wproj_d``
(reference API: oskar_grid_wproj.h:68-131) for the benchmark workloads:
station layouts, uvw tracks in wavelengths, complex64 visibilities, weights,
the support table and the oversampled W-kernel cube in the reference's
quadrant layout ``conv_func[2*((p*C + iy)*C + ix) + {0,1}]``
(oskar_grid_wproj.c:222-233, cube interleave main.cpp:197-201).

Everything is generated in float64 with numpy and cast once at the end, the
way the reference's readers hand float32 data to the gridder
(read_vis.c:49-73 scales uvw to wavelengths; main.cpp:197-201 interleaves the
kernel cube).

Parameters the workload descriptions leave open are fixed here:

* site latitude -27 deg, pointing declination -27 deg (source at zenith at
  transit), observation centred on transit (hour angle 0);
* "N km disk" is a disk of DIAMETER N km (radius N/2 km) for the uniform
  layout; the core + log-radial layout places half of the stations uniformly
  in area inside 1 km radius and the rest with log10(radius) uniform between
  1 km and 40 km; azimuths uniform; flat array (U = 0);
* baselines are ordered (i, j) with i < j, i outer, and samples are
  time-major ([time][baseline]), the order of an OSKAR visibility block;
* visibilities: real and imaginary parts each standard normal (sigma = 1
  per component), seed 1; weights seed 2; layout seed 0;
* w_scale = (P - 1)^2 / max|w| so that plane = round(sqrt(|w| * w_scale))
  spans [0, P - 1]; plane p corresponds to w_p = (p / (P - 1))^2 * max|w|;
* support S_p = S_min + round((S_max - S_min) * sqrt(p / (P - 1)));
* cube half-size C = h + S_max*ov + 1 with h = (ov + 1) // 2, the largest
  sub-cell offset |off| (the largest tap index read is |off + S*ov|);
* kernel for plane p at oversampled index (ix, iy), r^2 = (ix/ov)^2 +
  (iy/ov)^2 in pixels: exp(-r^2 / (2 sigma^2)) * exp(i pi w_p s r^2), with
  sigma = S_p / 2.5 and chirp scale s = 2 / (max|w| * S_max^2) (the top
  plane's phase reaches 2 pi at r = S_max);
* "unit sum": each plane is divided by the complex sum of its taps at zero
  sub-cell offset over the (2S+1)^2 footprint, so those taps sum to 1 + 0i
  in float64 before the cast to complex64.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

C_LIGHT = 299792458.0
OMEGA_EARTH = 7.2921150e-5  # sidereal rotation rate, rad/s


# --------------------------------------------------------------------------
# Array layouts and uvw tracks
# --------------------------------------------------------------------------

def layout_uniform_disk(n: int, radius_m: float, rng) -> np.ndarray:
    """n stations uniform in area inside a disk; returns ENU [n, 3] metres."""
    r = radius_m * np.sqrt(rng.random(n))
    az = 2.0 * np.pi * rng.random(n)
    return np.stack([r * np.sin(az), r * np.cos(az), np.zeros(n)], axis=1)


def layout_core_lograd(n: int, core_radius_m: float, outer_radius_m: float,
                       rng) -> np.ndarray:
    """Half of the stations uniform in the core disk, the rest log-radial."""
    n_core = n // 2
    n_out = n - n_core
    core = layout_uniform_disk(n_core, core_radius_m, rng)
    lr = rng.uniform(math.log10(core_radius_m), math.log10(outer_radius_m),
                     n_out)
    r = 10.0 ** lr
    az = 2.0 * np.pi * rng.random(n_out)
    outer = np.stack([r * np.sin(az), r * np.cos(az), np.zeros(n_out)], axis=1)
    return np.concatenate([core, outer], axis=0)


def baseline_vectors(enu: np.ndarray) -> np.ndarray:
    """Baselines s_j - s_i for i < j, i outer (OSKAR order); [B, 3]."""
    n = enu.shape[0]
    ii, jj = np.triu_indices(n, k=1)
    return enu[jj] - enu[ii]


def uvw_tracks(b_enu: np.ndarray, lat_rad: float, dec_rad: float,
               ha_rad: np.ndarray):
    """uvw in metres, time-major arrays of shape [T, B]."""
    e, n_, u_ = b_enu[:, 0], b_enu[:, 1], b_enu[:, 2]
    x = -math.sin(lat_rad) * n_ + math.cos(lat_rad) * u_
    y = e
    z = math.cos(lat_rad) * n_ + math.sin(lat_rad) * u_
    sh = np.sin(ha_rad)[:, None]
    ch = np.cos(ha_rad)[:, None]
    sd, cd = math.sin(dec_rad), math.cos(dec_rad)
    uu = sh * x + ch * y
    vv = -sd * ch * x + sd * sh * y + cd * z
    ww = cd * ch * x - cd * sh * y + sd * z
    return uu, vv, ww


# --------------------------------------------------------------------------
# W-kernels
# --------------------------------------------------------------------------

def support_table(num_planes: int, s_min: int, s_max: int) -> np.ndarray:
    if num_planes == 1:
        return np.array([s_min], dtype=np.int32)
    p = np.arange(num_planes, dtype=np.float64)
    s = s_min + np.round((s_max - s_min) * np.sqrt(p / (num_planes - 1)))
    return s.astype(np.int32)


def cube_half_size(oversample: int, s_max: int) -> int:
    return (oversample + 1) // 2 + s_max * oversample + 1


def wkernel_cube(support: np.ndarray, oversample: int, conv_size_half: int,
                 w_max: float) -> np.ndarray:
    """Oversampled W-kernel cube, complex128 [P, C, C] (quadrant layout)."""
    num_planes = len(support)
    s_max = int(support.max())
    c = conv_size_half
    ov = oversample
    idx = np.arange(c, dtype=np.float64) / ov
    r2 = idx[None, :] ** 2 + idx[:, None] ** 2          # [iy, ix]
    chirp_scale = 2.0 / (max(w_max, 1e-30) * max(s_max, 1) ** 2)
    cube = np.empty((num_planes, c, c), dtype=np.complex128)
    for p in range(num_planes):
        s = int(support[p])
        sigma = max(s, 1) / 2.5
        w_p = (p / (num_planes - 1)) ** 2 * w_max if num_planes > 1 else 0.0
        k = np.exp(-r2 / (2.0 * sigma * sigma)) * \
            np.exp(1j * np.pi * w_p * chirp_scale * r2)
        taps = np.abs(np.arange(-s, s + 1)) * ov
        total = k[np.ix_(taps, taps)].sum()
        cube[p] = k / total
    return cube


def cube_to_conv_func(cube: np.ndarray, dtype=np.float32) -> np.ndarray:
    """Interleave a complex cube into conv_func[2*((p*C+iy)*C+ix)+{0,1}]."""
    out = np.empty(cube.shape + (2,), dtype=dtype)
    out[..., 0] = cube.real
    out[..., 1] = cube.imag
    return out.reshape(-1)


# --------------------------------------------------------------------------
# Workloads
# --------------------------------------------------------------------------

@dataclass
class Workload:
    """All arguments of one oskar_grid_wproj_f call (float32 inputs)."""
    name: str
    num_w_planes: int
    support: np.ndarray          # int32 [P]
    oversample: int
    conv_size_half: int
    conv_func: np.ndarray        # float32 [2*P*C*C]
    uu: np.ndarray               # float32 [N] wavelengths
    vv: np.ndarray
    ww: np.ndarray
    vis: np.ndarray              # float32 [2N]
    weight: np.ndarray           # float32 [N]
    cell_size_rad: float         # value as passed (float32 in _f)
    w_scale: float
    grid_size: int
    params: dict = field(default_factory=dict)

    @property
    def num_points(self) -> int:
        return int(self.uu.shape[0])

    def conv_func_d(self) -> np.ndarray:
        return self.conv_func.astype(np.float64)


def _finish(name, uu_m, vv_m, ww_m, freq_hz, num_planes, s_min, s_max, ov,
            grid_size, fov_deg, weights, rng_vis, params) -> Workload:
    # Scale to wavelengths in float64, then cast (uvw "in wavelengths").
    scale = freq_hz / C_LIGHT
    uu = (uu_m * scale).astype(np.float32).reshape(-1)
    vv = (vv_m * scale).astype(np.float32).reshape(-1)
    ww = (ww_m * scale).astype(np.float32).reshape(-1)
    n = uu.shape[0]
    vis = rng_vis.standard_normal(2 * n).astype(np.float32)
    w_max = float(np.max(np.abs(ww.astype(np.float64))))
    support = support_table(num_planes, s_min, s_max)
    c = cube_half_size(ov, s_max)
    cube = wkernel_cube(support, ov, c, w_max)
    w_scale = np.float32((num_planes - 1) ** 2 / w_max) if w_max > 0 else \
        np.float32(0.0)
    cell = np.float32(math.radians(fov_deg) / grid_size)
    p = dict(params)
    p.update(freq_hz=freq_hz, fov_deg=fov_deg, w_max=w_max,
             chirp_scale=2.0 / (w_max * s_max * s_max) if w_max else 0.0,
             s_min=s_min, s_max=s_max)
    return Workload(name=name, num_w_planes=num_planes, support=support,
                    oversample=ov, conv_size_half=c,
                    conv_func=cube_to_conv_func(cube), uu=uu, vv=vv, ww=ww,
                    vis=vis, weight=weights.astype(np.float32),
                    cell_size_rad=float(cell), w_scale=float(w_scale),
                    grid_size=grid_size, params=p)


def snapshot_small(n_st=64, disk_diameter_m=2000.0, n_times=360, dt_s=10.0,
                   freq_hz=100e6, grid_size=2048, fov_deg=2.0, num_planes=32,
                   s_min=4, s_max=16, ov=4, seed=0, name="snapshot64"):
    """64-antenna 2 km disk, 1 h at 10 s, 100 MHz, 2048^2, 32 planes."""
    rng = np.random.default_rng(seed)
    enu = layout_uniform_disk(n_st, disk_diameter_m / 2.0, rng)
    b = baseline_vectors(enu)
    t = (np.arange(n_times) - (n_times - 1) / 2.0) * dt_s
    ha = t * OMEGA_EARTH
    lat = math.radians(-27.0)
    uu, vv, ww = uvw_tracks(b, lat, lat, ha)
    n = uu.size
    weights = np.ones(n, dtype=np.float32)
    return _finish(name, uu, vv, ww, freq_hz, num_planes, s_min, s_max, ov,
                   grid_size, fov_deg, weights, np.random.default_rng(seed + 1),
                   dict(n_stations=n_st, n_times=n_times, dt_s=dt_s,
                        n_baselines=b.shape[0], layout="uniform disk",
                        disk_diameter_m=disk_diameter_m, seed=seed))


def full_scale(n_st=512, n_times=100, dt_s=10.0, freq_hz=150e6,
               grid_size=8192, fov_deg=5.0, num_planes=128, s_min=4,
               s_max=36, ov=4, seed=0, name="ska512"):
    """512 stations (core + log-radial to 40 km), 100 x 10 s, 150 MHz."""
    rng = np.random.default_rng(seed)
    enu = layout_core_lograd(n_st, 1000.0, 40000.0, rng)
    b = baseline_vectors(enu)
    t = (np.arange(n_times) - (n_times - 1) / 2.0) * dt_s
    ha = t * OMEGA_EARTH
    lat = math.radians(-27.0)
    uu, vv, ww = uvw_tracks(b, lat, lat, ha)
    n = uu.size
    weights = np.random.default_rng(seed + 2).uniform(0.5, 1.0, n)
    return _finish(name, uu, vv, ww, freq_hz, num_planes, s_min, s_max, ov,
                   grid_size, fov_deg, weights, np.random.default_rng(seed + 1),
                   dict(n_stations=n_st, n_times=n_times, dt_s=dt_s,
                        n_baselines=b.shape[0], layout="core+lograd",
                        seed=seed))


def random_small(n=200, grid_size=32, num_planes=3, s_min=1, s_max=3, ov=4,
                 seed=0, edge_fraction=0.1, name="random"):
    """Small random workload for parity tests.

    Points are uniform in the uv square that maps onto the grid, with
    ``edge_fraction`` of them pushed close to the edges so the skip test
    (oskar_grid_wproj.c:207-212) fires; w is uniform in [-w_max, w_max]."""
    rng = np.random.default_rng(seed)
    cell = np.float32(math.radians(1.0) / grid_size)
    scale = float(grid_size) * float(cell)             # cells per wavelength^-1
    half = grid_size / 2.0
    pos = rng.uniform(-half + s_max + 1, half - s_max - 2, size=(n, 2))
    k = int(round(edge_fraction * n))
    if k:
        edge = rng.uniform(half - s_max - 2, half + 1.0, size=(k, 2))
        edge *= rng.choice([-1.0, 1.0], size=(k, 2))
        pos[:k] = edge
    uu = (-pos[:, 0] / scale).astype(np.float32)
    vv = (pos[:, 1] / scale).astype(np.float32)
    w_max = 1000.0
    ww = rng.uniform(-w_max, w_max, n).astype(np.float32)
    vis = rng.standard_normal(2 * n).astype(np.float32)
    support = support_table(num_planes, s_min, s_max)
    c = cube_half_size(ov, s_max)
    cube = wkernel_cube(support, ov, c, w_max)
    w_scale = np.float32((num_planes - 1) ** 2 / w_max) if num_planes > 1 \
        else np.float32(0.0)
    return Workload(name=name, num_w_planes=num_planes, support=support,
                    oversample=ov, conv_size_half=c,
                    conv_func=cube_to_conv_func(cube), uu=uu, vv=vv, ww=ww,
                    vis=vis, weight=np.ones(n, np.float32),
                    cell_size_rad=float(cell), w_scale=float(w_scale),
                    grid_size=grid_size,
                    params=dict(seed=seed, w_max=w_max))


def edge_cases(name="edges"):
    """Hand-built points that probe the index stage's corner cases.

    grid_scale is exactly 2^-10 (G = 64, cell = 2^-16 rad) and w_scale is 1,
    so positions and plane indices can be placed on exact ties:
    round-half-away-from-zero at pos = k + 0.5 (oskar_grid_wproj.c:175-176,
    :195-196) and at sqrt(|w|) = m + 0.5 (:173); footprints touching or
    crossing each grid edge (:207-208); w = 0 and -0 (conjugation only for
    w > 0, :172); denormal, NaN, +-inf and huge coordinates."""
    G = 64
    cell = np.float32(2.0 ** -16)
    gs = 2.0 ** -10                       # G * cell, exact in float32
    support = np.array([0, 1, 2, 3], np.int32)
    ov = 4
    pts = []                              # (pos_u, pos_v, w)
    for k in range(-6, 6):                # ties in u and in v
        pts.append((k + 0.5, 3.0, 1.0))
        pts.append((2.0, k + 0.5, -1.0))
        pts.append((k + 0.375, k - 0.625, 2.0))   # sub-cell ties (x ov)
        pts.append((k + 0.125, -k - 0.875, 0.5))
    for m in range(4):                    # plane ties: sqrt(|w|) = m + 0.5
        w = (m + 0.5) ** 2
        pts.append((1.0, 1.0, w))
        pts.append((1.0, 1.0, -w))
        pts.append((1.25, -1.25, np.nextafter(np.float32(w), np.float32(0))))
    for w in (0.0, -0.0, 1e-45, -1e-45, 1e-40, 20.0, 100.0, 1e6):
        pts.append((0.3, -0.7, w))
    c = G // 2
    for s_plane, w in ((0, 0.0), (1, 1.0), (2, 4.0), (3, 9.0)):
        s = int(support[s_plane])
        for gu in (s - 1, s, s + 1, G - 2 - s, G - 1 - s, G - s):
            pts.append((gu - c, 0.0, w))          # u edge
            pts.append((0.0, gu - c, w))          # v edge
    bad = [np.nan, np.inf, -np.inf, 3e9, -3e9, 1e38, -1e38, 2.1e9]
    for b in bad:
        pts.append((b, 0.0, 1.0))
        pts.append((0.0, b, 1.0))
        pts.append((1.0, 1.0, b))
    pos_u = np.array([p[0] for p in pts], np.float64)
    pos_v = np.array([p[1] for p in pts], np.float64)
    w = np.array([p[2] for p in pts], np.float64)
    # pos_u = -uu * gs  =>  uu = -pos_u / gs (exact for these values).
    with np.errstate(over="ignore", invalid="ignore"):
        uu = (-pos_u / gs).astype(np.float32)
        vv = (pos_v / gs).astype(np.float32)
    ww = w.astype(np.float32)
    n = uu.shape[0]
    rng = np.random.default_rng(7)
    vis = rng.standard_normal(2 * n).astype(np.float32)
    c_half = cube_half_size(ov, int(support.max()))
    cube = wkernel_cube(support, ov, c_half, 10.0)
    return Workload(name=name, num_w_planes=len(support), support=support,
                    oversample=ov, conv_size_half=c_half,
                    conv_func=cube_to_conv_func(cube), uu=uu, vv=vv, ww=ww,
                    vis=vis, weight=np.ones(n, np.float32),
                    cell_size_rad=float(cell), w_scale=1.0, grid_size=G,
                    params=dict(kind="edge cases"))


def workload_from_arrays(d, name="fixture"):
    """Rebuild a Workload from a dict/npz of its arrays and scalars."""
    return Workload(name=name, num_w_planes=int(d["num_w_planes"]),
                    support=np.asarray(d["support"], np.int32),
                    oversample=int(d["oversample"]),
                    conv_size_half=int(d["conv_size_half"]),
                    conv_func=np.asarray(d["conv_func"], np.float32),
                    uu=np.asarray(d["uu"], np.float32),
                    vv=np.asarray(d["vv"], np.float32),
                    ww=np.asarray(d["ww"], np.float32),
                    vis=np.asarray(d["vis"], np.float32),
                    weight=np.asarray(d["weight"], np.float32),
                    cell_size_rad=float(np.float32(d["cell_size_rad"])),
                    w_scale=float(np.float32(d["w_scale"])),
                    grid_size=int(d["grid_size"]))


def workload_arrays(wl):
    """The arrays and scalars of a Workload as a dict (for npz fixtures)."""
    return dict(num_w_planes=wl.num_w_planes, support=wl.support,
                oversample=wl.oversample, conv_size_half=wl.conv_size_half,
                conv_func=wl.conv_func, uu=wl.uu, vv=wl.vv, ww=wl.ww,
                vis=wl.vis, weight=wl.weight,
                cell_size_rad=np.float32(wl.cell_size_rad),
                w_scale=np.float32(wl.w_scale), grid_size=wl.grid_size)


@dataclass
class KernelSet:
    """A W-kernel set shared by several jobs (planes, support table, cube)."""
    num_w_planes: int
    support: np.ndarray
    oversample: int
    conv_size_half: int
    conv_func: np.ndarray
    w_scale: float


def batch_mix(n_jobs=64, seed=7, n_range=(10_000, 5_000_000),
              grids=(1024, 2048, 4096, 8192), n_kernel_sets=8,
              planes=(16, 128), s_max_range=(4, 36), ov=4, base=None,
              fov_deg_at_8192=5.0):
    """Independent gridding jobs of uneven size for batched launches.

    Each job takes a random subset (log-uniform size in n_range, kept in the
    base order) of the base workload's uvw and visibilities (default: the
    512-station workload), a grid side from ``grids`` and one of
    ``n_kernel_sets`` kernel sets (planes uniform in ``planes``, support
    s_min = min(4, s_max) .. s_max with s_max uniform in ``s_max_range``,
    kernels as in wkernel_cube, w_scale = (P-1)^2 / w_max).  The field of view
    is fov_deg_at_8192 * 8192 / G, so that every grid side covers the same
    fraction of the uv plane.  Returns (jobs: list of Workload, kernel set
    index per job, kernel sets)."""
    if base is None:
        base = full_scale()
    rng = np.random.default_rng(seed)
    w_max = float(np.max(np.abs(base.ww.astype(np.float64))))
    sets = []
    for _ in range(n_kernel_sets):
        P = int(rng.integers(planes[0], planes[1] + 1))
        s_max = int(rng.integers(s_max_range[0], s_max_range[1] + 1))
        s_min = min(4, s_max)
        support = support_table(P, s_min, s_max)
        c = cube_half_size(ov, s_max)
        cube = wkernel_cube(support, ov, c, w_max)
        sets.append(KernelSet(num_w_planes=P, support=support, oversample=ov,
                              conv_size_half=c,
                              conv_func=cube_to_conv_func(cube),
                              w_scale=float(np.float32((P - 1) ** 2 / w_max))))
    vis2 = base.vis.reshape(-1, 2)
    lo, hi = np.log(n_range[0]), np.log(n_range[1])
    jobs, which = [], []
    for j in range(n_jobs):
        k = int(rng.integers(0, n_kernel_sets))
        G = int(grids[int(rng.integers(0, len(grids)))])
        n = int(min(base.num_points, round(np.exp(rng.uniform(lo, hi)))))
        idx = np.sort(rng.choice(base.num_points, n, replace=False))
        ks = sets[k]
        fov = fov_deg_at_8192 * 8192.0 / G
        cell = float(np.float32(math.radians(fov) / G))
        jobs.append(Workload(
            name=f"job{j}", num_w_planes=ks.num_w_planes, support=ks.support,
            oversample=ov, conv_size_half=ks.conv_size_half,
            conv_func=ks.conv_func, uu=np.ascontiguousarray(base.uu[idx]),
            vv=np.ascontiguousarray(base.vv[idx]),
            ww=np.ascontiguousarray(base.ww[idx]),
            vis=np.ascontiguousarray(vis2[idx]).reshape(-1),
            weight=np.ones(n, np.float32), cell_size_rad=cell,
            w_scale=ks.w_scale, grid_size=G,
            params=dict(kernel_set=k, fov_deg=fov, base=base.name)))
        which.append(k)
    return jobs, which, sets
