// MI355X (gfx950) W-projection gridder: device kernels and the resident plan.
//
// Hot path of OxfordSKA/FPGA_Gridding: oskar_grid_wproj_f / _d
// (oskar_grid_wproj.c:39-243).  Each point's footprint of (2S+1)^2 taps is
// scattered onto the complex grid.  Pipeline, all on one HIP stream:
//
//   k_locate   per point: the reference's index math, bit-exact
//              (oskar_grid_wproj.c:169-212), skip count, per-(plane,
//              sub-cell) histogram, footprint bounding box, a sort key
//              (tile of the footprint origin, support rank, origin inside
//              the tile, plane, sub-cell table) and a 16/32-byte point
//              record;
//   sort       the library's stable LSD radix sort of the keys;
//   k_plan     points gridded, update count and norm increment (histogram x
//              per-(plane, sub-cell) tap sums, fixed summation order);
//   k_gather   sorted point records (origin, sub-cell table, plane; the
//              visibility pre-multiplied by the conjugation sign), one per
//              run of equal keys (same origin, plane and sub-cell table:
//              the run's visibility terms summed), and the heads of the
//              tile buckets;
//   scans      bucket starts and jobs (<= kJobPoints points of one tile),
//              then each job's support segments (k_item_segments);
//   k_grid     the scatter, one job per workgroup, into an FP64 LDS copy of
//              the tile's grid region, added to the grid in HBM once per job.

#include "gridder_internal.h"
#include "host_semantics.h"

#include <algorithm>
#include <cstdlib>
#include <climits>
#include <cmath>
#include <cstring>
#include <type_traits>

namespace gw {

void fail(int code, const std::string& detail)
{
    throw Error{code, detail};
}

void check_hip(hipError_t e, int code, const char* what)
{
    if (e != hipSuccess)
        fail(code, std::string(what) + ": " + hipGetErrorName(e) + " (" +
                hipGetErrorString(e) + ")");
}

__device__ __forceinline__ unsigned long long spread_bits16(unsigned x)
{
    unsigned long long v = x & 0xFFFFull;
    v = (v | (v << 8)) & 0x00FF00FFull;
    v = (v | (v << 4)) & 0x0F0F0F0Full;
    v = (v | (v << 2)) & 0x33333333ull;
    v = (v | (v << 1)) & 0x55555555ull;
    return v;
}

// Global-memory views of pointers whose address space the compiler cannot
// see (loaded from memory rather than passed as kernel arguments), so that
// accesses through them are global rather than flat instructions.
template <typename U>
using GlobalPtr = __attribute__((address_space(1))) U*;

template <typename U>
__device__ __forceinline__ GlobalPtr<U> as_global(U* p)
{
    return (GlobalPtr<U>)(uintptr_t)p;
}

// Hardware FP atomic add (no return) on global memory.
__device__ __forceinline__ void atomic_add_real(float* p, float v)
{
    (void)__builtin_amdgcn_global_atomic_fadd_f32(as_global(p), v);
}
__device__ __forceinline__ void atomic_add_real(double* p, double v)
{
    (void)__builtin_amdgcn_global_atomic_fadd_f64(as_global(p), v);
}

// A stored tap pair (TapT) at byte address p, from global memory.
__device__ __forceinline__ float2 ld_tap(const float2*, const char* p)
{
    const GlobalPtr<const float> q = as_global(
            reinterpret_cast<const float*>(p));
    return make_float2(q[0], q[1]);
}
__device__ __forceinline__ double2 ld_tap(const double2*, const char* p)
{
    const GlobalPtr<const double> q = as_global(
            reinterpret_cast<const double*>(p));
    return make_double2(q[0], q[1]);
}
__device__ __forceinline__ float2 ld_tap(const HalfPair*, const char* p)
{
    const GlobalPtr<const unsigned> q = as_global(
            reinterpret_cast<const unsigned*>(p));
    const unsigned u = q[0];
    HalfPair h;
    __builtin_memcpy(&h, &u, 4);
    return tap_pair(h);
}
__device__ __forceinline__ float2 ld_tap(const BF16Pair*, const char* p)
{
    const GlobalPtr<const unsigned> q = as_global(
            reinterpret_cast<const unsigned*>(p));
    const unsigned u = q[0];
    return make_float2(bf16_to_float((unsigned short)(u & 0xFFFFu)),
                       bf16_to_float((unsigned short)(u >> 16)));
}

// The same as loaded (raw) and as used (widened): the LDS scatter keeps
// the next table's taps raw while they are in flight, so that a 16-bit
// pair's conversion waits for its load only when the taps are taken.
__device__ __forceinline__ float2 ld_tap_raw(const float2*, const char* p)
{
    return ld_tap(static_cast<const float2*>(nullptr), p);
}
__device__ __forceinline__ float2 tap_widen(const float2*, float2 t)
{
    return t;
}
template <typename P16>
__device__ __forceinline__ unsigned ld_tap_raw(const P16*, const char* p)
{
    return *as_global(reinterpret_cast<const unsigned*>(p));
}
__device__ __forceinline__ float2 tap_widen(const HalfPair*, unsigned u)
{
    HalfPair h;
    __builtin_memcpy(&h, &u, 4);
    return tap_pair(h);
}
__device__ __forceinline__ float2 tap_widen(const BF16Pair*, unsigned u)
{
    return make_float2(bf16_to_float((unsigned short)(u & 0xFFFFu)),
                       bf16_to_float((unsigned short)(u >> 16)));
}

// One component of a compensated two-float grid cell, p[0] = hi and
// p[1] = lo (value hi + lo).  x is split into h = (float)x and
// l = (float)(x - h); h goes into hi with an atomic add that returns the
// old value o, whose rounding error e is then exact (Knuth's TwoSum of o
// and h: the atomic stored fl(o + h), IEEE round to nearest, for operands
// away from the denormal range), and l + e goes into lo.  Two atomics, one
// returning, no retries, whatever other adds interleave; the pair is not
// renormalised (hi + lo is the value).
__device__ __forceinline__ float two_sum_err(float o, float h)
{
    const float s = o + h;
    const float bp = s - o;
    return (o - (s - bp)) + (h - bp);
}

__device__ __forceinline__ void comp_add(float* p, double x)
{
    const float h = (float)x;
    const float l = (float)(x - (double)h);
    const float o = __builtin_amdgcn_global_atomic_fadd_f32(as_global(p), h);
    atomic_add_real(p + 1, l + two_sum_err(o, h));
}

// Adds (re, im) to cell g of the grid: interleaved (re, im) of type T, or
// with COMP the compensated two-float layout (re_hi, re_lo, im_hi, im_lo).
template <typename T, bool COMP>
__device__ __forceinline__ void grid_add(T* grid, size_t g, double re,
                                         double im)
{
    if constexpr (COMP)
    {
        comp_add(reinterpret_cast<float*>(grid) + 4 * g, re);
        comp_add(reinterpret_cast<float*>(grid) + 4 * g + 2, im);
    }
    else
    {
        atomic_add_real(grid + 2 * g, (T)re);
        atomic_add_real(grid + 2 * g + 1, (T)im);
    }
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ int wave_min(int v)
{
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ int wave_max(int v)
{
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

// -------------------------------------------------------------------------
// Kernels
// -------------------------------------------------------------------------

__device__ __forceinline__ void init_summary(Summary* __restrict__ s)
{
    s->skipped = 0;
        s->n_valid = 0;
        s->updates = 0;
        s->norm_delta = 0.0;
    s->v_min = INT_MAX;
    s->v_max = INT_MIN;
    s->u_min = INT_MAX;
    s->u_max = INT_MIN;
    s->nv32 = 0;
    s->n_buckets = 0;
    s->total_items = 0;
    s->flags = 0;
    for (int q = 0; q < kXcdGroups; ++q) s->item_next[q] = 0;
    s->nm32 = 0;
    s->run_updates = 0;
}

// The summaries s[0..ns) reset and combo[0..n_combo) cleared (one call's
// summary and histogram, or every job's of a batch at once).
__global__ void k_init(Summary* __restrict__ s, unsigned* __restrict__ combo,
                       size_t n_combo, int ns)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < (size_t)ns) init_summary(s + t);
    for (size_t i = t; i < n_combo; i += (size_t)gridDim.x * blockDim.x)
        combo[i] = 0u;
}

__global__ void k_fill_u64(unsigned long long* __restrict__ p, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

__global__ void k_fill_u32(unsigned* __restrict__ p, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

constexpr int kComboLds = 8192;   // per-block LDS histogram capacity
// k_locate: one 1024-thread workgroup per CU (grid-stride).  Every
// workgroup adds its (plane, sub-cell) histogram into the same global bins
// at the end, so fewer, larger workgroups mean fewer atomics on each bin
// (measured: the 64-job mix's graph 32.7 -> 31.0 ms against 256-thread
// workgroups, four per CU; the 512-station index stage unchanged).
constexpr int kLocateThreads = 1024;

// Index stage.  Follows oskar_grid_wproj.c:169-212 (FP32) and :71-101
// (FP64) operation by operation with FMA contraction disabled.  T is the
// arithmetic of the index math; coordinates of type TU and visibilities of
// type TV are widened to it (the reduced-precision plans: float
// coordinates, FP16/BF16 visibilities, FP64 index math, float records).
template <typename T, bool DIAG, typename TU = T, typename TV = T,
          typename Rec = typename Types<T>::Rec>
__global__ void __launch_bounds__(kLocateThreads) k_locate(
        size_t n, const TU* __restrict__ uu, const TU* __restrict__ vv,
        const TU* __restrict__ ww, const TV* __restrict__ vis, T grid_scale,
        T w_scale, int ov, int G, int P, const int* __restrict__ support,
        const int* __restrict__ srank, int srank_bits, int plane_bits,
        int h, int K, int tile, int tile_bits, int combo_bits, int key_bits,
        unsigned long long* __restrict__ keys, Rec* __restrict__ rec,
        unsigned* __restrict__ combo, Summary* __restrict__ summary,
        unsigned* __restrict__ gw_out,
        unsigned long long* __restrict__ diag_support,
        unsigned long long* __restrict__ diag_plane, int diag_plane_cap,
        unsigned long long key_tag)
{
#pragma clang fp contract(off)
    __shared__ unsigned s_combo[kComboLds];
    __shared__ int s_red[kLocateThreads / 64][5];
    const int n_combo = P * K * K;
    const bool lds_hist = n_combo <= kComboLds;
    if (lds_hist)
        for (int i = threadIdx.x; i < n_combo; i += blockDim.x) s_combo[i] = 0;
    __syncthreads();

    const int centre = G / 2;
    const T ovf = (T)ov;
    // Sort key, most significant first: tile (Z-order), support rank,
    // position inside the tile (Z-order), plane, sub-cell table.  Skipped
    // points sort last.  Points of one support share the footprint shape,
    // so the scatter takes them as one segment, and points at one position
    // share its running sums across planes and sub-cell tables.
    const unsigned long long skip_key = 1ull << key_bits;
    // srank[p] = the plane's support rank | its index among the planes of
    // that rank << 16: the key's plane field below the position need only
    // tell apart planes of one support, since the rank above it does the
    // rest (plane_bits wide).
    const int sh_plane = combo_bits;
    const int sh_pos = sh_plane + plane_bits;
    const int sh_srank = sh_pos + 2 * tile_bits;
    const int sh_tile = sh_srank + srank_bits;
    int nskip = 0;
    int vmin = INT_MAX, vmax = INT_MIN, umin = INT_MAX, umax = INT_MIN;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += stride)
    {
        const T pu = -(T)uu[i] * grid_scale;
        const T pv = (T)vv[i] * grid_scale;
        const T w = (T)ww[i];
        const T ru = d_round(pu);
        const T rv = d_round(pv);
        const unsigned long long grid_w =
                to_size_x86(d_round(d_sqrt(d_fabs(w * w_scale))));
        const long long gu = (long long)to_int_x86(ru) + centre;
        const long long gv = (long long)to_int_x86(rv) + centre;
        const int plane = grid_w < (unsigned long long)P ? (int)grid_w : P - 1;
        const int S = support[plane];
        const bool skip = gu + S >= G || gu - S < 0 || gv + S >= G ||
                gv - S < 0;
        if (DIAG)
        {
            atomicAdd(&diag_support[S], 1ull);
            if (grid_w < (unsigned long long)diag_plane_cap)
                atomicAdd(&diag_plane[grid_w], 1ull);
            else
                atomicOr(&summary->flags, 1u);
            gw_out[i] = grid_w < 0xFFFFFFFFull ? (unsigned)grid_w :
                    0xFFFFFFFFu;
        }
        unsigned long long key = skip_key;
        if (skip)
        {
            ++nskip;
        }
        else
        {
            const int off_u = to_int_x86(d_round((ru - pu) * ovf));
            const int off_v = to_int_x86(d_round((rv - pv) * ovf));
            const int u0 = (int)gu - S, v0 = (int)gv - S;
            const int conj = w > (T)0 ? 1 : 0;
            const int iu = min(max(off_u + h, 0), K - 1);
            const int iv = min(max(off_v + h, 0), K - 1);
            const int table = iv * K + iu;
            const int c = plane * K * K + table;
            // Tile (tu, tv) of side `tile` and the origin's place (ou, ov)
            // in it, each Z-ordered (for a power-of-two side, the origin's
            // own Z-order split at 2 tile_bits).
            const unsigned tu = (unsigned)u0 / (unsigned)tile;
            const unsigned tv = (unsigned)v0 / (unsigned)tile;
            const unsigned ou = (unsigned)u0 - tu * (unsigned)tile;
            const unsigned ov = (unsigned)v0 - tv * (unsigned)tile;
            const unsigned long long tz = spread_bits16(tu) |
                    (spread_bits16(tv) << 1);
            const unsigned long long pz = spread_bits16(ou) |
                    (spread_bits16(ov) << 1);
            const unsigned sr = (unsigned)srank[plane];
            key = (tz << sh_tile) |
                    ((unsigned long long)(sr & 0xFFFFu) << sh_srank) |
                    (pz << sh_pos) |
                    ((unsigned long long)(sr >> 16) << sh_plane) |
                    (unsigned long long)table;
            if (lds_hist) atomicAdd(&s_combo[c], 1u);
            else atomicAdd(&combo[c], 1u);
            vmin = min(vmin, v0);
            vmax = max(vmax, v0 + 2 * S);
            umin = min(umin, u0);
            umax = max(umax, u0 + 2 * S);
            Rec r;
            r.pos = (int)((unsigned)u0 | ((unsigned)v0 << 16));
            r.sub = (int)((unsigned)(off_u + 128) |
                    ((unsigned)(off_v + 128) << 8) | ((unsigned)conj << 16) |
                    ((unsigned)plane << 17));
            r.re = to_real(vis[2 * i]);
            r.im = to_real(vis[2 * i + 1]);
            rec[i] = r;
        }
        keys[i] = key | key_tag;
    }

    // Block reductions: one atomic per block for the counters and box.
    nskip = wave_sum(nskip);
    vmin = wave_min(vmin); vmax = wave_max(vmax);
    umin = wave_min(umin); umax = wave_max(umax);
    const int wid = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
    {
        s_red[wid][0] = nskip; s_red[wid][1] = vmin; s_red[wid][2] = vmax;
        s_red[wid][3] = umin; s_red[wid][4] = umax;
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        const int nw = blockDim.x >> 6;
        for (int k = 1; k < nw; ++k)
        {
            s_red[0][0] += s_red[k][0];
            s_red[0][1] = min(s_red[0][1], s_red[k][1]);
            s_red[0][2] = max(s_red[0][2], s_red[k][2]);
            s_red[0][3] = min(s_red[0][3], s_red[k][3]);
            s_red[0][4] = max(s_red[0][4], s_red[k][4]);
        }
        if (s_red[0][0]) atomicAdd(&summary->skipped,
                (unsigned long long)s_red[0][0]);
        if (s_red[0][1] != INT_MAX)
        {
            atomicMin(&summary->v_min, s_red[0][1]);
            atomicMax(&summary->v_max, s_red[0][2]);
            atomicMin(&summary->u_min, s_red[0][3]);
            atomicMax(&summary->u_max, s_red[0][4]);
        }
    }
    if (lds_hist)
        for (int i = threadIdx.x; i < n_combo; i += blockDim.x)
        {
            const unsigned v = s_combo[i];
            if (v) atomicAdd(&combo[i], v);
        }
}

// -------------------------------------------------------------------------
// Device-wide exclusive scan of u32 (three kernels; the element count is
// read from device memory so the sequence needs no host round trip)
// -------------------------------------------------------------------------

constexpr int kScanThreads = 1024;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanThreads * kScanItems;

// Exclusive scan of one value per thread across the block; *total gets the
// block sum.  s_tmp holds >= 17 words.
__device__ __forceinline__ unsigned block_exclusive_scan(unsigned v,
        unsigned* s_tmp, unsigned* total)
{
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    const int nwaves = blockDim.x >> 6;
    unsigned x = v;
    for (int o = 1; o < 64; o <<= 1)
    {
        const unsigned y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) s_tmp[wid] = x;
    __syncthreads();
    if (wid == 0)
    {
        const unsigned w = lane < nwaves ? s_tmp[lane] : 0u;
        unsigned wi = w;
        for (int o = 1; o < 64; o <<= 1)
        {
            const unsigned y = __shfl_up(wi, o, 64);
            if (lane >= o) wi += y;
        }
        if (lane < nwaves) s_tmp[lane] = wi - w;
        if (lane == nwaves - 1) s_tmp[16] = wi;
    }
    __syncthreads();
    const unsigned r = x - v + s_tmp[wid];
    *total = s_tmp[16];
    __syncthreads();
    return r;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_up(
        const unsigned* __restrict__ in, const unsigned* __restrict__ count,
        unsigned* __restrict__ partials)
{
    __shared__ unsigned s_tmp[17];
    const unsigned n = *count;
    const size_t base = (size_t)blockIdx.x * kScanTile +
            (size_t)threadIdx.x * kScanItems;
    unsigned sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (base + k < n) sum += in[base + k];
    unsigned total;
    (void)block_exclusive_scan(sum, s_tmp, &total);
    if (threadIdx.x == 0) partials[blockIdx.x] = total;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_mid(
        unsigned* __restrict__ partials, const unsigned* __restrict__ count,
        unsigned* __restrict__ total_out)
{
    __shared__ unsigned s_tmp[17];
    const unsigned n = *count;
    const unsigned nb = (n + kScanTile - 1) / kScanTile;
    unsigned carry = 0;
    for (unsigned base = 0; base < nb; base += kScanThreads)
    {
        const unsigned i = base + threadIdx.x;
        const unsigned v = i < nb ? partials[i] : 0u;
        unsigned total;
        const unsigned ex = block_exclusive_scan(v, s_tmp, &total);
        if (i < nb) partials[i] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) *total_out = carry;
}

// The whole scan in one workgroup, a tile at a time with a running carry
// (for short inputs: at most kScanOneTiles tiles).
constexpr unsigned kScanOneTiles = 4;
__global__ void __launch_bounds__(kScanThreads) k_scan_one(
        const unsigned* __restrict__ in, const unsigned* __restrict__ count,
        unsigned* __restrict__ out, unsigned* __restrict__ total_out)
{
    __shared__ unsigned s_tmp[17];
    const unsigned n = *count;
    unsigned carry = 0;
    for (unsigned t0 = 0; t0 < n; t0 += kScanTile)
    {
        const size_t base = (size_t)t0 + (size_t)threadIdx.x * kScanItems;
        unsigned v[kScanItems];
        unsigned sum = 0;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k)
        {
            v[k] = base + k < n ? in[base + k] : 0u;
            sum += v[k];
        }
        unsigned total;
        unsigned run = block_exclusive_scan(sum, s_tmp, &total) + carry;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k)
        {
            if (base + k < n) out[base + k] = run;
            run += v[k];
        }
        carry += total;
    }
    if (threadIdx.x == 0) *total_out = carry;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_down(
        const unsigned* __restrict__ in, const unsigned* __restrict__ count,
        const unsigned* __restrict__ partials, unsigned* __restrict__ out)
{
    __shared__ unsigned s_tmp[17];
    const unsigned n = *count;
    const size_t base = (size_t)blockIdx.x * kScanTile +
            (size_t)threadIdx.x * kScanItems;
    unsigned v[kScanItems];
    unsigned sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
    {
        v[k] = base + k < n ? in[base + k] : 0u;
        sum += v[k];
    }
    unsigned total;
    unsigned run = block_exclusive_scan(sum, s_tmp, &total) +
            partials[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
    {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
}

// -------------------------------------------------------------------------
// Planner, gather and work items
// -------------------------------------------------------------------------

// Points gridded, update count and norm increment.  One block of 1024.
__device__ __forceinline__ void plan_body(
        int P, int K2, const unsigned* __restrict__ combo,
        const PlaneInfo* __restrict__ pinfo,
        const double* __restrict__ normtab, Summary* __restrict__ summary)
{
    __shared__ double s_d[1024];
    __shared__ unsigned long long s_u[1024];
    const int t = threadIdx.x;
    unsigned long long pts = 0, upd = 0;
    for (int p = t; p < P; p += 1024)
    {
        unsigned cnt = 0;
        for (int c = 0; c < K2; ++c) cnt += combo[p * K2 + c];
        const unsigned long long W = (unsigned long long)pinfo[p].W;
        pts += cnt;
        upd += (unsigned long long)cnt * W * W;
    }
    // Norm increment: fixed-order reduction (deterministic).
    double part = 0.0;
    for (int i = t; i < P * K2; i += 1024)
        part += (double)combo[i] * normtab[i];
    s_d[t] = part;
    s_u[t] = upd;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1)
    {
        if (t < o)
        {
            s_d[t] += s_d[t + o];
            s_u[t] += s_u[t + o];
        }
        __syncthreads();
    }
    const unsigned long long upd_total = s_u[0];
    __syncthreads();
    s_u[t] = pts;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1)
    {
        if (t < o) s_u[t] += s_u[t + o];
        __syncthreads();
    }
    if (t == 0)
    {
        summary->n_valid = s_u[0];
        summary->nv32 = (unsigned)s_u[0];
        summary->updates = upd_total;
        summary->norm_delta = s_d[0];
    }
}

__global__ void __launch_bounds__(1024) k_plan(
        int P, int K2, const unsigned* __restrict__ combo,
        const PlaneInfo* __restrict__ pinfo,
        const double* __restrict__ normtab, Summary* __restrict__ summary)
{
    plan_body(P, K2, combo, pinfo, normtab, summary);
}

// k_plan for every job of a batch, one workgroup per job.
__global__ void __launch_bounds__(1024) k_plan_batch(
        const BatchJob* __restrict__ jobs,
        const unsigned* __restrict__ combo, Summary* __restrict__ sums)
{
    const BatchJob& d = jobs[blockIdx.x];
    plan_body(d.planes, d.sub_k * d.sub_k, combo + d.combo_off, d.pinfo,
              d.normtab, sums + blockIdx.x);
}

// Sorted point records.  meta = (u0 | v0 << 16, sub-cell table
// (off_v + h) * K + (off_u + h), plane, grid_w); vq = (re, -s*im, im, s*re)
// with s = -1 when ww > 0 (oskar_grid_wproj.c:172, :233), so that
//   grid_re += vq.x * c_re + vq.y * c_im,  grid_im += vq.z * c_re + vq.w * c_im
// with (c_re, c_im) the tap as stored in the kernel cube.  Bit (m & 31) of
// tb[m >> 5] is set where record m starts a new tile (key bits above
// tile_shift), tc[w] = the set bits of word w, summary->nwt32 = the words.
// meta.y of a point: bits 0-29 the element offset of its (plane, |off_v|,
// |off_u|) tap table (tables are below 4 GiB and taps at least 4 bytes),
// bit 30 set for off_u < 0, bit 31 for off_v < 0 (the scatter then reads
// the table at the mirrored footprint cell).
__device__ __forceinline__ unsigned table_ref(const PlaneInfo& pi, int off_u,
                                              int off_v, int h)
{
    const int M = h + 1;
    const unsigned W = (unsigned)pi.W;
    const unsigned t = (unsigned)(abs(off_v) * M + abs(off_u));
    return (pi.eoff + t * W * W) | (off_u < 0 ? 1u << 30 : 0u) |
           (off_v < 0 ? 1u << 31 : 0u);
}

// Control bits of a sorted point, in meta.z above the plane (bits 0-15)
// and the support rank (bits 16-29): the point's footprint origin differs
// from the previous sorted point's (the scatter adds the previous point's
// sums into the region before this point), and its tap table differs (the
// scatter loads other taps).  Within one stream's share of a segment the
// previous sorted point is the previous point the stream took.
constexpr unsigned kNewPos = 1u << 30;
constexpr unsigned kNewTab = 1u << 31;
constexpr unsigned kRankMask = 0x3FFFu;

// meta.x of a point for the LDS scatter: its origin's byte offset in the
// region of its tile (tiles start at multiples of the tile side T, region
// pitch R cells, FP64 cells); R == 0: the origin u | v << 16 itself.
__device__ __forceinline__ int region_offset(int pos, int R, int tile)
{
    if (R <= 0) return pos;
    const unsigned u = (unsigned)pos & 0xFFFFu;
    const unsigned v = (unsigned)pos >> 16;
    return ((int)(v % (unsigned)tile) * R + (int)(u % (unsigned)tile)) *
            (int)sizeof(double);
}

// Runs of equal sort keys, as bits: bit (s & 31) of hb[s >> 5] is set where
// sorted point s starts a run (merge: its key differs from the previous
// one; otherwise every point), hc[w] = the set bits of word w, and
// summary->nw32 = the words of the summary's nv32 points.  Each wave takes
// 64 consecutive points (two words) with one ballot.  Points of one run
// share the tile, footprint origin, plane and sub-cell table, so they use
// the same taps at the same cells: the scatter takes one record per run,
// its visibility terms summed (k_gather).
__global__ void __launch_bounds__(256) k_merge_heads(
        Summary* __restrict__ summary,
        const unsigned long long* __restrict__ keys,
        unsigned* __restrict__ hb, unsigned* __restrict__ hc, int merge)
{
    const unsigned nv = summary->nv32;
    const unsigned nw = (nv + 31) / 32;
    if (blockIdx.x == 0 && threadIdx.x == 0) summary->nw32 = nw;
    const int lane = threadIdx.x & 63;
    const unsigned end = (nv + 63) & ~63u;     // whole waves
    for (unsigned s = blockIdx.x * blockDim.x + threadIdx.x; s < end;
         s += gridDim.x * blockDim.x)
    {
        const bool h = s < nv &&
                (!merge || s == 0 || keys[s - 1] != keys[s]);
        const unsigned long long m = __ballot(h);
        const unsigned w = s >> 5;
        if ((lane & 31) == 0 && w < nw)
        {
            const unsigned bits = (unsigned)(lane == 0 ? m : m >> 32);
            hb[w] = bits;
            hc[w] = (unsigned)__popc(bits);
        }
    }
}

// rstart[m] = first sorted point of run m (hb, the run-start bits; hp, the
// exclusive scan of their per-word counts); rstart[runs] = the point count.
__global__ void __launch_bounds__(256) k_run_starts(
        const Summary* __restrict__ summary, const unsigned* __restrict__ hb,
        const unsigned* __restrict__ hp, unsigned* __restrict__ rstart)
{
    const unsigned nv = summary->nv32;
    for (unsigned s = blockIdx.x * blockDim.x + threadIdx.x; s < nv;
         s += gridDim.x * blockDim.x)
    {
        const unsigned bits = hb[s >> 5];
        const unsigned bit = 1u << (s & 31);
        if (bits & bit)
            rstart[hp[s >> 5] + (unsigned)__popc(bits & (bit - 1u))] = s;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
        rstart[summary->nm32] = nv;
}

// One record per run of equal sort keys, run m = sorted points
// [rstart[m], rstart[m + 1]).  The run's vq terms are summed component by
// component: each point's (re, -s*im, im, s*re) applies its own
// conjugation sign, so the sum is exact algebra for any mix of signs
// (summed in FP64 in sorted order, then stored in T).  mkeys[m] = the
// run's key; summary->run_updates = the footprint cells of all runs (the
// grid updates the scatter executes).
template <typename T>
__global__ void __launch_bounds__(256) k_gather(
        Summary* __restrict__ summary,
        const unsigned long long* __restrict__ keys,
        const unsigned* __restrict__ idx,
        const typename Types<T>::Rec* __restrict__ rec,
        const unsigned* __restrict__ gw, const PlaneInfo* __restrict__ pinfo,
        int h, int K, int tile_shift, int4* __restrict__ meta,
        typename Types<T>::T4* __restrict__ vq, unsigned* __restrict__ tb,
        unsigned* __restrict__ tc, int R, int tile, int pos_shift,
        const unsigned* __restrict__ rstart,
        unsigned long long* __restrict__ mkeys)
{
    typedef typename Types<T>::Rec Rec;
    const unsigned nm = summary->nm32;
    const unsigned nwt = (nm + 31) / 32;
    if (blockIdx.x == 0 && threadIdx.x == 0) summary->nwt32 = nwt;
    const int lane = threadIdx.x & 63;
    unsigned long long upd = 0;     // footprint cells of this thread's runs
    for (unsigned m = blockIdx.x * blockDim.x + threadIdx.x; m < nm;
         m += gridDim.x * blockDim.x)
    {
        const unsigned s = rstart[m], e = rstart[m + 1];
        const unsigned long long key = keys[s];
        mkeys[m] = key;
        // 64 consecutive records per wave (lanes past nm have left the
        // loop and count as zero bits).
        const unsigned long long tbits = __ballot(s == 0 ||
                (keys[s - 1] >> tile_shift) != (key >> tile_shift));
        if ((lane & 31) == 0 && (m >> 5) < nwt)
        {
            const unsigned bits = (unsigned)(lane == 0 ? tbits : tbits >> 32);
            tb[m >> 5] = bits;
            tc[m >> 5] = (unsigned)__popc(bits);
        }
        const unsigned i = idx[s];
        const Rec r = rec[i];
        const int plane = (unsigned)r.sub >> 17;
        const int off_u = (r.sub & 0xFF) - 128;
        const int off_v = ((r.sub >> 8) & 0xFF) - 128;
        const PlaneInfo& pi = pinfo[plane];
        const unsigned tref = table_ref(pi, off_u, off_v, h);
        upd += (unsigned long long)pi.W * (unsigned long long)pi.W;
        // From the previous sorted key (bits from the top: tile, support
        // rank, position, plane, sub-cell table): the same tile and
        // position, the same plane and sub-cell table (hence the same table
        // reference).  The previous point ends the previous run, whose
        // record carries its key.
        unsigned flags = kNewPos | kNewTab;
        if (s > 0)
        {
            const unsigned long long d = keys[s - 1] ^ key;
            if ((d >> pos_shift) == 0) flags &= ~kNewPos;
            if ((d & ((1ull << pos_shift) - 1)) == 0) flags &= ~kNewTab;
        }
        meta[m] = make_int4(region_offset(r.pos, R, tile), (int)tref,
                (int)((unsigned)plane | ((unsigned)pi.srank << 16) | flags),
                gw ? (int)min(gw[i], 0x7FFFFFFFu) : 0);
        double q0 = 0.0, q1 = 0.0, q2 = 0.0, q3 = 0.0;
        auto add = [&](const Rec& p) {
            const double sgn = ((p.sub >> 16) & 1) ? -1.0 : 1.0;
            q0 += (double)p.re;
            q1 += -sgn * (double)p.im;
            q2 += (double)p.im;
            q3 += sgn * (double)p.re;
        };
        add(r);
        // The rest of the run four points at a time: independent reads.
        unsigned t = s + 1;
        for (; t + 4 <= e; t += 4)
        {
            const unsigned i0 = idx[t], i1 = idx[t + 1], i2 = idx[t + 2],
                    i3 = idx[t + 3];
            const Rec p0 = rec[i0], p1 = rec[i1], p2 = rec[i2], p3 = rec[i3];
            add(p0);
            add(p1);
            add(p2);
            add(p3);
        }
        for (; t < e; ++t) add(rec[idx[t]]);
        typename Types<T>::T4 o;
        o.x = (T)q0;
        o.y = (T)q1;
        o.z = (T)q2;
        o.w = (T)q3;
        vq[m] = o;
    }
    upd = wave_sum(upd);
    if ((threadIdx.x & 63) == 0 && upd)
        atomicAdd(&summary->run_updates, upd);
}

// bstart[b] = first record of tile bucket b; bstart[n_buckets] = records.
__global__ void __launch_bounds__(256) k_buckets(
        const Summary* __restrict__ summary, const unsigned* __restrict__ tb,
        const unsigned* __restrict__ tp, unsigned* __restrict__ bstart)
{
    const unsigned nv = summary->nm32;
    for (unsigned s = blockIdx.x * blockDim.x + threadIdx.x; s < nv;
         s += gridDim.x * blockDim.x)
    {
        const unsigned bits = tb[s >> 5];
        const unsigned bit = 1u << (s & 31);
        if (bits & bit)
            bstart[tp[s >> 5] + (unsigned)__popc(bits & (bit - 1u))] = s;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
        bstart[summary->n_buckets] = nv;
}

// Jobs per tile bucket: ceil(points / job_points).
__global__ void __launch_bounds__(256) k_bucket_items(
        const Summary* __restrict__ summary,
        const unsigned* __restrict__ bstart, unsigned job_points,
        unsigned* __restrict__ icnt)
{
    const unsigned nb = summary->n_buckets;
    for (unsigned b = blockIdx.x * blockDim.x + threadIdx.x; b < nb;
         b += gridDim.x * blockDim.x)
        icnt[b] = (bstart[b + 1] - bstart[b] + job_points - 1) / job_points;
}

__device__ __forceinline__ unsigned compact_bits16(unsigned long long x)
{
    x &= 0x55555555ull;
    x = (x | (x >> 1)) & 0x33333333ull;
    x = (x | (x >> 2)) & 0x0F0F0F0Full;
    x = (x | (x >> 4)) & 0x00FF00FFull;
    x = (x | (x >> 8)) & 0x0000FFFFull;
    return (unsigned)x;
}

// Support segments of every work item, one wave per item: a segment starts
// at the item's first record and wherever the support rank (meta.z >> 16)
// changes (records are sorted by rank within a tile).  Segment g of item it
// (first record f) goes to segs[f + g] as its first record in the item |
// its plane << 12, and the count to bits 16-31 of items[it].z (which holds
// the item's records, <= kJobPoints, below).  The scatter then starts an
// item with a few scalar loads instead of a block scan.
__global__ void __launch_bounds__(256) k_item_segments(
        const Summary* __restrict__ summary, unsigned items_cap,
        int4* __restrict__ items, const int4* __restrict__ meta,
        unsigned* __restrict__ segs)
{
    const unsigned n = min(summary->total_items, items_cap);
    const int lane = threadIdx.x & 63;
    const unsigned long long below = (1ull << lane) - 1ull;
    const unsigned waves = gridDim.x * (blockDim.x >> 6);
    for (unsigned it = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
         it < n; it += waves)
    {
        const int4 job = items[it];
        const int f = job.y, len = job.z & 0xFFFF;
        unsigned cnt = 0;
        unsigned carry = 0;         // rank field of the previous chunk's end
        for (int b = 0; b < len; b += 4 * 64)
        {
            unsigned z[4];
#pragma unroll
            for (int u = 0; u < 4; ++u)
            {
                const int i = b + u * 64 + lane;
                z[u] = i < len ? (unsigned)meta[f + i].z : 0u;
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
            {
                const int i = b + u * 64 + lane;
                const unsigned r = (z[u] >> 16) & kRankMask;
                unsigned prev = (unsigned)__shfl_up((int)r, 1);
                if (lane == 0) prev = carry;
                const bool head = i < len && (i == 0 || prev != r);
                const unsigned long long m = __ballot(head);
                if (head)
                    segs[f + cnt + (unsigned)__popcll(m & below)] =
                            (unsigned)i | ((z[u] & 0xFFFFu) << 12);
                cnt += (unsigned)__popcll(m);
                carry = (unsigned)__shfl((int)r, 63);
            }
        }
        if (lane == 0) items[it].z = len | (int)(cnt << 16);
    }
}

// Job = (tile origin u | v << 16, first sorted point, points, 0).  Also
// resets the scatter's queue heads (the kernel runs right before it).
__global__ void __launch_bounds__(256) k_write_items(
        Summary* __restrict__ summary,
        const unsigned long long* __restrict__ keys,
        const unsigned* __restrict__ bstart,
        const unsigned* __restrict__ ioff, int tile, int tile_shift,
        unsigned job_points, unsigned items_cap, int4* __restrict__ items)
{
    if (blockIdx.x == 0 && threadIdx.x < kXcdGroups)
        summary->item_next[threadIdx.x] = 0;
    const unsigned nb = summary->n_buckets;
    for (unsigned b = blockIdx.x * blockDim.x + threadIdx.x; b < nb;
         b += gridDim.x * blockDim.x)
    {
        const unsigned start = bstart[b], end = bstart[b + 1];
        const unsigned long long tz = keys[start] >> tile_shift;
        const unsigned ux0 = compact_bits16(tz) * (unsigned)tile;
        const unsigned vy0 = compact_bits16(tz >> 1) * (unsigned)tile;
        const int origin = (int)(ux0 | (vy0 << 16));
        unsigned o = ioff[b];
        for (unsigned first = start; first < end; first += job_points, ++o)
            if (o < items_cap)
                items[o] = make_int4(origin, (int)first,
                        (int)min(job_points, end - first), 0);
    }
}

// -------------------------------------------------------------------------
// The scatter (oskar_grid_wproj.c:216-240), one job per workgroup
// -------------------------------------------------------------------------
//
// A job is up to kJobPoints sorted points of one T x T tile of footprint
// origins, ordered by support, then position, then plane, then sub-cell
// table.  The workgroup accumulates them into an LDS copy of the tile's
// (T + Wmax - 1)^2 grid region (planar real / imaginary arrays), across
// planes, and adds the region to the grid in HBM once at the end of the job.
//
// Inside a support segment the waves form streams of wpp waves; each stream
// takes a contiguous chunk of the segment's points (at least min_share
// points; chunk q goes to stream (q + first point of the segment) mod
// streams, so short segments spread over all streams).  Lane slot s of wave
// w (w-th in its stream) owns footprint cell c = w*64 + lane + s*64*wpp,
// i.e. offset (j, k) = (c / W, c mod W) from the footprint origin, whose tap
// for the point's sub-cell table is etab[table][c]
// (= conv_func[p][|off_v + (j-S)*ov|][|off_u + (k-S)*ov|]).  The tap stays in
// registers while consecutive points share the plane and table; the products
// with the visibilities are summed in registers while consecutive points
// share the position (whatever their planes and tables: all planes of a
// segment have the same footprint shape); when the position changes the
// slot adds its sum to the region with an LDS atomic.  For planes whose
// region would not fit in LDS the sums go straight to the grid in HBM.

// LDS of a scatter workgroup.  The imaginary array sits at a fixed byte
// offset (kImBytes < 64 KiB) from the real one, so both atomics of a cell
// share one address register and differ only in the instruction's offset.
struct GridLds
{
    double re[kRegionCells];      // region real parts, then the sink cells
    double im[kRegionCells];      // region imaginary parts, then the sinks
};
constexpr unsigned kImBytes = kRegionCells * sizeof(double);
static_assert(kImBytes < 65536, "imaginary offset must fit a DS offset");

// Read-only device data loaded through the constant address space: the
// compiler may then fetch uniform records with scalar loads even where
// other pointers (the grid) are written in the same loop.  Field-wise, so
// that the host compilation pass, which has no such address space for
// class types, accepts it.
template <typename U>
using ConstPtr = const __attribute__((address_space(4))) U*;

template <typename U>
__device__ __forceinline__ ConstPtr<U> as_const(const void* p)
{
    return (ConstPtr<U>)(uintptr_t)p;
}

__device__ __forceinline__ int4 ld_meta(const int4* p, int i)
{
    const ConstPtr<int> q = as_const<int>(p + i);
    return make_int4(q[0], q[1], q[2], q[3]);
}

template <typename T4>
__device__ __forceinline__ T4 ld_vq(const T4* p, int i)
{
    typedef decltype(T4().x) T;
    const ConstPtr<T> q = as_const<T>(p + i);
    T4 r;
    r.x = q[0];
    r.y = q[1];
    r.z = q[2];
    r.w = q[3];
    return r;
}

__device__ __forceinline__ PlaneInfo ld_pinfo(const PlaneInfo* p, int i)
{
    const ConstPtr<int> q = as_const<int>(p + i);
    PlaneInfo r;
    r.W = q[0];
    r.S = q[1];
    r.wpp = q[2];
    r.nslot = q[3];
    r.wmagic = (unsigned)q[4];
    r.sshift = q[5];
    r.eoff = (unsigned)q[6];
    r.srank = q[7];
    return r;
}

template <typename T>
struct ScatterCtx
{
    typedef typename Types<T>::T4 T4;
    const int4* __restrict__ meta;
    const T4* __restrict__ vq;
    const void* __restrict__ etab;   // tap pairs of type TapT
    const unsigned* __restrict__ segs; // per work item: its segments
    GridLds* lds;
    T* __restrict__ grid;
    unsigned* __restrict__ diag_count;
    unsigned long long* __restrict__ diag_wsum;
    int R, G, row0, ux0, vy0;
};

// The LDS region is FP64: on gfx950 ds_add_f64 sustains about 20 times the
// throughput of ds_add_f32 (measured: ~7 vs ~0.33 lane-ops per clock per
// CU), and the tile's sums are then carried in double until the one add
// into the grid per job.  c8 is the cell's byte offset in the real array.
__device__ __forceinline__ void lds_add_cell(GridLds* L, unsigned c8,
                                             double re, double im)
{
    char* const b = reinterpret_cast<char*>(L->re);
    atomicAdd(reinterpret_cast<double*>(b + c8), re);
    atomicAdd(reinterpret_cast<double*>(b + kImBytes + c8), im);
}

// One stream's share [a, b) of a plane segment: footprint cells
// [cb, cb + 64 * wpp * NS), NS cells per lane.  LDS selects the LDS region
// (else the grid in HBM).
//
// Arithmetic: products and the sums of a position run in T (FP32 for the
// float and 16-bit plans, FP64 for the double plan; kSumF64 switches the
// float sums to FP64, at fewer cells per lane), widened to FP64 when a run
// ends and is added into the FP64 region.  The first point of a position
// run writes its products with a multiply, so a flush leaves the sums as
// they are instead of clearing them.
//
// Point records are read two points ahead and the next sub-cell table's
// taps are loaded before the current point's arithmetic, so both latencies
// overlap work.  In the LDS path lanes past the footprint are not masked
// off: their sums go to a per-lane sink cell after the region (kSinkCells
// of them), which is never read.
template <typename T, bool DIAG, bool LDS, int NS, typename TapT, bool COMP>
__device__ __forceinline__ void scatter_stream(const ScatterCtx<T>& x,
        const PlaneInfo& pi, int wis, int a, int b, int cb)
{
    typedef typename Types<T>::T4 T4;
    const int lane = threadIdx.x & 63;
    const int W = pi.W, W2 = W * W;
    const int stride = 64 * pi.wpp;
    // meta.y: first tap of the point's (plane, sub-cell) table.
    const char* __restrict__ ptab = static_cast<const char*>(x.etab);
    const int cstride = LDS ? x.R : x.G;
    unsigned toff[NS];     // byte offset of the lane's tap in a table
    unsigned k2[NS];       // 2 k sizeof(TapT), k = the lane's footprint column
    unsigned loff[NS];     // LDS: byte offset from the footprint origin in
                           // the region; HBM: cell offset in the grid
    bool valid[NS];
    typedef typename std::conditional<kSumF64, double, T>::type A;
    A tre[NS], tim[NS];        // taps of the current table
    T nre[NS], nim[NS];        // taps of the next table, in flight
    A acc_re[NS], acc_im[NS];
    const int c0 = cb + wis * 64 + lane;
    if (pi.wmagic != 0u)
    {
#pragma unroll
        for (int s = 0; s < NS; ++s)
        {
            const int c = c0 + s * stride;
            valid[s] = c < W2;
            const unsigned cc = valid[s] ? (unsigned)c : 0u;
            const unsigned j = __umulhi(cc, pi.wmagic);
            const unsigned k = cc - j * (unsigned)W;
            loff[s] = j * (unsigned)cstride + k;
            toff[s] = cc * (unsigned)sizeof(TapT);
            k2[s] = 2 * k * (unsigned)sizeof(TapT);
        }
    }
    else
    {
#pragma unroll
        for (int s = 0; s < NS; ++s)
        {
            const int c = c0 + s * stride;
            valid[s] = c < W2;
            const unsigned cc = valid[s] ? (unsigned)c : 0u;
            const unsigned j = cc / (unsigned)W;
            const unsigned k = cc - j * (unsigned)W;
            loff[s] = j * (unsigned)cstride + k;
            toff[s] = cc * (unsigned)sizeof(TapT);
            k2[s] = 2 * k * (unsigned)sizeof(TapT);
        }
    }
    // Byte offsets in a table of the cell mirrored in u (k -> W-1-k), in v
    // (j -> W-1-j) and in both, for c = j W + k:
    //   u: c + (W-1) - 2k,  v: (W-1) W - c + 2k,  both: W^2 - 1 - c.
    const unsigned bu = (unsigned)(W - 1) * (unsigned)sizeof(TapT);
    const unsigned bv = (unsigned)((W - 1) * W) * (unsigned)sizeof(TapT);
    const unsigned bb = (unsigned)(W2 - 1) * (unsigned)sizeof(TapT);
    auto load_taps = [&](int table) {
        // table = meta.y: element offset in bits 0-29, mirror flags in
        // bits 30 (u) and 31 (v); the same for every lane.
        const unsigned e = (unsigned)table & 0x3FFFFFFFu;
        const unsigned flip = (unsigned)table >> 30;
        const char* tp = ptab + (size_t)e * sizeof(TapT);
        auto issue = [&](int s, unsigned o) {
            const auto t = ld_tap(static_cast<const TapT*>(nullptr), tp + o);
            nre[s] = (T)t.x;
            nim[s] = (T)t.y;
        };
        // One branch per point (flip is uniform), not a select per slot.
        if (flip == 0)
        {
#pragma unroll
            for (int s = 0; s < NS; ++s) issue(s, toff[s]);
        }
        else if (flip == 1)
        {
#pragma unroll
            for (int s = 0; s < NS; ++s) issue(s, toff[s] + bu - k2[s]);
        }
        else if (flip == 2)
        {
#pragma unroll
            for (int s = 0; s < NS; ++s) issue(s, bv - toff[s] + k2[s]);
        }
        else
        {
#pragma unroll
            for (int s = 0; s < NS; ++s) issue(s, bb - toff[s]);
        }
    };
    auto take_taps = [&]() {
#pragma unroll
        for (int s = 0; s < NS; ++s)
        {
            tre[s] = (A)nre[s];
            tim[s] = (A)nim[s];
        }
    };
    int4 m = ld_meta(x.meta, a);
    T4 v = ld_vq(x.vq, a);
    load_taps(m.y);
    if (LDS)
    {
#pragma unroll
        for (int s = 0; s < NS; ++s) loff[s] *= (unsigned)sizeof(double);
    }
    // LDS: meta.x is already the origin's byte offset in the region (the
    // gather stage's region_offset); otherwise the origin u | v << 16.
    auto origin = [&](int pos) -> unsigned {
        const int u0 = pos & 0xFFFF;
        const int v0 = (int)((unsigned)pos >> 16);
        return LDS ? (unsigned)pos : (unsigned)((v0 - x.row0) * x.G + u0);
    };
    unsigned cnt = 0;
    unsigned long long wsum = 0;
    const unsigned sink8 = (unsigned)(x.R * x.R + lane) *
            (unsigned)sizeof(double);
    auto flush = [&](unsigned base) {
#pragma unroll
        for (int s = 0; s < NS; ++s)
        {
            if (LDS)
            {
                const unsigned c8 = valid[s] ? base + loff[s] : sink8;
                lds_add_cell(x.lds, c8, (double)acc_re[s],
                             (double)acc_im[s]);
            }
            else if (valid[s])
            {
                const size_t g = (size_t)base + loff[s];
                grid_add<T, COMP>(x.grid, g, (double)acc_re[s],
                                  (double)acc_im[s]);
                if (DIAG)
                {
                    atomicAdd(&x.diag_count[g], cnt);
                    atomicAdd(&x.diag_wsum[g], wsum);
                }
            }
        }
    };

    int4 mn = m;
    T4 vn = v;
    if (a + 1 < b)
    {
        mn = ld_meta(x.meta, a + 1);
        vn = ld_vq(x.vq, a + 1);
    }
    unsigned base = origin(m.x);
    // The first table's taps are converted before the loop, so the loads
    // issued inside it stay in flight across a point's work.
    take_taps();
    bool fresh = true;
    for (int k = a; k < b; ++k)
    {
        const bool more = k + 1 < b;
        int4 mnn = mn;
        T4 vnn = vn;
        if (k + 2 < b)
        {
            mnn = ld_meta(x.meta, k + 2);
            vnn = ld_vq(x.vq, k + 2);
        }
        const bool new_table = more && mn.y != m.y;
        if (new_table) load_taps(mn.y);
        const A vx = (A)v.x, vy = (A)v.y;
        const A vz = (A)v.z, vw = (A)v.w;
        if (fresh)
        {
#pragma unroll
            for (int s = 0; s < NS; ++s)
            {
                acc_re[s] = fma(vy, tim[s], vx * tre[s]);
                acc_im[s] = fma(vw, tim[s], vz * tre[s]);
            }
        }
        else
        {
#pragma unroll
            for (int s = 0; s < NS; ++s)
            {
                acc_re[s] = fma(vx, tre[s], acc_re[s]);
                acc_re[s] = fma(vy, tim[s], acc_re[s]);
                acc_im[s] = fma(vz, tre[s], acc_im[s]);
                acc_im[s] = fma(vw, tim[s], acc_im[s]);
            }
        }
        fresh = false;
        if (DIAG)
        {
            cnt += 1;
            wsum += (unsigned long long)pi.S + (unsigned)m.w;
        }
        if (!more) break;
        if (mn.x != m.x)
        {
            flush(base);
            base = origin(mn.x);
            fresh = true;
            cnt = 0;
            wsum = 0;
        }
        if (new_table) take_taps();
        m = mn;
        v = vn;
        mn = mnn;
        vn = vnn;
    }
    flush(base);
}

// The LDS-region stream with the per-point control precomputed by the
// gather stage (meta.x the origin's byte offset in the region, meta.z's
// kNewPos / kNewTab bits): per point one record read (two scalar loads at
// one shared offset), three bit tests and the loop test on the scalar
// unit; the lane work as in scatter_stream (same lane map, taps, sums,
// flushes and results).  The loop is unrolled three times so that the
// records of the current, next and following points rotate through three
// register sets without moves.
template <int NS, typename TapT>
__device__ __forceinline__ void scatter_stream_lds(const ScatterCtx<float>& x,
        const PlaneInfo& pi, int wis, int a, int b, int cb)
{
    const int lane = threadIdx.x & 63;
    const int W = pi.W, W2 = W * W;
    const int stride = 64 * pi.wpp;
    const char* __restrict__ ptab = static_cast<const char*>(x.etab);
    unsigned toff[NS], k2[NS], loff[NS];
    bool valid[NS];
    const int c0 = cb + wis * 64 + lane;
#pragma unroll
    for (int s = 0; s < NS; ++s)
    {
        const int c = c0 + s * stride;
        valid[s] = c < W2;
        const unsigned cc = valid[s] ? (unsigned)c : 0u;
        const unsigned j = pi.wmagic ? __umulhi(cc, pi.wmagic) :
                cc / (unsigned)W;
        const unsigned k = cc - j * (unsigned)W;
        loff[s] = (j * (unsigned)x.R + k) * (unsigned)sizeof(double);
        toff[s] = cc * (unsigned)sizeof(TapT);
        k2[s] = 2 * k * (unsigned)sizeof(TapT);
    }
    const unsigned bu = (unsigned)(W - 1) * (unsigned)sizeof(TapT);
    const unsigned bv = (unsigned)((W - 1) * W) * (unsigned)sizeof(TapT);
    const unsigned bb = (unsigned)(W2 - 1) * (unsigned)sizeof(TapT);
    typedef decltype(ld_tap_raw(static_cast<const TapT*>(nullptr),
                                nullptr)) Raw;
    float tre[NS], tim[NS], acc_re[NS], acc_im[NS];
    Raw nraw[NS];           // the next table's taps, in flight
    auto load_taps = [&](unsigned table) {
        const unsigned e = table & 0x3FFFFFFFu;
        const unsigned flip = table >> 30;
        const char* tp = ptab + (size_t)e * sizeof(TapT);
        auto issue = [&](int s, unsigned o) {
            nraw[s] = ld_tap_raw(static_cast<const TapT*>(nullptr), tp + o);
        };
        if (flip == 0)
        {
#pragma unroll
            for (int s = 0; s < NS; ++s) issue(s, toff[s]);
        }
        else if (flip == 1)
        {
#pragma unroll
            for (int s = 0; s < NS; ++s) issue(s, toff[s] + bu - k2[s]);
        }
        else if (flip == 2)
        {
#pragma unroll
            for (int s = 0; s < NS; ++s) issue(s, bv - toff[s] + k2[s]);
        }
        else
        {
#pragma unroll
            for (int s = 0; s < NS; ++s) issue(s, bb - toff[s]);
        }
    };
    auto take_taps = [&]() {
#pragma unroll
        for (int s = 0; s < NS; ++s)
        {
            const float2 t = tap_widen(static_cast<const TapT*>(nullptr),
                                       nraw[s]);
            tre[s] = t.x;
            tim[s] = t.y;
        }
    };
    const unsigned sink8 = (unsigned)(x.R * x.R + lane) *
            (unsigned)sizeof(double);
    auto flush = [&](unsigned base) {
#pragma unroll
        for (int s = 0; s < NS; ++s)
            lds_add_cell(x.lds, valid[s] ? base + loff[s] : sink8,
                         (double)acc_re[s], (double)acc_im[s]);
    };
    // Records through the constant address space at a 32-bit byte offset
    // shared by the two arrays (16 bytes per point each).
    const ConstPtr<char> mbase = as_const<char>(x.meta);
    const ConstPtr<char> vbase = as_const<char>(x.vq);
    struct Rec
    {
        unsigned base, tref, z;
        float vx, vy, vz, vw;
    };
    auto load = [&](int i) {
        const unsigned o = (unsigned)i * 16u;
        const ConstPtr<unsigned> m =
                reinterpret_cast<ConstPtr<unsigned>>(mbase + o);
        const ConstPtr<float> v = reinterpret_cast<ConstPtr<float>>(vbase + o);
        Rec r;
        r.base = m[0];
        r.tref = m[1];
        r.z = m[2];
        r.vx = v[0];
        r.vy = v[1];
        r.vz = v[2];
        r.vw = v[3];
        return r;
    };
    Rec A = load(a), B = A, C = A;
    if (a + 1 < b) B = load(a + 1);
    load_taps(A.tref);
    take_taps();
    bool fresh = true;
    int k = a;
    // Point k from cur; nxt holds point k + 1, fut receives point k + 2.
    // Returns false after the last point.
    auto step = [&](const Rec& cur, const Rec& nxt, Rec& fut) -> bool {
        const bool more = k + 1 < b;
        if (k + 2 < b) fut = load(k + 2);
        const bool nt = more && (nxt.z & kNewTab);
        if (nt) load_taps(nxt.tref);
        if (fresh)
        {
#pragma unroll
            for (int s = 0; s < NS; ++s)
            {
                acc_re[s] = fmaf(cur.vy, tim[s], cur.vx * tre[s]);
                acc_im[s] = fmaf(cur.vw, tim[s], cur.vz * tre[s]);
            }
        }
        else
        {
#pragma unroll
            for (int s = 0; s < NS; ++s)
            {
                acc_re[s] = fmaf(cur.vx, tre[s], acc_re[s]);
                acc_re[s] = fmaf(cur.vy, tim[s], acc_re[s]);
                acc_im[s] = fmaf(cur.vz, tre[s], acc_im[s]);
                acc_im[s] = fmaf(cur.vw, tim[s], acc_im[s]);
            }
        }
        if (!more)
        {
            flush(cur.base);
            return false;
        }
        fresh = (nxt.z & kNewPos) != 0;
        if (fresh) flush(cur.base);
        if (nt) take_taps();
        ++k;
        return true;
    };
    for (;;)
    {
        if (!step(A, B, C)) break;
        if (!step(B, C, A)) break;
        if (!step(C, A, B)) break;
    }
}

template <typename T> constexpr int slots_upto(int n)
{
    return n < slots_max<T>() ? n : slots_max<T>();
}

template <typename T, bool DIAG, bool LDS, typename TapT, bool COMP>
__device__ __forceinline__ void scatter_dispatch(const ScatterCtx<T>& x,
        const PlaneInfo& pi, int wis, int a, int b, int cb)
{
    // The plan never sets nslot above slots_max<T>() (<= 8).
    static_assert(slots_max<T>() <= 8, "dispatch covers 1..8 slots");
#define GW_STREAM(n) do { if constexpr (LDS && !DIAG && sizeof(T) == 4) \
        scatter_stream_lds<slots_upto<T>(n), TapT>(x, pi, wis, a, b, cb); \
        else scatter_stream<T, DIAG, LDS, slots_upto<T>(n), TapT, COMP>( \
        x, pi, wis, a, b, cb); } while (0)
    switch (pi.nslot)
    {
    case 1: GW_STREAM(1); break;
    case 2: GW_STREAM(2); break;
    case 3: GW_STREAM(3); break;
    case 4: GW_STREAM(4); break;
    case 5: GW_STREAM(5); break;
    case 6: GW_STREAM(6); break;
    case 7: GW_STREAM(7); break;
    default: GW_STREAM(8); break;
    }
#undef GW_STREAM
}

// Region -> compensated grid (re_hi, re_lo, im_hi, im_lo per cell), kBatch
// cells per thread at a time, each component as in comp_add: the returning
// adds into hi of all of them are issued first, so their round trips
// overlap, then the adds into lo.  Clears the region.  y0, x0: grid row (relative to the rows held) and
// column of region cell 0.
__device__ __forceinline__ void flush_region_comp(GridLds& L, int R, int y0,
        int x0, int G, float* grid)
{
    constexpr int kBatch = 4;
    const int NT = blockDim.x;
    const int cells = R * R;
    for (int base = threadIdx.x; base < cells; base += NT * kBatch)
    {
        float* q[2 * kBatch];
        float hi[2 * kBatch], lo[2 * kBatch], old[2 * kBatch];
        bool live[2 * kBatch];
#pragma unroll
        for (int b = 0; b < kBatch; ++b)
        {
            const int i = base + b * NT;
            double re = 0.0, im = 0.0;
            if (i < cells)
            {
                re = L.re[i];
                im = L.im[i];
            }
            const bool nz = re != 0.0 || im != 0.0;
            live[2 * b] = live[2 * b + 1] = nz;
            hi[2 * b] = (float)re;
            lo[2 * b] = (float)(re - (double)hi[2 * b]);
            hi[2 * b + 1] = (float)im;
            lo[2 * b + 1] = (float)(im - (double)hi[2 * b + 1]);
            q[2 * b] = q[2 * b + 1] = nullptr;
            if (nz)
            {
                L.re[i] = 0.0;
                L.im[i] = 0.0;
                const int y = i / R;
                const int xx = i - y * R;
                const size_t g = (size_t)((long long)(y0 + y) * G + (x0 + xx));
                q[2 * b] = grid + 4 * g;
                q[2 * b + 1] = q[2 * b] + 2;
            }
        }
        // The returning adds of the batch first, then the low parts (the
        // same two atomics per component as comp_add, in flight together).
#pragma unroll
        for (int c = 0; c < 2 * kBatch; ++c)
            old[c] = live[c] ? __builtin_amdgcn_global_atomic_fadd_f32(
                    as_global(q[c]), hi[c]) : 0.0f;
#pragma unroll
        for (int c = 0; c < 2 * kBatch; ++c)
            if (live[c])
                atomic_add_real(q[c] + 1, lo[c] + two_sum_err(old[c], hi[c]));
    }
}

// One work item (up to kJobPoints sorted points of one tile, origin
// job.x = u | v << 16, points [job.y, job.y + (job.z & 0xFFFF))) by one
// workgroup: its support segments as k_item_segments listed them, streams
// of waves per segment, and
// with an LDS region (lds) the region added to the grid at the end and
// cleared, sink cells included (the next item may have another R).
template <typename T, bool DIAG, typename TapT, bool COMP, typename Pre>
__device__ __forceinline__ void grid_work_item(GridLds& L, ScatterCtx<T>& x,
        const PlaneInfo* __restrict__ pinfo, const int4& job, bool lds,
        int min_share, Pre&& pre)
{
    const int tid = threadIdx.x;
    const int NT = blockDim.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int R = x.R, G = x.G, row0 = x.row0;
    T* const grid = x.grid;
    x.ux0 = job.x & 0xFFFF;
    x.vy0 = (int)((unsigned)job.x >> 16);
    const int start = job.y;
    // job.z: the item's points, and its support segments above bit 16
    // (k_item_segments); segs[start + g] = segment g's first point in the
    // item | its plane << 12 (planes of a segment share the footprint
    // shape, hence the lane map).
    const int len = job.z & 0xFFFF;
    const unsigned nseg = (unsigned)job.z >> 16;
    const ConstPtr<unsigned> sg = as_const<unsigned>(x.segs + start);
    for (unsigned g = 0; g < nseg; ++g)
    {
        const unsigned sv = sg[g];
        const int plane = (int)(sv >> 12);
        const PlaneInfo pi = ld_pinfo(pinfo, plane);
        const int ra = (int)(sv & 0xFFFu);
        const int rb = g + 1 < nseg ? (int)(sg[g + 1] & 0xFFFu) : len;
        // streams = nwaves / wpp, a power of two (2^sshift).
        const int smask = (1 << pi.sshift) - 1;
        const int stream = wave / pi.wpp;
        if (stream > smask) continue;
        const int wis = wave - stream * pi.wpp;
        const int n = rb - ra;
        const int share = max(min_share, (n + smask) >> pi.sshift);
        const int q = (stream - ra) & smask;
        const int sa = start + ra + q * share;
        const int sb = min(start + rb, sa + share);
        if (sa >= sb) continue;
        const int pass = 64 * pi.wpp * pi.nslot;
        for (int cb = 0; cb < pi.W * pi.W; cb += pass)
        {
            if (lds && !DIAG)
                scatter_dispatch<T, DIAG, true, TapT, COMP>(x, pi, wis, sa,
                        sb, cb);
            else
                scatter_dispatch<T, DIAG, false, TapT, COMP>(x, pi, wis, sa,
                        sb, cb);
        }
    }
    // The next work item is requested as this one's points are done, so
    // that the queue round trip overlaps the region flush (see
    // for_each_item); earlier, more tiles would be in flight per XCD and
    // their tap tables would share its L2.
    pre();
    __syncthreads();
    if (lds)
    {
        // Region -> grid, coalesced along rows, clearing it for the next
        // item.  Cells no footprint reached hold zero and are skipped (they
        // may lie off the grid).
        if constexpr (COMP)
            flush_region_comp(L, R, x.vy0 - row0, x.ux0, G,
                              reinterpret_cast<float*>(grid));
        else
            for (int i = tid; i < R * R; i += NT)
            {
                const double re = L.re[i];
                const double im = L.im[i];
                if (re != 0.0 || im != 0.0)
                {
                    L.re[i] = 0.0;
                    L.im[i] = 0.0;
                    const int y = i / R;
                    const int xx = i - y * R;
                    const size_t g = (size_t)(x.vy0 - row0 + y) * (size_t)G +
                            (size_t)(x.ux0 + xx);
                    grid_add<T, COMP>(grid, g, re, im);
                }
            }
        for (int i = tid; i < kSinkCells; i += NT)
        {
            L.re[R * R + i] = 0.0;
            L.im[R * R + i] = 0.0;
        }
        __syncthreads();
    }
}

// Calls f(it, pre) for every scatter item it of [0, s->total_items), each
// item in exactly one workgroup (persistent grid).  xcds <= 1 (or more than
// the queues in Summary): grid-stride order.
// Otherwise: items are in tile Z-order, so a contiguous range of them covers
// one compact grid region, whose points use a narrower set of w-planes than
// the whole grid.  Workgroups are dispatched to the XCDs round robin
// (workgroup b runs on XCD b mod xcds); XCD x takes the x-th of xcds
// contiguous ranges from its own queue (s->item_next[x], zero at launch), so
// its L2 holds the tap tables of that region only, then helps the other
// queues in turn.  f calls pre() once, before its last barrier: lane 0 of
// the last wave then requests the next index of the same queue into LDS,
// and the index is read after f returns, so the atomic's round trip
// overlaps the end of the item instead of preceding the next one.  f has at
// least one barrier.  Every workgroup must call this (it synchronises).
// f is called from one place only, so that it is inlined once: an
// out-of-line f would receive the LDS region through a generic pointer and
// its atomics would become flat atomics.
template <typename F>
__device__ __forceinline__ void for_each_item(Summary* __restrict__ s,
        int xcds, unsigned* s_next, F&& f)
{
    // s_next[0..1]: the prefetched index by item parity (a slot is written
    // again two items later, after a barrier every wave passes once it has
    // read it); s_next[2]: take().
    const unsigned total = s->total_items;
    const bool queues = xcds > 1 && xcds <= kXcdGroups;
    const unsigned nq = queues ? (unsigned)xcds : 1u;
    const unsigned per = (total + nq - 1) / nq;
    const unsigned home = blockIdx.x % nq;
    const bool leader = (int)threadIdx.x == (int)blockDim.x - 64;
    unsigned q = 0;                 // queues: the q-th after the home queue
    // The next item of queue q on, synchronously (~0u once all are empty).
    auto take = [&]() -> unsigned {
        for (; q < nq; ++q)
        {
            const unsigned xq = (home + q) % nq;
            const unsigned lo = xq * per;
            const unsigned hi = min(total, lo + per);
            if (threadIdx.x == 0)
                s_next[2] = lo + atomicAdd(&s->item_next[xq], 1u);
            __syncthreads();
            const unsigned got = (unsigned)__builtin_amdgcn_readfirstlane(
                    (int)s_next[2]);
            __syncthreads();
            if (got < hi) return got;
        }
        return ~0u;
    };
    if (!queues)
    {
        for (unsigned it = blockIdx.x; it < total; it += gridDim.x)
            f(it, [] {});
        return;
    }
    unsigned it = take();
    unsigned par = 0;
    while (it != ~0u)
    {
        const unsigned xq = (home + q) % nq;
        const unsigned lo = xq * per;
        const unsigned hi = min(total, lo + per);
        f(it, [&] {
            if (leader) s_next[par] = lo + atomicAdd(&s->item_next[xq], 1u);
        });
        // f's last barrier orders the store above before this read; the
        // slot is stored again in the item after next, past the next
        // item's last barrier, so after every wave's read.
        const unsigned got = (unsigned)__builtin_amdgcn_readfirstlane(
                (int)s_next[par]);
        par ^= 1u;
        if (got < hi)
            it = got;
        else
        {
            // Every wave has read *s_next before take() writes it.
            __syncthreads();
            ++q;
            it = take();
        }
    }
}

// Grid-stride loop over the jobs (persistent: the plan launches as many
// workgroups as fit the device at once).  Taps are stored as TapT pairs;
// COMP selects the compensated two-float grid layout.
template <typename T, bool DIAG, typename TapT = typename Types<T>::T2,
          bool COMP = false>
__global__ void __launch_bounds__(kGridThreads) k_grid(
        Summary* __restrict__ summary, const int4* __restrict__ items,
        const PlaneInfo* __restrict__ pinfo, const int4* __restrict__ meta,
        const typename Types<T>::T4* __restrict__ vq,
        const TapT* __restrict__ etab, int G, int row0,
        int R, T* __restrict__ grid, int force_direct, int min_share,
        unsigned* __restrict__ diag_count,
        unsigned long long* __restrict__ diag_wsum, int xcds,
        const unsigned* __restrict__ segs)
{
    __shared__ GridLds L;
    __shared__ unsigned s_next[3];
    const int tid = threadIdx.x;
    const int NT = blockDim.x;
    const bool lds = !force_direct && R > 0;
    ScatterCtx<T> x;
    x.meta = meta;
    x.vq = vq;
    x.etab = etab;
    x.segs = segs;
    x.lds = &L;
    x.grid = grid;
    x.diag_count = diag_count;
    x.diag_wsum = diag_wsum;
    x.R = R;
    x.G = G;
    x.row0 = row0;
    const int rcells = lds ? R * R + kSinkCells : 0;
    for (int i = tid; i < rcells; i += NT)
    {
        L.re[i] = 0.0;
        L.im[i] = 0.0;
    }
    for_each_item(summary, xcds, s_next, [&](unsigned it, auto&& pre) {
        grid_work_item<T, DIAG, TapT, COMP>(L, x, pinfo, items[it], lds,
                                            min_share, pre);
    });
}

// -------------------------------------------------------------------------
// Streaming copy (profiler calibration and plain use)
// -------------------------------------------------------------------------

__global__ void __launch_bounds__(256) k_copy(uint4* __restrict__ dst,
        const uint4* __restrict__ src, size_t n16, unsigned char* dst_tail,
        const unsigned char* src_tail, int tail)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16;
         i += stride)
        dst[i] = src[i];
    if (blockIdx.x == 0 && (int)threadIdx.x < tail)
        dst_tail[threadIdx.x] = src_tail[threadIdx.x];
}

void device_copy(void* dst, const void* src, size_t bytes)
{
    if (bytes == 0) return;
    if ((reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src))
            & 15u)
        fail(kErrInvalidArgument, "gw_device_copy needs 16-byte aligned "
             "pointers");
    const size_t n16 = bytes / 16;
    const int tail = (int)(bytes - n16 * 16);
    int dev = 0, ncu = 0;
    check_hip(hipGetDevice(&dev), kErrComputeDevices, "hipGetDevice");
    check_hip(hipDeviceGetAttribute(&ncu,
            hipDeviceAttributeMultiprocessorCount, dev), kErrComputeDevices,
            "hipDeviceGetAttribute");
    const int blocks = (int)std::max<size_t>(1, std::min<size_t>(
            (n16 + 255) / 256, (size_t)ncu * 8));
    hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, nullptr,
            static_cast<uint4*>(dst), static_cast<const uint4*>(src), n16,
            static_cast<unsigned char*>(dst) + n16 * 16,
            static_cast<const unsigned char*>(src) + n16 * 16, tail);
    check_hip(hipGetLastError(), kErrKernelLaunch, "k_copy launch");
    check_hip(hipDeviceSynchronize(), kErrKernelLaunch, "k_copy");
}

// Compensated two-float grid -> float grid: out[2q + c] = hi + lo of
// component c of cell q (layout re_hi, re_lo, im_hi, im_lo per cell).
__global__ void __launch_bounds__(256) k_resolve(const float2* __restrict__ comp,
        float* __restrict__ out, size_t n_components)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
         i < n_components; i += stride)
    {
        const float2 p = comp[i];
        out[i] = p.x + p.y;
    }
}

void resolve_compensated(const float* comp, size_t num_cells, float* out,
                         hipStream_t stream)
{
    if (num_cells == 0) return;
    const size_t n = 2 * num_cells;
    const int blocks = (int)std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_resolve, dim3(blocks), dim3(256), 0, stream,
            reinterpret_cast<const float2*>(comp), out, n);
    check_hip(hipGetLastError(), kErrKernelLaunch, "k_resolve launch");
}

// -------------------------------------------------------------------------
// Plan
// -------------------------------------------------------------------------
namespace {

int bits_for(unsigned long long v)
{
    int b = 0;
    while (b < 64 && (v >> b) != 0) ++b;
    return b;
}

// Expanded tap table: for every plane p and sub-cell offset
// (off_u, off_v) >= 0 the (2S+1)^2 taps of one footprint in row-major (j, k)
// order,
//   etab[eoff_p + (off_v*M + off_u)*W^2 + (j+S)*W + (k+S)]
//     = conv_func[p*C^2 + |off_v + j*ov|*C + |off_u + k*ov|],   M = h + 1
// (the reference's flat cube index, oskar_grid_wproj.c:222-233).  Negative
// offsets need no table of their own: |-a + k*ov| = |a + (-k)*ov|, so the
// tap of (-off_u, off_v) at (j, k) is the tap of (off_u, off_v) at (j, -k),
// and likewise in v; the scatter reads the mirrored cell (the tables are
// (h+1)^2 / (2h+1)^2 = 9/25 of the full set at ov = 4).  Also the
// per-(plane, sub-cell) tap sums for all (2h+1)^2 offsets in the
// reference's order (rows j outer, columns k inner, double accumulation of
// float taps, :216-238), which are exactly the per-point `sum` the
// reference adds to the norm.
template <typename T>
void build_tables(const T* conv, size_t P, const int* support, int ov, int C,
                  int h, std::vector<PlaneInfo>& pinfo, std::vector<T>& etab,
                  std::vector<double>& normtab)
{
    const int K = 2 * h + 1;
    const int M = h + 1;
    size_t total = 0;
    for (size_t p = 0; p < P; ++p)
    {
        const size_t W = 2 * (size_t)support[p] + 1;
        pinfo[p].eoff = (unsigned)total;
        total += (size_t)M * M * W * W;
    }
    if (total * 2 * sizeof(T) >= (size_t)4 << 30)
        fail(kErrInvalidArgument,
             "support and oversample need 4 GiB or more of tap tables");
    etab.resize(2 * total);
    normtab.assign(P * K * K, 0.0);
    const size_t kdim = (size_t)C * C;
    for (size_t p = 0; p < P; ++p)
    {
        const int S = support[p];
        const size_t W = 2 * (size_t)S + 1;
        for (int iv = 0; iv < K; ++iv)
        {
            for (int iu = 0; iu < K; ++iu)
            {
                const int off_v = iv - h, off_u = iu - h;
                const bool keep = off_v >= 0 && off_u >= 0;
                T* dst = keep ? etab.data() + 2 * (pinfo[p].eoff +
                        (size_t)(off_v * M + off_u) * W * W) : nullptr;
                double sum = 0.0;
                for (int j = -S; j <= S; ++j)
                {
                    const size_t t1 = p * kdim +
                            (size_t)std::abs(off_v + j * ov) * C;
                    for (int k = -S; k <= S; ++k)
                    {
                        const size_t t = 2 * (t1 + std::abs(off_u + k * ov));
                        if (keep)
                        {
                            *dst++ = conv[t];
                            *dst++ = conv[t + 1];
                        }
                        sum += conv[t];
                    }
                }
                normtab[(p * K + iv) * K + iu] = sum;
            }
        }
    }
}


}  // namespace

// The expanded tap table and the tap-sum table built on the device from a
// float cube already in device memory: the same gathers as build_tables,
// and the same sums in the same order (rows j outer, columns k inner,
// double accumulation of float taps), so both tables are bit-identical to
// the host build of the same cube.
__global__ void __launch_bounds__(256) k_expand_taps(
        const float* __restrict__ conv, int C, int ov, int M,
        const int* __restrict__ support, const PlaneInfo* __restrict__ pinfo,
        float* __restrict__ etab)
{
    const int p = blockIdx.y;
    const int S = support[p];
    const int W = 2 * S + 1;
    const unsigned n = (unsigned)(M * M * W * W);
    const size_t base = (size_t)p * C * C;
    const size_t e0 = pinfo[p].eoff;
    for (unsigned t = blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += gridDim.x * blockDim.x)
    {
        const int combo = (int)(t / (unsigned)(W * W));
        const int jk = (int)(t - (unsigned)combo * W * W);
        const int iv = combo / M, iu = combo - iv * M;
        const int j = jk / W - S, k = jk - (jk / W) * W - S;
        const size_t src = 2 * (base + (size_t)abs(iv + j * ov) * C +
                (size_t)abs(iu + k * ov));
        etab[2 * (e0 + t)] = conv[src];
        etab[2 * (e0 + t) + 1] = conv[src + 1];
    }
}

__global__ void __launch_bounds__(256) k_tap_sums(
        const float* __restrict__ conv, int P, int C, int ov, int h, int K,
        const int* __restrict__ support, double* __restrict__ normtab)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P * K * K) return;
    const int p = i / (K * K);
    const int combo = i - p * K * K;
    const int iv = combo / K, iu = combo - iv * K;
    const int S = support[p];
    const size_t base = (size_t)p * C * C;
    double sum = 0.0;
    for (int j = -S; j <= S; ++j)
    {
        const size_t t1 = base + (size_t)abs(iv - h + j * ov) * C;
        for (int k = -S; k <= S; ++k)
            sum += conv[2 * (t1 + (size_t)abs(iu - h + k * ov))];
    }
    normtab[i] = sum;
}

void validate_args(size_t num_w_planes, const int* support, int oversample,
                   int conv_size_half, const void* conv_func, int grid_size)
{
    if (num_w_planes < 1 || num_w_planes > 32768)
        fail(kErrInvalidArgument, "num_w_planes must be in [1, 32768]");
    if (!support || !conv_func)
        fail(kErrInvalidArgument, "support and conv_func must not be NULL");
    if (oversample < 1 || oversample > 253)
        fail(kErrInvalidArgument, "oversample must be in [1, 253]");
    if (conv_size_half < 1)
        fail(kErrInvalidArgument, "conv_size_half must be >= 1");
    if (grid_size < 1 || grid_size > 65535)
        fail(kErrInvalidArgument, "grid_size must be in [1, 65535]");
    // Largest flat tap index the gridder can read, per plane:
    // p*C^2 + (h + S*ov)*C + (h + S*ov).  The reference reads the flat cube
    // the same way; it must stay inside the cube.
    const unsigned long long C = (unsigned long long)conv_size_half;
    const unsigned long long h = (unsigned long long)(oversample + 1) / 2;
    const unsigned long long cube_len = num_w_planes * C * C;
    for (size_t p = 0; p < num_w_planes; ++p)
    {
        const int S = support[p];
        if (S < 0 || S > 16383)
            fail(kErrInvalidArgument, "support values must be in [0, 16383]");
        const unsigned long long m = h + (unsigned long long)S * oversample;
        if ((unsigned long long)p * C * C + m * C + m >= cube_len)
            fail(kErrInvalidArgument,
                 "conv_size_half too small for the support and oversample "
                 "(taps would be read beyond the kernel cube)");
    }
}

Plan::Plan(int precision, size_t num_w_planes, const int* support,
           int oversample, int conv_size_half, const void* conv_func,
           double cell_size_rad, double w_scale, int grid_size,
           bool cube_on_device)
{
    if (precision < kPrecSingle || precision > kPrecSingleIndex64)
        fail(kErrInvalidArgument, "precision must be GW_SINGLE, GW_DOUBLE, "
             "GW_HALF, GW_BFLOAT16 or GW_SINGLE_INDEX64");
    validate_args(num_w_planes, support, oversample, conv_size_half,
            conv_func, grid_size);
    index64_ = precision == kPrecSingleIndex64;
    prec_ = index64_ ? kPrecSingle : precision;
    precision = prec_;
    P_ = num_w_planes;
    ov_ = oversample;
    C_ = conv_size_half;
    G_ = grid_size;
    h_ = (ov_ + 1) / 2;
    K_ = 2 * h_ + 1;
    support_.assign(support, support + P_);
    s_max_ = *std::max_element(support_.begin(), support_.end());
    cell_d_ = cell_size_rad;
    ws_d_ = w_scale;
    cell_f_ = (float)cell_size_rad;
    ws_f_ = (float)w_scale;
    // grid_scale as the reference forms it (oskar_grid_wproj.c:150, :61).
    gscale_f_ = (float)grid_size * cell_f_;
    gscale_d_ = (double)grid_size * cell_d_;
    coord_bits_ = std::max(1, bits_for((unsigned long long)(G_ - 1)));
    combo_bits_ = std::max(1, bits_for((unsigned long long)(K_ * K_ - 1)));

    // Support ranks (distinct S values in increasing order).
    std::vector<int> svals(support_);
    std::sort(svals.begin(), svals.end());
    svals.erase(std::unique(svals.begin(), svals.end()), svals.end());
    srank_.resize(P_);
    for (size_t p = 0; p < P_; ++p)
        srank_[p] = (int)(std::lower_bound(svals.begin(), svals.end(),
                support_[p]) - svals.begin());
    srank_bits_ = std::max(1, bits_for((unsigned long long)(svals.size() -
            1)));
    // Each plane's index among the planes of its support rank (in plane
    // order), packed with the rank for the index stage's keys.
    srank_key_.resize(P_);
    {
        std::vector<int> seen(svals.size(), 0);
        int most = 1;
        for (size_t p = 0; p < P_; ++p)
        {
            const int local = seen[srank_[p]]++;
            most = std::max(most, local + 1);
            srank_key_[p] = srank_[p] | (local << 16);
        }
        plane_bits_ = bits_for((unsigned long long)(most - 1));
    }
    // Tile and in-tile coordinates take at most one bit more per axis
    // than a grid coordinate (one each for a side that is not a power of
    // two); configure() sets the exact width for the chosen side.
    key_bits_max_ = 2 * coord_bits_ + 2 + combo_bits_ + plane_bits_ +
            srank_bits_;
    pinfo_.resize(P_);
    for (size_t p = 0; p < P_; ++p)
    {
        PlaneInfo& pi = pinfo_[p];
        pi.S = support_[p];
        pi.W = 2 * pi.S + 1;
        pi.eoff = 0;
        pi.srank = srank_[p];
    }
    configure(-1, 0, 0);

    check_hip(hipGetDevice(&device_), kErrComputeDevices, "hipGetDevice");
    check_hip(hipDeviceGetAttribute(&n_cu_,
            hipDeviceAttributeMultiprocessorCount, device_),
            kErrComputeDevices, "hipDeviceGetAttribute");
    int occ = 0;
    check_hip(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ,
            precision == kPrecDouble ? (const void*)k_grid<double, false> :
            (const void*)k_grid<float, false>, kGridThreads, 0),
            kErrComputeDevices, "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    grid_blocks_ = n_cu_ * std::max(1, occ);
    check_hip(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking),
            kErrComputeDevices, "hipStreamCreateWithFlags");
    own_stream_ = true;
    for (auto& e : ev_)
        check_hip(hipEventCreate(&e), kErrComputeDevices, "hipEventCreate");

    // Constant tables.
    if (cube_on_device)
    {
        build_tables_device(static_cast<const float*>(conv_func));
        return;
    }
    std::vector<double> normtab;
    std::vector<float> etab_f;
    std::vector<double> etab_d;
    const void* etab_host = nullptr;
    size_t etab_bytes = 0;
    std::vector<HalfPair> etab_h;
    std::vector<BF16Pair> etab_b;
    if (prec_ == kPrecSingle)
    {
        build_tables(static_cast<const float*>(conv_func), P_,
                support_.data(), ov_, C_, h_, pinfo_, etab_f, normtab);
        etab_host = etab_f.data();
        etab_bytes = etab_f.size() * sizeof(float);
        etab_taps_ = etab_f.size() / 2;
    }
    else if (prec_ == kPrecHalf || prec_ == kPrecBFloat16)
    {
        // Taps are stored in 16 bits: round the cube first, so that the
        // tables and the norm table hold exactly the stored values.
        const float* cf = static_cast<const float*>(conv_func);
        std::vector<float> cube(cf, cf + 2 * P_ * (size_t)C_ * C_);
        for (float& v : cube)
            v = prec_ == kPrecHalf ? (float)(half_t)v :
                    bf16_to_float(float_to_bf16(v));
        build_tables(cube.data(), P_, support_.data(), ov_, C_, h_, pinfo_,
                etab_f, normtab);
        etab_taps_ = etab_f.size() / 2;
        if (prec_ == kPrecHalf)
        {
            etab_h.resize(etab_taps_);
            for (size_t i = 0; i < etab_taps_; ++i)
            {
                etab_h[i].x = (half_t)etab_f[2 * i];
                etab_h[i].y = (half_t)etab_f[2 * i + 1];
            }
            etab_host = etab_h.data();
            etab_bytes = etab_h.size() * sizeof(HalfPair);
        }
        else
        {
            etab_b.resize(etab_taps_);
            for (size_t i = 0; i < etab_taps_; ++i)
            {
                etab_b[i].x = float_to_bf16(etab_f[2 * i]);
                etab_b[i].y = float_to_bf16(etab_f[2 * i + 1]);
            }
            etab_host = etab_b.data();
            etab_bytes = etab_b.size() * sizeof(BF16Pair);
        }
        std::vector<float>().swap(etab_f);
    }
    else
    {
        build_tables(static_cast<const double*>(conv_func), P_,
                support_.data(), ov_, C_, h_, pinfo_, etab_d, normtab);
        etab_host = etab_d.data();
        etab_bytes = etab_d.size() * sizeof(double);
        etab_taps_ = etab_d.size() / 2;
    }
    d_etab_ = dmalloc<char>(etab_bytes);
    table_bytes_ = etab_bytes + P_ * (sizeof(int) + sizeof(PlaneInfo)) +
            normtab.size() * sizeof(double);
    d_support_ = dmalloc<int>(P_);
    d_srank_ = dmalloc<int>(P_);
    d_pinfo_ = dmalloc<PlaneInfo>(P_);
    d_normtab_ = dmalloc<double>(normtab.size());
    d_combo_ = dmalloc<unsigned>(P_ * K_ * K_);
    d_summary_ = dmalloc<Summary>(1);
    check_hip(hipHostMalloc((void**)&h_summary_, sizeof(Summary),
            hipHostMallocDefault), kErrMemAlloc, "hipHostMalloc");
    std::memset(h_summary_, 0, sizeof(Summary));
    check_hip(hipMemcpy(d_etab_, etab_host, etab_bytes,
            hipMemcpyHostToDevice), kErrMemCopy, "upload tap tables");
    check_hip(hipMemcpy(d_support_, support_.data(), P_ * sizeof(int),
            hipMemcpyHostToDevice), kErrMemCopy, "upload support");
    check_hip(hipMemcpy(d_srank_, srank_key_.data(), P_ * sizeof(int),
            hipMemcpyHostToDevice), kErrMemCopy, "upload support ranks");
    check_hip(hipMemcpy(d_pinfo_, pinfo_.data(), P_ * sizeof(PlaneInfo),
            hipMemcpyHostToDevice), kErrMemCopy, "upload plane table");
    check_hip(hipMemcpy(d_normtab_, normtab.data(),
            normtab.size() * sizeof(double), hipMemcpyHostToDevice),
            kErrMemCopy, "upload norm table");
}

void Plan::build_tables_device(const float* d_conv)
{
    if (prec_ != kPrecSingle)
        fail(kErrInvalidArgument, "a kernel cube in device memory needs a "
             "GW_SINGLE or GW_SINGLE_INDEX64 plan");
    size_t total = 0;
    size_t most = 0;
    for (size_t p = 0; p < P_; ++p)
    {
        const size_t W = 2 * (size_t)support_[p] + 1;
        pinfo_[p].eoff = (unsigned)total;
        total += (size_t)(h_ + 1) * (h_ + 1) * W * W;
        most = std::max(most, (size_t)(h_ + 1) * (h_ + 1) * W * W);
    }
    if (total * 2 * sizeof(float) >= (size_t)4 << 30)
        fail(kErrInvalidArgument,
             "support and oversample need 4 GiB or more of tap tables");
    etab_taps_ = total;
    const size_t etab_bytes = total * 2 * sizeof(float);
    const size_t ntab = P_ * (size_t)K_ * K_;
    d_etab_ = dmalloc<char>(etab_bytes);
    table_bytes_ = etab_bytes + P_ * (sizeof(int) + sizeof(PlaneInfo)) +
            ntab * sizeof(double);
    d_support_ = dmalloc<int>(P_);
    d_srank_ = dmalloc<int>(P_);
    d_pinfo_ = dmalloc<PlaneInfo>(P_);
    d_normtab_ = dmalloc<double>(ntab);
    d_combo_ = dmalloc<unsigned>(ntab);
    d_summary_ = dmalloc<Summary>(1);
    check_hip(hipHostMalloc((void**)&h_summary_, sizeof(Summary),
            hipHostMallocDefault), kErrMemAlloc, "hipHostMalloc");
    std::memset(h_summary_, 0, sizeof(Summary));
    check_hip(hipMemcpy(d_support_, support_.data(), P_ * sizeof(int),
            hipMemcpyHostToDevice), kErrMemCopy, "upload support");
    check_hip(hipMemcpy(d_srank_, srank_key_.data(), P_ * sizeof(int),
            hipMemcpyHostToDevice), kErrMemCopy, "upload support ranks");
    check_hip(hipMemcpy(d_pinfo_, pinfo_.data(), P_ * sizeof(PlaneInfo),
            hipMemcpyHostToDevice), kErrMemCopy, "upload plane table");
    const unsigned bx = (unsigned)std::min<size_t>((most + 255) / 256, 1024);
    hipLaunchKernelGGL(k_expand_taps, dim3(bx, (unsigned)P_), dim3(256), 0,
            stream_, d_conv, C_, ov_, h_ + 1, (const int*)d_support_,
            (const PlaneInfo*)d_pinfo_, reinterpret_cast<float*>(d_etab_));
    hipLaunchKernelGGL(k_tap_sums, dim3((unsigned)((ntab + 255) / 256)),
            dim3(256), 0, stream_, d_conv, (int)P_, C_, ov_, h_, K_,
            (const int*)d_support_, d_normtab_);
    check_hip(hipGetLastError(), kErrKernelLaunch, "table build launch");
    check_hip(hipStreamSynchronize(stream_), kErrKernelLaunch,
            "stream synchronize");
}

int Plan::max_tile_bits() const
{
    const long long wmax = 2ll * s_max_ + 1;
    for (int t = std::min(6, coord_bits_); t >= 0; --t)
    {
        const long long R = (1ll << t) + wmax - 1;
        if (R * R + kSinkCells <= kRegionCells) return t;
    }
    return -1;
}

int Plan::max_tile() const
{
    const long long wmax = 2ll * s_max_ + 1;
    for (int t = 1 << std::min(6, coord_bits_); t >= 1; --t)
    {
        const long long R = t + wmax - 1;
        if (R * R + kSinkCells <= kRegionCells) return t;
    }
    return -1;
}

void Plan::configure(int tile_bits, int slot_cap, int min_share)
{
    if (tile_bits > std::min(6, coord_bits_))
        fail(kErrInvalidArgument, "tile_log2 must be in [0, " +
             std::to_string(std::min(6, coord_bits_)) + "]");
    configure_side(tile_bits < 0 ? -1 : 1 << tile_bits, slot_cap, min_share);
}

void Plan::configure_side(int tile, int slot_cap, int min_share)
{
    const int smax_all = prec_ == kPrecDouble ? slots_max<double>() :
            slots_max<float>();
    if (slot_cap < 0 || slot_cap > smax_all)
        fail(kErrInvalidArgument, "slot_cap must be in [0, " +
             std::to_string(smax_all) + "] (0: the largest)");
    if (min_share < 0 || min_share > kJobPoints)
        fail(kErrInvalidArgument, "min_share must be in [0, " +
             std::to_string(kJobPoints) + "] (0: the default)");
    // Tile side T (<= 64) of footprint origins: by default the largest whose
    // LDS region (T + Wmax - 1)^2 plus the sink cells fits each region
    // array, a power of two or not (a side that fills the region makes
    // fewer, fuller work items where points are sparse).  If none fits the
    // scatter adds straight into the grid in HBM.
    const int tcap = 1 << std::min(6, coord_bits_);
    const int tmax = max_tile();
    const long long wmax = 2ll * s_max_ + 1;
    if (tile < 0)
    {
        tile_ = tmax >= 1 ? tmax : tcap;
        region_ = tmax >= 1 ? (int)(tmax + wmax - 1) : 0;
    }
    else
    {
        if (tile < 1 || tile > tcap)
            fail(kErrInvalidArgument, "tile side must be in [1, " +
                 std::to_string(tcap) + "]");
        tile_ = tile;
        const long long R = tile + wmax - 1;
        region_ = R * R + kSinkCells <= kRegionCells ? (int)R : 0;
    }
    tile_bits_ = bits_for((unsigned long long)(tile_ - 1));
    tile_coord_bits_ = bits_for((unsigned long long)((G_ - 1) / tile_));
    tile_shift_ = 2 * tile_bits_ + combo_bits_ + plane_bits_ + srank_bits_;
    key_bits_ = tile_shift_ + 2 * tile_coord_bits_;
    slot_cap_ = slot_cap > 0 ? slot_cap : smax_all;
    min_share_ = min_share > 0 ? min_share : kMinShare;

    // Streams per plane: wpp waves cover the footprint with <= slot_cap_
    // cells per lane (larger footprints take several passes).
    const int smax = slot_cap_;
    for (size_t p = 0; p < P_; ++p)
    {
        PlaneInfo& pi = pinfo_[p];
        const long long cells = (long long)pi.W * pi.W;
        pi.wpp = 1;
        while (pi.wpp < kGridThreads / 64 &&
                (long long)pi.wpp * 64 * smax < cells)
            pi.wpp *= 2;
        pi.nslot = (int)std::min<long long>(smax,
                (cells + 64ll * pi.wpp - 1) / (64ll * pi.wpp));
        pi.sshift = 0;
        while ((kGridThreads / 64 / pi.wpp) >> (pi.sshift + 1)) ++pi.sshift;
        // ceil(2^32 / W) divides exactly over c < W^2 when
        // (W^2 - 1) * (m * W - 2^32) < 2^32 (W == 1: only c == 0 occurs).
        const unsigned long long two32 = 1ull << 32;
        const unsigned long long wm = pi.W == 1 ? 0xFFFFFFFFull :
                (two32 + (unsigned long long)pi.W - 1) /
                (unsigned long long)pi.W;
        const unsigned long long err = wm * (unsigned long long)pi.W - two32;
        pi.wmagic = (pi.W == 1 ||
                (unsigned long long)(cells - 1) * err < two32) ?
                (unsigned)wm : 0u;
    }
    if (d_pinfo_)
    {
        if (stream_)
            check_hip(hipStreamSynchronize(stream_), kErrKernelLaunch,
                      "stream synchronize");
        check_hip(hipMemcpy(d_pinfo_, pinfo_.data(), P_ * sizeof(PlaneInfo),
                hipMemcpyHostToDevice), kErrMemCopy, "upload plane table");
    }
}

double Plan::autotune(size_t n, const void* uu, const void* vv,
                      const void* ww, const void* vis, void* scratch_grid,
                      int reps)
{
    if (reps < 1) reps = 1;
    auto measure = [&]() {
        double best = 1e300;
        for (int r = 0; r < reps + 1; ++r)   // first run warms up
        {
            locate(n, uu, vv, ww, vis, nullptr);
            scatter(n, scratch_grid, 0, nullptr);
            double ms[5];
            if (timings(ms, 5) == 5 && r > 0) best = std::min(best, ms[4]);
        }
        return best;
    };
    const int smax_all = prec_ == kPrecDouble ? slots_max<double>() :
            slots_max<float>();
    int bt = tile_, bs = slot_cap_, bm = min_share_;
    configure_side(bt, bs, bm);
    double best = measure();
    auto consider = [&](int t, int sc, int ms) {
        configure_side(t, sc, ms);
        const double v = measure();
        if (v < best)
        {
            best = v;
            bt = t;
            bs = sc;
            bm = ms;
        }
    };
    // Tile sides: the largest that fits, then the three largest powers of
    // two not above it.
    const int tfit = max_tile();
    if (tfit >= 1)
    {
        int p2 = 1;
        while (p2 * 2 <= tfit) p2 *= 2;
        if (tfit != bt) consider(tfit, bs, bm);
        for (int k = 0; k < 3 && p2 >= 1; ++k, p2 /= 2)
            if (p2 != tfit && p2 != bt) consider(p2, bs, bm);
    }
    for (int sc : {4, 6, 8})
        if (sc <= smax_all && sc != bs) consider(bt, sc, bm);
    for (int ms : {4, 8, 16})
        if (ms != bm) consider(bt, bs, ms);
    configure_side(bt, bs, bm);
    return best;
}

Plan::~Plan()
{
    if (stream_) (void)hipStreamSynchronize(stream_);
    free_workspace();
    dfree(d_etab_);
    dfree(d_support_);
    dfree(d_srank_);
    dfree(d_pinfo_);
    dfree(d_normtab_);
    dfree(d_combo_);
    dfree(d_summary_);
    if (h_summary_) (void)hipHostFree(h_summary_);
    for (auto& e : ev_)
        if (e) (void)hipEventDestroy(e);
    if (own_stream_ && stream_) (void)hipStreamDestroy(stream_);
}

void Plan::set_stream(hipStream_t s)
{
    if (s == stream_ && s) return;
    // Calls still queued on the previous stream use the plan's workspace:
    // finish them first, whoever owns that stream.
    if (stream_ && (s || !own_stream_)) (void)hipStreamSynchronize(stream_);
    if (s)
    {
        if (own_stream_ && stream_) (void)hipStreamDestroy(stream_);
        stream_ = s;
        own_stream_ = false;
    }
    else if (!own_stream_)
    {
        check_hip(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking),
                kErrComputeDevices, "hipStreamCreateWithFlags");
        own_stream_ = true;
    }
}

Plan::Workspace Plan::save_workspace() const
{
    Workspace w;
    w.cap = cap_;
    w.items_cap = items_cap_;
    for (int i = 0; i < 2; ++i)
    {
        w.keys[i] = d_keys_[i];
        w.idx[i] = d_idx_[i];
    }
    w.rec = d_rec_;
    w.meta = d_meta_;
    w.vq = d_vq_;
    w.gw = d_gw_;
    w.sort_tmp = d_sort_tmp_;
    w.sort_tmp_bytes = sort_tmp_bytes_;
    w.head = d_head_;
    w.bpos = d_bpos_;
    w.bstart = d_bstart_;
    w.scan_part = d_scan_part_;
    w.items = d_items_;
    return w;
}

void Plan::set_workspace(const Workspace& w)
{
    cap_ = w.cap;
    items_cap_ = w.items_cap;
    for (int i = 0; i < 2; ++i)
    {
        d_keys_[i] = w.keys[i];
        d_idx_[i] = w.idx[i];
    }
    d_rec_ = w.rec;
    d_meta_ = w.meta;
    d_vq_ = w.vq;
    d_gw_ = w.gw;
    d_sort_tmp_ = w.sort_tmp;
    sort_tmp_bytes_ = w.sort_tmp_bytes;
    d_head_ = w.head;
    d_bpos_ = w.bpos;
    d_bstart_ = w.bstart;
    d_scan_part_ = w.scan_part;
    d_items_ = w.items;
}

// Scratch of the plan's key sort for n points: digit histograms and their
// offsets (radix_hist_words(n) words each) and the 256 digit totals.
static size_t sort_scratch_bytes(size_t n)
{
    return (2 * radix_hist_words(n) + 256) * sizeof(unsigned);
}

// Carves a workspace for cap points (the layout of ensure_capacity) from
// [base, base + bytes) with 256-byte aligned parts; with base == nullptr
// only counts.  Returns the bytes used.
static size_t carve_workspace(char* base, size_t cap, size_t rec_bytes,
        size_t real_size, size_t sort_tmp_bytes, size_t items_cap,
        unsigned long long** keys, unsigned** idx, void** rec, int4** meta,
        void** vq, unsigned** gw, void** sort_tmp, unsigned** head,
        unsigned** bpos, unsigned** bstart, unsigned** scan_part,
        int4** items)
{
    size_t off = 0;
    auto take = [&](size_t n) -> char* {
        char* p = base ? base + off : nullptr;
        off += (n + 255) & ~(size_t)255;
        return p;
    };
    for (int i = 0; i < 2; ++i)
    {
        keys[i] = reinterpret_cast<unsigned long long*>(take(cap * 8));
        idx[i] = reinterpret_cast<unsigned*>(take(cap * 4));
    }
    *rec = take(cap * rec_bytes);
    *meta = reinterpret_cast<int4*>(take(cap * sizeof(int4)));
    *vq = take(cap * 4 * real_size);
    *gw = reinterpret_cast<unsigned*>(take(cap * 4));
    *head = reinterpret_cast<unsigned*>(take(cap * 4));
    *bpos = reinterpret_cast<unsigned*>(take(cap * 4));
    *bstart = reinterpret_cast<unsigned*>(take((cap + 1) * 4));
    *scan_part = reinterpret_cast<unsigned*>(take((cap / kScanTile + 2) * 4));
    *sort_tmp = take(sort_tmp_bytes);
    *items = reinterpret_cast<int4*>(take(items_cap * sizeof(int4)));
    return off;
}

size_t Plan::external_workspace_bytes(size_t n) const
{
    const size_t cap = std::max<size_t>(n, 1024);
    const size_t sort_bytes = sort_scratch_bytes(cap);
    Workspace w{};
    return carve_workspace(nullptr, cap, prec_ == kPrecDouble ?
            sizeof(RecD) : sizeof(RecF), prec_ == kPrecDouble ?
            sizeof(double) : sizeof(float), sort_bytes,
            cap + cap / kJobPoints + 2, w.keys, w.idx, &w.rec, &w.meta,
            &w.vq, &w.gw, &w.sort_tmp, &w.head, &w.bpos, &w.bstart,
            &w.scan_part, &w.items);
}

void Plan::use_external_workspace(void* base, size_t bytes, size_t n)
{
    if (external_) fail(kErrInvalidArgument, "workspace already swapped");
    if (n >= 0x7FFFFFFFull)
        fail(kErrInvalidArgument, "num_points must be below 2^31 per call");
    const size_t need = external_workspace_bytes(n);
    if (need > bytes)
        fail(kErrOutOfRange, "external workspace too small");
    const size_t cap = std::max<size_t>(n, 1024);
    Workspace w{};
    w.cap = cap;
    w.items_cap = cap + cap / kJobPoints + 2;
    w.sort_tmp_bytes = sort_scratch_bytes(cap);
    carve_workspace(static_cast<char*>(base), cap, prec_ == kPrecDouble ?
            sizeof(RecD) : sizeof(RecF), prec_ == kPrecDouble ?
            sizeof(double) : sizeof(float), w.sort_tmp_bytes, w.items_cap,
            w.keys, w.idx, &w.rec, &w.meta, &w.vq, &w.gw, &w.sort_tmp,
            &w.head, &w.bpos, &w.bstart, &w.scan_part, &w.items);
    saved_ = save_workspace();
    set_workspace(w);
    external_ = true;
}

void Plan::restore_workspace()
{
    if (!external_) return;
    set_workspace(saved_);
    external_ = false;
}

void Plan::free_workspace()
{
    for (int i = 0; i < 2; ++i)
    {
        dfree(d_keys_[i]);
        dfree(d_idx_[i]);
        d_keys_[i] = nullptr;
        d_idx_[i] = nullptr;
    }
    void** bufs[] = {&d_rec_, &d_vq_, &d_sort_tmp_};
    for (void** b : bufs)
    {
        dfree(*b);
        *b = nullptr;
    }
    dfree(d_meta_);
    dfree(d_gw_);
    dfree(d_head_);
    dfree(d_bpos_);
    dfree(d_bstart_);
    dfree(d_scan_part_);
    dfree(d_items_);
    d_meta_ = nullptr;
    d_gw_ = d_head_ = d_bpos_ = d_bstart_ = d_scan_part_ = nullptr;
    d_items_ = nullptr;
    cap_ = items_cap_ = 0;
    sort_tmp_bytes_ = 0;
}

void Plan::ensure_capacity(size_t n)
{
    // The planning kernels write the scan partials and the bucket table
    // even for zero points, so the workspace exists from the first call.
    if (cap_ > 0 && n <= cap_) return;
    if (n >= 0x7FFFFFFFull)
        fail(kErrInvalidArgument, "num_points must be below 2^31 per call");
    free_workspace();
    const size_t cap = std::max<size_t>(n, 1024);
    const size_t real_size = prec_ == kPrecDouble ? sizeof(double) :
            sizeof(float);
    for (int i = 0; i < 2; ++i)
    {
        d_keys_[i] = dmalloc<unsigned long long>(cap);
        d_idx_[i] = dmalloc<unsigned>(cap);
    }
    d_rec_ = dmalloc<char>(cap *
            (prec_ == kPrecDouble ? sizeof(RecD) : sizeof(RecF)));
    d_meta_ = dmalloc<int4>(cap);
    d_vq_ = dmalloc<char>(cap * 4 * real_size);
    d_gw_ = dmalloc<unsigned>(cap);
    d_head_ = dmalloc<unsigned>(cap);
    d_bpos_ = dmalloc<unsigned>(cap);
    d_bstart_ = dmalloc<unsigned>(cap + 1);
    d_scan_part_ = dmalloc<unsigned>(cap / kScanTile + 2);
    const size_t bytes = sort_scratch_bytes(cap);
    d_sort_tmp_ = dmalloc<char>(bytes);
    sort_tmp_bytes_ = bytes;
    // Jobs: one per tile plus one per kJobPoints points.
    items_cap_ = cap + cap / kJobPoints + 2;
    d_items_ = dmalloc<int4>(items_cap_);
    cap_ = cap;
}

// Exclusive scan of *count (<= max_n) u32 values on `stream`; *total gets
// their sum.  part holds max_n / kScanTile + 2 words.  Up to kScanOneTiles
// tiles one workgroup scans them in turn (one launch instead of three).
void scan_u32_enqueue(const unsigned* in, unsigned* out, const unsigned* count,
                      size_t max_n, unsigned* total, unsigned* part,
                      hipStream_t stream)
{
    const unsigned nb = (unsigned)std::max<size_t>(1,
            (max_n + kScanTile - 1) / kScanTile);
    if (nb <= kScanOneTiles)
    {
        hipLaunchKernelGGL(k_scan_one, dim3(1), dim3(kScanThreads), 0, stream,
                in, count, out, total);
        check_hip(hipGetLastError(), kErrKernelLaunch, "scan launch");
        return;
    }
    hipLaunchKernelGGL(k_scan_up, dim3(nb), dim3(kScanThreads), 0, stream,
            in, count, part);
    hipLaunchKernelGGL(k_scan_mid, dim3(1), dim3(kScanThreads), 0, stream,
            part, count, total);
    hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(kScanThreads), 0, stream,
            in, count, (const unsigned*)part, out);
    check_hip(hipGetLastError(), kErrKernelLaunch, "scan launch");
}

void Plan::scan_u32(const unsigned* in, unsigned* out, const unsigned* count,
                    size_t max_n, unsigned* total)
{
    scan_u32_enqueue(in, out, count, max_n, total, d_scan_part_, stream_);
}

template <typename T, typename TU, typename TV, typename Rec>
void Plan::locate_t(size_t n, const TU* uu, const TU* vv, const TU* ww,
                    const TV* vis, const DiagBuffers* diag)
{
    ensure_capacity(n);
    timed_ = true;
    check_hip(hipEventRecord(ev_[0], stream_), kErrKernelLaunch, "event");
    const int n_combo = (int)(P_ * K_ * K_);
    hipLaunchKernelGGL(k_init, dim3(std::max(1, std::min(1024,
            (n_combo + 255) / 256))), dim3(256), 0, stream_, d_summary_,
            d_combo_, (size_t)n_combo, 1);
    const int blocks = (int)std::max<size_t>(1, std::min<size_t>(
            (n + kLocateThreads - 1) / kLocateThreads, (size_t)n_cu_));
    const T gscale = sizeof(T) == 4 ? (T)gscale_f_ : (T)gscale_d_;
    const T wsc = sizeof(T) == 4 ? (T)ws_f_ : (T)ws_d_;
    if (diag)
        hipLaunchKernelGGL((k_locate<T, true, TU, TV, Rec>), dim3(blocks),
                dim3(kLocateThreads), 0,
                stream_, n, uu, vv, ww, vis, gscale, wsc, ov_, G_, (int)P_,
                d_support_, (const int*)d_srank_, srank_bits_, plane_bits_,
                h_, K_,
                tile_, tile_bits_, combo_bits_, key_bits_,
                d_keys_[0],
                (Rec*)d_rec_, d_combo_, d_summary_, d_gw_, diag->per_support,
                diag->per_plane, diag->plane_cap, 0ull);
    else
        hipLaunchKernelGGL((k_locate<T, false, TU, TV, Rec>), dim3(blocks),
                dim3(kLocateThreads), 0,
                stream_, n, uu, vv, ww, vis, gscale, wsc, ov_, G_, (int)P_,
                d_support_, (const int*)d_srank_, srank_bits_, plane_bits_,
                h_, K_,
                tile_, tile_bits_, combo_bits_, key_bits_,
                d_keys_[0],
                (Rec*)d_rec_, d_combo_, d_summary_, (unsigned*)nullptr,
                (unsigned long long*)nullptr, (unsigned long long*)nullptr,
                0, 0ull);
    check_hip(hipGetLastError(), kErrKernelLaunch, "k_locate launch");
    check_hip(hipEventRecord(ev_[1], stream_), kErrKernelLaunch, "event");
}

template <typename T, typename TapT, bool COMP>
void Plan::scatter_t(size_t n, T* grid, int row0, const DiagBuffers* diag)
{
    typedef typename Types<T>::Rec Rec;
    typedef typename Types<T>::T2 T2;
    typedef typename Types<T>::T4 T4;
    // Stable LSD sort of the keys (one more bit than the valid keys: the
    // skipped points' key); the result lands in pair `cur`, and the other
    // pair is free for the stages below.
    const size_t hw = radix_hist_words(n);
    unsigned* const hist = static_cast<unsigned*>(d_sort_tmp_);
    unsigned* const offs = hist + hw;
    unsigned long long* kk[2] = {d_keys_[0], d_keys_[1]};
    unsigned* vv[2] = {d_idx_[0], d_idx_[1]};
    const int cur = radix_sort_enqueue(n, key_bits_ + 1, kk, vv, hist, offs,
            offs + hw, stream_);
    const unsigned long long* const skeys = kk[cur];
    const unsigned* const sidx = vv[cur];
    check_hip(hipEventRecord(ev_[2], stream_), kErrKernelLaunch, "event");
    hipLaunchKernelGGL(k_plan, dim3(1), dim3(1024), 0, stream_, (int)P_,
            K_ * K_, (const unsigned*)d_combo_, (const PlaneInfo*)d_pinfo_,
            (const double*)d_normtab_, d_summary_);
    const int blocks = (int)std::max<size_t>(1, std::min<size_t>(
            (n + 255) / 256, (size_t)n_cu_ * 4));
    // Runs of equal keys -> one record each (not for the diagnostics, which
    // count points per cell): the run-start bits, their counts and the
    // counts' scan into the sort's spent value buffer, the runs' keys into
    // its spent key buffer.
    unsigned* const mh = vv[cur ^ 1];
    unsigned long long* const mkeys = kk[cur ^ 1];
    const size_t nwmax = n / 32 + 1;
    unsigned* const hb = mh;                 // run-start bits
    unsigned* const hc = hb + nwmax;         // their counts per word
    unsigned* const hp = hc + nwmax;         // the counts' exclusive scan
    hipLaunchKernelGGL(k_merge_heads, dim3(blocks), dim3(256), 0, stream_,
            d_summary_, skeys, hb, hc, diag ? 0 : 1);
    scan_u32(hc, hp, &d_summary_->nw32, nwmax, &d_summary_->nm32);
    // Run starts into the bucket table (free until the tile scan below).
    hipLaunchKernelGGL(k_run_starts, dim3(blocks), dim3(256), 0, stream_,
            (const Summary*)d_summary_, (const unsigned*)hb,
            (const unsigned*)hp, d_bstart_);
    unsigned* const tb = d_head_;            // tile-start bits
    unsigned* const tc = tb + nwmax;         // their counts per word
    hipLaunchKernelGGL((k_gather<T>), dim3(blocks), dim3(256), 0, stream_,
            d_summary_, skeys, sidx, (const Rec*)d_rec_, (const unsigned*)(diag ? d_gw_ : nullptr),
            (const PlaneInfo*)d_pinfo_, h_, K_, tile_shift_, d_meta_,
            (T4*)d_vq_, tb, tc, diag ? 0 : region_, tile_,
            combo_bits_ + plane_bits_, (const unsigned*)d_bstart_, mkeys);
    check_hip(hipGetLastError(), kErrKernelLaunch, "k_gather launch");
    // Tiles: exclusive scan of the tile-start bits' counts, then the
    // tiles' first records.
    scan_u32(tc, d_bpos_, &d_summary_->nwt32, nwmax, &d_summary_->n_buckets);
    hipLaunchKernelGGL(k_buckets, dim3(blocks), dim3(256), 0, stream_,
            (const Summary*)d_summary_, (const unsigned*)tb,
            (const unsigned*)d_bpos_, d_bstart_);
    // Jobs per tile (reusing head / bpos), scanned to offsets.
    hipLaunchKernelGGL(k_bucket_items, dim3(blocks), dim3(256), 0, stream_,
            (const Summary*)d_summary_, (const unsigned*)d_bstart_,
            (unsigned)kJobPoints, d_head_);
    scan_u32(d_head_, d_bpos_, &d_summary_->n_buckets,
            std::min<size_t>(n, (size_t)1 << (2 * tile_coord_bits_)),
            &d_summary_->total_items);
    hipLaunchKernelGGL(k_write_items, dim3(blocks), dim3(256), 0, stream_,
            d_summary_,
            (const unsigned long long*)mkeys, (const unsigned*)d_bstart_,
            (const unsigned*)d_bpos_, tile_, tile_shift_,
            (unsigned)kJobPoints, (unsigned)items_cap_, d_items_);
    // Each item's support segments, into the run-head buffer (spent).
    unsigned* const segs = mh;
    hipLaunchKernelGGL(k_item_segments, dim3(blocks), dim3(256), 0, stream_,
            (const Summary*)d_summary_, (unsigned)items_cap_, d_items_,
            (const int4*)d_meta_, segs);
    check_hip(hipGetLastError(), kErrKernelLaunch, "job launch");
    check_hip(hipEventRecord(ev_[3], stream_), kErrKernelLaunch, "event");
    if (diag)
        hipLaunchKernelGGL((k_grid<T, true>), dim3(grid_blocks_),
                dim3(kGridThreads), 0, stream_,
                d_summary_, (const int4*)d_items_,
                (const PlaneInfo*)d_pinfo_, (const int4*)d_meta_,
                (const T4*)d_vq_, (const T2*)d_etab_, G_, row0, 0, grid, 1,
                min_share_, diag->per_cell, diag->wsum_cell, xcds_,
                (const unsigned*)segs);
    else
        hipLaunchKernelGGL((k_grid<T, false, TapT, COMP>), dim3(grid_blocks_),
                dim3(kGridThreads), 0, stream_,
                d_summary_, (const int4*)d_items_,
                (const PlaneInfo*)d_pinfo_, (const int4*)d_meta_,
                (const T4*)d_vq_, (const TapT*)d_etab_, G_, row0, region_, grid,
                0, min_share_, (unsigned*)nullptr,
                (unsigned long long*)nullptr, xcds_, (const unsigned*)segs);
    check_hip(hipGetLastError(), kErrKernelLaunch, "k_grid launch");
    check_hip(hipEventRecord(ev_[4], stream_), kErrKernelLaunch, "event");
}

void Plan::locate(size_t n, const void* uu, const void* vv, const void* ww,
                  const void* vis, const DiagBuffers* diag)
{
    if (prec_ == kPrecSingle && index64_)
    {
        if (diag)
            fail(kErrInvalidArgument, "diagnostics need a GW_SINGLE plan");
        locate_t<double, float, float, RecF>(n, (const float*)uu,
                (const float*)vv, (const float*)ww, (const float*)vis,
                nullptr);
    }
    else if (prec_ == kPrecSingle)
        locate_t<float>(n, (const float*)uu, (const float*)vv,
                (const float*)ww, (const float*)vis, diag);
    else if (prec_ == kPrecDouble)
        locate_t<double>(n, (const double*)uu, (const double*)vv,
                (const double*)ww, (const double*)vis, diag);
    else
    {
        if (diag)
            fail(kErrInvalidArgument, "diagnostics need a GW_SINGLE plan");
        if (prec_ == kPrecHalf)
            locate_t<double, float, half_t, RecF>(n, (const float*)uu,
                    (const float*)vv, (const float*)ww, (const half_t*)vis,
                    nullptr);
        else
            locate_t<double, float, bf16_t, RecF>(n, (const float*)uu,
                    (const float*)vv, (const float*)ww, (const bf16_t*)vis,
                    nullptr);
    }
}

void Plan::scatter(size_t n, void* grid, int row0, const DiagBuffers* diag)
{
    if (prec_ == kPrecSingle)
        scatter_t<float>(n, (float*)grid, row0, diag);
    else if (prec_ == kPrecDouble)
        scatter_t<double>(n, (double*)grid, row0, diag);
    else if (prec_ == kPrecHalf)
        scatter_t<float, HalfPair, true>(n, (float*)grid, row0, nullptr);
    else
        scatter_t<float, BF16Pair, true>(n, (float*)grid, row0, nullptr);
}

const Summary& Plan::read_summary()
{
    check_hip(hipMemcpyAsync(h_summary_, d_summary_, sizeof(Summary),
            hipMemcpyDeviceToHost, stream_), kErrMemCopy, "read summary");
    check_hip(hipStreamSynchronize(stream_), kErrKernelLaunch,
            "stream synchronize");
    return *h_summary_;
}

int Plan::timings(double* ms, int max_n)
{
    if (!timed_ || max_n <= 0) return 0;
    check_hip(hipEventSynchronize(ev_[4]), kErrKernelLaunch, "event sync");
    float t[5] = {0, 0, 0, 0, 0};
    check_hip(hipEventElapsedTime(&t[0], ev_[0], ev_[1]), kErrKernelLaunch,
            "elapsed");
    check_hip(hipEventElapsedTime(&t[1], ev_[1], ev_[2]), kErrKernelLaunch,
            "elapsed");
    check_hip(hipEventElapsedTime(&t[2], ev_[2], ev_[3]), kErrKernelLaunch,
            "elapsed");
    check_hip(hipEventElapsedTime(&t[3], ev_[3], ev_[4]), kErrKernelLaunch,
            "elapsed");
    check_hip(hipEventElapsedTime(&t[4], ev_[0], ev_[4]), kErrKernelLaunch,
            "elapsed");
    const int n = std::min(max_n, 5);
    for (int i = 0; i < n; ++i) ms[i] = t[i];
    return n;
}

// -------------------------------------------------------------------------
// Batched jobs: one pipeline, one persistent scatter launch
// -------------------------------------------------------------------------
//
// Point keys carry their job index in bits 48-63 (the plan's own key below,
// < 2^48).  The global tile index of a point is its job's tile_base plus
// its plan-level tile (the key bits above tile_shift).

constexpr int kBatchJobShift = 48;
static_assert(offsetof(BatchJob, etab) == 8 && offsetof(BatchJob, grid) == 16
        && offsetof(BatchJob, G) == 24 && offsetof(BatchJob, R) == 28 &&
        offsetof(BatchJob, min_share) == 44, "BatchJob word layout");
static_assert(sizeof(PlaneInfo) == 32 && offsetof(PlaneInfo, eoff) == 24 &&
        offsetof(PlaneInfo, srank) == 28,
        "PlaneInfo word layout");
constexpr unsigned long long kBatchLocalMask = (1ull << kBatchJobShift) - 1;

// ------------------------------------------------------------------------
// Batch ordering.  Every point's key becomes (global tile, in-tile key):
//   (tile_base + tile) << lmax | (local key & in-tile bits)
// with lmax the widest in-tile key of the batch; skipped points get
// tiles_total << lmax and sort last.  A stable LSD radix sort of these keys
// (8-bit digits, as many passes as the key has bits) gives each tile's
// points in the order the single-job path sorts them (plane, position in
// the tile, sub-cell table), so a batch job's work items are the same as
// the job's own.  The sort is built from plain kernels and the scan above
// (no memset, no host round trip), so the sequence can be captured.
// ------------------------------------------------------------------------

constexpr int kRadixThreads = 512;
constexpr int kRadixKeys = 8;                   // keys per lane
constexpr int kRadixWaves = kRadixThreads / 64;
constexpr int kRadixChunk = kRadixThreads * kRadixKeys;
static_assert(kRadixChunk == 4096, "chunk of the histogram layout");

__global__ void __launch_bounds__(256) k_batch_key(size_t n,
        const unsigned long long* __restrict__ keys,
        const BatchJob* __restrict__ jobs, int lmax,
        unsigned long long sentinel, unsigned long long* __restrict__ out,
        Summary* __restrict__ bsum)
{
    if (blockIdx.x == 0 && threadIdx.x == 0)
    {
        bsum->nv32 = 0;
        for (int q = 0; q < kXcdGroups; ++q) bsum->item_next[q] = 0;
    }
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
    {
        const unsigned long long key = keys[i];
        const BatchJob& d = jobs[key >> kBatchJobShift];
        const unsigned long long local = key & kBatchLocalMask;
        unsigned long long c = sentinel;
        if (!(local >> d.key_bits))
            c = ((unsigned long long)(d.tile_base +
                    (unsigned)(local >> d.tile_shift)) << lmax) |
                (local & ((1ull << d.tile_shift) - 1));
        out[i] = c;
    }
}

// Digit histogram of each chunk of kRadixChunk keys, digit-major:
// hist[digit * nblocks + block].  Wave w counts keys [w * 512, w * 512 +
// 512) of the chunk into its own LDS histogram (all its loads first); lanes
// with equal digits are found by ballots and counted by one add, since
// neighbouring keys often share a digit (the high digits of sorted keys).
__global__ void __launch_bounds__(kRadixThreads) k_radix_hist(size_t n,
        const unsigned long long* __restrict__ keys, int shift,
        unsigned nblocks, unsigned* __restrict__ hist)
{
    __shared__ unsigned h[kRadixWaves][256];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int w = tid >> 6;
    for (int i = tid; i < kRadixWaves * 256; i += kRadixThreads)
        (&h[0][0])[i] = 0;
    const size_t base = (size_t)blockIdx.x * kRadixChunk +
            (size_t)w * (64 * kRadixKeys) + lane;
    unsigned long long k[kRadixKeys];
#pragma unroll
    for (int j = 0; j < kRadixKeys; ++j)
    {
        const size_t i = base + (size_t)j * 64;
        k[j] = i < n ? keys[i] : 0ull;
    }
    __syncthreads();
    const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll
    for (int j = 0; j < kRadixKeys; ++j)
    {
        const bool ok = base + (size_t)j * 64 < n;
        const unsigned d = (unsigned)(k[j] >> shift) & 255u;
        unsigned long long peers = __ballot(ok);
#pragma unroll
        for (int b = 0; b < 8; ++b)
        {
            const unsigned long long m = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? m : ~m;
        }
        if (ok && (peers & below) == 0)
            h[w][d] += (unsigned)__popcll(peers);
    }
    __syncthreads();
    if (tid < 256)
    {
        unsigned c = 0;
#pragma unroll
        for (int q = 0; q < kRadixWaves; ++q) c += h[q][tid];
        hist[(size_t)tid * nblocks + blockIdx.x] = c;
    }
}

// Stable scatter of each chunk by digit; offs[d * nblocks + b] is the
// exclusive scan of digit d's chunk counts over the chunks and tot[d] the
// digit's total, so the digit's keys start at the exclusive scan of tot.
// Wave w ranks keys [w * 512, w * 512 + 512) of the chunk in order
// (lanes with equal digits found by ballots, a running count per digit in
// the wave's own LDS row), the waves' counts give each wave's start per
// digit inside the chunk, the chunk is staged in LDS in digit order, and
// each digit's run is written out contiguously at its global offset.
__global__ void __launch_bounds__(kRadixThreads) k_radix_scatter(size_t n,
        const unsigned long long* __restrict__ keys,
        const unsigned* __restrict__ vals, int shift, unsigned nblocks,
        const unsigned* __restrict__ offs, const unsigned* __restrict__ tot,
        unsigned long long* __restrict__ keys_out,
        unsigned* __restrict__ vals_out)
{
    __shared__ unsigned long long sk[kRadixChunk];
    __shared__ unsigned sv[kRadixChunk];
    __shared__ unsigned wh[kRadixWaves][256];
    __shared__ unsigned gof[256];
    __shared__ unsigned wsum[4];
    __shared__ unsigned tsum[4];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int w = tid >> 6;
    const size_t c0 = (size_t)blockIdx.x * kRadixChunk;
    const size_t base = c0 + (size_t)w * (64 * kRadixKeys) + lane;
    for (int i = tid; i < kRadixWaves * 256; i += kRadixThreads)
        (&wh[0][0])[i] = 0;
    unsigned long long k[kRadixKeys];
    unsigned v[kRadixKeys], dg[kRadixKeys], rk[kRadixKeys];
#pragma unroll
    for (int j = 0; j < kRadixKeys; ++j)
    {
        const size_t i = base + (size_t)j * 64;
        k[j] = i < n ? keys[i] : 0ull;
        v[j] = i < n ? (vals ? vals[i] : (unsigned)i) : 0u;
    }
    __syncthreads();
    const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll
    for (int j = 0; j < kRadixKeys; ++j)
    {
        const bool ok = base + (size_t)j * 64 < n;
        const unsigned d = (unsigned)(k[j] >> shift) & 255u;
        unsigned long long peers = __ballot(ok);
#pragma unroll
        for (int b = 0; b < 8; ++b)
        {
            const unsigned long long m = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? m : ~m;
        }
        dg[j] = d;
        rk[j] = 0;
        if (ok)
        {
            // This wave's keys of digit d before this one: the running
            // count, then the lanes below with the same digit; the lowest
            // such lane advances the count.
            const unsigned pre = wh[w][d];
            rk[j] = pre + (unsigned)__popcll(peers & below);
            if ((peers & below) == 0)
                wh[w][d] = pre + (unsigned)__popcll(peers);
        }
    }
    __syncthreads();
    // Digit tid: the waves' starts (exclusive over waves), the chunk's
    // start of the digit (exclusive over digits), and the digit's start in
    // the output (exclusive scan of the totals) plus this chunk's offset.
    unsigned acc = 0, inc = 0, t = 0, tinc = 0;
    if (tid < 256)
    {
#pragma unroll
        for (int q = 0; q < kRadixWaves; ++q)
        {
            const unsigned c = wh[q][tid];
            wh[q][tid] = acc;
            acc += c;
        }
        inc = acc;
        t = tot[tid];
        tinc = t;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1)
        {
            const unsigned y = __shfl_up(inc, o);
            const unsigned z = __shfl_up(tinc, o);
            if (lane >= o)
            {
                inc += y;
                tinc += z;
            }
        }
        if (lane == 63)
        {
            wsum[w] = inc;
            tsum[w] = tinc;
        }
    }
    __syncthreads();
    if (tid < 256)
    {
        unsigned ls = inc - acc, ds = tinc - t;
        for (int q = 0; q < w; ++q)
        {
            ls += wsum[q];
            ds += tsum[q];
        }
#pragma unroll
        for (int q = 0; q < kRadixWaves; ++q) wh[q][tid] += ls;
        gof[tid] = ds + offs[(size_t)tid * nblocks + blockIdx.x] - ls;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kRadixKeys; ++j)
        if (base + (size_t)j * 64 < n)
        {
            const unsigned p = wh[w][dg[j]] + rk[j];
            sk[p] = k[j];
            sv[p] = v[j];
        }
    __syncthreads();
    const unsigned cnt = (unsigned)min((size_t)kRadixChunk, n - c0);
    for (unsigned p = tid; p < cnt; p += kRadixThreads)
    {
        const unsigned long long kk = sk[p];
        const unsigned g = gof[(unsigned)(kk >> shift) & 255u] + p;
        keys_out[g] = kk;
        vals_out[g] = sv[p];
    }
}

size_t radix_hist_words(size_t n)
{
    return 256 * ((n + kRadixChunk - 1) / kRadixChunk);
}

// Digit d per workgroup: exclusive scan of its chunk counts hist[d * nb +
// b] over the chunks b into offs, its total into tot[d].
__global__ void __launch_bounds__(kScanThreads) k_radix_scan(
        const unsigned* __restrict__ hist, unsigned nb,
        unsigned* __restrict__ offs, unsigned* __restrict__ tot)
{
    __shared__ unsigned s_tmp[17];
    const unsigned* const h = hist + (size_t)blockIdx.x * nb;
    unsigned* const o = offs + (size_t)blockIdx.x * nb;
    unsigned carry = 0;
    for (unsigned t0 = 0; t0 < nb; t0 += kScanTile)
    {
        const unsigned base = t0 + threadIdx.x * kScanItems;
        unsigned v[kScanItems];
        unsigned sum = 0;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k)
        {
            v[k] = base + k < nb ? h[base + k] : 0u;
            sum += v[k];
        }
        unsigned total;
        unsigned run = block_exclusive_scan(sum, s_tmp, &total) + carry;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k)
        {
            if (base + k < nb) o[base + k] = run;
            run += v[k];
        }
        carry += total;
    }
    if (threadIdx.x == 0) tot[blockIdx.x] = carry;
}

int radix_sort_enqueue(size_t n, int bits, unsigned long long* keys[2],
                       unsigned* vals[2], unsigned* hist, unsigned* offs,
                       unsigned* tot, hipStream_t s)
{
    const unsigned nchunk = (unsigned)((n + kRadixChunk - 1) / kRadixChunk);
    if (nchunk == 0) return 0;
    int cur = 0;
    for (int shift = 0; shift < bits; shift += 8)
    {
        hipLaunchKernelGGL(k_radix_hist, dim3(nchunk), dim3(kRadixThreads),
                0, s, n, (const unsigned long long*)keys[cur], shift, nchunk,
                hist);
        hipLaunchKernelGGL(k_radix_scan, dim3(256), dim3(kScanThreads), 0, s,
                (const unsigned*)hist, nchunk, offs, tot);
        // The first pass reads no values: they start as 0..n-1.
        hipLaunchKernelGGL(k_radix_scatter, dim3(nchunk),
                dim3(kRadixThreads), 0, s, n,
                (const unsigned long long*)keys[cur],
                shift == 0 ? (const unsigned*)nullptr :
                (const unsigned*)vals[cur], shift, nchunk,
                (const unsigned*)offs, (const unsigned*)tot, keys[cur ^ 1],
                vals[cur ^ 1]);
        cur ^= 1;
    }
    check_hip(hipGetLastError(), kErrKernelLaunch, "radix sort launch");
    return cur;
}

// Tile ranges of the sorted keys: tstart[t] = first point, tend[t] = one
// past the last (both zero for empty tiles); bsum->nv32 = points gridded.
// Sorted points gridded (all but the skipped ones, whose keys carry the
// sentinel tile and sort last) -> bsum->nv32.
__global__ void __launch_bounds__(256) k_batch_count(size_t n,
        const unsigned long long* __restrict__ keys, int lmax,
        unsigned sentinel_tile, Summary* __restrict__ bsum)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
    {
        if ((unsigned)(keys[i] >> lmax) == sentinel_tile) continue;
        if (i + 1 == n || (unsigned)(keys[i + 1] >> lmax) == sentinel_tile)
            bsum->nv32 = (unsigned)(i + 1);
    }
}

// Record range [tstart[t], tend[t]) of every global tile t that has
// records (the bsum->nm32 run records, in key order).
__global__ void __launch_bounds__(256) k_tile_bounds(
        const unsigned long long* __restrict__ keys, int lmax,
        unsigned* __restrict__ tstart, unsigned* __restrict__ tend,
        const Summary* __restrict__ bsum)
{
    const unsigned nm = bsum->nm32;
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < nm;
         i += gridDim.x * blockDim.x)
    {
        const unsigned t = (unsigned)(keys[i] >> lmax);
        if (i == 0 || (unsigned)(keys[i - 1] >> lmax) != t) tstart[t] = i;
        if (i + 1 == nm || (unsigned)(keys[i + 1] >> lmax) != t)
            tend[t] = i + 1;
    }
}

// Work items per tile: ceil(points / kJobPoints).
__global__ void __launch_bounds__(256) k_tile_items(
        const unsigned* __restrict__ ntiles, const unsigned* __restrict__ tstart,
        const unsigned* __restrict__ tend, unsigned* __restrict__ icnt)
{
    const unsigned nt = *ntiles;
    for (unsigned t = blockIdx.x * blockDim.x + threadIdx.x; t < nt;
         t += gridDim.x * blockDim.x)
        icnt[t] = (tend[t] - tstart[t] + kJobPoints - 1) / kJobPoints;
}

// Job owning global tile t: the last with tile_base <= t.
__device__ __forceinline__ int batch_job_of_tile(const BatchJob* jobs,
                                                 int njobs, unsigned t)
{
    int lo = 0, hi = njobs - 1;
    while (lo < hi)
    {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].tile_base <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Work items = (tile origin u | v << 16, first sorted point, points, job).
__global__ void __launch_bounds__(256) k_write_tile_items(unsigned ntiles,
        const unsigned* __restrict__ tstart, const unsigned* __restrict__ tend,
        const unsigned* __restrict__ ioff, const BatchJob* __restrict__ jobs,
        int njobs, unsigned items_cap, int4* __restrict__ items)
{
    for (unsigned t = blockIdx.x * blockDim.x + threadIdx.x; t < ntiles;
         t += gridDim.x * blockDim.x)
    {
        const unsigned c = tend[t] - tstart[t];
        if (c == 0) continue;
        const int j = batch_job_of_tile(jobs, njobs, t);
        const BatchJob& d = jobs[j];
        const unsigned lt = t - d.tile_base;
        const unsigned ux0 = compact_bits16(lt) * (unsigned)d.tile;
        const unsigned vy0 = compact_bits16((unsigned long long)lt >> 1) *
                (unsigned)d.tile;
        const int origin = (int)(ux0 | (vy0 << 16));
        unsigned o = ioff[t];
        for (unsigned first = 0; first < c; first += kJobPoints, ++o)
            if (o < items_cap)
                items[o] = make_int4(origin, (int)(tstart[t] + first),
                        (int)min((unsigned)kJobPoints, c - first), j);
    }
}

// Run records of all jobs (as k_gather, with each record's job): run m =
// sorted points [rstart[m], rstart[m + 1]), its visibility terms summed in
// FP64; mkeys[m] = the run's key.
__global__ void __launch_bounds__(256) k_gather_batch(
        const Summary* __restrict__ bsum,
        const unsigned long long* __restrict__ keys,
        const unsigned* __restrict__ idx, const RecF* __restrict__ rec,
        const BatchJob* __restrict__ jobs, int njobs, int lmax,
        int4* __restrict__ meta, float4* __restrict__ vq,
        const unsigned* __restrict__ rstart,
        unsigned long long* __restrict__ mkeys)
{
    const unsigned nm = bsum->nm32;
    for (unsigned m = blockIdx.x * blockDim.x + threadIdx.x; m < nm;
         m += gridDim.x * blockDim.x)
    {
        const unsigned s = rstart[m], e = rstart[m + 1];
        const unsigned long long key = keys[s];
        mkeys[m] = key;
        const int jd = batch_job_of_tile(jobs, njobs,
                (unsigned)(key >> lmax));
        const BatchJob& d = jobs[jd];
        const RecF r = rec[idx[s]];
        const int plane = (unsigned)r.sub >> 17;
        const int off_u = (r.sub & 0xFF) - 128;
        const int off_v = ((r.sub >> 8) & 0xFF) - 128;
        const PlaneInfo& pi = d.pinfo[plane];
        const unsigned tref = table_ref(pi, off_u, off_v, d.sub_h);
        // From the previous sorted key when it is of the same global tile
        // (hence job): below lmax the job's own key bits without the tile
        // (support rank, position, plane, sub-cell table).
        unsigned flags = kNewPos | kNewTab;
        if (s > 0)
        {
            const unsigned long long dk = keys[s - 1] ^ key;
            if ((dk >> d.pos_shift) == 0) flags &= ~kNewPos;
            if ((dk & ((1ull << d.pos_shift) - 1)) == 0) flags &= ~kNewTab;
        }
        meta[m] = make_int4(region_offset(r.pos, d.R, d.tile), (int)tref,
                (int)((unsigned)plane | ((unsigned)pi.srank << 16) | flags),
                0);
        double q0 = 0.0, q1 = 0.0, q2 = 0.0, q3 = 0.0;
        auto add = [&](const RecF& p) {
            const double sgn = ((p.sub >> 16) & 1) ? -1.0 : 1.0;
            q0 += (double)p.re;
            q1 += -sgn * (double)p.im;
            q2 += (double)p.im;
            q3 += sgn * (double)p.re;
        };
        add(r);
        unsigned t = s + 1;
        for (; t + 4 <= e; t += 4)
        {
            const unsigned i0 = idx[t], i1 = idx[t + 1], i2 = idx[t + 2],
                    i3 = idx[t + 3];
            const RecF p0 = rec[i0], p1 = rec[i1], p2 = rec[i2],
                    p3 = rec[i3];
            add(p0);
            add(p1);
            add(p2);
            add(p3);
        }
        for (; t < e; ++t) add(rec[idx[t]]);
        vq[m] = make_float4((float)q0, (float)q1, (float)q2, (float)q3);
    }
}

// The scatter of every job's work items, one item per workgroup at a time
// (persistent); each item takes its job's tables, grid and region side.
__global__ void __launch_bounds__(kGridThreads) k_grid_batch(
        Summary* __restrict__ bsum, const int4* __restrict__ items,
        const BatchJob* __restrict__ jobs, const int4* __restrict__ meta,
        const float4* __restrict__ vq, int xcds,
        const unsigned* __restrict__ segs)
{
    __shared__ GridLds L;
    __shared__ unsigned s_next[3];
    for (int i = threadIdx.x; i < kRegionCells; i += blockDim.x)
    {
        L.re[i] = 0.0;
        L.im[i] = 0.0;
    }
    __syncthreads();
    ScatterCtx<float> x;
    x.meta = meta;
    x.vq = vq;
    x.segs = segs;
    x.lds = &L;
    x.diag_count = nullptr;
    x.diag_wsum = nullptr;
    x.row0 = 0;
    for_each_item(bsum, xcds, s_next, [&](unsigned it, auto&& pre) {
        const int4 item = items[it];
        // BatchJob fields by word: pinfo, etab, grid (2 words each), G, R,
        // tile_bits, tile_shift, key_bits, min_share, ...
        const ConstPtr<unsigned> w = as_const<unsigned>(jobs + item.w);
        auto ptr_at = [&](int k) {
            return (uintptr_t)w[k] | ((uintptr_t)w[k + 1] << 32);
        };
        const PlaneInfo* pinfo = reinterpret_cast<const PlaneInfo*>(ptr_at(0));
        x.etab = reinterpret_cast<const void*>(ptr_at(2));
        x.grid = reinterpret_cast<float*>(ptr_at(4));
        x.G = (int)w[6];
        x.R = (int)w[7];
        const int min_share = (int)w[11];
        grid_work_item<float, false, float2, false>(L, x, pinfo, item,
                x.R > 0, min_share, pre);
    });
}

Batch::Batch()
{
    int dev = 0;
    check_hip(hipGetDevice(&dev), kErrComputeDevices, "hipGetDevice");
    check_hip(hipDeviceGetAttribute(&n_cu_,
            hipDeviceAttributeMultiprocessorCount, dev), kErrComputeDevices,
            "hipDeviceGetAttribute");
    int occ = 0;
    check_hip(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ,
            (const void*)k_grid_batch, kGridThreads, 0), kErrComputeDevices,
            "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    grid_blocks_ = n_cu_ * std::max(1, occ);
    check_hip(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking),
            kErrComputeDevices, "hipStreamCreateWithFlags");
    own_stream_ = true;
    for (auto& e : ev_)
        check_hip(hipEventCreate(&e), kErrComputeDevices, "hipEventCreate");
    d_bsum_ = dmalloc<Summary>(1);
    d_ntiles_ = dmalloc<unsigned>(1);
    check_hip(hipHostMalloc((void**)&h_ntiles_, sizeof(unsigned),
            hipHostMallocDefault), kErrMemAlloc, "hipHostMalloc");
    check_hip(hipEventCreateWithFlags(&ev_desc_, hipEventDisableTiming),
            kErrComputeDevices, "hipEventCreate");
}

Batch::~Batch()
{
    if (stream_) (void)hipStreamSynchronize(stream_);
    drop_graph();
    free_buffers();
    dfree(d_bsum_);
    dfree(d_ntiles_);
    if (h_desc_) (void)hipHostFree(h_desc_);
    if (h_ntiles_) (void)hipHostFree(h_ntiles_);
    if (ev_desc_) (void)hipEventDestroy(ev_desc_);
    for (auto& e : ev_)
        if (e) (void)hipEventDestroy(e);
    if (own_stream_ && stream_) (void)hipStreamDestroy(stream_);
}

void Batch::drop_graph()
{
    // A replay of the graph may still run on the stream.
    if (exec_ && stream_) (void)hipStreamSynchronize(stream_);
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    exec_ = nullptr;
    graph_ = nullptr;
}

void Batch::free_buffers()
{
    for (void* p : {(void*)d_desc_, (void*)d_sum_, (void*)d_combo_,
                    (void*)d_keys_[0], (void*)d_keys_[1], (void*)d_idx_[0],
                    (void*)d_idx_[1], d_rec_, (void*)d_meta_, d_vq_,
                    (void*)d_tstart_, (void*)d_tend_, (void*)d_icnt_,
                    (void*)d_ioff_, (void*)d_part_, (void*)d_rhist_,
                    (void*)d_roff_, (void*)d_items_,
                    (void*)d_rstart_})
        dfree(p);
    if (h_sum_) (void)hipHostFree(h_sum_);
    d_desc_ = nullptr;
    d_sum_ = nullptr;
    d_combo_ = nullptr;
    d_keys_[0] = d_keys_[1] = nullptr;
    d_idx_[0] = d_idx_[1] = nullptr;
    d_rec_ = nullptr;
    d_meta_ = nullptr;
    d_vq_ = nullptr;
    d_tstart_ = d_tend_ = d_icnt_ = d_ioff_ = d_part_ = nullptr;
    d_rhist_ = d_roff_ = nullptr;
    d_rstart_ = nullptr;
    d_items_ = nullptr;
    h_sum_ = nullptr;
    cap_points_ = cap_tiles_ = cap_combo_ = cap_jobs_ = 0;
}

int Batch::add(Plan* plan, size_t n, const float* uu, const float* vv,
               const float* ww, const float* vis, float* grid)
{
    if (!plan) fail(kErrInvalidArgument, "plan is NULL");
    if (plan->precision() != kPrecSingle)
        fail(kErrInvalidArgument, "batched jobs need GW_SINGLE plans");
    if (!grid || (n && (!uu || !vv || !ww || !vis)))
        fail(kErrInvalidArgument, "device pointers must not be NULL");
    if (plan->key_bits_max() + 1 > kBatchJobShift)
        fail(kErrInvalidArgument, "grid too large for a batched job");
    if (jobs_.size() >= 65535)
        fail(kErrInvalidArgument, "a batch holds at most 65535 jobs");
    jobs_.push_back(Entry{plan, n, uu, vv, ww, vis, grid});
    dirty_ = true;
    return (int)jobs_.size() - 1;
}

void Batch::clear()
{
    jobs_.clear();
    dirty_ = true;
}

void Batch::set_stream(hipStream_t s)
{
    if (!s || s == stream_) return;
    // Work still queued on the previous stream (a run or a graph replay)
    // reads the job table and the workspace that the next run on s
    // rewrites: finish it first, whoever owns that stream.
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (own_stream_ && stream_) (void)hipStreamDestroy(stream_);
    stream_ = s;
    own_stream_ = false;
    dirty_ = true;
}

void Batch::prepare()
{
    const size_t J = jobs_.size();
    desc_.resize(J);
    combo_off_.resize(J + 1);
    size_t pts = 0, tiles = 0, combos = 0;
    for (size_t j = 0; j < J; ++j)
    {
        const Entry& e = jobs_[j];
        const Plan& pl = *e.plan;
        BatchJob& d = desc_[j];
        d.pinfo = pl.device_pinfo();
        d.etab = pl.device_etab();
        d.grid = e.grid;
        d.G = pl.grid_size();
        d.R = pl.region();
        d.tile = pl.tile();
        d.planes = (int)pl.num_w_planes();
        d.normtab = pl.device_normtab();
        d.combo_off = (unsigned)combos;
        d.tile_bits = pl.tile_bits();
        d.tile_shift = pl.tile_shift();
        d.key_bits = pl.key_bits();
        d.min_share = pl.min_share();
        d.sub_h = pl.half_sub();
        d.sub_k = pl.sub_tables();
        d.pos_shift = pl.tile_shift() - 2 * d.tile_bits - pl.srank_bits();
        d.tile_base = (unsigned)tiles;
        d.ntiles = 1u << (2 * pl.tile_coord_bits());
        d.point_base = (unsigned)pts;
        d.n = (unsigned)e.n;
        combo_off_[j] = combos;
        pts += e.n;
        tiles += d.ntiles;
        combos += pl.num_w_planes() * (size_t)pl.sub_tables() *
                pl.sub_tables();
        if (pts >= 0x7FFFFFFFull || tiles >= 0x7FFFFFFFull)
            fail(kErrInvalidArgument, "batch too large (2^31 points or "
                 "tiles)");
    }
    combo_off_[J] = combos;
    n_total_ = pts;
    tiles_total_ = tiles;
    combo_total_ = combos;
    // Batch key: global tile above the widest in-tile key; the sentinel
    // tile (skipped points) is tiles_total.
    lmax_ = 0;
    for (const BatchJob& d : desc_) lmax_ = std::max(lmax_, d.tile_shift);
    int tbits = 0;
    while ((tiles >> tbits) != 0) ++tbits;
    radix_bits_ = lmax_ + tbits;
    if (radix_bits_ > 64)
        fail(kErrInvalidArgument, "batch key exceeds 64 bits");
    const size_t pcap = std::max<size_t>(pts, 1024);
    items_cap_ = pcap + pcap / kJobPoints + 2;
    // Allocated on first use even for an empty batch: the scans write
    // their partial sums whatever the counts.
    if (!d_keys_[0] || pts > cap_points_ || tiles > cap_tiles_ ||
            combos > cap_combo_ || J > cap_jobs_)
    {
        // The previous run may still use the buffers being replaced.
        check_hip(hipStreamSynchronize(stream_), kErrKernelLaunch,
                "stream synchronize");
        drop_graph();
        free_buffers();
        cap_points_ = pcap;
        cap_tiles_ = std::max<size_t>(tiles, 1);
        cap_combo_ = std::max<size_t>(combos, 1);
        cap_jobs_ = std::max<size_t>(J, 1);
        for (int i = 0; i < 2; ++i)
        {
            d_keys_[i] = dmalloc<unsigned long long>(cap_points_);
            d_idx_[i] = dmalloc<unsigned>(cap_points_);
        }
        d_rec_ = dmalloc<RecF>(cap_points_);
        d_meta_ = dmalloc<int4>(cap_points_);
        d_vq_ = dmalloc<float4>(cap_points_);
        d_tstart_ = dmalloc<unsigned>(cap_tiles_);
        d_tend_ = dmalloc<unsigned>(cap_tiles_);
        d_rstart_ = dmalloc<unsigned>(cap_points_ + 1);
        const size_t rh = radix_hist_words(cap_points_);
        d_rhist_ = dmalloc<unsigned>(rh);
        d_roff_ = dmalloc<unsigned>(rh);
        d_icnt_ = dmalloc<unsigned>(cap_tiles_);
        d_ioff_ = dmalloc<unsigned>(cap_tiles_);
        // Scan partial sums; also the sort's 256 digit totals.
        d_part_ = dmalloc<unsigned>(std::max<size_t>(256,
                std::max(cap_points_, cap_tiles_) / kScanTile + 2));
        d_items_ = dmalloc<int4>(cap_points_ + cap_points_ / kJobPoints + 2);
        d_desc_ = dmalloc<BatchJob>(cap_jobs_);
        d_sum_ = dmalloc<Summary>(cap_jobs_);
        d_combo_ = dmalloc<unsigned>(cap_combo_);
        check_hip(hipHostMalloc((void**)&h_sum_, cap_jobs_ * sizeof(Summary),
                hipHostMallocDefault), kErrMemAlloc, "hipHostMalloc");
    }
    items_cap_ = cap_points_ + cap_points_ / kJobPoints + 2;
    // The job table and tile count go up on the batch's stream, after the
    // previous run's kernels (which read them) and before this run's: no
    // host wait for the stream, so uploads queued on it by the caller are
    // not waited for either.  Only the staging copy of the previous
    // upload is waited for before it is overwritten.
    if (desc_pending_)
    {
        check_hip(hipEventSynchronize(ev_desc_), kErrKernelLaunch,
                "event synchronize");
        desc_pending_ = false;
    }
    if (J > h_desc_cap_)
    {
        if (h_desc_) (void)hipHostFree(h_desc_);
        h_desc_ = nullptr;
        h_desc_cap_ = 0;
        check_hip(hipHostMalloc((void**)&h_desc_, J * sizeof(BatchJob),
                hipHostMallocDefault), kErrMemAlloc, "hipHostMalloc");
        h_desc_cap_ = J;
    }
    if (J)
    {
        std::memcpy(h_desc_, desc_.data(), J * sizeof(BatchJob));
        check_hip(hipMemcpyAsync(d_desc_, h_desc_, J * sizeof(BatchJob),
                hipMemcpyHostToDevice, stream_), kErrMemCopy,
                "upload job table");
    }
    *h_ntiles_ = (unsigned)tiles;
    check_hip(hipMemcpyAsync(d_ntiles_, h_ntiles_, sizeof(unsigned),
            hipMemcpyHostToDevice, stream_), kErrMemCopy,
            "upload tile count");
    check_hip(hipEventRecord(ev_desc_, stream_), kErrKernelLaunch, "event");
    desc_pending_ = true;
}

void Batch::enqueue()
{
    const hipStream_t s = stream_;
    const size_t J = jobs_.size();
    // Every job's summary and (plane, sub-cell) histogram cleared at once.
    const size_t combos = J ? combo_off_[J] : 0;
    if (J)
        hipLaunchKernelGGL(k_init, dim3((unsigned)std::max<size_t>(1,
                std::min<size_t>(1024, (std::max(combos, J) + 255) / 256))),
                dim3(256), 0, s, d_sum_, d_combo_, combos, (int)J);
    for (size_t j = 0; j < J; ++j)
    {
        const Entry& e = jobs_[j];
        const Plan& pl = *e.plan;
        const BatchJob& d = desc_[j];
        const int P = (int)pl.num_w_planes();
        const int K = pl.sub_tables();
        unsigned* combo = d_combo_ + combo_off_[j];
        if (e.n)
        {
            const int blocks = (int)std::max<size_t>(1, std::min<size_t>(
                    (e.n + kLocateThreads - 1) / kLocateThreads,
                    (size_t)n_cu_));
            if (pl.index64())
                hipLaunchKernelGGL(
                        (k_locate<double, false, float, float, RecF>),
                        dim3(blocks), dim3(kLocateThreads), 0, s, e.n, e.uu, e.vv, e.ww,
                        e.vis, pl.grid_scale_d(), pl.w_scale_d(),
                        pl.oversample(), d.G, P, pl.device_support(),
                        pl.device_srank(), pl.srank_bits(), pl.plane_bits(),
                        d.sub_h,
                        K, d.tile, d.tile_bits, pl.combo_bits(), d.key_bits,
                        d_keys_[0] + d.point_base,
                        static_cast<RecF*>(d_rec_) + d.point_base, combo,
                        d_sum_ + j, (unsigned*)nullptr,
                        (unsigned long long*)nullptr,
                        (unsigned long long*)nullptr, 0,
                        (unsigned long long)j << kBatchJobShift);
            else
                hipLaunchKernelGGL((k_locate<float, false, float, float, RecF>),
                        dim3(blocks), dim3(kLocateThreads), 0, s, e.n, e.uu, e.vv, e.ww,
                        e.vis, pl.grid_scale_f(), pl.w_scale_f(),
                        pl.oversample(), d.G, P, pl.device_support(),
                        pl.device_srank(), pl.srank_bits(), pl.plane_bits(),
                        d.sub_h,
                        K, d.tile, d.tile_bits, pl.combo_bits(), d.key_bits,
                        d_keys_[0] + d.point_base,
                        static_cast<RecF*>(d_rec_) + d.point_base, combo,
                        d_sum_ + j, (unsigned*)nullptr,
                        (unsigned long long*)nullptr,
                        (unsigned long long*)nullptr, 0,
                        (unsigned long long)j << kBatchJobShift);
        }
    }
    // Every job's points gridded, updates and norm increment, one
    // workgroup per job.
    if (J)
        hipLaunchKernelGGL(k_plan_batch, dim3((unsigned)J), dim3(1024), 0, s,
                (const BatchJob*)d_desc_, (const unsigned*)d_combo_, d_sum_);
    check_hip(hipGetLastError(), kErrKernelLaunch, "batch index launch");
    const int pblocks = (int)std::max<size_t>(1, std::min<size_t>(
            (n_total_ + 255) / 256, (size_t)n_cu_ * 8));
    const int tblocks = (int)std::max<size_t>(1, std::min<size_t>(
            (tiles_total_ + 255) / 256, (size_t)n_cu_ * 8));
    const unsigned long long sentinel =
            (unsigned long long)tiles_total_ << lmax_;
    fill_zero_u32(d_tstart_, tiles_total_, s);
    fill_zero_u32(d_tend_, tiles_total_, s);
    hipLaunchKernelGGL(k_batch_key, dim3(pblocks), dim3(256), 0, s,
            n_total_, (const unsigned long long*)d_keys_[0],
            (const BatchJob*)d_desc_, lmax_, sentinel, d_keys_[1], d_bsum_);
    // Stable LSD radix sort of (batch key, point index).
    unsigned long long* kk[2] = {d_keys_[1], d_keys_[0]};
    unsigned* vv[2] = {d_idx_[0], d_idx_[1]};
    const int cur = radix_sort_enqueue(n_total_, radix_bits_, kk, vv,
            d_rhist_, d_roff_, d_part_, s);
    const unsigned long long* kin = kk[cur];
    const unsigned* vin = vv[cur];
    // Runs of equal keys (same job, tile, origin, plane and sub-cell table)
    // -> one record each: the run-start bits, their counts and the counts'
    // scan into the sort's spent value buffer, the runs' keys into its spent
    // key buffer; tiles over the run records.
    unsigned* const mh = vv[1 - cur];
    unsigned long long* const mkeys = kk[1 - cur];
    hipLaunchKernelGGL(k_batch_count, dim3(pblocks), dim3(256), 0, s,
            n_total_, kin, lmax_, (unsigned)tiles_total_, d_bsum_);
    const size_t nwmax = n_total_ / 32 + 1;
    unsigned* const hb = mh;                 // run-start bits
    unsigned* const hc = hb + nwmax;         // their counts per word
    unsigned* const hp = hc + nwmax;         // the counts' exclusive scan
    hipLaunchKernelGGL(k_merge_heads, dim3(pblocks), dim3(256), 0, s,
            d_bsum_, kin, hb, hc, 1);
    scan_u32_enqueue(hc, hp, &d_bsum_->nw32, nwmax, &d_bsum_->nm32,
            d_part_, s);
    hipLaunchKernelGGL(k_run_starts, dim3(pblocks), dim3(256), 0, s,
            (const Summary*)d_bsum_, (const unsigned*)hb,
            (const unsigned*)hp, d_rstart_);
    hipLaunchKernelGGL(k_gather_batch, dim3(pblocks), dim3(256), 0, s,
            (const Summary*)d_bsum_, kin, vin, (const RecF*)d_rec_,
            (const BatchJob*)d_desc_, (int)J, lmax_, d_meta_,
            (float4*)d_vq_, (const unsigned*)d_rstart_, mkeys);
    hipLaunchKernelGGL(k_tile_bounds, dim3(pblocks), dim3(256), 0, s,
            (const unsigned long long*)mkeys, lmax_, d_tstart_, d_tend_,
            (const Summary*)d_bsum_);
    hipLaunchKernelGGL(k_tile_items, dim3(tblocks), dim3(256), 0, s,
            (const unsigned*)d_ntiles_, (const unsigned*)d_tstart_,
            (const unsigned*)d_tend_, d_icnt_);
    scan_u32_enqueue(d_icnt_, d_ioff_, d_ntiles_, tiles_total_,
            &d_bsum_->total_items, d_part_, s);
    hipLaunchKernelGGL(k_write_tile_items, dim3(tblocks), dim3(256), 0, s,
            (unsigned)tiles_total_, (const unsigned*)d_tstart_,
            (const unsigned*)d_tend_, (const unsigned*)d_ioff_,
            (const BatchJob*)d_desc_, (int)J, (unsigned)items_cap_, d_items_);
    // Each item's support segments, into the run-head buffer (spent).
    hipLaunchKernelGGL(k_item_segments, dim3(pblocks), dim3(256), 0, s,
            (const Summary*)d_bsum_, (unsigned)items_cap_, d_items_,
            (const int4*)d_meta_, mh);
    hipLaunchKernelGGL(k_grid_batch, dim3(grid_blocks_), dim3(kGridThreads),
            0, s, d_bsum_, (const int4*)d_items_,
            (const BatchJob*)d_desc_, (const int4*)d_meta_,
            (const float4*)d_vq_, kXcdGroups, (const unsigned*)mh);
    check_hip(hipGetLastError(), kErrKernelLaunch, "batch launch");
}

void Batch::run()
{
    prepare();
    dirty_ = false;
    drop_graph();
    check_hip(hipEventRecord(ev_[0], stream_), kErrKernelLaunch, "event");
    enqueue();
    check_hip(hipEventRecord(ev_[1], stream_), kErrKernelLaunch, "event");
    timed_ = true;
}

void Batch::capture()
{
    prepare();
    drop_graph();
    check_hip(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal),
            kErrKernelLaunch, "hipStreamBeginCapture");
    try
    {
        enqueue();
    }
    catch (...)
    {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(stream_, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    check_hip(hipStreamEndCapture(stream_, &graph_), kErrKernelLaunch,
            "hipStreamEndCapture");
    check_hip(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0),
            kErrKernelLaunch, "hipGraphInstantiate");
    dirty_ = false;
}

void Batch::launch()
{
    if (dirty_ || !exec_) capture();
    check_hip(hipEventRecord(ev_[0], stream_), kErrKernelLaunch, "event");
    check_hip(hipGraphLaunch(exec_, stream_), kErrKernelLaunch,
            "hipGraphLaunch");
    check_hip(hipEventRecord(ev_[1], stream_), kErrKernelLaunch, "event");
    timed_ = true;
}

void Batch::results(int job, size_t* skipped, double* norm_delta,
                    unsigned long long* updates)
{
    if (job < 0 || (size_t)job >= jobs_.size())
        fail(kErrInvalidArgument, "job index out of range");
    check_hip(hipMemcpyAsync(h_sum_, d_sum_, jobs_.size() * sizeof(Summary),
            hipMemcpyDeviceToHost, stream_), kErrMemCopy, "read summaries");
    check_hip(hipStreamSynchronize(stream_), kErrKernelLaunch,
            "stream synchronize");
    const Summary& sm = h_sum_[job];
    if (skipped) *skipped = (size_t)sm.skipped;
    if (norm_delta) *norm_delta = sm.norm_delta;
    if (updates) *updates = sm.updates;
}

size_t Batch::device_bytes() const
{
    const size_t items = cap_points_ + cap_points_ / kJobPoints + 2;
    return cap_points_ * (2 * sizeof(unsigned long long) + 2 * sizeof(unsigned)
                          + sizeof(RecF) + sizeof(int4) + sizeof(float4)) +
           cap_tiles_ * 5 * sizeof(unsigned) + items * sizeof(int4) +
           cap_jobs_ * (sizeof(BatchJob) + sizeof(Summary)) +
           cap_combo_ * sizeof(unsigned);
}

size_t Batch::workspace_estimate(size_t points)
{
    const size_t pcap = std::max<size_t>(points, 1024);
    return pcap * (2 * sizeof(unsigned long long) + 2 * sizeof(unsigned) +
                   sizeof(RecF) + sizeof(int4) + sizeof(float4)) +
           (pcap + pcap / kJobPoints + 2) * sizeof(int4) + (1u << 20);
}

void Batch::release_workspace()
{
    if (stream_) (void)hipStreamSynchronize(stream_);
    drop_graph();
    free_buffers();
    dirty_ = true;
}

double Batch::last_ms()
{
    if (!timed_) return 0.0;
    check_hip(hipEventSynchronize(ev_[1]), kErrKernelLaunch, "event sync");
    float ms = 0.0f;
    check_hip(hipEventElapsedTime(&ms, ev_[0], ev_[1]), kErrKernelLaunch,
            "elapsed");
    return ms;
}

// Zero-fill helpers used by the C ABI layer (a fill kernel rather than
// hipMemsetAsync, so that the same enqueue sequence is capture-safe).
void fill_zero_u64(unsigned long long* p, size_t n, hipStream_t s)
{
    if (!n) return;
    const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill_u64, dim3(blocks), dim3(256), 0, s, p, n);
    check_hip(hipGetLastError(), kErrKernelLaunch, "fill launch");
}

void fill_zero_u32(unsigned* p, size_t n, hipStream_t s)
{
    if (!n) return;
    const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks), dim3(256), 0, s, p, n);
    check_hip(hipGetLastError(), kErrKernelLaunch, "fill launch");
}

}  // namespace gw
