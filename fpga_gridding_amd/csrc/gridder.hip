// MI355X (gfx950) W-projection gridder: device kernels and the resident plan.
//
// Hot path of OxfordSKA/FPGA_Gridding: oskar_grid_wproj_f / _d
// (oskar_grid_wproj.c:39-243).  Each point's footprint of (2S+1)^2 taps is
// scattered onto the complex grid.  Pipeline, all on one HIP stream:
//
//   k_locate   per point: the reference's index math, bit-exact
//              (oskar_grid_wproj.c:169-212), skip count, per-(plane,
//              sub-cell) histogram, footprint bounding box, sort key
//              (plane, Morton(u0, v0)) and a 16/32-byte point record;
//   sort       radix sort of the keys (rocPRIM), points grouped by plane and
//              ordered along a Z-curve inside each plane;
//   k_gather   sorted point records (position, position mod W, sub-cell,
//              visibility pre-multiplied by the conjugation sign);
//   k_plan     per-plane counts and offsets, norm increment
//              (sum of histogram x per-(plane, sub-cell) tap sums), number of
//              grid updates;
//   k_items    wave work items: (plane, run of sorted points, block of
//              footprint cells);
//   k_grid     the scatter.  Each wave64 owns a block of the (2S+1)^2
//              footprint-cell residues of one plane (Romein's scheme): lane
//              slot (rx, ry) always feeds the grid cell congruent to
//              (rx, ry) mod W inside the current footprint, so it keeps its
//              sum in registers while consecutive points leave that cell in
//              place and flushes with a float atomic add only when it moves.
//              Consecutive points are Z-curve neighbours, so flushes are a
//              small fraction of the taps.
#include "gridder_internal.h"

#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>

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

// -------------------------------------------------------------------------
// Scalar helpers with the reference's (x86-64 host) semantics
// -------------------------------------------------------------------------

__device__ __forceinline__ float d_round(float x) { return roundf(x); }
__device__ __forceinline__ double d_round(double x) { return round(x); }
__device__ __forceinline__ float d_sqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ double d_sqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float d_fabs(float x) { return fabsf(x); }
__device__ __forceinline__ double d_fabs(double x) { return fabs(x); }

// (int)x of an integral-valued x: cvttss2si / cvttsd2si return INT_MIN for
// NaN and for values outside the int range.
__device__ __forceinline__ int to_int_x86(float x)
{
    return (x >= -2147483648.0f && x < 2147483648.0f) ? (int)x : INT_MIN;
}
__device__ __forceinline__ int to_int_x86(double x)
{
    return (x >= -2147483648.0 && x < 2147483648.0) ? (int)x : INT_MIN;
}

// (size_t)x for x >= 0 or NaN, as gcc lowers it on x86-64: values below
// 2^63 convert directly, NaN gives 2^63, larger values go through
// x - 2^63 with bit 63 flipped, and the indefinite result of an out-of-range
// x - 2^63 (e.g. +inf) flips to 0.
template <typename T>
__device__ __forceinline__ unsigned long long to_size_x86(T x)
{
    const T two63 = (T)9223372036854775808.0;
    if (x < two63) return (unsigned long long)(long long)x;
    if (x != x) return 0x8000000000000000ull;
    const T t = x - two63;
    if (t < two63)
        return ((unsigned long long)(long long)t) ^ 0x8000000000000000ull;
    return 0ull;
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

__device__ __forceinline__ void atomic_add_real(float* p, float v)
{
    unsafeAtomicAdd(p, v);
}
__device__ __forceinline__ void atomic_add_real(double* p, double v)
{
    unsafeAtomicAdd(p, v);
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

__global__ void k_init(Summary* __restrict__ s, unsigned* __restrict__ combo,
                       int n_combo)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0)
    {
        s->skipped = 0;
        s->n_valid = 0;
        s->updates = 0;
        s->norm_delta = 0.0;
        s->v_min = INT_MAX;
        s->v_max = INT_MIN;
        s->u_min = INT_MAX;
        s->u_max = INT_MIN;
        s->total_items = 0;
        s->flags = 0;
    }
    for (int i = t; i < n_combo; i += gridDim.x * blockDim.x) combo[i] = 0u;
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

// Index stage.  Follows oskar_grid_wproj.c:169-212 (FP32) and :71-101
// (FP64) operation by operation with FMA contraction disabled.
template <typename T, bool DIAG>
__global__ void __launch_bounds__(256) k_locate(
        size_t n, const T* __restrict__ uu, const T* __restrict__ vv,
        const T* __restrict__ ww, const T* __restrict__ vis, T grid_scale,
        T w_scale, int ov, int G, int P, const int* __restrict__ support,
        int h, int K, int coord_bits,
        unsigned long long* __restrict__ keys, unsigned* __restrict__ idx,
        typename Types<T>::Rec* __restrict__ rec,
        unsigned* __restrict__ combo, Summary* __restrict__ summary,
        unsigned* __restrict__ gw_out,
        unsigned long long* __restrict__ diag_support,
        unsigned long long* __restrict__ diag_plane, int diag_plane_cap)
{
#pragma clang fp contract(off)
    __shared__ unsigned s_combo[kComboLds];
    __shared__ int s_red[4][5];
    const int n_combo = P * K * K;
    const bool lds_hist = n_combo <= kComboLds;
    if (lds_hist)
        for (int i = threadIdx.x; i < n_combo; i += blockDim.x) s_combo[i] = 0;
    __syncthreads();

    const int centre = G / 2;
    const T ovf = (T)ov;
    const unsigned long long skip_key =
            (unsigned long long)P << (2 * coord_bits);
    int nskip = 0;
    int vmin = INT_MAX, vmax = INT_MIN, umin = INT_MAX, umax = INT_MIN;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += stride)
    {
        const T pu = -uu[i] * grid_scale;
        const T pv = vv[i] * grid_scale;
        const T w = ww[i];
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
            key = ((unsigned long long)plane << (2 * coord_bits)) |
                    spread_bits16((unsigned)u0) |
                    (spread_bits16((unsigned)v0) << 1);
            const int iu = min(max(off_u + h, 0), K - 1);
            const int iv = min(max(off_v + h, 0), K - 1);
            const int c = (plane * K + iv) * K + iu;
            if (lds_hist) atomicAdd(&s_combo[c], 1u);
            else atomicAdd(&combo[c], 1u);
            vmin = min(vmin, v0);
            vmax = max(vmax, v0 + 2 * S);
            umin = min(umin, u0);
            umax = max(umax, u0 + 2 * S);
            typename Types<T>::Rec r;
            r.pos = (int)((unsigned)u0 | ((unsigned)v0 << 16));
            r.sub = (int)((unsigned)(off_u + 128) |
                    ((unsigned)(off_v + 128) << 8) | ((unsigned)conj << 16) |
                    ((unsigned)plane << 17));
            r.re = vis[2 * i];
            r.im = vis[2 * i + 1];
            rec[i] = r;
        }
        keys[i] = key;
        idx[i] = (unsigned)i;
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

// Sorted point records.  meta = (u0 | v0 << 16, u0 mod W | v0 mod W << 16,
// sub-cell/conj/plane, grid_w); vq = (re, -s*im, im, s*re) with s = -1 when
// ww > 0 (oskar_grid_wproj.c:172, :233), so that
//   grid_re += vq.x * c_re + vq.y * c_im,  grid_im += vq.z * c_re + vq.w * c_im
// with (c_re, c_im) read from the cube unmodified.
template <typename T>
__global__ void __launch_bounds__(256) k_gather(
        size_t n, const Summary* __restrict__ summary,
        const unsigned* __restrict__ idx,
        const typename Types<T>::Rec* __restrict__ rec,
        const PlaneInfo* __restrict__ pinfo, const unsigned* __restrict__ gw,
        int4* __restrict__ meta, typename Types<T>::T4* __restrict__ vq)
{
    const size_t nv = n - summary->skipped;
    for (size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x; s < nv;
         s += (size_t)gridDim.x * blockDim.x)
    {
        const unsigned i = idx[s];
        const typename Types<T>::Rec r = rec[i];
        const int plane = (unsigned)r.sub >> 17;
        const unsigned W = (unsigned)pinfo[plane].W;
        const unsigned u0 = (unsigned)r.pos & 0xFFFFu;
        const unsigned v0 = (unsigned)r.pos >> 16;
        const T sgn = ((r.sub >> 16) & 1) ? (T)-1 : (T)1;
        meta[s] = make_int4(r.pos, (int)((u0 % W) | ((v0 % W) << 16)), r.sub,
                gw ? (int)min(gw[i], 0x7FFFFFFFu) : 0);
        typename Types<T>::T4 q;
        q.x = r.re;
        q.y = -sgn * r.im;
        q.z = r.im;
        q.w = sgn * r.re;
        vq[s] = q;
    }
}

// Per-plane offsets, work-item counts, update count and norm increment.
// One block of 1024 threads; tiles of 1024 planes with a running carry.
__global__ void __launch_bounds__(1024) k_plan(
        int P, int K2, const unsigned* __restrict__ combo,
        const PlaneInfo* __restrict__ pinfo, const double* __restrict__ normtab,
        unsigned* __restrict__ plane_start, unsigned* __restrict__ item_off,
        unsigned items_cap, Summary* __restrict__ summary)
{
    __shared__ unsigned s_a[1024];
    __shared__ unsigned s_b[1024];
    __shared__ double s_d[1024];
    __shared__ unsigned long long s_u[1024];
    const int t = threadIdx.x;
    unsigned carry_pts = 0, carry_items = 0;
    unsigned long long upd = 0;
    for (int base = 0; base < P; base += 1024)
    {
        const int p = base + t;
        unsigned cnt = 0, items = 0;
        if (p < P)
        {
            for (int c = 0; c < K2; ++c) cnt += combo[p * K2 + c];
            const PlaneInfo pi = pinfo[p];
            items = ((cnt + pi.L - 1) / pi.L) * pi.nw;
            upd += (unsigned long long)cnt * pi.W * pi.W;
        }
        s_a[t] = cnt;
        s_b[t] = items;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1)
        {
            const unsigned a = t >= o ? s_a[t - o] : 0u;
            const unsigned b = t >= o ? s_b[t - o] : 0u;
            __syncthreads();
            s_a[t] += a;
            s_b[t] += b;
            __syncthreads();
        }
        if (p < P)
        {
            plane_start[p] = carry_pts + s_a[t] - cnt;
            item_off[p] = carry_items + s_b[t] - items;
        }
        carry_pts += s_a[1023];
        carry_items += s_b[1023];
        __syncthreads();
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
    if (t == 0)
    {
        item_off[P] = carry_items;
        summary->n_valid = carry_pts;
        summary->updates = s_u[0];
        summary->norm_delta = s_d[0];
        if (carry_items > items_cap)
        {
            summary->flags |= 2u;
            summary->total_items = items_cap;
        }
        else
        {
            summary->total_items = carry_items;
        }
    }
}

// Expand per-plane work into wave items (plane, first point, points,
// first footprint cell).  One block per plane.
__global__ void __launch_bounds__(256) k_items(
        int P, const PlaneInfo* __restrict__ pinfo,
        const unsigned* __restrict__ plane_start,
        const unsigned* __restrict__ item_off,
        const Summary* __restrict__ summary, unsigned items_cap,
        int4* __restrict__ items)
{
    const int p = blockIdx.x;
    if (p >= P) return;
    const PlaneInfo pi = pinfo[p];
    const unsigned start = plane_start[p];
    const unsigned end = (p + 1 < P) ? plane_start[p + 1] :
            (unsigned)summary->n_valid;
    const unsigned cnt = end - start;
    const unsigned off = item_off[p];
    const unsigned n_items = item_off[p + 1] - off;
    for (unsigned j = threadIdx.x; j < n_items; j += blockDim.x)
    {
        const unsigned chunk = j / pi.nw;
        const unsigned wv = j - chunk * pi.nw;
        const unsigned first = chunk * pi.L;
        const unsigned len = min((unsigned)pi.L, cnt - first);
        if (off + j < items_cap)
            items[off + j] = make_int4(p, (int)(start + first), (int)len,
                    (int)(wv * pi.cpw));
    }
}

// The scatter (oskar_grid_wproj.c:216-240).  Persistent: each wave walks
// the work items with a grid-wide stride.  A wave item covers footprint
// cells [c0, c0 + cpw) of plane p; lane slot s owns cell
// c = c0 + lane + 64*s, i.e. residue (rx, ry) = (c mod W, c / W).  For a
// point whose footprint starts at (u0, v0), that residue lands on grid cell
// (u0 + dx, v0 + dy) with dx = (rx - u0) mod W, dy = (ry - v0) mod W, tap
// offset k = dx - S, j = dy - S, tap
//   conv_func[p][|off_v + j*ov|][|off_u + k*ov|].
// The slot accumulates while its grid cell stays the same and flushes with
// an atomic add when the footprint moves it elsewhere.
template <typename T, bool DIAG>
__global__ void __launch_bounds__(256) k_grid(
        const Summary* __restrict__ summary, const int4* __restrict__ items,
        const PlaneInfo* __restrict__ pinfo, const int4* __restrict__ meta,
        const typename Types<T>::T4* __restrict__ vq,
        const typename Types<T>::T2* __restrict__ cube, int C, int ov, int G,
        int row0, T* __restrict__ grid, unsigned* __restrict__ diag_count,
        unsigned long long* __restrict__ diag_wsum)
{
    typedef typename Types<T>::T2 T2;
    typedef typename Types<T>::T4 T4;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int waves_per_block = blockDim.x >> 6;
    const unsigned total = summary->total_items;
    const unsigned stride = gridDim.x * waves_per_block;
    for (unsigned it = blockIdx.x * waves_per_block + wave; it < total;
         it += stride)
    {
        const int4 item = items[it];
        const int p = item.x;
        const PlaneInfo pi = pinfo[p];
        const int W = pi.W, S = pi.S, nslot = pi.nslot;
        const int cend = min(item.w + pi.cpw, W * W);
        const T2* __restrict__ taps = cube + (size_t)p * C * C;

        int rx[kSlotsMax], ry[kSlotsMax];
        bool valid[kSlotsMax];
        T acc_re[kSlotsMax], acc_im[kSlotsMax];
        unsigned cur[kSlotsMax];
        unsigned cnt[kSlotsMax];
        unsigned long long wsum[kSlotsMax];
        int kxo[kSlotsMax], kyo[kSlotsMax];
#pragma unroll
        for (int s = 0; s < kSlotsMax; ++s)
        {
            const int c = item.w + lane + 64 * s;
            valid[s] = (s < nslot) && (c < cend);
            ry[s] = c / W;
            rx[s] = c - ry[s] * W;
            acc_re[s] = (T)0;
            acc_im[s] = (T)0;
            cur[s] = kNoCell;
            cnt[s] = 0;
            wsum[s] = 0;
            kxo[s] = 0;
            kyo[s] = 0;
        }

        unsigned prev = 0xFFFFFFFFu;
        const int k_end = item.y + item.z;
        for (int k = item.y; k < k_end; ++k)
        {
            const int4 m = meta[k];
            const T4 v = vq[k];
            if ((unsigned)m.x != prev)
            {
                prev = (unsigned)m.x;
                const int u0 = m.x & 0xFFFF;
                const int v0 = (int)((unsigned)m.x >> 16);
                const int u0m = m.y & 0xFFFF;
                const int v0m = (int)((unsigned)m.y >> 16);
                const int vrow = v0 - row0;
#pragma unroll
                for (int s = 0; s < kSlotsMax; ++s)
                {
                    if (s < nslot && valid[s])
                    {
                        int dx = rx[s] - u0m;
                        dx += dx < 0 ? W : 0;
                        int dy = ry[s] - v0m;
                        dy += dy < 0 ? W : 0;
                        const unsigned g = (unsigned)(vrow + dy) * (unsigned)G +
                                (unsigned)(u0 + dx);
                        if (g != cur[s])
                        {
                            if (cur[s] != kNoCell)
                            {
                                atomic_add_real(grid + 2 * (size_t)cur[s],
                                        acc_re[s]);
                                atomic_add_real(grid + 2 * (size_t)cur[s] + 1,
                                        acc_im[s]);
                                if (DIAG)
                                {
                                    atomicAdd(&diag_count[cur[s]], cnt[s]);
                                    atomicAdd(&diag_wsum[cur[s]], wsum[s]);
                                }
                            }
                            cur[s] = g;
                            acc_re[s] = (T)0;
                            acc_im[s] = (T)0;
                            cnt[s] = 0;
                            wsum[s] = 0;
                        }
                        kxo[s] = (dx - S) * ov;
                        kyo[s] = (dy - S) * ov;
                    }
                }
            }
            const int off_u = (m.z & 0xFF) - 128;
            const int off_v = ((m.z >> 8) & 0xFF) - 128;
#pragma unroll
            for (int s = 0; s < kSlotsMax; ++s)
            {
                if (s < nslot && valid[s])
                {
                    const int ix = abs(off_u + kxo[s]);
                    const int iy = abs(off_v + kyo[s]);
                    const T2 c = taps[iy * C + ix];
                    acc_re[s] = fma(v.x, c.x, acc_re[s]);
                    acc_re[s] = fma(v.y, c.y, acc_re[s]);
                    acc_im[s] = fma(v.z, c.x, acc_im[s]);
                    acc_im[s] = fma(v.w, c.y, acc_im[s]);
                    if (DIAG)
                    {
                        cnt[s] += 1;
                        wsum[s] += (unsigned long long)S + (unsigned)m.w;
                    }
                }
            }
        }
#pragma unroll
        for (int s = 0; s < kSlotsMax; ++s)
        {
            if (s < nslot && valid[s] && cur[s] != kNoCell)
            {
                atomic_add_real(grid + 2 * (size_t)cur[s], acc_re[s]);
                atomic_add_real(grid + 2 * (size_t)cur[s] + 1, acc_im[s]);
                if (DIAG)
                {
                    atomicAdd(&diag_count[cur[s]], cnt[s]);
                    atomicAdd(&diag_wsum[cur[s]], wsum[s]);
                }
            }
        }
    }
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

template <typename T>
void compute_normtab(const T* conv, size_t P, const int* support, int ov,
                     int C, int h, std::vector<double>& tab)
{
    const int K = 2 * h + 1;
    tab.assign(P * K * K, 0.0);
    const size_t kdim = (size_t)C * C;
    for (size_t p = 0; p < P; ++p)
    {
        const int S = support[p];
        for (int iv = 0; iv < K; ++iv)
        {
            for (int iu = 0; iu < K; ++iu)
            {
                // Same order as the reference's per-point `sum`
                // (oskar_grid_wproj.c:216-238): rows j outer, columns k.
                const int off_v = iv - h, off_u = iu - h;
                double sum = 0.0;
                for (int j = -S; j <= S; ++j)
                {
                    const size_t t1 = p * kdim +
                            (size_t)std::abs(off_v + j * ov) * C;
                    for (int k = -S; k <= S; ++k)
                        sum += conv[2 * (t1 + std::abs(off_u + k * ov))];
                }
                tab[(p * K + iv) * K + iu] = sum;
            }
        }
    }
}

template <typename T>
T* dmalloc(size_t count)
{
    void* p = nullptr;
    if (count == 0) count = 1;
    check_hip(hipMalloc(&p, count * sizeof(T)), kErrMemAlloc, "hipMalloc");
    return static_cast<T*>(p);
}

void dfree(void* p)
{
    if (p) (void)hipFree(p);
}

}  // namespace

Plan::Plan(int precision, size_t num_w_planes, const int* support,
           int oversample, int conv_size_half, const void* conv_func,
           double cell_size_rad, double w_scale, int grid_size)
{
    if (precision != 0x01 && precision != 0x02)
        fail(kErrInvalidArgument, "precision must be GW_SINGLE or GW_DOUBLE");
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
    prec_ = precision;
    P_ = num_w_planes;
    ov_ = oversample;
    C_ = conv_size_half;
    G_ = grid_size;
    h_ = (ov_ + 1) / 2;
    K_ = 2 * h_ + 1;
    support_.assign(support, support + P_);
    s_max_ = 0;
    // Largest flat tap index the gridder can read, per plane:
    // p*C^2 + (h + S*ov)*C + (h + S*ov).  The reference reads the flat cube
    // the same way; it must stay inside the cube.
    const unsigned long long cube_len = (unsigned long long)P_ * C_ * C_;
    for (size_t p = 0; p < P_; ++p)
    {
        const int S = support_[p];
        if (S < 0 || S > 16383)
            fail(kErrInvalidArgument, "support values must be in [0, 16383]");
        s_max_ = std::max(s_max_, S);
        const unsigned long long m = (unsigned long long)h_ +
                (unsigned long long)S * ov_;
        if ((unsigned long long)p * C_ * C_ + m * C_ + m >= cube_len)
            fail(kErrInvalidArgument,
                 "conv_size_half too small for the support and oversample "
                 "(taps would be read beyond the kernel cube)");
    }
    cell_d_ = cell_size_rad;
    ws_d_ = w_scale;
    cell_f_ = (float)cell_size_rad;
    ws_f_ = (float)w_scale;
    // grid_scale as the reference forms it (oskar_grid_wproj.c:150, :61).
    gscale_f_ = (float)grid_size * cell_f_;
    gscale_d_ = (double)grid_size * cell_d_;
    coord_bits_ = std::max(1, bits_for((unsigned long long)(G_ - 1)));
    key_bits_ = 2 * coord_bits_ + bits_for((unsigned long long)P_);

    // Work decomposition per plane.
    pinfo_.resize(P_);
    items_per_point_ = 0.0;
    items_fixed_ = 1;
    for (size_t p = 0; p < P_; ++p)
    {
        PlaneInfo& pi = pinfo_[p];
        pi.S = support_[p];
        pi.W = 2 * pi.S + 1;
        const long long cells = (long long)pi.W * pi.W;
        pi.nw = (int)((cells + 64 * kSlotsMax - 1) / (64 * kSlotsMax));
        pi.cpw = (int)((cells + pi.nw - 1) / pi.nw);
        pi.nslot = (pi.cpw + 63) / 64;
        pi.L = std::max(1, std::min(65535, kTapsPerItem / pi.cpw));
        pi.pad0 = pi.pad1 = 0;
        items_per_point_ = std::max(items_per_point_, (double)pi.nw / pi.L);
        items_fixed_ += pi.nw;
    }

    check_hip(hipGetDevice(&device_), kErrComputeDevices, "hipGetDevice");
    check_hip(hipDeviceGetAttribute(&n_cu_, hipDeviceAttributeMultiprocessorCount,
            device_), kErrComputeDevices, "hipDeviceGetAttribute");
    int occ = 0;
    check_hip(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ,
            precision == 0x01 ? (const void*)k_grid<float, false> :
            (const void*)k_grid<double, false>, 256, 0),
            kErrComputeDevices, "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    grid_blocks_ = n_cu_ * std::max(1, occ);
    check_hip(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking),
            kErrComputeDevices, "hipStreamCreateWithFlags");
    own_stream_ = true;
    for (auto& e : ev_)
        check_hip(hipEventCreate(&e), kErrComputeDevices, "hipEventCreate");

    // Constant tables.
    std::vector<double> normtab;
    const size_t real_size = prec_ == 0x01 ? sizeof(float) : sizeof(double);
    const size_t cube_bytes = 2 * cube_len * real_size;
    if (prec_ == 0x01)
        compute_normtab(static_cast<const float*>(conv_func), P_,
                support_.data(), ov_, C_, h_, normtab);
    else
        compute_normtab(static_cast<const double*>(conv_func), P_,
                support_.data(), ov_, C_, h_, normtab);
    d_cube_ = dmalloc<char>(cube_bytes);
    d_support_ = dmalloc<int>(P_);
    d_pinfo_ = dmalloc<PlaneInfo>(P_);
    d_normtab_ = dmalloc<double>(normtab.size());
    d_combo_ = dmalloc<unsigned>(P_ * K_ * K_);
    d_plane_start_ = dmalloc<unsigned>(P_);
    d_item_off_ = dmalloc<unsigned>(P_ + 1);
    d_summary_ = dmalloc<Summary>(1);
    check_hip(hipHostMalloc((void**)&h_summary_, sizeof(Summary),
            hipHostMallocDefault), kErrMemAlloc, "hipHostMalloc");
    std::memset(h_summary_, 0, sizeof(Summary));
    check_hip(hipMemcpy(d_cube_, conv_func, cube_bytes, hipMemcpyHostToDevice),
            kErrMemCopy, "upload kernel cube");
    check_hip(hipMemcpy(d_support_, support_.data(), P_ * sizeof(int),
            hipMemcpyHostToDevice), kErrMemCopy, "upload support");
    check_hip(hipMemcpy(d_pinfo_, pinfo_.data(), P_ * sizeof(PlaneInfo),
            hipMemcpyHostToDevice), kErrMemCopy, "upload plane table");
    check_hip(hipMemcpy(d_normtab_, normtab.data(),
            normtab.size() * sizeof(double), hipMemcpyHostToDevice),
            kErrMemCopy, "upload norm table");
}

Plan::~Plan()
{
    if (stream_) (void)hipStreamSynchronize(stream_);
    free_workspace();
    dfree(d_cube_);
    dfree(d_support_);
    dfree(d_pinfo_);
    dfree(d_normtab_);
    dfree(d_combo_);
    dfree(d_plane_start_);
    dfree(d_item_off_);
    dfree(d_summary_);
    if (h_summary_) (void)hipHostFree(h_summary_);
    for (auto& e : ev_)
        if (e) (void)hipEventDestroy(e);
    if (own_stream_ && stream_) (void)hipStreamDestroy(stream_);
}

void Plan::set_stream(hipStream_t s)
{
    if (s)
    {
        if (own_stream_ && stream_)
        {
            (void)hipStreamSynchronize(stream_);
            (void)hipStreamDestroy(stream_);
        }
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

void Plan::free_workspace()
{
    for (int i = 0; i < 2; ++i)
    {
        dfree(d_keys_[i]);
        dfree(d_idx_[i]);
        d_keys_[i] = nullptr;
        d_idx_[i] = nullptr;
    }
    dfree(d_rec_);
    dfree(d_meta_);
    dfree(d_vq_);
    dfree(d_gw_);
    dfree(d_sort_tmp_);
    dfree(d_items_);
    d_rec_ = d_vq_ = d_sort_tmp_ = nullptr;
    d_meta_ = nullptr;
    d_gw_ = nullptr;
    d_items_ = nullptr;
    cap_ = items_cap_ = 0;
    sort_tmp_bytes_ = 0;
}

void Plan::ensure_capacity(size_t n)
{
    if (n <= cap_) return;
    if (n > 0xFFFFFFFFull)
        fail(kErrInvalidArgument, "num_points must be below 2^32 per call");
    free_workspace();
    const size_t cap = std::max<size_t>(n, 1024);
    const size_t real_size = prec_ == 0x01 ? sizeof(float) : sizeof(double);
    for (int i = 0; i < 2; ++i)
    {
        d_keys_[i] = dmalloc<unsigned long long>(cap);
        d_idx_[i] = dmalloc<unsigned>(cap);
    }
    d_rec_ = dmalloc<char>(cap * (prec_ == 0x01 ? sizeof(RecF) : sizeof(RecD)));
    d_meta_ = dmalloc<int4>(cap);
    d_vq_ = dmalloc<char>(cap * 4 * real_size);
    d_gw_ = dmalloc<unsigned>(cap);
    size_t bytes = 0;
    check_hip(rocprim::radix_sort_pairs(nullptr, bytes, d_keys_[0], d_keys_[1],
            d_idx_[0], d_idx_[1], cap, 0u, (unsigned)key_bits_, stream_),
            kErrComputeDevices, "radix sort size query");
    d_sort_tmp_ = dmalloc<char>(bytes);
    sort_tmp_bytes_ = bytes;
    const double bound = (double)cap * items_per_point_ + (double)items_fixed_;
    if (bound > 2147483647.0)
        fail(kErrInvalidArgument, "too many work items for one call");
    items_cap_ = (size_t)bound + 1;
    d_items_ = dmalloc<int4>(items_cap_);
    cap_ = cap;
}

template <typename T>
void Plan::locate_t(size_t n, const T* uu, const T* vv, const T* ww,
                    const T* vis, const DiagBuffers* diag)
{
    ensure_capacity(n);
    timed_ = true;
    check_hip(hipEventRecord(ev_[0], stream_), kErrKernelLaunch, "event");
    const int n_combo = (int)(P_ * K_ * K_);
    hipLaunchKernelGGL(k_init, dim3(std::max(1, std::min(1024,
            (n_combo + 255) / 256))), dim3(256), 0, stream_, d_summary_,
            d_combo_, n_combo);
    const int blocks = (int)std::min<size_t>((n + 255) / 256,
            (size_t)n_cu_ * 4);
    const T gscale = sizeof(T) == 4 ? (T)gscale_f_ : (T)gscale_d_;
    const T wsc = sizeof(T) == 4 ? (T)ws_f_ : (T)ws_d_;
    typedef typename Types<T>::Rec Rec;
    if (diag)
        hipLaunchKernelGGL((k_locate<T, true>), dim3(std::max(blocks, 1)),
                dim3(256), 0, stream_, n, uu, vv, ww, vis, gscale, wsc, ov_,
                G_, (int)P_, d_support_, h_, K_, coord_bits_, d_keys_[0],
                d_idx_[0], (Rec*)d_rec_, d_combo_, d_summary_, d_gw_,
                diag->per_support, diag->per_plane, diag->plane_cap);
    else
        hipLaunchKernelGGL((k_locate<T, false>), dim3(std::max(blocks, 1)),
                dim3(256), 0, stream_, n, uu, vv, ww, vis, gscale, wsc, ov_,
                G_, (int)P_, d_support_, h_, K_, coord_bits_, d_keys_[0],
                d_idx_[0], (Rec*)d_rec_, d_combo_, d_summary_,
                (unsigned*)nullptr, (unsigned long long*)nullptr,
                (unsigned long long*)nullptr, 0);
    check_hip(hipGetLastError(), kErrKernelLaunch, "k_locate launch");
    check_hip(hipEventRecord(ev_[1], stream_), kErrKernelLaunch, "event");
}

template <typename T>
void Plan::scatter_t(size_t n, T* grid, int row0, const DiagBuffers* diag)
{
    typedef typename Types<T>::Rec Rec;
    typedef typename Types<T>::T2 T2;
    typedef typename Types<T>::T4 T4;
    size_t bytes = sort_tmp_bytes_;
    check_hip(rocprim::radix_sort_pairs(d_sort_tmp_, bytes, d_keys_[0],
            d_keys_[1], d_idx_[0], d_idx_[1], n, 0u, (unsigned)key_bits_,
            stream_), kErrKernelLaunch, "radix sort");
    check_hip(hipEventRecord(ev_[2], stream_), kErrKernelLaunch, "event");
    const int blocks = (int)std::max<size_t>(1, std::min<size_t>(
            (n + 255) / 256, (size_t)n_cu_ * 4));
    hipLaunchKernelGGL((k_gather<T>), dim3(blocks), dim3(256), 0, stream_, n,
            (const Summary*)d_summary_, (const unsigned*)d_idx_[1],
            (const Rec*)d_rec_, (const PlaneInfo*)d_pinfo_,
            (const unsigned*)(diag ? d_gw_ : nullptr), d_meta_, (T4*)d_vq_);
    check_hip(hipGetLastError(), kErrKernelLaunch, "k_gather launch");
    hipLaunchKernelGGL(k_plan, dim3(1), dim3(1024), 0, stream_, (int)P_,
            K_ * K_, (const unsigned*)d_combo_, (const PlaneInfo*)d_pinfo_,
            (const double*)d_normtab_, d_plane_start_, d_item_off_,
            (unsigned)items_cap_, d_summary_);
    check_hip(hipGetLastError(), kErrKernelLaunch, "k_plan launch");
    hipLaunchKernelGGL(k_items, dim3((unsigned)P_), dim3(256), 0, stream_,
            (int)P_, (const PlaneInfo*)d_pinfo_,
            (const unsigned*)d_plane_start_, (const unsigned*)d_item_off_,
            (const Summary*)d_summary_, (unsigned)items_cap_, d_items_);
    check_hip(hipGetLastError(), kErrKernelLaunch, "k_items launch");
    check_hip(hipEventRecord(ev_[3], stream_), kErrKernelLaunch, "event");
    if (diag)
        hipLaunchKernelGGL((k_grid<T, true>), dim3(grid_blocks_), dim3(256),
                0, stream_, (const Summary*)d_summary_,
                (const int4*)d_items_, (const PlaneInfo*)d_pinfo_,
                (const int4*)d_meta_, (const T4*)d_vq_, (const T2*)d_cube_,
                C_, ov_, G_, row0, grid, diag->per_cell, diag->wsum_cell);
    else
        hipLaunchKernelGGL((k_grid<T, false>), dim3(grid_blocks_), dim3(256),
                0, stream_, (const Summary*)d_summary_,
                (const int4*)d_items_, (const PlaneInfo*)d_pinfo_,
                (const int4*)d_meta_, (const T4*)d_vq_, (const T2*)d_cube_,
                C_, ov_, G_, row0, grid, (unsigned*)nullptr,
                (unsigned long long*)nullptr);
    check_hip(hipGetLastError(), kErrKernelLaunch, "k_grid launch");
    check_hip(hipEventRecord(ev_[4], stream_), kErrKernelLaunch, "event");
}

void Plan::locate(size_t n, const void* uu, const void* vv, const void* ww,
                  const void* vis, const DiagBuffers* diag)
{
    if (prec_ == 0x01)
        locate_t<float>(n, (const float*)uu, (const float*)vv,
                (const float*)ww, (const float*)vis, diag);
    else
        locate_t<double>(n, (const double*)uu, (const double*)vv,
                (const double*)ww, (const double*)vis, diag);
}

void Plan::scatter(size_t n, void* grid, int row0, const DiagBuffers* diag)
{
    if (prec_ == 0x01)
        scatter_t<float>(n, (float*)grid, row0, diag);
    else
        scatter_t<double>(n, (double*)grid, row0, diag);
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
