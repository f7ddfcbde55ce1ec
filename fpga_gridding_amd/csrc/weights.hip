// Uniform-weighting step on the device: the weights grid (sum of the
// weights of the points whose nearest cell it is) and the per-point read
// back, weight_out = weight_in / density (utils/oskar_grid_weights.c:38-172).
//
// Both results are bit-exact with the reference's sequential loops:
//
// * write: points are ordered by cell with a stable radix sort (input order
//   kept inside a cell), then one thread per occupied cell adds the cell's
//   weights to the existing grid value one after the other, in input order:
//   the same float (or double) additions in the same order as the loop at
//   oskar_grid_weights.c:126-137 (:58-69 for double).
// * read: one division per point, IEEE-rounded (:160-171, :92-103).
//
// The nearest-cell index follows the reference's arithmetic in the
// precision of the call: grid_scale = grid_size * cell_size_rad (:114 /
// :46), u = (int)round(-uu * grid_scale) + grid_size / 2 with the x86-64
// conversion of out-of-range values, and the same edge test.
#include "gridder_internal.h"
#include "host_semantics.h"

#include <algorithm>

namespace gw {

namespace {

constexpr int kThreads = 256;

// Nearest cell of a point (u, v in the reference's int arithmetic; the
// addition of the centre wraps as the host's does), or -1 off the grid.
template <typename T>
__device__ __forceinline__ long long weights_cell(T uu, T vv, T scale, int G)
{
    const int c = G / 2;
    const int u = (int)((unsigned)to_int_x86(d_round(-uu * scale)) +
                        (unsigned)c);
    const int v = (int)((unsigned)to_int_x86(d_round(vv * scale)) +
                        (unsigned)c);
    if (u >= G || u < 0 || v >= G || v < 0) return -1;
    return (long long)v * G + u;
}

__device__ __forceinline__ void add_skipped(unsigned count,
                                            unsigned long long* skipped)
{
    __shared__ unsigned s_count;
    if (threadIdx.x == 0) s_count = 0;
    __syncthreads();
    if (count) atomicAdd(&s_count, count);
    __syncthreads();
    if (threadIdx.x == 0 && s_count)
        atomicAdd(skipped, (unsigned long long)s_count);
}

// Sort keys: the cell index, or `sentinel` (= G^2, after every cell) for
// points off the grid; values: the point index.
template <typename T>
__global__ void __launch_bounds__(kThreads) k_weights_key(size_t n,
        const T* __restrict__ uu, const T* __restrict__ vv, T scale, int G,
        unsigned long long sentinel, unsigned long long* __restrict__ keys,
        unsigned long long* __restrict__ skipped)
{
    unsigned count = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
    {
        const long long q = weights_cell(uu[i], vv[i], scale, G);
        count += q < 0;
        keys[i] = q < 0 ? sentinel : (unsigned long long)q;
    }
    add_skipped(count, skipped);
}

// One thread per run of equal keys (an occupied cell): the cell's weights
// added to the grid value one at a time, in input order.
template <typename T>
__global__ void __launch_bounds__(kThreads) k_weights_sum(size_t n,
        const unsigned long long* __restrict__ keys,
        const unsigned* __restrict__ idx, const T* __restrict__ weight,
        unsigned long long sentinel, T* __restrict__ grid)
{
    for (size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x; s < n;
         s += (size_t)gridDim.x * blockDim.x)
    {
        const unsigned long long q = keys[s];
        if (q == sentinel || (s > 0 && keys[s - 1] == q)) continue;
        T acc = grid[q];
        for (size_t r = s; r < n && keys[r] == q; ++r) acc += weight[idx[r]];
        grid[q] = acc;
    }
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_weights_read(size_t n,
        const T* __restrict__ uu, const T* __restrict__ vv,
        const T* __restrict__ w_in, T* __restrict__ w_out, T scale, int G,
        const T* __restrict__ grid, unsigned long long* __restrict__ skipped)
{
    unsigned count = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
    {
        const long long q = weights_cell(uu[i], vv[i], scale, G);
        if (q < 0)
        {
            ++count;
            continue;
        }
        const T g = grid[q];
        w_out[i] = g != (T)0 ? w_in[i] / g : (T)0;
    }
    add_skipped(count, skipped);
}

int blocks_for(size_t n, int n_cu)
{
    return (int)std::max<size_t>(1, std::min<size_t>((n + kThreads - 1) /
            kThreads, (size_t)n_cu * 8));
}

}  // namespace

Weights::Weights()
{
    int dev = 0;
    check_hip(hipGetDevice(&dev), kErrComputeDevices, "hipGetDevice");
    check_hip(hipDeviceGetAttribute(&n_cu_,
            hipDeviceAttributeMultiprocessorCount, dev), kErrComputeDevices,
            "hipDeviceGetAttribute");
    check_hip(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking),
            kErrComputeDevices, "hipStreamCreateWithFlags");
    own_stream_ = true;
    d_small_ = dmalloc<unsigned long long>(1);
    check_hip(hipHostMalloc((void**)&h_skipped_, sizeof(unsigned long long),
            hipHostMallocDefault), kErrMemAlloc, "hipHostMalloc");
}

Weights::~Weights()
{
    if (stream_) (void)hipStreamSynchronize(stream_);
    free_buffers();
    dfree(d_small_);
    if (h_skipped_) (void)hipHostFree(h_skipped_);
    if (own_stream_ && stream_) (void)hipStreamDestroy(stream_);
}

void Weights::set_stream(hipStream_t s)
{
    if (!s) return;
    if (own_stream_ && stream_)
    {
        (void)hipStreamSynchronize(stream_);
        (void)hipStreamDestroy(stream_);
    }
    stream_ = s;
    own_stream_ = false;
}

void Weights::free_buffers()
{
    for (void* p : {(void*)d_keys_[0], (void*)d_keys_[1], (void*)d_idx_[0],
                    (void*)d_idx_[1], (void*)d_hist_, (void*)d_offs_,
                    (void*)d_part_})
        dfree(p);
    d_keys_[0] = d_keys_[1] = nullptr;
    d_idx_[0] = d_idx_[1] = nullptr;
    d_hist_ = d_offs_ = d_part_ = nullptr;
    cap_ = 0;
}

void Weights::ensure_capacity(size_t n)
{
    if (n <= cap_ && d_keys_[0]) return;
    // Every call waits for its own work, so no kernel still uses the
    // buffers here.
    free_buffers();
    const size_t cap = std::max<size_t>(n, 4096);
    for (int i = 0; i < 2; ++i)
    {
        d_keys_[i] = dmalloc<unsigned long long>(cap);
        d_idx_[i] = dmalloc<unsigned>(cap);
    }
    const size_t hw = radix_hist_words(cap);
    d_hist_ = dmalloc<unsigned>(hw);
    d_offs_ = dmalloc<unsigned>(hw);
    d_part_ = dmalloc<unsigned>(256);        // the sort's digit totals
    cap_ = cap;
}

size_t Weights::device_bytes() const
{
    const size_t hw = cap_ ? radix_hist_words(cap_) : 0;
    return cap_ * 2 * (sizeof(unsigned long long) + sizeof(unsigned)) +
           (2 * hw + 256) * sizeof(unsigned);
}

static void check_weights_args(size_t n, const void* uu, const void* vv,
                               const void* w, int G, const void* grid)
{
    if (G < 1 || G > 65535)
        fail(kErrInvalidArgument, "grid_size must be in [1, 65535]");
    if (n >= 0xFFFFFFFFull)
        fail(kErrInvalidArgument, "num_points must be below 2^32 - 1");
    if (!grid || (n && (!uu || !vv || !w)))
        fail(kErrInvalidArgument, "pointers must not be NULL");
}

size_t Weights::finish_skipped(hipStream_t st)
{
    check_hip(hipMemcpyAsync(h_skipped_, d_small_, sizeof(unsigned long long),
            hipMemcpyDeviceToHost, st), kErrMemCopy, "read skipped");
    check_hip(hipStreamSynchronize(st), kErrKernelLaunch,
            "stream synchronize");
    return (size_t)*h_skipped_;
}

template <typename T>
size_t Weights::write_t(size_t n, const T* uu, const T* vv, const T* weight,
                        T cell, int G, T* grid, hipStream_t st)
{
    check_weights_args(n, uu, vv, weight, G, grid);
    fill_zero_u64(d_small_, 1, st);
    if (n == 0) return finish_skipped(st);
    ensure_capacity(n);
    const T scale = (T)G * cell;
    const unsigned long long cells = (unsigned long long)G * G;
    int bits = 0;
    while ((cells >> bits) != 0) ++bits;
    const int blocks = blocks_for(n, n_cu_);
    hipLaunchKernelGGL(k_weights_key<T>, dim3(blocks), dim3(kThreads), 0,
            st, n, uu, vv, scale, G, cells, d_keys_[0], d_small_);
    const int cur = radix_sort_enqueue(n, bits, d_keys_, d_idx_, d_hist_,
            d_offs_, d_part_, st);
    hipLaunchKernelGGL(k_weights_sum<T>, dim3(blocks), dim3(kThreads), 0,
            st, n, (const unsigned long long*)d_keys_[cur],
            (const unsigned*)d_idx_[cur], weight, cells, grid);
    check_hip(hipGetLastError(), kErrKernelLaunch, "weights launch");
    return finish_skipped(st);
}

template <typename T>
size_t Weights::read_t(size_t n, const T* uu, const T* vv, const T* w_in,
                       T* w_out, T cell, int G, const T* grid, hipStream_t st)
{
    check_weights_args(n, uu, vv, w_in, G, grid);
    if (n && !w_out) fail(kErrInvalidArgument, "weight_out must not be NULL");
    fill_zero_u64(d_small_, 1, st);
    if (n == 0) return finish_skipped(st);
    const T scale = (T)G * cell;
    hipLaunchKernelGGL(k_weights_read<T>, dim3(blocks_for(n, n_cu_)),
            dim3(kThreads), 0, st, n, uu, vv, w_in, w_out, scale, G,
            grid, d_small_);
    check_hip(hipGetLastError(), kErrKernelLaunch, "weights launch");
    return finish_skipped(st);
}

size_t Weights::write(int prec, size_t n, const void* uu, const void* vv,
                      const void* weight, double cell, int G, void* grid,
                      hipStream_t s)
{
    const hipStream_t st = s ? s : stream_;
    if (prec == kPrecSingle)
        return write_t<float>(n, (const float*)uu, (const float*)vv,
                (const float*)weight, (float)cell, G, (float*)grid, st);
    if (prec == kPrecDouble)
        return write_t<double>(n, (const double*)uu, (const double*)vv,
                (const double*)weight, cell, G, (double*)grid, st);
    fail(kErrInvalidArgument, "weights need GW_SINGLE or GW_DOUBLE");
}

size_t Weights::read(int prec, size_t n, const void* uu, const void* vv,
                     const void* w_in, void* w_out, double cell, int G,
                     const void* grid, hipStream_t s)
{
    const hipStream_t st = s ? s : stream_;
    if (prec == kPrecSingle)
        return read_t<float>(n, (const float*)uu, (const float*)vv,
                (const float*)w_in, (float*)w_out, (float)cell, G,
                (const float*)grid, st);
    if (prec == kPrecDouble)
        return read_t<double>(n, (const double*)uu, (const double*)vv,
                (const double*)w_in, (double*)w_out, cell, G,
                (const double*)grid, st);
    fail(kErrInvalidArgument, "weights need GW_SINGLE or GW_DOUBLE");
}

}  // namespace gw
