// C ABI of the MI355X W-projection gridder.
//
// Drop-in entry points with the reference's signatures
// (oskar_grid_wproj.h:68-131, oskar_grid_wproj.c:245-267), OSKAR-style
// status variants, the resident plan API and device-memory helpers.  The
// drop-in calls take HOST pointers: each call uploads the points, runs the
// device pipeline of gridder.hip, and moves only the grid rows that the
// footprints touch across PCIe (read-modify-write of the caller's
// accumulating grid).  A per-thread cache keeps the plan (device W-kernel
// cube and tables) and the device buffers between calls with the same
// kernel set, so repeated calls per visibility block (main.cpp:413-535)
// upload the cube once.
#include "gridder_internal.h"
#include "server.h"

#include "fpga_gridding_amd.h"
#include "oskar_grid_weights.h"

#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

namespace {

thread_local int t_last_error = 0;
thread_local std::string t_last_detail;

void set_error(int code, const std::string& detail)
{
    t_last_error = code;
    t_last_detail = detail;
}

// Device buffer that only grows.
struct DevBuf
{
    void* ptr = nullptr;
    size_t bytes = 0;
    void* get(size_t need)
    {
        if (need > bytes)
        {
            if (ptr) (void)hipFree(ptr);
            ptr = nullptr;
            bytes = 0;
            gw::check_hip(hipMalloc(&ptr, std::max<size_t>(need, 256)),
                    gw::kErrMemAlloc, "hipMalloc");
            bytes = std::max<size_t>(need, 256);
        }
        return ptr;
    }
    void release()
    {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
};

// Page-locked host buffer that only grows.
struct PinBuf
{
    void* ptr = nullptr;
    size_t bytes = 0;
    void* get(size_t need)
    {
        if (need > bytes)
        {
            if (ptr) (void)hipHostFree(ptr);
            ptr = nullptr;
            bytes = 0;
            gw::check_hip(hipHostMalloc(&ptr, need, hipHostMallocDefault),
                    gw::kErrMemAlloc, "hipHostMalloc");
            bytes = need;
        }
        return ptr;
    }
    void release()
    {
        if (ptr) (void)hipHostFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
};

// A fixed team of host threads for the drop-in's copies between pageable
// and page-locked memory and its grid additions: run(f) calls f(t, nt) for
// t = 0 .. nt-1 (t = 0 on the calling thread) and returns when all are
// done.  Process-wide, created on first use; its threads are detached and
// never stopped (they idle on a condition variable).  Calls from several
// host threads take turns.
class HostPool
{
public:
    static HostPool& get()
    {
        static HostPool* pool = new HostPool();
        return *pool;
    }
    int size() const { return nt_; }
    void run(const std::function<void(int, int)>& f)
    {
        std::lock_guard<std::mutex> turn(turn_);
        if (nt_ == 1)
        {
            f(0, 1);
            return;
        }
        {
            std::lock_guard<std::mutex> lk(m_);
            job_ = &f;
            pending_ = nt_ - 1;
            ++gen_;
        }
        cv_.notify_all();
        f(0, nt_);
        std::unique_lock<std::mutex> lk(m_);
        done_.wait(lk, [&] { return pending_ == 0; });
        job_ = nullptr;
    }

private:
    HostPool()
    {
        const unsigned hw = std::thread::hardware_concurrency();
        nt_ = (int)std::max(1u, std::min(hw ? hw : 1u, 16u));
        for (int t = 1; t < nt_; ++t)
            std::thread([this, t] { work(t); }).detach();
    }
    void work(int t)
    {
        unsigned seen = 0;
        for (;;)
        {
            const std::function<void(int, int)>* f;
            {
                std::unique_lock<std::mutex> lk(m_);
                cv_.wait(lk, [&] { return gen_ != seen; });
                seen = gen_;
                f = job_;
            }
            (*f)(t, nt_);
            std::lock_guard<std::mutex> lk(m_);
            if (--pending_ == 0) done_.notify_one();
        }
    }
    int nt_ = 1;
    std::mutex turn_, m_;
    std::condition_variable cv_, done_;
    const std::function<void(int, int)>* job_ = nullptr;
    unsigned gen_ = 0;
    int pending_ = 0;
};

// memcpy of n bytes split over the pool's threads.
void pool_memcpy(void* dst, const void* src, size_t n)
{
    HostPool& pool = HostPool::get();
    if (n < ((size_t)1 << 20) || pool.size() == 1)
    {
        std::memcpy(dst, src, n);
        return;
    }
    pool.run([&](int t, int nt) {
        const size_t per = (n / nt + 63) & ~(size_t)63;
        const size_t a = std::min(n, per * t), b = std::min(n, a + per);
        if (b > a)
            std::memcpy(static_cast<char*>(dst) + a,
                        static_cast<const char*>(src) + a, b - a);
    });
}

// memcmp == 0 of n bytes split over the pool's threads.
bool pool_equal(const void* x, const void* y, size_t n)
{
    HostPool& pool = HostPool::get();
    if (n < ((size_t)1 << 20) || pool.size() == 1)
        return std::memcmp(x, y, n) == 0;
    std::vector<char> diff(pool.size(), 0);
    pool.run([&](int t, int nt) {
        const size_t per = (n / nt + 63) & ~(size_t)63;
        const size_t a = std::min(n, per * t), b = std::min(n, a + per);
        if (b > a && std::memcmp(static_cast<const char*>(x) + a,
                                 static_cast<const char*>(y) + a, b - a))
            diff[t] = 1;
    });
    for (char d : diff)
        if (d) return false;
    return true;
}

// Per-thread state of the drop-in calls.  Deliberately never destroyed at
// thread or process exit (the HIP runtime may be gone by then); released
// explicitly by gw_release_cached().
struct DropinCache
{
    std::unique_ptr<gw::Plan> plan;
    int prec = 0;
    size_t P = 0;
    std::vector<int> support;
    int ov = 0, C = 0, G = 0;
    double cell = 0, ws = 0;
    std::vector<char> cube;
    int device = -1;
    DevBuf in, grid, d_support_hist, d_plane_hist, d_cell_count, d_cell_wsum;
    // Page-locked staging of the drop-in copies (two slots taking turns)
    // and the events that free a slot.
    PinBuf stage[2];
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
    bool stage_live[2] = {false, false};
    // Waits until slot k is free (its last DMA done).
    void stage_wait(int k)
    {
        if (stage_live[k])
            gw::check_hip(hipEventSynchronize(stage_ev[k]),
                    gw::kErrKernelLaunch, "event synchronize");
    }
    // Marks slot k busy until the work queued on s so far is done.
    void stage_mark(int k, hipStream_t s)
    {
        if (!stage_ev[k])
            gw::check_hip(hipEventCreateWithFlags(&stage_ev[k],
                    hipEventDisableTiming), gw::kErrComputeDevices,
                    "hipEventCreate");
        gw::check_hip(hipEventRecord(stage_ev[k], s), gw::kErrKernelLaunch,
                "event record");
        stage_live[k] = true;
    }
    // Uniform weighting.
    std::unique_ptr<gw::Weights> weights;
    int weights_device = -1;
    DevBuf w_in, w_grid;

    gw::Weights& weights_for(int dev)
    {
        if (!weights || weights_device != dev)
        {
            weights.reset();
            weights.reset(new gw::Weights());
            weights_device = dev;
        }
        return *weights;
    }

    bool matches(int prec_, size_t P_, const int* support_, int ov_, int C_,
                 const void* conv, size_t cube_bytes, double cell_,
                 double ws_, int G_, int dev) const
    {
        return plan && prec == prec_ && P == P_ && ov == ov_ && C == C_ &&
                G == G_ && cell == cell_ && ws == ws_ && device == dev &&
                cube.size() == cube_bytes &&
                std::equal(support.begin(), support.end(), support_) &&
                pool_equal(cube.data(), conv, cube_bytes);
    }
};

thread_local DropinCache* t_cache = nullptr;

DropinCache& cache()
{
    if (!t_cache) t_cache = new DropinCache();
    return *t_cache;
}

int current_device()
{
    int dev = 0;
    gw::check_hip(hipGetDevice(&dev), gw::kErrComputeDevices, "hipGetDevice");
    return dev;
}

// oskar_grid_weights_write_* on host arrays: points and the caller's grid
// uploaded, the grid accumulated on the device and copied back.
template <typename T>
void weights_write_host(int prec, size_t n, const T* uu, const T* vv,
                        const T* weight, T cell, int G, size_t* num_skipped,
                        T* grid)
{
    if (!num_skipped || !grid)
        gw::fail(gw::kErrInvalidArgument, "num_skipped and grid must not be "
                 "NULL");
    if (n && (!uu || !vv || !weight))
        gw::fail(gw::kErrInvalidArgument, "uu, vv and weight must not be "
                 "NULL");
    if (G < 1 || G > 65535)
        gw::fail(gw::kErrInvalidArgument, "grid_size must be in [1, 65535]");
    DropinCache& c = cache();
    gw::Weights& w = c.weights_for(current_device());
    const hipStream_t s = w.stream();
    const size_t gbytes = (size_t)G * G * sizeof(T);
    T* d_in = static_cast<T*>(c.w_in.get(3 * std::max<size_t>(n, 1) *
                                         sizeof(T)));
    T* d_grid = static_cast<T*>(c.w_grid.get(gbytes));
    if (n)
    {
        gw::check_hip(hipMemcpyAsync(d_in, uu, n * sizeof(T),
                hipMemcpyHostToDevice, s), gw::kErrMemCopy, "upload uu");
        gw::check_hip(hipMemcpyAsync(d_in + n, vv, n * sizeof(T),
                hipMemcpyHostToDevice, s), gw::kErrMemCopy, "upload vv");
        gw::check_hip(hipMemcpyAsync(d_in + 2 * n, weight, n * sizeof(T),
                hipMemcpyHostToDevice, s), gw::kErrMemCopy, "upload weight");
    }
    gw::check_hip(hipMemcpyAsync(d_grid, grid, gbytes, hipMemcpyHostToDevice,
            s), gw::kErrMemCopy, "upload grid");
    const size_t sk = w.write(prec, n, d_in, d_in + n, d_in + 2 * n,
                              (double)cell, G, d_grid);
    gw::check_hip(hipMemcpyAsync(grid, d_grid, gbytes, hipMemcpyDeviceToHost,
            s), gw::kErrMemCopy, "download grid");
    gw::check_hip(hipStreamSynchronize(s), gw::kErrKernelLaunch,
            "stream synchronize");
    *num_skipped = sk;
}

// oskar_grid_weights_read_*: weight_out is uploaded too, so that entries of
// skipped points come back unchanged.
template <typename T>
void weights_read_host(int prec, size_t n, const T* uu, const T* vv,
                       const T* w_in, T* w_out, T cell, int G,
                       size_t* num_skipped, const T* grid)
{
    if (!num_skipped || !grid)
        gw::fail(gw::kErrInvalidArgument, "num_skipped and grid must not be "
                 "NULL");
    if (n && (!uu || !vv || !w_in || !w_out))
        gw::fail(gw::kErrInvalidArgument, "uu, vv, weight_in and weight_out "
                 "must not be NULL");
    if (G < 1 || G > 65535)
        gw::fail(gw::kErrInvalidArgument, "grid_size must be in [1, 65535]");
    DropinCache& c = cache();
    gw::Weights& w = c.weights_for(current_device());
    const hipStream_t s = w.stream();
    const size_t gbytes = (size_t)G * G * sizeof(T);
    T* d_in = static_cast<T*>(c.w_in.get(4 * std::max<size_t>(n, 1) *
                                         sizeof(T)));
    T* d_grid = static_cast<T*>(c.w_grid.get(gbytes));
    if (n)
    {
        gw::check_hip(hipMemcpyAsync(d_in, uu, n * sizeof(T),
                hipMemcpyHostToDevice, s), gw::kErrMemCopy, "upload uu");
        gw::check_hip(hipMemcpyAsync(d_in + n, vv, n * sizeof(T),
                hipMemcpyHostToDevice, s), gw::kErrMemCopy, "upload vv");
        gw::check_hip(hipMemcpyAsync(d_in + 2 * n, w_in, n * sizeof(T),
                hipMemcpyHostToDevice, s), gw::kErrMemCopy,
                "upload weight_in");
        gw::check_hip(hipMemcpyAsync(d_in + 3 * n, w_out, n * sizeof(T),
                hipMemcpyHostToDevice, s), gw::kErrMemCopy,
                "upload weight_out");
    }
    gw::check_hip(hipMemcpyAsync(d_grid, grid, gbytes, hipMemcpyHostToDevice,
            s), gw::kErrMemCopy, "upload grid");
    const size_t sk = w.read(prec, n, d_in, d_in + n, d_in + 2 * n,
                             d_in + 3 * n, (double)cell, G, d_grid);
    if (n)
        gw::check_hip(hipMemcpyAsync(w_out, d_in + 3 * n, n * sizeof(T),
                hipMemcpyDeviceToHost, s), gw::kErrMemCopy,
                "download weight_out");
    gw::check_hip(hipStreamSynchronize(s), gw::kErrKernelLaunch,
            "stream synchronize");
    *num_skipped = sk;
}

// Staging slot size of the drop-in copies.
constexpr size_t kStageBytes = (size_t)32 << 20;

// Host arrays to device through the two page-locked slots: each chunk is
// copied into a slot by the host pool while the previous chunk's DMA runs.
struct Upload
{
    void* dst;
    const void* src;
    size_t bytes;
};

void upload_staged(DropinCache& c, hipStream_t s,
                   std::initializer_list<Upload> parts)
{
    int k = 0;
    for (const Upload& u : parts)
    {
        for (size_t off = 0; off < u.bytes; off += kStageBytes, ++k)
        {
            const size_t len = std::min(kStageBytes, u.bytes - off);
            const int slot = k & 1;
            c.stage_wait(slot);
            void* pin = c.stage[slot].get(kStageBytes);
            pool_memcpy(pin, static_cast<const char*>(u.src) + off, len);
            gw::check_hip(hipMemcpyAsync(static_cast<char*>(u.dst) + off, pin,
                    len, hipMemcpyHostToDevice, s), gw::kErrMemCopy,
                    "upload points");
            c.stage_mark(slot, s);
        }
    }
}

// The device grid rows (a delta gridded from zero, pitch G cells) added
// into the caller's grid: only columns col0 .. col0 + cols - 1 come back,
// in chunks of rows through the two slots, and the host pool adds each
// chunk's nonzero cells while the next chunk's DMA runs (cells nothing
// reached stay as the caller left them).
template <typename T>
void download_add(DropinCache& c, hipStream_t s, T* grid, const T* d_grid,
                  int row0, int rows, int col0, int cols, int G)
{
    const size_t rb = (size_t)cols * 2 * sizeof(T);
    const int per = (int)std::max<size_t>(1, kStageBytes / rb);
    const int chunks = (rows + per - 1) / per;
    auto issue = [&](int k) {
        const int slot = k & 1;
        c.stage_wait(slot);
        void* pin = c.stage[slot].get(kStageBytes);
        const int r = k * per;
        const int nr = std::min(per, rows - r);
        gw::check_hip(hipMemcpy2DAsync(pin, rb,
                d_grid + ((size_t)r * G + col0) * 2, (size_t)G * 2 * sizeof(T),
                rb, (size_t)nr, hipMemcpyDeviceToHost, s), gw::kErrMemCopy,
                "download grid rows");
        c.stage_mark(slot, s);
    };
    issue(0);
    if (chunks > 1) issue(1);
    for (int k = 0; k < chunks; ++k)
    {
        const int slot = k & 1;
        c.stage_wait(slot);
        const T* pin = static_cast<const T*>(c.stage[slot].ptr);
        const int r0 = k * per;
        const int nr = std::min(per, rows - r0);
        HostPool::get().run([&](int t, int nt) {
            for (int r = t; r < nr; r += nt)
            {
                T* g = grid + ((size_t)(row0 + r0 + r) * G + col0) * 2;
                const T* d = pin + (size_t)r * cols * 2;
                for (int i = 0; i < cols; ++i)
                {
                    const T re = d[2 * i], im = d[2 * i + 1];
                    if (re != (T)0 || im != (T)0)
                    {
                        g[2 * i] += re;
                        g[2 * i + 1] += im;
                    }
                }
            }
        });
        if (k + 2 < chunks) issue(k + 2);
    }
}

// Caller-side diagnostics arrays of oskar_grid_wproj_f_diagnostics.
struct DiagHost
{
    long long* per_support;
    long long* per_plane;
    long long* per_cell;
    float* wsum_cell;
};

template <typename T>
void run_dropin(int prec, size_t P, const int* support, int ov, int C,
                const T* conv, size_t n, const T* uu, const T* vv,
                const T* ww, const T* vis, T cell, T ws, int G,
                size_t* num_skipped, double* norm, T* grid,
                const DiagHost* dh)
{
    if (!num_skipped || !norm || !grid)
        gw::fail(gw::kErrInvalidArgument,
                "num_skipped, norm and grid must not be NULL");
    if (n > 0 && (!uu || !vv || !ww || !vis))
        gw::fail(gw::kErrInvalidArgument, "uu, vv, ww and vis must not be NULL");
    if (dh && (!dh->per_support || !dh->per_plane || !dh->per_cell ||
            !dh->wsum_cell))
        gw::fail(gw::kErrInvalidArgument, "diagnostics arrays must not be NULL");
    gw::validate_args(P, support, ov, C, conv, G);

    int dev = 0;
    gw::check_hip(hipGetDevice(&dev), gw::kErrComputeDevices, "hipGetDevice");
    DropinCache& c = cache();
    const size_t cube_bytes = 2 * P * (size_t)C * C * sizeof(T);
    if (!c.matches(prec, P, support, ov, C, conv, cube_bytes, (double)cell,
            (double)ws, G, dev))
    {
        c.plan.reset();
        c.plan.reset(new gw::Plan(prec, P, support, ov, C, conv, (double)cell,
                (double)ws, G));
        c.prec = prec;
        c.P = P;
        c.support.assign(support, support + P);
        c.ov = ov;
        c.C = C;
        c.G = G;
        c.cell = (double)cell;
        c.ws = (double)ws;
        c.device = dev;
        c.cube.assign((const char*)conv, (const char*)conv + cube_bytes);
    }
    gw::Plan& plan = *c.plan;
    hipStream_t s = plan.stream();

    if (n == 0)
    {
        // The reference resets the count and grids nothing (:153-155).
        *num_skipped = 0;
        return;
    }

    // Inputs: uu | vv | ww | vis, one device allocation, staged through
    // page-locked memory.
    T* d_in = (T*)c.in.get(5 * n * sizeof(T));
    upload_staged(c, s, {{d_in, uu, n * sizeof(T)},
                         {d_in + n, vv, n * sizeof(T)},
                         {d_in + 2 * n, ww, n * sizeof(T)},
                         {d_in + 3 * n, vis, 2 * n * sizeof(T)}});

    gw::DiagBuffers diag;
    if (dh)
    {
        diag.max_support = plan.max_support();
        diag.plane_cap = (int)P + 1024;
        diag.per_support = (unsigned long long*)c.d_support_hist.get(
                (diag.max_support + 1) * sizeof(unsigned long long));
        diag.per_plane = (unsigned long long*)c.d_plane_hist.get(
                diag.plane_cap * sizeof(unsigned long long));
        gw::fill_zero_u64(diag.per_support, diag.max_support + 1, s);
        gw::fill_zero_u64(diag.per_plane, diag.plane_cap, s);
    }

    plan.locate(n, d_in, d_in + n, d_in + 2 * n, d_in + 3 * n,
            dh ? &diag : nullptr);
    gw::Summary sm = plan.read_summary();
    if (dh && (sm.flags & 1u))
        gw::fail(gw::kErrOutOfRange,
                "diagnostics: a plane index exceeds num_w_planes + 1023");

    const bool any = sm.skipped < n;
    int row0 = 0, rows = 0;
    if (any)
    {
        row0 = sm.v_min;
        rows = sm.v_max - sm.v_min + 1;
        const size_t row_vals = 2 * (size_t)G;
        T* d_grid = (T*)c.grid.get((size_t)rows * row_vals * sizeof(T));
        // Diagnostics: the caller's rows go up and come back (the
        // reference's accumulation on the device).  Otherwise the rows are
        // gridded from zero and the delta added on the host.
        if (dh)
            gw::check_hip(hipMemcpyAsync(d_grid, grid + (size_t)row0 *
                    row_vals, (size_t)rows * row_vals * sizeof(T),
                    hipMemcpyHostToDevice, s), gw::kErrMemCopy,
                    "upload grid rows");
        else
            gw::check_hip(hipMemsetAsync(d_grid, 0, (size_t)rows * row_vals *
                    sizeof(T), s), gw::kErrMemCopy, "clear grid rows");
        if (dh)
        {
            const size_t cells = (size_t)rows * G;
            diag.per_cell = (unsigned*)c.d_cell_count.get(
                    cells * sizeof(unsigned));
            diag.wsum_cell = (unsigned long long*)c.d_cell_wsum.get(
                    cells * sizeof(unsigned long long));
            gw::fill_zero_u32(diag.per_cell, cells, s);
            gw::fill_zero_u64(diag.wsum_cell, cells, s);
        }
        plan.scatter(n, d_grid, row0, dh ? &diag : nullptr);
        sm = plan.read_summary();
        if (sm.flags & 2u)
            gw::fail(gw::kErrOutOfRange, "internal: work item table overflow");
        if (dh)
            gw::check_hip(hipMemcpyAsync(grid + (size_t)row0 * row_vals,
                    d_grid, (size_t)rows * row_vals * sizeof(T),
                    hipMemcpyDeviceToHost, s), gw::kErrMemCopy,
                    "download grid rows");
        else
            download_add(c, s, grid, d_grid, row0, rows, sm.u_min,
                         sm.u_max - sm.u_min + 1, G);
    }
    if (dh)
    {
        std::vector<unsigned long long> hs(diag.max_support + 1),
                hp(diag.plane_cap);
        gw::check_hip(hipMemcpyAsync(hs.data(), diag.per_support,
                hs.size() * sizeof(unsigned long long),
                hipMemcpyDeviceToHost, s), gw::kErrMemCopy, "download hist");
        gw::check_hip(hipMemcpyAsync(hp.data(), diag.per_plane,
                hp.size() * sizeof(unsigned long long),
                hipMemcpyDeviceToHost, s), gw::kErrMemCopy, "download hist");
        std::vector<unsigned> hc;
        std::vector<unsigned long long> hw;
        if (any)
        {
            const size_t cells = (size_t)rows * G;
            hc.resize(cells);
            hw.resize(cells);
            gw::check_hip(hipMemcpyAsync(hc.data(), diag.per_cell,
                    cells * sizeof(unsigned), hipMemcpyDeviceToHost, s),
                    gw::kErrMemCopy, "download cell counts");
            gw::check_hip(hipMemcpyAsync(hw.data(), diag.wsum_cell,
                    cells * sizeof(unsigned long long),
                    hipMemcpyDeviceToHost, s), gw::kErrMemCopy,
                    "download cell sums");
        }
        gw::check_hip(hipStreamSynchronize(s), gw::kErrKernelLaunch,
                "stream synchronize");
        // Touch only the entries the reference would touch.
        for (size_t i = 0; i < hs.size(); ++i)
            if (hs[i]) dh->per_support[i] += (long long)hs[i];
        for (size_t i = 0; i < hp.size(); ++i)
            if (hp[i]) dh->per_plane[i] += (long long)hp[i];
        for (size_t i = 0; i < hc.size(); ++i)
        {
            if (!hc[i]) continue;
            const size_t cell_i = (size_t)row0 * G + i;
            dh->per_cell[cell_i] += (long long)hc[i];
            dh->wsum_cell[cell_i] += (float)hw[i];
        }
    }
    else
    {
        gw::check_hip(hipStreamSynchronize(s), gw::kErrKernelLaunch,
                "stream synchronize");
    }
    *num_skipped = (size_t)sm.skipped;
    *norm += sm.norm_delta;
}

template <typename F>
void guarded(int* status, F&& f)
{
    try
    {
        f();
    }
    catch (const gw::Error& e)
    {
        *status = e.code;
        set_error(e.code, e.detail);
    }
    catch (const std::bad_alloc&)
    {
        *status = gw::kErrMemAlloc;
        set_error(gw::kErrMemAlloc, "host allocation failed");
    }
    catch (...)
    {
        *status = gw::kErrFunctionNotAvailable;
        set_error(gw::kErrFunctionNotAvailable, "unexpected exception");
    }
}

}  // namespace

struct gw_plan
{
    gw::Plan* impl;
};

struct gw_batch
{
    gw::Batch* impl;
};

struct gw_server
{
    gw::Server* impl;
};

extern "C" {

GW_EXPORT
void oskar_grid_wproj_f_status(const size_t num_w_planes, const int* support,
        const int oversample, const int conv_size_half,
        const float* conv_func, const size_t num_points, const float* uu,
        const float* vv, const float* ww, const float* vis,
        const float* weight, const float cell_size_rad, const float w_scale,
        const int grid_size, size_t* num_skipped, double* norm, float* grid,
        int* status)
{
    int local = 0;
    if (!status) status = &local;
    if (*status) return;
    (void)weight;
    guarded(status, [&] {
        run_dropin<float>(GW_SINGLE, num_w_planes, support, oversample,
                conv_size_half, conv_func, num_points, uu, vv, ww, vis,
                cell_size_rad, w_scale, grid_size, num_skipped, norm, grid,
                nullptr);
    });
}

GW_EXPORT
void oskar_grid_wproj_d_status(const size_t num_w_planes, const int* support,
        const int oversample, const int conv_size_half,
        const double* conv_func, const size_t num_points, const double* uu,
        const double* vv, const double* ww, const double* vis,
        const double* weight, const double cell_size_rad,
        const double w_scale, const int grid_size, size_t* num_skipped,
        double* norm, double* grid, int* status)
{
    int local = 0;
    if (!status) status = &local;
    if (*status) return;
    (void)weight;
    guarded(status, [&] {
        run_dropin<double>(GW_DOUBLE, num_w_planes, support, oversample,
                conv_size_half, conv_func, num_points, uu, vv, ww, vis,
                cell_size_rad, w_scale, grid_size, num_skipped, norm, grid,
                nullptr);
    });
}

GW_EXPORT
void oskar_grid_wproj_f(const size_t num_w_planes, const int* support,
        const int oversample, const int conv_size_half,
        const float* conv_func, const size_t num_points, const float* uu,
        const float* vv, const float* ww, const float* vis,
        const float* weight, const float cell_size_rad, const float w_scale,
        const int grid_size, size_t* num_skipped, double* norm, float* grid)
{
    int status = 0;
    t_last_error = 0;
    t_last_detail.clear();
    oskar_grid_wproj_f_status(num_w_planes, support, oversample,
            conv_size_half, conv_func, num_points, uu, vv, ww, vis, weight,
            cell_size_rad, w_scale, grid_size, num_skipped, norm, grid,
            &status);
    t_last_error = status;
}

GW_EXPORT
void oskar_grid_wproj_d(const size_t num_w_planes, const int* support,
        const int oversample, const int conv_size_half,
        const double* conv_func, const size_t num_points, const double* uu,
        const double* vv, const double* ww, const double* vis,
        const double* weight, const double cell_size_rad,
        const double w_scale, const int grid_size, size_t* num_skipped,
        double* norm, double* grid)
{
    int status = 0;
    t_last_error = 0;
    t_last_detail.clear();
    oskar_grid_wproj_d_status(num_w_planes, support, oversample,
            conv_size_half, conv_func, num_points, uu, vv, ww, vis, weight,
            cell_size_rad, w_scale, grid_size, num_skipped, norm, grid,
            &status);
    t_last_error = status;
}

GW_EXPORT
void oskar_grid_wproj_f_diagnostics(const size_t num_w_planes,
        const int* support, const int oversample, const int conv_size_half,
        const float* conv_func, const size_t num_points, const float* uu,
        const float* vv, const float* ww, const float* vis,
        const float* weight, const float cell_size_rad, const float w_scale,
        const int grid_size, size_t* num_skipped, double* norm, float* grid,
        long long int* num_vis_per_wsupport, long long int* num_vis_per_wplane,
        long long int* num_vis_per_grid, float* avg_wsupport_per_grid,
        float* avg_wplane_per_grid)
{
    int status = 0;
    t_last_error = 0;
    t_last_detail.clear();
    (void)weight;
    (void)avg_wplane_per_grid;
    DiagHost dh{num_vis_per_wsupport, num_vis_per_wplane, num_vis_per_grid,
                avg_wsupport_per_grid};
    guarded(&status, [&] {
        run_dropin<float>(GW_SINGLE, num_w_planes, support, oversample,
                conv_size_half, conv_func, num_points, uu, vv, ww, vis,
                cell_size_rad, w_scale, grid_size, num_skipped, norm, grid,
                &dh);
    });
    t_last_error = status;
}

GW_EXPORT int oskar_grid_wproj_last_error(void)
{
    return t_last_error;
}

GW_EXPORT const char* gw_last_error_detail(void)
{
    return t_last_detail.c_str();
}

GW_EXPORT const char* gw_error_string(int status)
{
    switch (status)
    {
    case GW_OK: return "no error";
    case GW_ERR_INVALID_ARGUMENT: return "invalid function argument";
    case GW_ERR_FUNCTION_NOT_AVAILABLE: return "function not available";
    case GW_ERR_OUT_OF_RANGE: return "value out of range";
    case GW_ERR_MEMORY_ALLOC_FAILURE: return "memory allocation failure";
    case GW_ERR_MEMORY_COPY_FAILURE: return "memory copy failure";
    case GW_ERR_KERNEL_LAUNCH_FAILURE: return "kernel launch failure";
    case GW_ERR_COMPUTE_DEVICES: return "compute device error";
    default: return "unknown error";
    }
}

GW_EXPORT void gw_release_cached(void)
{
    if (t_cache)
    {
        t_cache->plan.reset();
        t_cache->weights.reset();
        t_cache->weights_device = -1;
        t_cache->w_in.release();
        t_cache->w_grid.release();
        t_cache->in.release();
        t_cache->grid.release();
        t_cache->d_support_hist.release();
        t_cache->d_plane_hist.release();
        t_cache->d_cell_count.release();
        t_cache->d_cell_wsum.release();
        for (int k = 0; k < 2; ++k)
        {
            if (t_cache->stage_ev[k])
            {
                (void)hipEventSynchronize(t_cache->stage_ev[k]);
                (void)hipEventDestroy(t_cache->stage_ev[k]);
            }
            t_cache->stage_live[k] = false;
            t_cache->stage[k].release();
        }
        delete t_cache;
        t_cache = nullptr;
    }
}

GW_EXPORT
gw_plan* gw_plan_create(int precision, size_t num_w_planes,
        const int* support, int oversample, int conv_size_half,
        const void* conv_func, double cell_size_rad, double w_scale,
        int grid_size, int* status)
{
    int local = 0;
    if (!status) status = &local;
    if (*status) return nullptr;
    gw_plan* out = nullptr;
    guarded(status, [&] {
        std::unique_ptr<gw::Plan> p(new gw::Plan(precision, num_w_planes,
                support, oversample, conv_size_half, conv_func,
                cell_size_rad, w_scale, grid_size));
        out = new gw_plan{p.release()};
    });
    return out;
}

GW_EXPORT
gw_plan* gw_plan_create_device_cube(int precision, size_t num_w_planes,
        const int* support, int oversample, int conv_size_half,
        const float* d_conv_func, double cell_size_rad, double w_scale,
        int grid_size, int* status)
{
    int local = 0;
    if (!status) status = &local;
    if (*status) return nullptr;
    gw_plan* out = nullptr;
    guarded(status, [&] {
        std::unique_ptr<gw::Plan> p(new gw::Plan(precision, num_w_planes,
                support, oversample, conv_size_half, d_conv_func,
                cell_size_rad, w_scale, grid_size, true));
        out = new gw_plan{p.release()};
    });
    return out;
}

GW_EXPORT
void gw_wkernel_generate(size_t num_w_planes, const int* support,
        int oversample, int conv_size_half, double w_max, float* d_conv_func,
        void* hip_stream, int* status)
{
    int local = 0;
    if (!status) status = &local;
    if (*status) return;
    guarded(status, [&] {
        gw::generate_wkernels(num_w_planes, support, oversample,
                conv_size_half, w_max, d_conv_func,
                static_cast<hipStream_t>(hip_stream));
    });
}

GW_EXPORT void gw_plan_free(gw_plan* plan)
{
    if (!plan) return;
    delete plan->impl;
    delete plan;
}

GW_EXPORT void gw_plan_set_stream(gw_plan* plan, void* hip_stream)
{
    if (!plan) return;
    int status = 0;
    guarded(&status, [&] {
        plan->impl->set_stream(static_cast<hipStream_t>(hip_stream));
    });
}

GW_EXPORT
void gw_grid_device(gw_plan* plan, size_t num_points, const void* d_uu,
        const void* d_vv, const void* d_ww, const void* d_vis, void* d_grid,
        int* status)
{
    int local = 0;
    if (!status) status = &local;
    if (*status) return;
    guarded(status, [&] {
        if (!plan) gw::fail(gw::kErrInvalidArgument, "plan is NULL");
        if (!d_grid || (num_points && (!d_uu || !d_vv || !d_ww || !d_vis)))
            gw::fail(gw::kErrInvalidArgument, "device pointers must not be NULL");
        plan->impl->locate(num_points, d_uu, d_vv, d_ww, d_vis, nullptr);
        plan->impl->scatter(num_points, d_grid, 0, nullptr);
    });
}

// -------------------------------------------------------------------------
// Allocation-free gridding call
// -------------------------------------------------------------------------

namespace {

// The per-call workspace of gw_grid_device_noalloc (sort keys and values,
// point records, tile tables, work items, sort temporaries): one copy of
// this array per GPU the library runs on, held until the process ends.
// In host code its name is a stand-in for the symbol; the device address
// comes from hipGetSymbolAddress (once per device) and only that address is
// passed to kernels.
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device` (looked up with that device
// current on its first use, the caller's current device restored).
void* scratch_on(int device)
{
    if (device < 0 || device >= kMaxDevices)
        gw::fail(gw::kErrComputeDevices, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device])
    {
        int cur = 0;
        gw::check_hip(hipGetDevice(&cur), gw::kErrComputeDevices,
                      "hipGetDevice");
        if (cur != device)
            gw::check_hip(hipSetDevice(device), gw::kErrComputeDevices,
                          "hipSetDevice");
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        gw::check_hip(e, gw::kErrComputeDevices, "hipGetSymbolAddress");
        if (!p)
            gw::fail(gw::kErrComputeDevices, "scratch has no device address");
        g_scratch_ptr[device] = p;
    }
    return g_scratch_ptr[device];
}

// Most points of one gw_grid_device_noalloc call with this plan.
size_t noalloc_max_points(const gw::Plan& p)
{
    size_t lo = 0, hi = kNoAllocScratchBytes;
    while (lo < hi)
    {
        const size_t mid = lo + (hi - lo + 1) / 2;
        if (p.external_workspace_bytes(mid) <= kNoAllocScratchBytes) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

}  // namespace

GW_EXPORT
void gw_grid_device_noalloc(gw_plan* plan, size_t num_points,
        const void* d_uu, const void* d_vv, const void* d_ww,
        const void* d_vis, void* d_grid, int* status)
{
    int local = 0;
    if (!status) status = &local;
    if (*status) return;
    guarded(status, [&] {
        if (!plan) gw::fail(gw::kErrInvalidArgument, "plan is NULL");
        if (!d_grid || (num_points && (!d_uu || !d_vv || !d_ww || !d_vis)))
            gw::fail(gw::kErrInvalidArgument, "device pointers must not be NULL");
        gw::Plan& p = *plan->impl;
        if (p.external_workspace_bytes(num_points) > kNoAllocScratchBytes)
            gw::fail(gw::kErrOutOfRange, "gw_grid_device_noalloc: " +
                     std::to_string(num_points) + " points exceed its " +
                     std::to_string(kNoAllocScratchBytes >> 20) +
                     " MiB scratch (at most " +
                     std::to_string(noalloc_max_points(p)) +
                     " points per call with this plan)");
        void* base = scratch_on(p.device());
        struct Restore
        {
            gw::Plan& p;
            ~Restore() { p.restore_workspace(); }
        } restore{p};
        p.use_external_workspace(base, kNoAllocScratchBytes, num_points);
        p.locate(num_points, d_uu, d_vv, d_ww, d_vis, nullptr);
        p.scatter(num_points, d_grid, 0, nullptr);
    });
}

GW_EXPORT
void gw_plan_set_tuning(gw_plan* plan, int tile_log2, int slot_cap,
        int min_share, int* status)
{
    int local = 0;
    if (!status) status = &local;
    if (*status) return;
    guarded(status, [&] {
        if (!plan) gw::fail(gw::kErrInvalidArgument, "plan is NULL");
        plan->impl->configure(tile_log2, slot_cap, min_share);
    });
}

GW_EXPORT
int gw_plan_get_tuning(gw_plan* plan, int* values, int max_n)
{
    if (!plan || !values || max_n <= 0) return 0;
    const gw::Plan& p = *plan->impl;
    const int t = p.tile();
    const int v[7] = {(t & (t - 1)) == 0 ? p.tile_bits() : -1, p.slot_cap(),
                      p.min_share(), p.region(), p.max_tile_bits(), t,
                      p.max_tile()};
    const int k = max_n < 7 ? max_n : 7;
    for (int i = 0; i < k; ++i) values[i] = v[i];
    return k;
}

GW_EXPORT
void gw_plan_autotune(gw_plan* plan, size_t num_points, const void* d_uu,
        const void* d_vv, const void* d_ww, const void* d_vis,
        void* d_scratch_grid, int reps, double* best_ms, int* status)
{
    int local = 0;
    if (!status) status = &local;
    if (*status) return;
    guarded(status, [&] {
        if (!plan) gw::fail(gw::kErrInvalidArgument, "plan is NULL");
        if (!d_scratch_grid || (num_points && (!d_uu || !d_vv || !d_ww ||
                !d_vis)))
            gw::fail(gw::kErrInvalidArgument, "device pointers must not be NULL");
        const double ms = plan->impl->autotune(num_points, d_uu, d_vv, d_ww,
                d_vis, d_scratch_grid, reps);
        if (best_ms) *best_ms = ms;
    });
}

GW_EXPORT
void gw_plan_results(gw_plan* plan, size_t* num_skipped, double* norm_delta,
        unsigned long long* num_updates, int* status)
{
    int local = 0;
    if (!status) status = &local;
    if (*status) return;
    guarded(status, [&] {
        if (!plan) gw::fail(gw::kErrInvalidArgument, "plan is NULL");
        const gw::Summary& sm = plan->impl->read_summary();
        if (sm.flags & 2u)
            gw::fail(gw::kErrOutOfRange, "internal: work item table overflow");
        if (num_skipped) *num_skipped = (size_t)sm.skipped;
        if (norm_delta) *norm_delta = sm.norm_delta;
        if (num_updates) *num_updates = sm.updates;
    });
}

GW_EXPORT
void gw_plan_records(gw_plan* plan, unsigned long long* num_records,
        unsigned long long* record_updates, int* status)
{
    int local = 0;
    if (!status) status = &local;
    if (*status) return;
    guarded(status, [&] {
        if (!plan) gw::fail(gw::kErrInvalidArgument, "plan is NULL");
        const gw::Summary& sm = plan->impl->read_summary();
        if (num_records) *num_records = sm.nm32;
        if (record_updates) *record_updates = sm.run_updates;
    });
}

GW_EXPORT int gw_plan_timings(gw_plan* plan, double* ms, int max_n)
{
    if (!plan || !ms) return 0;
    int n = 0, status = 0;
    guarded(&status, [&] { n = plan->impl->timings(ms, max_n); });
    return status ? 0 : n;
}

GW_EXPORT int gw_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

GW_EXPORT void* gw_device_malloc(size_t bytes, int* status)
{
    int local = 0;
    if (!status) status = &local;
    if (*status) return nullptr;
    void* p = nullptr;
    guarded(status, [&] {
        gw::check_hip(hipMalloc(&p, bytes ? bytes : 1), gw::kErrMemAlloc,
                "hipMalloc");
    });
    return p;
}

GW_EXPORT void gw_device_free(void* ptr)
{
    if (ptr) (void)hipFree(ptr);
}

GW_EXPORT void* gw_host_malloc(size_t bytes, int* status)
{
    int local = 0;
    if (!status) status = &local;
    if (*status) return nullptr;
    void* p = nullptr;
    guarded(status, [&] {
        gw::check_hip(hipHostMalloc(&p, bytes ? bytes : 1,
                hipHostMallocDefault), gw::kErrMemAlloc, "hipHostMalloc");
    });
    return p;
}

GW_EXPORT void gw_host_free(void* ptr)
{
    if (ptr) (void)hipHostFree(ptr);
}

GW_EXPORT void gw_memcpy_h2d(void* dst, const void* src, size_t bytes,
        int* status)
{
    int local = 0;
    if (!status) status = &local;
    if (*status) return;
    guarded(status, [&] {
        gw::check_hip(hipDeviceSynchronize(), gw::kErrKernelLaunch,
                "hipDeviceSynchronize");
        gw::check_hip(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice),
                gw::kErrMemCopy, "hipMemcpy H2D");
    });
}

GW_EXPORT void gw_memcpy_d2h(void* dst, const void* src, size_t bytes,
        int* status)
{
    int local = 0;
    if (!status) status = &local;
    if (*status) return;
    guarded(status, [&] {
        gw::check_hip(hipDeviceSynchronize(), gw::kErrKernelLaunch,
                "hipDeviceSynchronize");
        gw::check_hip(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost),
                gw::kErrMemCopy, "hipMemcpy D2H");
    });
}

GW_EXPORT void gw_memset_device(void* ptr, int value, size_t bytes,
        int* status)
{
    int local = 0;
    if (!status) status = &local;
    if (*status) return;
    guarded(status, [&] {
        gw::check_hip(hipDeviceSynchronize(), gw::kErrKernelLaunch,
                "hipDeviceSynchronize");
        gw::check_hip(hipMemset(ptr, value, bytes), gw::kErrMemCopy,
                "hipMemset");
        gw::check_hip(hipDeviceSynchronize(), gw::kErrKernelLaunch,
                "hipDeviceSynchronize");
    });
}

GW_EXPORT gw_batch* gw_batch_create(int* status)
{
    int local = 0;
    if (!status) status = &local;
    if (*status) return nullptr;
    gw_batch* out = nullptr;
    guarded(status, [&] {
        std::unique_ptr<gw::Batch> b(new gw::Batch());
        out = new gw_batch{b.release()};
    });
    return out;
}

GW_EXPORT void gw_batch_free(gw_batch* batch)
{
    if (!batch) return;
    delete batch->impl;
    delete batch;
}

GW_EXPORT int gw_batch_add(gw_batch* batch, gw_plan* plan, size_t num_points,
        const float* d_uu, const float* d_vv, const float* d_ww,
        const float* d_vis, float* d_grid, int* status)
{
    int local = 0;
    if (!status) status = &local;
    if (*status) return -1;
    int job = -1;
    guarded(status, [&] {
        if (!batch || !plan)
            gw::fail(gw::kErrInvalidArgument, "batch and plan must not be NULL");
        job = batch->impl->add(plan->impl, num_points, d_uu, d_vv, d_ww,
                               d_vis, d_grid);
    });
    return job;
}

GW_EXPORT void gw_batch_clear(gw_batch* batch)
{
    if (batch) batch->impl->clear();
}

GW_EXPORT void gw_batch_set_stream(gw_batch* batch, void* hip_stream)
{
    if (!batch) return;
    int status = 0;
    guarded(&status, [&] {
        batch->impl->set_stream(static_cast<hipStream_t>(hip_stream));
    });
}

GW_EXPORT void gw_batch_run(gw_batch* batch, int* status)
{
    int local = 0;
    if (!status) status = &local;
    if (*status) return;
    guarded(status, [&] {
        if (!batch) gw::fail(gw::kErrInvalidArgument, "batch is NULL");
        batch->impl->run();
    });
}

GW_EXPORT void gw_batch_capture(gw_batch* batch, int* status)
{
    int local = 0;
    if (!status) status = &local;
    if (*status) return;
    guarded(status, [&] {
        if (!batch) gw::fail(gw::kErrInvalidArgument, "batch is NULL");
        batch->impl->capture();
    });
}

GW_EXPORT void gw_batch_launch(gw_batch* batch, int* status)
{
    int local = 0;
    if (!status) status = &local;
    if (*status) return;
    guarded(status, [&] {
        if (!batch) gw::fail(gw::kErrInvalidArgument, "batch is NULL");
        batch->impl->launch();
    });
}

GW_EXPORT void gw_batch_results(gw_batch* batch, int job, size_t* num_skipped,
        double* norm_delta, unsigned long long* num_updates, int* status)
{
    int local = 0;
    if (!status) status = &local;
    if (*status) return;
    guarded(status, [&] {
        if (!batch) gw::fail(gw::kErrInvalidArgument, "batch is NULL");
        batch->impl->results(job, num_skipped, norm_delta, num_updates);
    });
}

GW_EXPORT double gw_batch_last_ms(gw_batch* batch)
{
    if (!batch) return 0.0;
    int status = 0;
    double ms = 0.0;
    guarded(&status, [&] { ms = batch->impl->last_ms(); });
    return ms;
}

GW_EXPORT
void oskar_grid_weights_write_f(const size_t num_points, const float* uu,
        const float* vv, const float* weight, const float cell_size_rad,
        const int grid_size, size_t* num_skipped, float* grid)
{
    int status = 0;
    t_last_error = 0;
    t_last_detail.clear();
    guarded(&status, [&] {
        weights_write_host<float>(GW_SINGLE, num_points, uu, vv, weight,
                cell_size_rad, grid_size, num_skipped, grid);
    });
    t_last_error = status;
}

GW_EXPORT
void oskar_grid_weights_write_d(const size_t num_points, const double* uu,
        const double* vv, const double* weight, const double cell_size_rad,
        const int grid_size, size_t* num_skipped, double* grid)
{
    int status = 0;
    t_last_error = 0;
    t_last_detail.clear();
    guarded(&status, [&] {
        weights_write_host<double>(GW_DOUBLE, num_points, uu, vv, weight,
                cell_size_rad, grid_size, num_skipped, grid);
    });
    t_last_error = status;
}

GW_EXPORT
void oskar_grid_weights_read_f(const size_t num_points, const float* uu,
        const float* vv, const float* weight_in, float* weight_out,
        const float cell_size_rad, const int grid_size,
        size_t* num_skipped, const float* grid)
{
    int status = 0;
    t_last_error = 0;
    t_last_detail.clear();
    guarded(&status, [&] {
        weights_read_host<float>(GW_SINGLE, num_points, uu, vv, weight_in,
                weight_out, cell_size_rad, grid_size, num_skipped, grid);
    });
    t_last_error = status;
}

GW_EXPORT
void oskar_grid_weights_read_d(const size_t num_points, const double* uu,
        const double* vv, const double* weight_in, double* weight_out,
        const double cell_size_rad, const int grid_size,
        size_t* num_skipped, const double* grid)
{
    int status = 0;
    t_last_error = 0;
    t_last_detail.clear();
    guarded(&status, [&] {
        weights_read_host<double>(GW_DOUBLE, num_points, uu, vv, weight_in,
                weight_out, cell_size_rad, grid_size, num_skipped, grid);
    });
    t_last_error = status;
}

GW_EXPORT void gw_grid_weights_write_device(int precision, size_t num_points,
        const void* d_uu, const void* d_vv, const void* d_weight,
        double cell_size_rad, int grid_size, size_t* num_skipped,
        void* d_grid, void* hip_stream, int* status)
{
    int local = 0;
    if (!status) status = &local;
    if (*status) return;
    guarded(status, [&] {
        if (!num_skipped)
            gw::fail(gw::kErrInvalidArgument, "num_skipped must not be NULL");
        gw::Weights& w = cache().weights_for(current_device());
        *num_skipped = w.write(precision, num_points, d_uu, d_vv, d_weight,
                cell_size_rad, grid_size, d_grid,
                static_cast<hipStream_t>(hip_stream));
    });
}

GW_EXPORT void gw_grid_weights_read_device(int precision, size_t num_points,
        const void* d_uu, const void* d_vv, const void* d_weight_in,
        void* d_weight_out, double cell_size_rad, int grid_size,
        size_t* num_skipped, const void* d_grid, void* hip_stream,
        int* status)
{
    int local = 0;
    if (!status) status = &local;
    if (*status) return;
    guarded(status, [&] {
        if (!num_skipped)
            gw::fail(gw::kErrInvalidArgument, "num_skipped must not be NULL");
        gw::Weights& w = cache().weights_for(current_device());
        *num_skipped = w.read(precision, num_points, d_uu, d_vv, d_weight_in,
                d_weight_out, cell_size_rad, grid_size, d_grid,
                static_cast<hipStream_t>(hip_stream));
    });
}

GW_EXPORT gw_server* gw_server_create(size_t hbm_budget_bytes,
        int max_batch_jobs, int num_streams, int precision, int* status)
{
    int local = 0;
    if (!status) status = &local;
    if (*status) return nullptr;
    gw_server* out = nullptr;
    guarded(status, [&] {
        std::unique_ptr<gw::Server> sv(new gw::Server(hbm_budget_bytes,
                max_batch_jobs, num_streams, precision));
        out = new gw_server{sv.release()};
    });
    return out;
}

GW_EXPORT void gw_server_free(gw_server* server)
{
    if (!server) return;
    delete server->impl;
    delete server;
}

GW_EXPORT int gw_server_add_kernel_set(gw_server* server, size_t num_w_planes,
        const int* support, int oversample, int conv_size_half,
        const float* conv_func, int* status)
{
    int local = 0;
    if (!status) status = &local;
    if (*status) return -1;
    int id = -1;
    guarded(status, [&] {
        if (!server) gw::fail(gw::kErrInvalidArgument, "server is NULL");
        id = server->impl->add_kernel_set(num_w_planes, support, oversample,
                                          conv_size_half, conv_func);
    });
    return id;
}

GW_EXPORT int gw_server_add_wkernel_set(gw_server* server,
        size_t num_w_planes, const int* support, int oversample,
        int conv_size_half, double w_max, int* status)
{
    int local = 0;
    if (!status) status = &local;
    if (*status) return -1;
    int id = -1;
    guarded(status, [&] {
        if (!server) gw::fail(gw::kErrInvalidArgument, "server is NULL");
        id = server->impl->add_wkernel_set(num_w_planes, support, oversample,
                                           conv_size_half, w_max);
    });
    return id;
}

GW_EXPORT int gw_server_create_grid(gw_server* server, int grid_size,
        int* status)
{
    int local = 0;
    if (!status) status = &local;
    if (*status) return -1;
    int id = -1;
    guarded(status, [&] {
        if (!server) gw::fail(gw::kErrInvalidArgument, "server is NULL");
        id = server->impl->create_grid(grid_size);
    });
    return id;
}

GW_EXPORT void gw_server_read_grid(gw_server* server, int grid_id,
        float* grid, int* status)
{
    int local = 0;
    if (!status) status = &local;
    if (*status) return;
    guarded(status, [&] {
        if (!server || !grid)
            gw::fail(gw::kErrInvalidArgument, "server and grid must not be "
                     "NULL");
        server->impl->read_grid(grid_id, grid);
    });
}

GW_EXPORT void gw_server_free_grid(gw_server* server, int grid_id)
{
    if (!server) return;
    int status = 0;
    guarded(&status, [&] { server->impl->free_grid(grid_id); });
}

GW_EXPORT long long gw_server_submit(gw_server* server, int kernel_set,
        int grid_size, double cell_size_rad, double w_scale,
        size_t num_points, const float* uu, const float* vv, const float* ww,
        const float* vis, int grid_id, float* grid_out, int* status)
{
    int local = 0;
    if (!status) status = &local;
    if (*status) return -1;
    long long id = -1;
    guarded(status, [&] {
        if (!server) gw::fail(gw::kErrInvalidArgument, "server is NULL");
        gw::Server::Request r;
        r.kernel_set = kernel_set;
        r.grid_size = grid_size;
        r.cell_size_rad = cell_size_rad;
        r.w_scale = w_scale;
        r.n = num_points;
        r.uu = uu;
        r.vv = vv;
        r.ww = ww;
        r.vis = vis;
        r.grid_id = grid_id;
        r.grid_out = grid_out;
        id = server->impl->submit(r);
    });
    return id;
}

GW_EXPORT int gw_server_wait(gw_server* server, long long request,
        double timeout_s, gw_request_result* result, int* status)
{
    int local = 0;
    if (!status) status = &local;
    if (*status) return 0;
    int done = 0;
    guarded(status, [&] {
        if (!server) gw::fail(gw::kErrInvalidArgument, "server is NULL");
        gw::Server::Result r;
        done = server->impl->wait(request, timeout_s, &r) ? 1 : 0;
        if (done && result)
        {
            result->status = r.status;
            result->num_skipped = r.skipped;
            result->norm_delta = r.norm_delta;
            result->num_updates = r.updates;
            result->t_submit = r.t_submit;
            result->t_admit = r.t_admit;
            result->t_start = r.t_start;
            result->t_done = r.t_done;
        }
    });
    return done;
}

GW_EXPORT void gw_server_get_stats(gw_server* server, gw_server_stats* out)
{
    if (!server || !out) return;
    int status = 0;
    guarded(&status, [&] {
        const gw::Server::Stats st = server->impl->stats();
        out->completed = st.completed;
        out->batches = st.batches;
        out->max_batch = st.max_batch;
        out->plans_evicted = st.plans_evicted;
        out->points = st.points;
        out->device_bytes = st.device_bytes;
        out->resident_plans = st.resident_plans;
        out->pending = st.pending;
    });
}

GW_EXPORT void gw_compensated_resolve(const float* d_comp, size_t num_cells,
        float* d_out, void* hip_stream, int* status)
{
    int local = 0;
    if (!status) status = &local;
    if (*status) return;
    guarded(status, [&] {
        if (num_cells && (!d_comp || !d_out))
            gw::fail(gw::kErrInvalidArgument, "device pointers must not be NULL");
        gw::resolve_compensated(d_comp, num_cells, d_out,
                static_cast<hipStream_t>(hip_stream));
    });
}

GW_EXPORT void gw_device_copy(void* dst, const void* src, size_t bytes,
        int* status)
{
    int local = 0;
    if (!status) status = &local;
    if (*status) return;
    guarded(status, [&] { gw::device_copy(dst, src, bytes); });
}

GW_EXPORT void gw_device_synchronize(int* status)
{
    int local = 0;
    if (!status) status = &local;
    if (*status) return;
    guarded(status, [&] {
        gw::check_hip(hipDeviceSynchronize(), gw::kErrKernelLaunch,
                "hipDeviceSynchronize");
    });
}

}  // extern "C"
