// Serving loop for gridding requests: admission by an HBM budget, resident
// kernel sets (plans) and grids, continuous batching onto the batched
// pipeline (gw::Batch), several batches in flight on their own streams.
//
// The reference grids one job per call (oskar_grid_wproj.c:128-243) and
// has no serving layer; this is the request loop around that call.
#include "server.h"

#include <algorithm>
#include <chrono>
#include <cstring>

namespace gw {

namespace {

double now_s()
{
    using clock = std::chrono::steady_clock;
    return std::chrono::duration<double>(clock::now().time_since_epoch())
            .count();
}

size_t align_up(size_t b)
{
    return (b + 255) & ~(size_t)255;
}

void* dalloc(size_t bytes)
{
    void* p = nullptr;
    check_hip(hipMalloc(&p, bytes ? bytes : 1), kErrMemAlloc, "hipMalloc");
    return p;
}

}  // namespace

Server::Server(size_t hbm_budget, int max_batch_jobs, int n_streams,
               int precision)
    : budget_(hbm_budget),
      precision_(precision),
      max_batch_(std::max(1, max_batch_jobs)),
      slots_(std::max(1, std::min(n_streams, 8)))
{
    if (precision != kPrecSingle && precision != kPrecSingleIndex64)
        fail(kErrInvalidArgument, "the server grids GW_SINGLE or "
             "GW_SINGLE_INDEX64 requests");
    check_hip(hipGetDevice(&device_), kErrComputeDevices, "hipGetDevice");
    for (int i = 0; i < slots_; ++i)
        {
        slot_.emplace_back();
        slot_.back().batch = std::make_unique<Batch>();
    }
    worker_ = std::thread([this] { loop(); });
}

Server::~Server()
{
    {
        std::lock_guard<std::mutex> g(mu_);
        stop_ = true;
    }
    cv_.notify_all();
    if (worker_.joinable()) worker_.join();
    for (auto& kv : grids_)
        if (kv.second.ptr) (void)hipFree(kv.second.ptr);
    for (Slot& s : slot_)
        if (s.stage) (void)hipFree(s.stage);
    plans_.clear();
    for (KernelSet& k : ksets_) dfree(k.d_cube);
}

int Server::add_kernel_set(size_t num_w_planes, const int* support,
                           int oversample, int conv_size_half,
                           const float* conv_func)
{
    validate_args(num_w_planes, support, oversample, conv_size_half,
                  conv_func, 1);
    KernelSet k;
    k.P = num_w_planes;
    k.support.assign(support, support + num_w_planes);
    k.ov = oversample;
    k.C = conv_size_half;
    k.cube.assign(conv_func, conv_func + 2 * num_w_planes *
                  (size_t)conv_size_half * conv_size_half);
    std::lock_guard<std::mutex> g(mu_);
    ksets_.push_back(std::move(k));
    return (int)ksets_.size() - 1;
}

int Server::add_wkernel_set(size_t num_w_planes, const int* support,
                            int oversample, int conv_size_half, double w_max)
{
    validate_args(num_w_planes, support, oversample, conv_size_half,
                  support, 1);
    const size_t bytes = 2 * num_w_planes * (size_t)conv_size_half *
            conv_size_half * sizeof(float);
    {
        std::lock_guard<std::mutex> g(mu_);
        if (used_ + bytes > budget_)
            fail(kErrMemAlloc, "kernel cube exceeds the HBM budget");
        used_ += bytes;
    }
    KernelSet k;
    k.P = num_w_planes;
    k.support.assign(support, support + num_w_planes);
    k.ov = oversample;
    k.C = conv_size_half;
    try
    {
        (void)hipSetDevice(device_);
        k.d_cube = dmalloc<float>(bytes / sizeof(float));
        generate_wkernels(num_w_planes, support, oversample, conv_size_half,
                          w_max, k.d_cube, nullptr);
    }
    catch (...)
    {
        dfree(k.d_cube);
        std::lock_guard<std::mutex> g(mu_);
        used_ -= bytes;
        throw;
    }
    std::lock_guard<std::mutex> g(mu_);
    ksets_.push_back(std::move(k));
    return (int)ksets_.size() - 1;
}

int Server::create_grid(int grid_size)
{
    if (grid_size < 1 || grid_size > 65535)
        fail(kErrInvalidArgument, "grid_size must be in [1, 65535]");
    const size_t bytes = 8 * (size_t)grid_size * grid_size;
    std::unique_lock<std::mutex> g(mu_);
    if (used_ + bytes > budget_)
        fail(kErrMemAlloc, "resident grid exceeds the HBM budget");
    used_ += bytes;
    g.unlock();
    int dev_saved = 0;
    (void)hipGetDevice(&dev_saved);
    (void)hipSetDevice(device_);
    void* p = nullptr;
    const hipError_t e1 = hipMalloc(&p, bytes);
    const hipError_t e2 = e1 == hipSuccess ? hipMemset(p, 0, bytes) : e1;
    const hipError_t e3 = e2 == hipSuccess ? hipDeviceSynchronize() : e2;
    (void)hipSetDevice(dev_saved);
    g.lock();
    if (e3 != hipSuccess)
    {
        used_ -= bytes;
        if (p) (void)hipFree(p);
        fail(kErrMemAlloc, "resident grid allocation failed");
    }
    const int id = next_grid_++;
    grids_[id] = Grid{p, grid_size, 0};
    return id;
}

void Server::read_grid(int grid_id, float* host_out)
{
    std::unique_lock<std::mutex> g(mu_);
    auto it = grids_.find(grid_id);
    if (it == grids_.end()) fail(kErrInvalidArgument, "unknown grid id");
    // Wait until no queued or running request writes this grid.
    cv_done_.wait(g, [&] { return it->second.users == 0; });
    void* p = it->second.ptr;
    const size_t bytes = 8 * (size_t)it->second.G * it->second.G;
    g.unlock();
    (void)hipSetDevice(device_);
    check_hip(hipDeviceSynchronize(), kErrKernelLaunch, "synchronize");
    check_hip(hipMemcpy(host_out, p, bytes, hipMemcpyDeviceToHost),
              kErrMemCopy, "read grid");
}

void Server::free_grid(int grid_id)
{
    std::unique_lock<std::mutex> g(mu_);
    auto it = grids_.find(grid_id);
    if (it == grids_.end()) return;
    cv_done_.wait(g, [&] { return it->second.users == 0; });
    used_ -= 8 * (size_t)it->second.G * it->second.G;
    void* p = it->second.ptr;
    grids_.erase(it);
    g.unlock();
    (void)hipSetDevice(device_);
    (void)hipFree(p);
}

long long Server::submit(const Request& r)
{
    if (r.n && (!r.uu || !r.vv || !r.ww || !r.vis))
        fail(kErrInvalidArgument, "host arrays must not be NULL");
    std::lock_guard<std::mutex> g(mu_);
    if (r.kernel_set < 0 || r.kernel_set >= (int)ksets_.size())
        fail(kErrInvalidArgument, "unknown kernel set");
    if (r.grid_id >= 0)
    {
        auto it = grids_.find(r.grid_id);
        if (it == grids_.end()) fail(kErrInvalidArgument, "unknown grid id");
        if (it->second.G != r.grid_size)
            fail(kErrInvalidArgument, "grid_size differs from the grid's");
        ++it->second.users;
    }
    else if (!r.grid_out)
        fail(kErrInvalidArgument, "grid_out must not be NULL without a "
             "resident grid");
    const long long id = next_req_++;
    Job j;
    j.id = id;
    j.req = r;
    j.res.t_submit = now_s();
    jobs_[id] = j;
    pending_.push_back(id);
    cv_.notify_all();
    return id;
}

bool Server::wait(long long id, double timeout_s, Result* out)
{
    std::unique_lock<std::mutex> g(mu_);
    auto done = [&] {
        auto it = jobs_.find(id);
        return it == jobs_.end() || it->second.done;
    };
    if (timeout_s < 0)
        cv_done_.wait(g, done);
    else if (!cv_done_.wait_for(g, std::chrono::duration<double>(timeout_s),
                                done))
        return false;
    auto it = jobs_.find(id);
    if (it == jobs_.end()) fail(kErrInvalidArgument, "unknown request id");
    if (out) *out = it->second.res;
    jobs_.erase(it);
    return true;
}

Server::Stats Server::stats()
{
    std::lock_guard<std::mutex> g(mu_);
    Stats s = stats_;
    s.device_bytes = used_;
    s.resident_plans = plans_.size();
    s.pending = pending_.size();
    return s;
}

// ------------------------------------------------------------------------
// Worker
// ------------------------------------------------------------------------

Plan* Server::plan_for(const Request& r, size_t* new_bytes)
{
    const PlanKey key = plan_key(r);
    auto it = plans_.find(key);
    if (it != plans_.end())
    {
        it->second.last_use = ++use_clock_;
        *new_bytes = 0;
        return it->second.plan.get();
    }
    const KernelSet& k = ksets_[r.kernel_set];
    std::unique_ptr<Plan> p(k.d_cube ?
            new Plan(precision_, k.P, k.support.data(), k.ov, k.C, k.d_cube,
                     r.cell_size_rad, r.w_scale, r.grid_size, true) :
            new Plan(precision_, k.P, k.support.data(), k.ov, k.C,
                     k.cube.data(), r.cell_size_rad, r.w_scale,
                     r.grid_size));
    *new_bytes = p->table_bytes();
    Plan* raw = p.get();
    plans_[key] = CachedPlan{std::move(p), ++use_clock_, 0};
    return raw;
}

// Evicts idle plans, least recently used first, until `need` more bytes
// fit the budget; returns whether they do.
bool Server::make_room(size_t need)
{
    while (used_ + need > budget_)
    {
        auto victim = plans_.end();
        for (auto it = plans_.begin(); it != plans_.end(); ++it)
            if (it->second.in_use == 0 &&
                    (victim == plans_.end() ||
                     it->second.last_use < victim->second.last_use))
                victim = it;
        if (victim == plans_.end()) return false;
        used_ -= victim->second.plan->table_bytes();
        plans_.erase(victim);
        ++stats_.plans_evicted;
    }
    return true;
}

size_t Server::stage_bytes(const Request& r)
{
    // Points (uu, vv, ww, vis: 20 B each) and, without a resident grid, a
    // temporary one; 256-byte aligned pieces.
    size_t b = align_up(r.n * 20);
    if (r.grid_id < 0) b += align_up(8 * (size_t)r.grid_size * r.grid_size);
    return b;
}

size_t Server::slot_bytes(const Slot& s)
{
    return s.stage_cap + s.batch->device_bytes();
}

// Frees the staging and workspace of idle slots other than `keep` until
// `need` more bytes fit the budget; returns whether they do.
bool Server::trim_slots(const Slot* keep, size_t need)
{
    for (Slot& s : slot_)
    {
        if (used_ + need <= budget_) break;
        if (&s == keep || s.busy) continue;
        const size_t b = slot_bytes(s);
        if (!b) continue;
        if (s.stage) (void)hipFree(s.stage);
        s.stage = nullptr;
        s.stage_cap = 0;
        s.batch->release_workspace();
        used_ -= b;
    }
    return used_ + need <= budget_;
}

void Server::loop()
{
    (void)hipSetDevice(device_);
    std::unique_lock<std::mutex> g(mu_);
    for (;;)
    {
        // 1. Completed batches.
        for (Slot& s : slot_)
        {
            if (!s.busy) continue;
            g.unlock();
            const hipError_t q = hipEventQuery(s.batch->done_event());
            g.lock();
            if (q == hipErrorNotReady) continue;
            finish(s, q);
        }
        // 2. Admit and launch on a free slot.
        Slot* free_slot = nullptr;
        for (Slot& s : slot_)
            if (!s.busy)
            {
                free_slot = &s;
                break;
            }
        bool launched = false;
        if (free_slot && !pending_.empty()) launched = launch(*free_slot, g);
        bool busy = false;
        for (const Slot& s : slot_) busy = busy || s.busy;
        if (stop_ && pending_.empty() && !busy) break;
        if (!launched)
        {
            // Nothing new could start: sleep until a submission arrives or
            // a little while (running batches are polled).
            cv_.wait_for(g, std::chrono::microseconds(busy ? 50 : 2000));
        }
    }
}

bool Server::launch(Slot& s, std::unique_lock<std::mutex>& g)
{
    std::vector<long long> take;
    size_t stage_total = 0, points = 0, reserved = 0;
    Batch& b = *s.batch;
    while (!pending_.empty() && (int)take.size() < max_batch_)
    {
        Job& j = jobs_[pending_.front()];
        size_t plan_bytes = 0;
        Plan* p = nullptr;
        try
        {
            p = plan_for(j.req, &plan_bytes);
        }
        catch (const Error& e)
        {
            j.res.status = e.code;
            release(j);
            complete(j);
            pending_.pop_front();
            continue;
        }
        used_ += plan_bytes;
        CachedPlan& cp = plans_[plan_key(j.req)];
        ++cp.in_use;                    // pinned until the job completes
        // Growth of this slot's staging and workspace the batch would need.
        const size_t st = stage_total + stage_bytes(j.req);
        const size_t ws = Batch::workspace_estimate(points + j.req.n);
        const size_t cur_ws = b.device_bytes();
        const size_t grow = (st > s.stage_cap ? st - s.stage_cap : 0) +
                            (ws > cur_ws ? ws - cur_ws : 0);
        const size_t extra = grow > reserved ? grow - reserved : 0;
        const bool fits = trim_slots(&s, extra) || make_room(extra);
        if (!fits)
        {
            --cp.in_use;
            if (take.empty() && !any_busy())
            {
                // Nothing else holds memory that could be freed: this
                // request can never fit the budget.
                j.res.status = kErrMemAlloc;
                release(j);
                complete(j);
                pending_.pop_front();
                continue;
            }
            break;                      // FIFO: wait for memory to free
        }
        used_ += extra;
        reserved += extra;
        stage_total = st;
        points += j.req.n;
        j.plan = p;
        j.res.t_admit = now_s();
        take.push_back(j.id);
        pending_.pop_front();
    }
    if (take.empty()) return false;
    stats_.max_batch = std::max<size_t>(stats_.max_batch, take.size());
    const size_t before = slot_bytes(s);
    std::vector<Job*> jp(take.size());
    for (size_t i = 0; i < take.size(); ++i) jp[i] = &jobs_[take[i]];
    g.unlock();
    // Stage and enqueue outside the lock (submitters are not blocked).  A
    // job's records in jobs_ stay put: it is not done, so wait() keeps it.
    std::vector<int> status(take.size(), 0);
    b.clear();
    const hipStream_t st = b.stream();
    int staged = 0;
    try
    {
        if (stage_total > s.stage_cap)
        {
            if (s.stage) (void)hipFree(s.stage);
            s.stage = nullptr;
            s.stage_cap = 0;
            s.stage = dalloc(stage_total);
            s.stage_cap = stage_total;
        }
        staged = 1;
    }
    catch (const Error& e)
    {
        std::fill(status.begin(), status.end(), e.code);
    }
    char* cursor = static_cast<char*>(s.stage);
    for (size_t i = 0; staged && i < take.size(); ++i)
    {
        Job& j = *jp[i];
        const Request& r = j.req;
        const size_t n = r.n;
        float* d_uu = reinterpret_cast<float*>(cursor);
        float* d_vv = d_uu + n;
        float* d_ww = d_vv + n;
        float* d_vis = d_ww + n;
        cursor += align_up(n * 20);
        try
        {
            if (n)
            {
                check_hip(hipMemcpyAsync(d_uu, r.uu, n * 4,
                        hipMemcpyHostToDevice, st), kErrMemCopy, "upload");
                check_hip(hipMemcpyAsync(d_vv, r.vv, n * 4,
                        hipMemcpyHostToDevice, st), kErrMemCopy, "upload");
                check_hip(hipMemcpyAsync(d_ww, r.ww, n * 4,
                        hipMemcpyHostToDevice, st), kErrMemCopy, "upload");
                check_hip(hipMemcpyAsync(d_vis, r.vis, n * 8,
                        hipMemcpyHostToDevice, st), kErrMemCopy, "upload");
            }
            float* grid = nullptr;
            if (r.grid_id >= 0)
            {
                std::lock_guard<std::mutex> lk(mu_);
                grid = static_cast<float*>(grids_[r.grid_id].ptr);
            }
            else
            {
                const size_t cells = (size_t)r.grid_size * r.grid_size;
                grid = reinterpret_cast<float*>(cursor);
                cursor += align_up(8 * cells);
                fill_zero_u32(reinterpret_cast<unsigned*>(grid), 2 * cells,
                              st);
                j.d_grid = grid;
            }
            j.batch_index = b.add(j.plan, n, d_uu, d_vv, d_ww, d_vis, grid);
        }
        catch (const Error& e)
        {
            status[i] = e.code;
        }
    }
    try
    {
        int ok = 0;
        for (int v : status) ok += v == 0;
        if (ok) b.run();
    }
    catch (const Error& e)
    {
        for (int& v : status)
            if (v == 0) v = e.code;
    }
    const double t0 = now_s();
    g.lock();
    // Settle the reservation against what the slot now holds.
    used_ = used_ - reserved + (slot_bytes(s) - std::min(before,
            slot_bytes(s)));
    s.jobs.clear();
    for (size_t i = 0; i < take.size(); ++i)
    {
        Job& j = *jp[i];
        j.res.t_start = t0;
        if (status[i] != 0)
        {
            j.res.status = status[i];
            release(j);
            complete(j);
        }
        else
            s.jobs.push_back(j.id);
    }
    s.busy = !s.jobs.empty();
    ++stats_.batches;
    return true;
}

bool Server::any_busy() const
{
    for (const Slot& s : slot_)
        if (s.busy) return true;
    return false;
}

void Server::finish(Slot& s, hipError_t q)
{
    Batch& b = *s.batch;
    std::vector<long long> ids = s.jobs;
    s.jobs.clear();
    s.busy = false;
    for (long long id : ids)
    {
        Job& j = jobs_[id];
        if (q != hipSuccess)
            j.res.status = kErrKernelLaunch;
        else
        {
            try
            {
                b.results(j.batch_index, &j.res.skipped, &j.res.norm_delta,
                          &j.res.updates);
                if (j.req.grid_id < 0)
                    check_hip(hipMemcpy(j.req.grid_out, j.d_grid,
                            8 * (size_t)j.req.grid_size * j.req.grid_size,
                            hipMemcpyDeviceToHost), kErrMemCopy,
                            "read grid");
            }
            catch (const Error& e)
            {
                j.res.status = e.code;
            }
        }
        stats_.points += j.req.n;
        release(j);
        complete(j);
    }
}

void Server::release(Job& j)
{
    j.d_grid = nullptr;
    auto it = plans_.find(plan_key(j.req));
    if (j.plan && it != plans_.end() && it->second.in_use > 0)
        --it->second.in_use;
    j.plan = nullptr;
    if (j.req.grid_id >= 0)
    {
        auto gi = grids_.find(j.req.grid_id);
        if (gi != grids_.end() && gi->second.users > 0) --gi->second.users;
    }
}

void Server::complete(Job& j)
{
    j.res.t_done = now_s();
    j.done = true;
    ++stats_.completed;
    cv_done_.notify_all();
}

}  // namespace gw
