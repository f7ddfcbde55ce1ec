// Serving loop for gridding requests (see server.cpp).
#pragma once

#include "gridder_internal.h"

#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <tuple>

namespace gw {

class Server
{
public:
    // One gridding request (FP32, the reference's oskar_grid_wproj_f
    // arguments): host arrays the caller keeps alive until the request is
    // waited for; the grid is a resident one (grid_id >= 0, accumulated)
    // or a fresh zeroed one copied to grid_out (grid_id < 0).
    struct Request
    {
        int kernel_set = -1;
        int grid_size = 0;
        double cell_size_rad = 0.0;
        double w_scale = 0.0;
        size_t n = 0;
        const float* uu = nullptr;
        const float* vv = nullptr;
        const float* ww = nullptr;
        const float* vis = nullptr;
        int grid_id = -1;
        float* grid_out = nullptr;
    };
    struct Result
    {
        int status = 0;
        size_t skipped = 0;
        double norm_delta = 0.0;
        unsigned long long updates = 0;
        // Steady-clock seconds: submitted, admitted (memory reserved),
        // batch enqueued, completed (results copied back).
        double t_submit = 0, t_admit = 0, t_start = 0, t_done = 0;
    };
    struct Stats
    {
        size_t completed = 0, batches = 0, max_batch = 0, plans_evicted = 0;
        size_t points = 0;
        size_t device_bytes = 0, resident_plans = 0, pending = 0;
    };

    // hbm_budget: device bytes the server may hold (tables, grids, inputs,
    // workspace); max_batch_jobs: most requests per batch; n_streams:
    // batches in flight at once; precision: kPrecSingle (the index math of
    // oskar_grid_wproj_f) or kPrecSingleIndex64 (FP64 index math).
    Server(size_t hbm_budget, int max_batch_jobs, int n_streams,
           int precision = kPrecSingle);
    ~Server();
    int add_kernel_set(size_t num_w_planes, const int* support,
                       int oversample, int conv_size_half,
                       const float* conv_func);
    // A kernel set generated on the device (generate_wkernels); its cube
    // stays in device memory and plans are built from it there.
    int add_wkernel_set(size_t num_w_planes, const int* support,
                        int oversample, int conv_size_half, double w_max);
    int create_grid(int grid_size);
    void read_grid(int grid_id, float* host_out);
    void free_grid(int grid_id);
    long long submit(const Request& r);
    // Waits (timeout_s < 0: without limit); false on timeout.  The request
    // is forgotten once its result has been returned.
    bool wait(long long id, double timeout_s, Result* out);
    Stats stats();

private:
    struct KernelSet
    {
        size_t P = 0;
        std::vector<int> support;
        int ov = 0, C = 0;
        std::vector<float> cube;     // host cube, or
        float* d_cube = nullptr;     // a cube in device memory
    };
    typedef std::tuple<int, int, double, double> PlanKey;
    struct CachedPlan
    {
        std::unique_ptr<Plan> plan;
        unsigned long long last_use = 0;
        int in_use = 0;
    };
    struct Grid
    {
        void* ptr = nullptr;
        int G = 0;
        int users = 0;
    };
    struct Job
    {
        long long id = 0;
        Request req;
        Result res;
        Plan* plan = nullptr;
        void* d_grid = nullptr;         // temporary grid inside the stage
        int batch_index = -1;
        bool done = false;
    };
    struct Slot
    {
        std::unique_ptr<Batch> batch;
        std::vector<long long> jobs;
        bool busy = false;
        void* stage = nullptr;          // inputs and temporary grids
        size_t stage_cap = 0;
    };

    void loop();
    bool launch(Slot& s, std::unique_lock<std::mutex>& g);
    void finish(Slot& s, hipError_t q);
    void release(Job& j);
    void complete(Job& j);
    bool any_busy() const;
    Plan* plan_for(const Request& r, size_t* new_bytes);
    bool make_room(size_t need);
    bool trim_slots(const Slot* keep, size_t need);
    static size_t stage_bytes(const Request& r);
    static size_t slot_bytes(const Slot& s);
    static PlanKey plan_key(const Request& r)
    {
        return PlanKey{r.kernel_set, r.grid_size, r.cell_size_rad,
                       r.w_scale};
    }

    size_t budget_;
    int precision_;
    int max_batch_;
    int slots_;
    int device_ = 0;
    size_t used_ = 0;
    unsigned long long use_clock_ = 0;
    long long next_req_ = 1;
    int next_grid_ = 0;
    bool stop_ = false;
    std::vector<KernelSet> ksets_;
    std::map<PlanKey, CachedPlan> plans_;
    std::map<int, Grid> grids_;
    std::map<long long, Job> jobs_;
    std::deque<long long> pending_;
    std::vector<Slot> slot_;
    Stats stats_;
    std::mutex mu_;
    std::condition_variable cv_;
    std::condition_variable cv_done_;
    std::thread worker_;
};

}  // namespace gw
