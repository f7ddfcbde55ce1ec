// Internal declarations shared by the gridder's device code (gridder.hip)
// and its C ABI layer (capi.cpp).  Not installed; the public interface is
// include/oskar_grid_wproj.h and include/fpga_gridding_amd.h.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace gw {

// Threads of a scatter workgroup (16 waves), most footprint cells one lane
// owns at a time (registers per slot: the current taps and the running
// sums in FP64, the next taps in the plan's precision, complex; at 6 slots
// an FP64 wave needs 128 VGPRs and spills a few, and is still faster than
// at 5, measured), and the most points in one scatter job
// (the per-job plane segment table lives in LDS).
constexpr int kGridThreads = 1024;
constexpr bool kSumF64 = false;
template <typename T> constexpr int slots_max()
{
    return (kSumF64 || sizeof(T) == 8) ? 6 : 8;
}
constexpr int kJobPoints = 2 * kGridThreads;
// LDS bytes for the job's grid region: planar FP64 real and imaginary
// arrays of kRegionCells each, holding the R x R region and then the sink
// cells (88 x 88 cells fit).
constexpr int kRegionBytes = 124 * 1024;
constexpr int kRegionCells = kRegionBytes / 16;
// Cells after each region array that lanes past a footprint add into (one
// per lane of a wave; never read).
constexpr int kSinkCells = 64;
// Default least number of points a scatter stream takes from a plane
// segment (shorter chunks pay the stream's set-up for too little work).
constexpr int kMinShare = 8;
// XCDs of the MI355X (8 per device, 32 CUs each): scatter items are split
// into this many contiguous ranges, one per XCD.
constexpr int kXcdGroups = 8;

// Per-plane constants, fixed by the support table (host-computed at plan
// creation, uploaded once).
struct PlaneInfo
{
    int W;       // footprint side 2S+1
    int S;       // support
    int wpp;     // waves per stream (1, 2, 4, 8 or 16)
    int nslot;   // footprint cells per lane per pass (<= kSlotsMax)
    // c / W == umulhi(c, wmagic) for every cell index c < W*W, or 0 if no
    // 32-bit multiplier is exact over that range (then plain division).
    unsigned wmagic;
    int sshift;  // log2 of the streams per workgroup, kGridThreads/64/wpp
    // First tap of this plane in the expanded tap table: K*K sub-cell
    // tables of W*W taps each, tap (j, k) at [(iv*K + iu)*W*W + j*W + k]
    // (tables are limited to 4 GiB, so byte offsets fit 32 bits).
    unsigned eoff;
    // Rank of S among the distinct supports: planes of one rank share the
    // footprint shape, and the scatter takes their points as one segment.
    int srank;
};

// Device-side per-call summary (written by the kernels, read back once).
struct Summary
{
    unsigned long long skipped;     // points failing the edge test
    unsigned long long n_valid;     // points gridded
    unsigned long long updates;     // sum over gridded points of W^2
    double norm_delta;              // sum over gridded points of tap sums
    int v_min, v_max;               // grid rows touched (inclusive)
    int u_min, u_max;               // grid columns touched (inclusive)
    unsigned nv32;                  // n_valid as a 32-bit count
    unsigned n_buckets;             // tile buckets
    unsigned total_items;           // work items of the scatter
    unsigned flags;                 // bit 0: diagnostics plane index beyond
                                    // its table; bit 1: work items overflow
    unsigned item_next[kXcdGroups]; // scatter queue heads, one per XCD
    unsigned nm32;                  // scatter records: runs of equal sort
                                    // keys (nv32 without merging)
    unsigned long long run_updates; // sum over records of W^2: the grid
                                    // updates the scatter executes
    unsigned nw32;                  // 32-point words of run-start bits
    unsigned nwt32;                 // 32-record words of tile-start bits
};

// Per-point record written by the locate kernel in input order.
struct alignas(16) RecF { int pos; int sub; float re; float im; };
struct alignas(16) RecD { int pos; int sub; int pad0, pad1; double re, im; };

// Plan precisions (values of gw_precision in include/fpga_gridding_amd.h).
enum
{
    kPrecSingle = 0x01,     // FP32 throughout (oskar_grid_wproj_f)
    kPrecDouble = 0x02,     // FP64 throughout (oskar_grid_wproj_d)
    kPrecHalf = 0x03,       // FP16 taps and visibilities, FP64 index math,
                            // FP32 products, compensated two-float grid
    kPrecBFloat16 = 0x04,   // as kPrecHalf with BF16 storage
    kPrecSingleIndex64 = 0x05   // kPrecSingle with FP64 index math (the
                                // plan's precision() stays kPrecSingle)
};

// 16-bit storage types.  FP16 is the compiler's _Float16; BF16 is the top
// half of an FP32 (conversion to float is a shift, from float rounds to
// nearest even).
typedef _Float16 half_t;
struct bf16_t { unsigned short bits; };
struct alignas(4) HalfPair { half_t x, y; };
struct alignas(4) BF16Pair { unsigned short x, y; };

__host__ __device__ inline float bf16_to_float(unsigned short b)
{
    const unsigned u = (unsigned)b << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

inline unsigned short float_to_bf16(float f)
{
    unsigned u;
    __builtin_memcpy(&u, &f, 4);
    if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu))
        return (unsigned short)((u >> 16) | 0x0040u);      // quiet NaN
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

// Real value of an input element.
__host__ __device__ inline float to_real(float x) { return x; }
__host__ __device__ inline double to_real(double x) { return x; }
__host__ __device__ inline float to_real(half_t x) { return (float)x; }
__host__ __device__ inline float to_real(bf16_t x)
{
    return bf16_to_float(x.bits);
}

// A stored tap pair as float2 / double2.
__host__ __device__ inline float2 tap_pair(const float2& t) { return t; }
__host__ __device__ inline double2 tap_pair(const double2& t) { return t; }
__host__ __device__ inline float2 tap_pair(const HalfPair& t)
{
    return make_float2((float)t.x, (float)t.y);
}
__host__ __device__ inline float2 tap_pair(const BF16Pair& t)
{
    return make_float2(bf16_to_float(t.x), bf16_to_float(t.y));
}

template <typename T> struct Types;
template <> struct Types<float>
{
    typedef float2 T2; typedef float4 T4; typedef RecF Rec;
};
template <> struct Types<double>
{
    typedef double2 T2; typedef double4 T4; typedef RecD Rec;
};

// Status codes (values of utils/oskar_global.h:46-64).
enum
{
    kOk = 0,
    kErrInvalidArgument = -3,
    kErrFunctionNotAvailable = -4,
    kErrOutOfRange = -5,
    kErrMemAlloc = -6,
    kErrMemCopy = -7,
    kErrKernelLaunch = -18,
    kErrComputeDevices = -19
};

// Error raised inside the library and turned into a status code at the C
// ABI boundary.
struct Error
{
    int code;
    std::string detail;
};

[[noreturn]] void fail(int code, const std::string& detail);
void check_hip(hipError_t e, int code, const char* what);

// Device allocation that raises kErrMemAlloc (count 0 allocates one
// element, so the pointer is never null).
template <typename T>
T* dmalloc(size_t count)
{
    void* p = nullptr;
    if (count == 0) count = 1;
    check_hip(hipMalloc(&p, count * sizeof(T)), kErrMemAlloc, "hipMalloc");
    return static_cast<T*>(p);
}

inline void dfree(void* p)
{
    if (p) (void)hipFree(p);
}

// Compensated two-float grid (re_hi, re_lo, im_hi, im_lo per cell) to a
// float grid (re, im per cell), enqueued on `stream`.
void resolve_compensated(const float* comp, size_t num_cells, float* out,
                         hipStream_t stream);

// Streaming device-to-device copy (k_copy) on the null stream; waits.
void device_copy(void* dst, const void* src, size_t bytes);

// Host-side argument checks of a gridding call (no device access): raises
// kErrInvalidArgument for sizes the gridder cannot handle, including a
// kernel cube too small for support and oversample.
void validate_args(size_t num_w_planes, const int* support, int oversample,
                   int conv_size_half, const void* conv_func, int grid_size);

// Diagnostics outputs (device buffers) of one call.
struct DiagBuffers
{
    unsigned long long* per_support = nullptr;   // [max_support + 1]
    unsigned long long* per_plane = nullptr;     // [plane_cap]
    unsigned* per_cell = nullptr;                // [rows * G]
    unsigned long long* wsum_cell = nullptr;     // [rows * G]
    int max_support = 0;
    int plane_cap = 0;
};

// A resident gridding plan: device W-kernel cube, derived tables, workspace.
class Plan
{
public:
    Plan(int precision, size_t num_w_planes, const int* support,
         int oversample, int conv_size_half, const void* conv_func,
         double cell_size_rad, double w_scale, int grid_size,
         bool cube_on_device = false);
    ~Plan();

    // Stage 1 (enqueue): index math, skip count, histograms, sort keys.
    void locate(size_t n, const void* uu, const void* vv, const void* ww,
                const void* vis, const DiagBuffers* diag);
    // Stages 2-4 (enqueue): sort, gather + plan, scatter into the grid
    // whose first row held in memory is row0.
    void scatter(size_t n, void* grid, int row0, const DiagBuffers* diag);
    // Copy the device summary to the host (synchronises the stream).
    const Summary& read_summary();
    void set_stream(hipStream_t s);
    hipStream_t stream() const { return stream_; }
    int precision() const { return prec_; }
    int grid_size() const { return G_; }
    size_t num_w_planes() const { return P_; }
    int max_support() const { return s_max_; }
    int timings(double* ms, int max_n);
    // Scatter shape: tile side 2^tile_bits (-1: the largest side, of any
    // length up to 64, whose LDS region fits), at most slot_cap footprint
    // cells per lane (0: the largest the precision allows), at least
    // min_share points per stream chunk (0: the default).  Waits for the
    // plan's stream before changing its tables.
    void configure(int tile_bits, int slot_cap, int min_share);
    // The same with the tile side given directly (1..64; -1 as above).
    void configure_side(int tile, int slot_cap, int min_share);
    int tile() const { return tile_; }
    // Bits of a coordinate inside a tile (the side is at most 2^tile_bits).
    int tile_bits() const { return tile_bits_; }
    // Bits of a tile coordinate (tiles per axis at most 2^tile_coord_bits).
    int tile_coord_bits() const { return tile_coord_bits_; }
    int slot_cap() const { return slot_cap_; }
    int min_share() const { return min_share_; }
    int region() const { return region_; }
    int max_tile_bits() const;
    int max_tile() const;
    // Autotuner: runs locate + scatter on the given device points (grid
    // into scratch_grid, which is overwritten) for scatter shapes around
    // the current one -- tile side, then slots per lane, then chunk minimum,
    // each kept if faster (best of reps timed runs) -- and keeps the
    // fastest.  Returns its device milliseconds per call.
    double autotune(size_t n, const void* uu, const void* vv, const void* ww,
                    const void* vis, void* scratch_grid, int reps);
    // Device tables and geometry for batched launches.
    const PlaneInfo* device_pinfo() const { return d_pinfo_; }
    const void* device_etab() const { return d_etab_; }
    const double* device_normtab() const { return d_normtab_; }
    const int* device_support() const { return d_support_; }
    const int* device_srank() const { return d_srank_; }
    int srank_bits() const { return srank_bits_; }
    int plane_bits() const { return plane_bits_; }
    int coord_bits() const { return coord_bits_; }
    int combo_bits() const { return combo_bits_; }
    int tile_shift() const { return tile_shift_; }
    int key_bits() const { return key_bits_; }
    int key_bits_max() const { return key_bits_max_; }
    int oversample() const { return ov_; }
    int half_sub() const { return h_; }
    int sub_tables() const { return K_; }
    float grid_scale_f() const { return gscale_f_; }
    float w_scale_f() const { return ws_f_; }
    double grid_scale_d() const { return gscale_d_; }
    double w_scale_d() const { return ws_d_; }
    // FP64 index math on FP32 data (created as kPrecSingleIndex64).
    bool index64() const { return index64_; }
    int grid_blocks() const { return grid_blocks_; }
    int device() const { return device_; }
    // Device bytes of the plan's constant tables (kernel set).
    size_t table_bytes() const { return table_bytes_; }
    // Per-call workspace carved from device memory the caller provides
    // (no allocation): bytes a call of n points needs, and swapping such a
    // workspace in for one call (restore_workspace() puts the plan's own
    // back; the kernels already enqueued keep the carved one).  The
    // carved memory must not be in use by another call while this one runs.
    size_t external_workspace_bytes(size_t n) const;
    void use_external_workspace(void* base, size_t bytes, size_t n);
    void restore_workspace();

private:
    void ensure_capacity(size_t n);
    void build_tables_device(const float* d_conv);
    void free_workspace();
    template <typename T, typename TU = T, typename TV = T,
              typename Rec = typename Types<T>::Rec>
    void locate_t(size_t n, const TU* uu, const TU* vv, const TU* ww,
                  const TV* vis, const DiagBuffers* diag);
    template <typename T, typename TapT = typename Types<T>::T2,
              bool COMP = false>
    void scatter_t(size_t n, T* grid, int row0, const DiagBuffers* diag);
    void scan_u32(const unsigned* in, unsigned* out, const unsigned* count,
                  size_t max_n, unsigned* total);

    int prec_;
    bool index64_ = false;
    size_t P_;
    int ov_, C_, G_, h_, K_;
    int s_max_;
    float cell_f_, ws_f_, gscale_f_;
    double cell_d_, ws_d_, gscale_d_;
    int coord_bits_;     // bits of a grid coordinate
    int combo_bits_;     // bits of a sub-cell table index
    int plane_bits_;     // bits of a plane's index among its rank's planes
    int srank_bits_;     // bits of a support rank
    int tile_;           // tile side T (1..64)
    int tile_bits_;      // bits of a coordinate inside a tile (T <= 2^bits)
    int tile_coord_bits_; // bits of a tile coordinate
    int tile_shift_;     // key bits below the tile index
    int key_bits_;       // valid keys are < 2^key_bits_
    int key_bits_max_;   // key_bits_ for any tile side (sort scratch size)
    int region_;         // LDS region side T + Wmax - 1 (0: none)
    std::vector<int> support_;
    std::vector<int> srank_;
    std::vector<int> srank_key_;  // srank_ | index among the rank's planes << 16
    std::vector<PlaneInfo> pinfo_;
    int min_share_ = kMinShare; // least points per scatter stream
    int xcds_ = kXcdGroups;     // scatter item queues (1: grid-stride order)
    int slot_cap_ = 8;          // most footprint cells per lane
    int device_ = 0;
    int n_cu_ = 0;
    int grid_blocks_ = 0;
    hipStream_t stream_ = nullptr;
    bool own_stream_ = false;

    // Constant device tables.
    void* d_etab_ = nullptr;             // expanded taps (complex, T)
    size_t table_bytes_ = 0;
    size_t etab_taps_ = 0;
    int* d_support_ = nullptr;
    int* d_srank_ = nullptr;
    PlaneInfo* d_pinfo_ = nullptr;
    double* d_normtab_ = nullptr;

    // Workspace (grown on demand), or carved from external memory while
    // external_ is set (the plan's own pointers then wait in saved_).
    struct Workspace
    {
        size_t cap, items_cap;
        unsigned long long* keys[2];
        unsigned* idx[2];
        void* rec;
        int4* meta;
        void* vq;
        unsigned* gw;
        void* sort_tmp;
        size_t sort_tmp_bytes;
        unsigned *head, *bpos, *bstart, *scan_part;
        int4* items;
    };
    Workspace save_workspace() const;
    void set_workspace(const Workspace& w);
    Workspace saved_{};
    bool external_ = false;
    size_t cap_ = 0;
    size_t items_cap_ = 0;
    unsigned long long* d_keys_[2] = {nullptr, nullptr};
    unsigned* d_idx_[2] = {nullptr, nullptr};
    void* d_rec_ = nullptr;
    int4* d_meta_ = nullptr;
    void* d_vq_ = nullptr;
    unsigned* d_gw_ = nullptr;
    void* d_sort_tmp_ = nullptr;
    size_t sort_tmp_bytes_ = 0;
    unsigned* d_combo_ = nullptr;        // [P * K * K]
    unsigned* d_head_ = nullptr;         // [N] bucket heads / item counts
    unsigned* d_bpos_ = nullptr;         // [N] bucket index / item offsets
    unsigned* d_bstart_ = nullptr;       // [N + 1] bucket starts
    unsigned* d_scan_part_ = nullptr;    // scan partial sums
    int4* d_items_ = nullptr;
    Summary* d_summary_ = nullptr;
    Summary* h_summary_ = nullptr;       // pinned
    hipEvent_t ev_[5] = {};
    bool timed_ = false;
};

// One gridding job of a batch, as the batch kernels read it.
struct BatchJob
{
    const PlaneInfo* pinfo;   // the job's plan tables
    const void* etab;
    float* grid;              // 2*G*G floats
    int G, R, tile_bits, tile_shift;
    int key_bits, min_share, sub_h, sub_k;
    unsigned tile_base;       // first global tile of the job
    unsigned ntiles;          // 4^tile_coord_bits
    unsigned point_base;      // first global point of the job
    unsigned n;
    int pos_shift;            // position bits start here in the job's keys
    int tile;                 // tile side
    int planes;               // w-planes of the job's plan
    const double* normtab;    // its per-(plane, sub-cell) tap sums
    unsigned combo_off;       // its histogram in the batch's combo array
};

// Many independent FP32 gridding jobs (each with its own plan: kernel set,
// grid geometry and scatter shape; its own device points and grid) run as
// one pipeline: per-job index stages, then a counting sort of all points by
// (job, tile), work items sorted in LDS, and ONE persistent scatter launch
// over the work items of every job.  The enqueue sequence uses only the
// library's own kernels (no memset, no library sort), so it can be captured
// once into a HIP graph and replayed.
class Batch
{
public:
    Batch();
    ~Batch();
    // Adds a job; returns its index.  plan must be GW_SINGLE and outlive
    // the batch's use of it; pointers are device memory.
    int add(Plan* plan, size_t n, const float* uu, const float* vv,
            const float* ww, const float* vis, float* grid);
    void clear();
    size_t size() const { return jobs_.size(); }
    void set_stream(hipStream_t s);
    hipStream_t stream() const { return stream_; }
    // Enqueue the whole pipeline on the stream (eager launches).
    void run();
    // Capture the pipeline of the current jobs into a graph (instantiated).
    void capture();
    // Replay the captured graph (captures first if the jobs changed).
    void launch();
    // Waits for the stream; the job's skip count, norm increment and
    // update count of the last run / launch.
    void results(int job, size_t* skipped, double* norm_delta,
                 unsigned long long* updates);
    // Device milliseconds of the last run / launch (waits).
    double last_ms();
    // Device bytes held by the batch's workspace.
    size_t device_bytes() const;
    // Workspace bytes a batch of `points` points needs (tile and job
    // tables approximated by a fixed allowance).
    static size_t workspace_estimate(size_t points);
    // Frees the workspace (reallocated by the next run / capture).
    void release_workspace();
    // Event recorded after the last run / launch (for polling).
    hipEvent_t done_event() const { return ev_[1]; }

private:
    struct Entry
    {
        Plan* plan;
        size_t n;
        const float *uu, *vv, *ww, *vis;
        float* grid;
    };
    void prepare();
    void enqueue();
    void free_buffers();
    void drop_graph();

    std::vector<Entry> jobs_;
    std::vector<BatchJob> desc_;
    std::vector<size_t> combo_off_;
    bool dirty_ = true;
    size_t n_total_ = 0, tiles_total_ = 0, combo_total_ = 0;
    size_t items_cap_ = 0;
    size_t cap_points_ = 0, cap_tiles_ = 0, cap_combo_ = 0, cap_jobs_ = 0;
    BatchJob* d_desc_ = nullptr;
    Summary* d_sum_ = nullptr;        // per job
    Summary* d_bsum_ = nullptr;       // batch totals (nv32, total_items)
    Summary* h_sum_ = nullptr;        // pinned copy of d_sum_
    unsigned* d_combo_ = nullptr;
    unsigned long long* d_keys_[2] = {nullptr, nullptr};
    unsigned* d_idx_[2] = {nullptr, nullptr};
    void* d_rec_ = nullptr;
    int4* d_meta_ = nullptr;
    void* d_vq_ = nullptr;
    unsigned* d_tstart_ = nullptr;    // per global tile: first record
    unsigned* d_tend_ = nullptr;      // one past the last
                                      // (runs of equal keys: one record)
    unsigned* d_rstart_ = nullptr;    // per run: first sorted point
    unsigned* d_icnt_ = nullptr;
    unsigned* d_ioff_ = nullptr;
    unsigned* d_part_ = nullptr;
    unsigned* d_rhist_ = nullptr;     // radix digit counts per chunk
    unsigned* d_roff_ = nullptr;      // their exclusive scan
    unsigned* d_ntiles_ = nullptr;    // tiles (the tile scan's length)
    // Pinned staging of the job table and tile count, copied on the
    // batch's stream (stream order keeps the previous run's reads first);
    // ev_desc_ marks the copy done before the staging is reused.
    BatchJob* h_desc_ = nullptr;
    size_t h_desc_cap_ = 0;
    unsigned* h_ntiles_ = nullptr;
    hipEvent_t ev_desc_ = nullptr;
    bool desc_pending_ = false;
    int lmax_ = 0;                    // in-tile key bits of the batch key
    int radix_bits_ = 0;              // significant bits of the batch key
    int4* d_items_ = nullptr;
    hipStream_t stream_ = nullptr;
    bool own_stream_ = false;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
    hipEvent_t ev_[2] = {};
    bool timed_ = false;
    int grid_blocks_ = 0;
    int n_cu_ = 0;
};

// Stable LSD radix sort of n (key, value) pairs on the low `bits` bits of
// the keys (8-bit digits, one pass per digit: histogram, per-digit scan,
// scatter), capture-safe.  keys[0] holds the input; the values start as
// 0..n-1 (each key's input position; vals[0] is not read); keys[1] /
// vals[1] are scratch of the same size; returns which pair holds the
// result.  hist and offs hold radix_hist_words(n) words, tot 256 words.
size_t radix_hist_words(size_t n);
int radix_sort_enqueue(size_t n, int bits, unsigned long long* keys[2],
                       unsigned* vals[2], unsigned* hist, unsigned* offs,
                       unsigned* tot, hipStream_t s);

// Uniform-weighting step (oskar_grid_weights_write / _read) on device
// pointers, FP32 (kPrecSingle) or FP64 (kPrecDouble); each call runs on `s`
// (nullptr: the object's own stream), waits, and returns the number of
// points off the grid.
class Weights
{
public:
    Weights();
    ~Weights();
    void set_stream(hipStream_t s);
    hipStream_t stream() const { return stream_; }
    size_t write(int prec, size_t n, const void* uu, const void* vv,
                 const void* weight, double cell, int G, void* grid,
                 hipStream_t s = nullptr);
    size_t read(int prec, size_t n, const void* uu, const void* vv,
                const void* w_in, void* w_out, double cell, int G,
                const void* grid, hipStream_t s = nullptr);
    size_t device_bytes() const;

private:
    template <typename T>
    size_t write_t(size_t n, const T* uu, const T* vv, const T* weight,
                   T cell, int G, T* grid, hipStream_t st);
    template <typename T>
    size_t read_t(size_t n, const T* uu, const T* vv, const T* w_in,
                  T* w_out, T cell, int G, const T* grid, hipStream_t st);
    void ensure_capacity(size_t n);
    void free_buffers();
    size_t finish_skipped(hipStream_t st);

    hipStream_t stream_ = nullptr;
    bool own_stream_ = false;
    int n_cu_ = 0;
    size_t cap_ = 0;
    unsigned long long* d_keys_[2] = {nullptr, nullptr};
    unsigned* d_idx_[2] = {nullptr, nullptr};
    unsigned* d_hist_ = nullptr;
    unsigned* d_offs_ = nullptr;
    unsigned* d_part_ = nullptr;
    unsigned long long* d_small_ = nullptr;   // skipped points
    unsigned long long* h_skipped_ = nullptr;
};

// W-kernel cube generated on the device (float, the reference cube's
// layout) from the plane supports, oversample, cube half size and w_max;
// enqueued on `stream`, waits.
void generate_wkernels(size_t num_w_planes, const int* support,
                       int oversample, int conv_size_half, double w_max,
                       float* d_out, hipStream_t stream);

// Zero-fill with a kernel (capture-safe, unlike hipMemsetAsync).
void fill_zero_u64(unsigned long long* p, size_t n, hipStream_t s);
void fill_zero_u32(unsigned* p, size_t n, hipStream_t s);

}  // namespace gw
