/*
 * CPU oracle for the W-projection gridder.  TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load
 * this library, and only as the checker (or the timed CPU baseline).  The
 * product library (fpga_gridding_amd/libfpga_gridding_amd.so) never links,
 * loads or calls it.
 *
 * It restates the reference algorithm of OxfordSKA/FPGA_Gridding:
 *   gridref_oskar_grid_wproj_f  <- oskar_grid_wproj.c:128-243 (oskar_grid_wproj_f)
 *                                  without the debug truncation to 10000 points
 *                                  and the progress printf at :154-157, i.e. the
 *                                  full-length loop of :276 and :65;
 *   gridref_oskar_grid_wproj_d  <- oskar_grid_wproj.c:39-126
 *   gridref_oskar_grid_wproj_f_diagnostics <- oskar_grid_wproj.c:245-370
 *   gridref_grid_weights_*      <- utils/oskar_grid_weights.c:38-172
 *
 * Floating-point semantics follow the reference exactly so that results are
 * bit-identical to a gcc build of the reference with the same flags
 * (-O2 -ffp-contract=off): per-point float index math (:169-196), weight
 * fixed to 1 (:185), double `sum` of float taps (:165, :238), footprint
 * rows j outer and columns k inner (:216-240), `norm += sum * 1.0f` (:241).
 *
 * Extra entry points (gridref_locate_f/_d, gridref_cells_f/_d) expose the
 * per-point index stage and single-cell brute force for tests.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define EXPORT __attribute__((visibility("default")))

/* ---------------------------------------------------------------------- */
/* Index stage (oskar_grid_wproj.c:169-212, FP32 and :71-101, FP64)        */
/* ---------------------------------------------------------------------- */

typedef struct
{
    int grid_u, grid_v;     /* nearest grid cell of the point          */
    int off_u, off_v;       /* sub-cell offset in oversampled units    */
    size_t grid_w;          /* unclamped plane index (diagnostics use)  */
    size_t plane;           /* plane index clamped to num_w_planes - 1  */
    int support;            /* support[plane]                          */
    int skip;               /* footprint crosses the grid edge          */
    int conj;               /* 1 when ww > 0 (tap imaginary part negated) */
} locate_t;

static void locate_f(const float uu, const float vv, const float ww,
        const float grid_scale, const float w_scale, const int oversample,
        const int grid_size, const size_t num_w_planes, const int* support,
        locate_t* r)
{
    const int centre = grid_size / 2;
    const float pu = -uu * grid_scale;
    const float pv = vv * grid_scale;
    const float ru = roundf(pu);
    const float rv = roundf(pv);
    r->conj = (ww > 0.0f);
    r->grid_w = (size_t)roundf(sqrtf(fabsf(ww * w_scale)));
    r->grid_u = (int)ru + centre;
    r->grid_v = (int)rv + centre;
    r->off_u = (int)roundf((ru - pu) * oversample);
    r->off_v = (int)roundf((rv - pv) * oversample);
    r->plane = r->grid_w < num_w_planes ? r->grid_w : num_w_planes - 1;
    r->support = support[r->plane];
    r->skip = (r->grid_u + r->support >= grid_size ||
            r->grid_u - r->support < 0 ||
            r->grid_v + r->support >= grid_size ||
            r->grid_v - r->support < 0);
}

static void locate_d(const double uu, const double vv, const double ww,
        const double grid_scale, const double w_scale, const int oversample,
        const int grid_size, const size_t num_w_planes, const int* support,
        locate_t* r)
{
    const int centre = grid_size / 2;
    const double pu = -uu * grid_scale;
    const double pv = vv * grid_scale;
    const double ru = round(pu);
    const double rv = round(pv);
    r->conj = (ww > 0.0);
    r->grid_w = (size_t)round(sqrt(fabs(ww * w_scale)));
    r->grid_u = (int)ru + centre;
    r->grid_v = (int)rv + centre;
    r->off_u = (int)round((ru - pu) * oversample);
    r->off_v = (int)round((rv - pv) * oversample);
    r->plane = r->grid_w < num_w_planes ? r->grid_w : num_w_planes - 1;
    r->support = support[r->plane];
    r->skip = (r->grid_u + r->support >= grid_size ||
            r->grid_u - r->support < 0 ||
            r->grid_v + r->support >= grid_size ||
            r->grid_v - r->support < 0);
}

/* ---------------------------------------------------------------------- */
/* Footprint scatter (oskar_grid_wproj.c:216-241)                          */
/* ---------------------------------------------------------------------- */

/* Scatter one located point, FP32 grid.  Returns the double tap sum. */
static double scatter_f(const locate_t* r, const float v_re, const float v_im,
        const int oversample, const int conv_size_half,
        const float* conv_func, const int grid_size, float* grid)
{
    const size_t kdim = (size_t)conv_size_half * conv_size_half;
    const size_t kbase = r->plane * kdim;
    const float sgn = r->conj ? -1.0f : 1.0f;
    const int s = r->support;
    double sum = 0.0;
    int j, k;
    for (j = -s; j <= s; ++j)
    {
        const size_t row = ((size_t)(r->grid_v + j)) * grid_size + r->grid_u;
        const size_t trow = kbase +
                (size_t)abs(r->off_v + j * oversample) * conv_size_half;
        for (k = -s; k <= s; ++k)
        {
            const size_t t = (trow + abs(r->off_u + k * oversample)) << 1;
            const float c_re = conv_func[t];
            const float c_im = conv_func[t + 1] * sgn;
            const size_t g = (row + k) << 1;
            grid[g]     += (v_re * c_re - v_im * c_im);
            grid[g + 1] += (v_im * c_re + v_re * c_im);
            sum += c_re;
        }
    }
    return sum;
}

static double scatter_d(const locate_t* r, const double v_re,
        const double v_im, const int oversample, const int conv_size_half,
        const double* conv_func, const int grid_size, double* grid)
{
    const size_t kdim = (size_t)conv_size_half * conv_size_half;
    const size_t kbase = r->plane * kdim;
    const double sgn = r->conj ? -1.0 : 1.0;
    const int s = r->support;
    double sum = 0.0;
    int j, k;
    for (j = -s; j <= s; ++j)
    {
        const size_t row = ((size_t)(r->grid_v + j)) * grid_size + r->grid_u;
        const size_t trow = kbase +
                (size_t)abs(r->off_v + j * oversample) * conv_size_half;
        for (k = -s; k <= s; ++k)
        {
            const size_t t = (trow + abs(r->off_u + k * oversample)) << 1;
            const double c_re = conv_func[t];
            const double c_im = conv_func[t + 1] * sgn;
            const size_t g = (row + k) << 1;
            grid[g]     += (v_re * c_re - v_im * c_im);
            grid[g + 1] += (v_im * c_re + v_re * c_im);
            sum += c_re;
        }
    }
    return sum;
}

/* ---------------------------------------------------------------------- */
/* Reference-equivalent entry points                                       */
/* ---------------------------------------------------------------------- */

EXPORT void gridref_oskar_grid_wproj_f(const size_t num_w_planes,
        const int* support, const int oversample, const int conv_size_half,
        const float* conv_func, const size_t num_points, const float* uu,
        const float* vv, const float* ww, const float* vis,
        const float* weight, const float cell_size_rad, const float w_scale,
        const int grid_size, size_t* num_skipped, double* norm, float* grid)
{
    const float grid_scale = grid_size * cell_size_rad;
    const float weight_i = 1.0f;          /* weights ignored (:185) */
    size_t i;
    (void)weight;
    *num_skipped = 0;
    for (i = 0; i < num_points; ++i)
    {
        locate_t r;
        locate_f(uu[i], vv[i], ww[i], grid_scale, w_scale, oversample,
                grid_size, num_w_planes, support, &r);
        if (r.skip) { *num_skipped += 1; continue; }
        *norm += scatter_f(&r, weight_i * vis[2 * i],
                weight_i * vis[2 * i + 1], oversample, conv_size_half,
                conv_func, grid_size, grid) * weight_i;
    }
}

EXPORT void gridref_oskar_grid_wproj_d(const size_t num_w_planes,
        const int* support, const int oversample, const int conv_size_half,
        const double* conv_func, const size_t num_points, const double* uu,
        const double* vv, const double* ww, const double* vis,
        const double* weight, const double cell_size_rad,
        const double w_scale, const int grid_size, size_t* num_skipped,
        double* norm, double* grid)
{
    const double grid_scale = grid_size * cell_size_rad;
    const double weight_i = 1.0;          /* weights ignored (:81) */
    size_t i;
    (void)weight;
    *num_skipped = 0;
    for (i = 0; i < num_points; ++i)
    {
        locate_t r;
        locate_d(uu[i], vv[i], ww[i], grid_scale, w_scale, oversample,
                grid_size, num_w_planes, support, &r);
        if (r.skip) { *num_skipped += 1; continue; }
        *norm += scatter_d(&r, weight_i * vis[2 * i],
                weight_i * vis[2 * i + 1], oversample, conv_size_half,
                conv_func, grid_size, grid) * weight_i;
    }
}

/* The exact value of what oskar_grid_wproj_f computes: the FP32 index math
 * of the reference (same plane, cell, sub-cell and taps), but products and
 * accumulation carried in float64 into a float64 grid.  The FP32 gridders
 * (the reference's sequential one and any reordered one) differ from this
 * only by float rounding of their sums, so it measures both sides' error. */
EXPORT void gridref_oskar_grid_wproj_f_exact(const size_t num_w_planes,
        const int* support, const int oversample, const int conv_size_half,
        const float* conv_func, const size_t num_points, const float* uu,
        const float* vv, const float* ww, const float* vis,
        const float cell_size_rad, const float w_scale, const int grid_size,
        size_t* num_skipped, double* norm, double* grid)
{
    const float grid_scale = grid_size * cell_size_rad;
    const size_t kdim = (size_t)conv_size_half * conv_size_half;
    size_t i;
    *num_skipped = 0;
    for (i = 0; i < num_points; ++i)
    {
        locate_t r;
        int j, k;
        double sum = 0.0;
        locate_f(uu[i], vv[i], ww[i], grid_scale, w_scale, oversample,
                grid_size, num_w_planes, support, &r);
        if (r.skip) { *num_skipped += 1; continue; }
        for (j = -r.support; j <= r.support; ++j)
        {
            const size_t row = ((size_t)(r.grid_v + j)) * grid_size +
                    r.grid_u;
            const size_t trow = r.plane * kdim +
                    (size_t)abs(r.off_v + j * oversample) * conv_size_half;
            for (k = -r.support; k <= r.support; ++k)
            {
                const size_t t = (trow + abs(r.off_u + k * oversample)) << 1;
                const double c_re = conv_func[t];
                const double c_im = (double)conv_func[t + 1] *
                        (r.conj ? -1.0 : 1.0);
                const size_t g = (row + k) << 1;
                grid[g]     += (double)vis[2 * i] * c_re -
                        (double)vis[2 * i + 1] * c_im;
                grid[g + 1] += (double)vis[2 * i + 1] * c_re +
                        (double)vis[2 * i] * c_im;
                sum += c_re;
            }
        }
        *norm += sum;
    }
}

/* oskar_grid_wproj.c:245-370.  Histograms are counted before the skip test
 * (:317-318); num_vis_per_wplane uses the UNCLAMPED plane index; both
 * w_support and grid_w are added into avg_wsupport_per_grid (:359-360) and
 * avg_wplane_per_grid is never written. */
EXPORT void gridref_oskar_grid_wproj_f_diagnostics(const size_t num_w_planes,
        const int* support, const int oversample, const int conv_size_half,
        const float* conv_func, const size_t num_points, const float* uu,
        const float* vv, const float* ww, const float* vis,
        const float* weight, const float cell_size_rad, const float w_scale,
        const int grid_size, size_t* num_skipped, double* norm, float* grid,
        long long* num_vis_per_wsupport, long long* num_vis_per_wplane,
        long long* num_vis_per_grid, float* avg_wsupport_per_grid,
        float* avg_wplane_per_grid)
{
    const float grid_scale = grid_size * cell_size_rad;
    const float weight_i = 1.0f;
    size_t i;
    (void)weight;
    (void)avg_wplane_per_grid;
    *num_skipped = 0;
    for (i = 0; i < num_points; ++i)
    {
        locate_t r;
        int j, k;
        locate_f(uu[i], vv[i], ww[i], grid_scale, w_scale, oversample,
                grid_size, num_w_planes, support, &r);
        num_vis_per_wsupport[r.support] += 1;
        num_vis_per_wplane[r.grid_w] += 1;
        if (r.skip) { *num_skipped += 1; continue; }
        *norm += scatter_f(&r, weight_i * vis[2 * i],
                weight_i * vis[2 * i + 1], oversample, conv_size_half,
                conv_func, grid_size, grid) * weight_i;
        for (j = -r.support; j <= r.support; ++j)
        {
            for (k = -r.support; k <= r.support; ++k)
            {
                const size_t c = ((size_t)(r.grid_v + j)) * grid_size +
                        r.grid_u + k;
                num_vis_per_grid[c]++;
                avg_wsupport_per_grid[c] += r.support;
                avg_wsupport_per_grid[c] += r.grid_w;
            }
        }
    }
}

/* ---------------------------------------------------------------------- */
/* Per-point index stage and single-cell brute force (test helpers)        */
/* ---------------------------------------------------------------------- */

/* out[i*8 + {0..7}] = grid_u, grid_v, off_u, off_v, plane, support, skip,
 * conj; grid_w_out[i] = unclamped plane index.  Returns U = sum over
 * non-skipped points of (2S+1)^2. */
EXPORT unsigned long long gridref_locate_f(const size_t num_w_planes,
        const int* support, const int oversample, const size_t num_points,
        const float* uu, const float* vv, const float* ww,
        const float cell_size_rad, const float w_scale, const int grid_size,
        int* out, unsigned long long* grid_w_out)
{
    const float grid_scale = grid_size * cell_size_rad;
    unsigned long long updates = 0;
    size_t i;
    for (i = 0; i < num_points; ++i)
    {
        locate_t r;
        locate_f(uu[i], vv[i], ww[i], grid_scale, w_scale, oversample,
                grid_size, num_w_planes, support, &r);
        if (out)
        {
            int* o = out + 8 * i;
            o[0] = r.grid_u; o[1] = r.grid_v; o[2] = r.off_u; o[3] = r.off_v;
            o[4] = (int)r.plane; o[5] = r.support; o[6] = r.skip;
            o[7] = r.conj;
        }
        if (grid_w_out) grid_w_out[i] = (unsigned long long)r.grid_w;
        if (!r.skip)
            updates += (unsigned long long)(2 * r.support + 1) *
                    (2 * r.support + 1);
    }
    return updates;
}

EXPORT unsigned long long gridref_locate_d(const size_t num_w_planes,
        const int* support, const int oversample, const size_t num_points,
        const double* uu, const double* vv, const double* ww,
        const double cell_size_rad, const double w_scale, const int grid_size,
        int* out, unsigned long long* grid_w_out)
{
    const double grid_scale = grid_size * cell_size_rad;
    unsigned long long updates = 0;
    size_t i;
    for (i = 0; i < num_points; ++i)
    {
        locate_t r;
        locate_d(uu[i], vv[i], ww[i], grid_scale, w_scale, oversample,
                grid_size, num_w_planes, support, &r);
        if (out)
        {
            int* o = out + 8 * i;
            o[0] = r.grid_u; o[1] = r.grid_v; o[2] = r.off_u; o[3] = r.off_v;
            o[4] = (int)r.plane; o[5] = r.support; o[6] = r.skip;
            o[7] = r.conj;
        }
        if (grid_w_out) grid_w_out[i] = (unsigned long long)r.grid_w;
        if (!r.skip)
            updates += (unsigned long long)(2 * r.support + 1) *
                    (2 * r.support + 1);
    }
    return updates;
}

/* Value of a set of grid cells, in float64, summed over every point whose
 * footprint covers the cell (the same taps and products as the scatter, with
 * the products and their sum carried in double).  cells[c] = v*G + u;
 * out[2c], out[2c+1] = re, im contribution (excluding the grid's prior
 * contents).  Cost O(num_points * num_cells); used to spot-check large grids. */
EXPORT void gridref_cells_f(const size_t num_w_planes, const int* support,
        const int oversample, const int conv_size_half,
        const float* conv_func, const size_t num_points, const float* uu,
        const float* vv, const float* ww, const float* vis,
        const float cell_size_rad, const float w_scale, const int grid_size,
        const size_t num_cells, const long long* cells, double* out)
{
    const float grid_scale = grid_size * cell_size_rad;
    const size_t kdim = (size_t)conv_size_half * conv_size_half;
    size_t i, c;
    memset(out, 0, 2 * num_cells * sizeof(double));
    for (i = 0; i < num_points; ++i)
    {
        locate_t r;
        locate_f(uu[i], vv[i], ww[i], grid_scale, w_scale, oversample,
                grid_size, num_w_planes, support, &r);
        if (r.skip) continue;
        for (c = 0; c < num_cells; ++c)
        {
            const long long cu = cells[c] % grid_size;
            const long long cv = cells[c] / grid_size;
            const long long k = cu - r.grid_u, j = cv - r.grid_v;
            size_t t;
            double c_re, c_im;
            if (k < -r.support || k > r.support ||
                    j < -r.support || j > r.support)
                continue;
            t = (r.plane * kdim +
                    (size_t)abs(r.off_v + (int)j * oversample) *
                    conv_size_half +
                    (size_t)abs(r.off_u + (int)k * oversample)) << 1;
            c_re = conv_func[t];
            c_im = (double)conv_func[t + 1] * (r.conj ? -1.0 : 1.0);
            out[2 * c]     += (double)vis[2 * i] * c_re -
                    (double)vis[2 * i + 1] * c_im;
            out[2 * c + 1] += (double)vis[2 * i + 1] * c_re +
                    (double)vis[2 * i] * c_im;
        }
    }
}

/* The same brute force with the FP64 index math and arithmetic of
 * oskar_grid_wproj_d (oskar_grid_wproj.c:71-112): double coordinates,
 * cube and visibilities.  Used to spot-check the FP64-index plans (the
 * reduced-precision ones on their stored inputs) on large grids. */
EXPORT void gridref_cells_d(const size_t num_w_planes, const int* support,
        const int oversample, const int conv_size_half,
        const double* conv_func, const size_t num_points, const double* uu,
        const double* vv, const double* ww, const double* vis,
        const double cell_size_rad, const double w_scale, const int grid_size,
        const size_t num_cells, const long long* cells, double* out)
{
    const double grid_scale = grid_size * cell_size_rad;
    const size_t kdim = (size_t)conv_size_half * conv_size_half;
    size_t i, c;
    memset(out, 0, 2 * num_cells * sizeof(double));
    for (i = 0; i < num_points; ++i)
    {
        locate_t r;
        locate_d(uu[i], vv[i], ww[i], grid_scale, w_scale, oversample,
                grid_size, num_w_planes, support, &r);
        if (r.skip) continue;
        for (c = 0; c < num_cells; ++c)
        {
            const long long cu = cells[c] % grid_size;
            const long long cv = cells[c] / grid_size;
            const long long k = cu - r.grid_u, j = cv - r.grid_v;
            size_t t;
            double c_re, c_im;
            if (k < -r.support || k > r.support ||
                    j < -r.support || j > r.support)
                continue;
            t = (r.plane * kdim +
                    (size_t)abs(r.off_v + (int)j * oversample) *
                    conv_size_half +
                    (size_t)abs(r.off_u + (int)k * oversample)) << 1;
            c_re = conv_func[t];
            c_im = conv_func[t + 1] * (r.conj ? -1.0 : 1.0);
            out[2 * c]     += vis[2 * i] * c_re - vis[2 * i + 1] * c_im;
            out[2 * c + 1] += vis[2 * i + 1] * c_re + vis[2 * i] * c_im;
        }
    }
}

/* ---------------------------------------------------------------------- */
/* Uniform weighting (utils/oskar_grid_weights.c:38-172)                   */
/* ---------------------------------------------------------------------- */

EXPORT void gridref_grid_weights_write_f(const size_t num_points,
        const float* uu, const float* vv, const float* weight,
        const float cell_size_rad, const int grid_size, size_t* num_skipped,
        float* grid)
{
    const int centre = grid_size / 2;
    const float grid_scale = grid_size * cell_size_rad;
    size_t i;
    *num_skipped = 0;
    for (i = 0; i < num_points; ++i)
    {
        const int gu = (int)roundf(-uu[i] * grid_scale) + centre;
        const int gv = (int)roundf(vv[i] * grid_scale) + centre;
        if (gu >= grid_size || gu < 0 || gv >= grid_size || gv < 0)
        {
            *num_skipped += 1;
            continue;
        }
        grid[(size_t)gv * grid_size + gu] += weight[i];
    }
}

EXPORT void gridref_grid_weights_read_f(const size_t num_points,
        const float* uu, const float* vv, const float* weight_in,
        float* weight_out, const float cell_size_rad, const int grid_size,
        size_t* num_skipped, const float* grid)
{
    const int centre = grid_size / 2;
    const float grid_scale = grid_size * cell_size_rad;
    size_t i;
    *num_skipped = 0;
    for (i = 0; i < num_points; ++i)
    {
        const int gu = (int)roundf(-uu[i] * grid_scale) + centre;
        const int gv = (int)roundf(vv[i] * grid_scale) + centre;
        float d;
        if (gu >= grid_size || gu < 0 || gv >= grid_size || gv < 0)
        {
            *num_skipped += 1;
            continue;
        }
        d = grid[(size_t)gv * grid_size + gu];
        weight_out[i] = (d != 0.0) ? weight_in[i] / d : 0.0;
    }
}

EXPORT void gridref_grid_weights_write_d(const size_t num_points,
        const double* uu, const double* vv, const double* weight,
        const double cell_size_rad, const int grid_size, size_t* num_skipped,
        double* grid)
{
    const int centre = grid_size / 2;
    const double grid_scale = grid_size * cell_size_rad;
    size_t i;
    *num_skipped = 0;
    for (i = 0; i < num_points; ++i)
    {
        const int gu = (int)round(-uu[i] * grid_scale) + centre;
        const int gv = (int)round(vv[i] * grid_scale) + centre;
        if (gu >= grid_size || gu < 0 || gv >= grid_size || gv < 0)
        {
            *num_skipped += 1;
            continue;
        }
        grid[(size_t)gv * grid_size + gu] += weight[i];
    }
}

EXPORT void gridref_grid_weights_read_d(const size_t num_points,
        const double* uu, const double* vv, const double* weight_in,
        double* weight_out, const double cell_size_rad, const int grid_size,
        size_t* num_skipped, const double* grid)
{
    const int centre = grid_size / 2;
    const double grid_scale = grid_size * cell_size_rad;
    size_t i;
    *num_skipped = 0;
    for (i = 0; i < num_points; ++i)
    {
        const int gu = (int)round(-uu[i] * grid_scale) + centre;
        const int gv = (int)round(vv[i] * grid_scale) + centre;
        double d;
        if (gu >= grid_size || gu < 0 || gv >= grid_size || gv < 0)
        {
            *num_skipped += 1;
            continue;
        }
        d = grid[(size_t)gv * grid_size + gu];
        weight_out[i] = (d != 0.0) ? weight_in[i] / d : 0.0;
    }
}
