// W-kernel cubes generated on the device.
//
// The reference reads precomputed kernels from FITS (main.cpp:183-201,
// read_kernel.c:36-118); a server that takes kernel sets by their
// parameters generates them instead.  The construction is the one the
// benchmark workloads use: per plane p of P, with support S_p and
// oversample ov, on the quadrant grid (ix, iy) in [0, C)^2 at radius
// r^2 = (ix^2 + iy^2) / ov^2 (pixels),
//
//   k_p(r) = exp(-r^2 / (2 sigma_p^2)) * exp(i pi w_p s r^2),
//   sigma_p = max(S_p, 1) / 2.5,  w_p = (p / (P-1))^2 w_max,
//   s = 2 / (w_max * max(S)^2),
//
// normalised by its complex sum over the (2 S_p + 1)^2 footprint taps at
// offset zero (samples |j| ov, |k| ov), all in FP64, then rounded to
// float.  The layout is the reference cube's,
// conv_func[2 * ((p * C + iy) * C + ix) + {0, 1}] (oskar_grid_wproj.c:
// 222-233).
#include "gridder_internal.h"

#include <algorithm>
#include <cmath>

namespace gw {

namespace {

struct WkParams
{
    int P, ov, C, s_max;
    double w_max, chirp;
};

__device__ __forceinline__ void wk_value(const WkParams& q, int p, int S,
                                         int iy, int ix, double& re,
                                         double& im)
{
    const double fy = (double)iy / q.ov, fx = (double)ix / q.ov;
    const double r2 = fx * fx + fy * fy;
    const double sigma = (double)max(S, 1) / 2.5;
    const double fp = q.P > 1 ? (double)p / (q.P - 1) : 0.0;
    const double w = q.P > 1 ? fp * fp * q.w_max : 0.0;
    const double g = exp(-r2 / (2.0 * sigma * sigma));
    const double th = M_PI * w * q.chirp * r2;
    re = g * cos(th);
    im = g * sin(th);
}

// Complex sum of plane p's footprint taps at offset zero (one block per
// plane; FP64 block reduction).
__global__ void __launch_bounds__(256) k_wk_sum(WkParams q,
        const int* __restrict__ support, double2* __restrict__ total)
{
    __shared__ double s_re[256], s_im[256];
    const int p = blockIdx.x;
    const int S = support[p];
    const int W = 2 * S + 1;
    double re = 0.0, im = 0.0;
    for (int t = threadIdx.x; t < W * W; t += blockDim.x)
    {
        const int j = t / W - S, k = t % W - S;
        double a, b;
        wk_value(q, p, S, abs(j) * q.ov, abs(k) * q.ov, a, b);
        re += a;
        im += b;
    }
    s_re[threadIdx.x] = re;
    s_im[threadIdx.x] = im;
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1)
    {
        if (threadIdx.x < h)
        {
            s_re[threadIdx.x] += s_re[threadIdx.x + h];
            s_im[threadIdx.x] += s_im[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) total[p] = make_double2(s_re[0], s_im[0]);
}

// Every cube value divided by its plane's sum, rounded to float.
__global__ void __launch_bounds__(256) k_wk_fill(WkParams q,
        const int* __restrict__ support, const double2* __restrict__ total,
        float* __restrict__ out)
{
    const size_t n = (size_t)q.P * q.C * q.C;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
    {
        const int p = (int)(i / ((size_t)q.C * q.C));
        const int r = (int)(i - (size_t)p * q.C * q.C);
        const int iy = r / q.C, ix = r - iy * q.C;
        double a, b;
        wk_value(q, p, support[p], iy, ix, a, b);
        // (a + ib) / (c + id) by Smith's method (Smith, CACM 5(8), 1962).
        const double c = total[p].x, d = total[p].y;
        double re, im;
        if (fabs(c) >= fabs(d))
        {
            const double rat = d / c, scl = 1.0 / (c + d * rat);
            re = (a + b * rat) * scl;
            im = (b - a * rat) * scl;
        }
        else
        {
            const double rat = c / d, scl = 1.0 / (d + c * rat);
            re = (a * rat + b) * scl;
            im = (b * rat - a) * scl;
        }
        out[2 * i] = (float)re;
        out[2 * i + 1] = (float)im;
    }
}

}  // namespace

void generate_wkernels(size_t num_w_planes, const int* support,
                       int oversample, int conv_size_half, double w_max,
                       float* d_out, hipStream_t stream)
{
    validate_args(num_w_planes, support, oversample, conv_size_half,
                  d_out, 1);
    if (!(w_max >= 0.0) || !std::isfinite(w_max))
        fail(kErrInvalidArgument, "w_max must be finite and >= 0");
    WkParams q;
    q.P = (int)num_w_planes;
    q.ov = oversample;
    q.C = conv_size_half;
    q.s_max = *std::max_element(support, support + num_w_planes);
    q.w_max = w_max;
    const int sm = std::max(q.s_max, 1);
    q.chirp = 2.0 / (std::max(w_max, 1e-30) * (double)(sm * sm));
    int* d_support = dmalloc<int>(num_w_planes);
    double2* d_total = dmalloc<double2>(num_w_planes);
    try
    {
        check_hip(hipMemcpyAsync(d_support, support,
                num_w_planes * sizeof(int), hipMemcpyHostToDevice, stream),
                kErrMemCopy, "upload support");
        hipLaunchKernelGGL(k_wk_sum, dim3((unsigned)num_w_planes), dim3(256),
                0, stream, q, (const int*)d_support, d_total);
        const size_t n = num_w_planes * (size_t)conv_size_half *
                conv_size_half;
        const int blocks = (int)std::max<size_t>(1, std::min<size_t>(
                (n + 255) / 256, 4096));
        hipLaunchKernelGGL(k_wk_fill, dim3(blocks), dim3(256), 0, stream, q,
                (const int*)d_support, (const double2*)d_total, d_out);
        check_hip(hipGetLastError(), kErrKernelLaunch, "kernel generation");
        check_hip(hipStreamSynchronize(stream), kErrKernelLaunch,
                "stream synchronize");
    }
    catch (...)
    {
        (void)hipStreamSynchronize(stream);
        dfree(d_support);
        dfree(d_total);
        throw;
    }
    dfree(d_support);
    dfree(d_total);
}

}  // namespace gw
