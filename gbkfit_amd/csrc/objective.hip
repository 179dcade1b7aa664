// Objective stages (core/objective/objective.hpp:7-33; host reductions
// gbkfit_host/include/gbkfit/host/kernels.hpp:425-514).
//
// Reductions are deterministic: every block reduces a fixed slice in FP64,
// partials land in a per-block slot, and one block sums the slots in order,
// so results do not depend on scheduling (the reference's OpenMP FP32
// reduction does).
#include "common.hpp"

namespace gbkhip {

constexpr int kRedBlock = 256;
constexpr int kRedBlocks = 1024;

__global__ void residual_kernel(const float* __restrict__ od,
                                const float* __restrict__ oe,
                                const float* __restrict__ om,
                                const float* __restrict__ md,
                                const float* __restrict__ mw,
                                const float* __restrict__ mm, long long n,
                                float weight, float* __restrict__ res)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        float r = md[i] - od[i];
        if (oe) r /= oe[i];
        if (mw) r *= mw[i];
        // a zero mask entry gives a zero residual even where the data (or
        // the model, after a mask with apply) hold the NaN that marks the
        // pixel as masked; the reference multiplies, NaN * 0 = NaN
        if (om) r = om[i] != 0.0f ? r * om[i] : 0.0f;
        if (mm) r = mm[i] != 0.0f ? r * mm[i] : 0.0f;
        res[i] = r * weight;
    }
}

__device__ __forceinline__ double block_sum(double v, double* sh)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += sh[i];
    __syncthreads();
    return t;
}

// Stage 1: block b sums elements [b*chunk, (b+1)*chunk) in FP64.
__global__ __launch_bounds__(kRedBlock) void sum_partials(
    const float* __restrict__ r, long long n, long long chunk, int squared,
    double* __restrict__ partial)
{
    __shared__ double sh[kRedBlock / 64];
    const long long b0 = blockIdx.x * chunk;
    const long long b1 = min(n, b0 + chunk);
    double acc = 0.0;
    for (long long i = b0 + threadIdx.x; i < b1; i += blockDim.x) {
        const double v = (double)r[i];
        acc += squared ? v * v : fabs(v);
    }
    const double t = block_sum(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// Stage 2: one block adds the partials in a fixed order.
__global__ __launch_bounds__(kRedBlock) void sum_final(
    const double* __restrict__ partial, int nparts, float* out_f,
    double* out_d)
{
    __shared__ double sh[kRedBlock / 64];
    double acc = 0.0;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) acc += partial[i];
    const double t = block_sum(acc, sh);
    if (threadIdx.x == 0) {
        if (out_f) out_f[0] = (float)t;
        if (out_d) out_d[0] = t;
    }
}

__global__ void count_kernel(const float* __restrict__ a,
                             const float* __restrict__ b, long long n,
                             float eps, int* counts)
{
    int c1 = 0, c2 = 0, cb = 0;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const bool h1 = fabsf(a[i]) > eps, h2 = fabsf(b[i]) > eps;
        c1 += h1 && !h2;
        c2 += !h1 && h2;
        cb += h1 && h2;
    }
    for (int o = 32; o > 0; o >>= 1) {
        c1 += __shfl_xor(c1, o, 64);
        c2 += __shfl_xor(c2, o, 64);
        cb += __shfl_xor(cb, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        if (c1) atomicAdd(&counts[0], c1);
        if (c2) atomicAdd(&counts[1], c2);
        if (cb) atomicAdd(&counts[2], cb);
    }
}

void launch_residual(Context& c, const float* od, const float* oe,
                     const float* om, const float* md, const float* mw,
                     const float* mm, long long n, float weight, float* res)
{
    if (n <= 0) return;
    long long blocks = std::min<long long>((n + 255) / 256, 65536LL * 4);
    hipLaunchKernelGGL(residual_kernel, dim3((unsigned)blocks), dim3(256), 0,
                       c.stream, od, oe, om, md, mw, mm, n, weight, res);
    GBK_HIP(hipGetLastError());
}

void launch_sum(Context& c, const float* r, long long n, int squared,
                float* out_f, double* out_d)
{
    c.reduce_scratch.ensure(kRedBlocks * sizeof(double));
    double* part = c.reduce_scratch.as<double>();
    const long long chunk = std::max<long long>(
        kRedBlock, (n + kRedBlocks - 1) / kRedBlocks);
    const int nparts = n > 0 ? (int)((n + chunk - 1) / chunk) : 0;
    if (n > 0) {
        hipLaunchKernelGGL(sum_partials, dim3(nparts), dim3(kRedBlock), 0,
                           c.stream, r, n, chunk, squared, part);
        GBK_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(sum_final, dim3(1), dim3(kRedBlock), 0, c.stream, part,
                       nparts, out_f, out_d);
    GBK_HIP(hipGetLastError());
}

void launch_count(Context& c, const float* a, const float* b, long long n,
                  float eps, int* counts)
{
    if (n <= 0) return;
    long long blocks = std::min<long long>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(count_kernel, dim3((unsigned)blocks), dim3(256), 0,
                       c.stream, a, b, n, eps, counts);
    GBK_HIP(hipGetLastError());
}

}  // namespace gbkhip
