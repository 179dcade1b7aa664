// mcdisk: Monte Carlo cloud disk (core/gmodel/gmodels.hpp:325-643; host
// loop gbkfit_host/include/gbkfit/host/kernels.hpp:183-306).  The per-cloud
// body (Philox stream per (seed, cloud id), samplers, projection, int64
// fixed-point channel deposits) is in mcdisk.hpp.  The reference seeds every
// OpenMP thread with 42 (kernels.hpp:239-242), so its output depends on the
// thread count; this stream depends only on (seed, cloud id).  The int64
// cube is converted and added to the FP32 cube afterwards: integer addition
// is associative, so reruns are bit-identical.
#include "mcdisk.hpp"

namespace gbkhip {

__global__ __launch_bounds__(256) void mcdisk_kernel(
    McParams mp, gbkhip_disk d, gbkhip_grid g, gbkhip_outputs o,
    unsigned long long* icube)
{
    extern __shared__ int s_csum[];
    __shared__ float s_nodes[kMcNodesLds];
    mp.csum = mc_stage_csum(mp.csum, mp.csum_len, s_csum);
    d.rnodes = mc_stage_nodes(d.rnodes, d.nrnodes, s_nodes);
    for (long long ci = blockIdx.x * (long long)blockDim.x + threadIdx.x;
         ci < mp.nclouds; ci += (long long)gridDim.x * blockDim.x)
        mc_cloud(ci, mp, d, g, o, icube);
}

// cube += icube / 2^S, and icube <- 0 for the next call.
__global__ void fixed_to_float(unsigned long long* icube, float* cube,
                               long long n, double inv_scale)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const long long q = (long long)icube[i];
        if (q) {
            cube[i] += (float)((double)q * inv_scale);
            icube[i] = 0ull;
        }
    }
}

__global__ void zero_u64(unsigned long long* p, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

// Fixed-point exponent: every deposit is |flux| * N(.; v, s) <= |flux| /
// (s_min sqrt(2 pi)) with s_min = 1e-3 channel widths, and at most nclouds
// deposits can land on one voxel, so nclouds * that * 2^S < 2^62.
double mcdisk_fixed_scale(float cflux, int nclouds, const gbkhip_grid& g)
{
    const double flux = std::fabs((double)cflux) /
        ((double)g.spat_step[0] * (double)g.spat_step[1]);
    const double smin = 1e-3 * std::fabs((double)g.spec_step);
    const double peak = flux / (smin * 2.5066282746310002);
    const double bound = std::max(1.0, (double)nclouds) * std::max(peak,
                                                                    1e-30);
    const int e = (int)std::floor(62.0 - std::log2(bound));
    return std::ldexp(1.0, std::min(e, 900));
}

void launch_mcdisk(Context& c, float cflux, int nclouds, const int* csum,
                   int csum_len, const unsigned char* hoi,
                   const gbkhip_disk& d, const gbkhip_grid& g,
                   const gbkhip_outputs& o, uint64_t seed)
{
    if (nclouds <= 0) return;
    GBK_CHECK(d.rp.uids && d.rh.uids, GBKHIP_ERR_ARG,
              "mcdisk needs density polar and height traits");
    const long long ncube = (long long)g.spat_size[0] * g.spat_size[1] *
                            g.spec_size;
    unsigned long long* icube = nullptr;
    if (o.scube && ncube > 0) {
        const size_t bytes = (size_t)ncube * sizeof(unsigned long long);
        if (c.int_scratch.bytes < bytes) {
            c.int_scratch.alloc(bytes);
            hipLaunchKernelGGL(zero_u64, dim3(2048), dim3(256), 0, c.stream,
                               c.int_scratch.as<unsigned long long>(), ncube);
            GBK_HIP(hipGetLastError());
        }
        icube = c.int_scratch.as<unsigned long long>();
    }
    McParams mp;
    mp.cflux = cflux;
    mp.nclouds = nclouds;
    mp.csum = csum;
    mp.csum_len = csum_len;
    mp.hoi = hoi;
    mp.key0 = (uint32_t)seed;
    mp.key1 = (uint32_t)(seed >> 32);
    mp.qscale = mcdisk_fixed_scale(cflux, nclouds, g);
    const int blocks = std::max(1, std::min(ceil_div(nclouds, 256), 8192));
    const size_t lds = csum_len <= kMcCsumLds ? csum_len * sizeof(int) : 0;
    hipLaunchKernelGGL(mcdisk_kernel, dim3(blocks), dim3(256), lds, c.stream,
                       mp, d, g, o, icube);
    GBK_HIP(hipGetLastError());
    if (icube) {
        hipLaunchKernelGGL(fixed_to_float, dim3(4096), dim3(256), 0, c.stream,
                           icube, o.scube, ncube, 1.0 / mp.qscale);
        GBK_HIP(hipGetLastError());
    }
}

}  // namespace gbkhip
