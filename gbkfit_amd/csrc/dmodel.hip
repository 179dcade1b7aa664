// Data-model stages: downscale, mask, moment maps (core/dmodel/dmodels.hpp),
// plus the direct convolution and spectral weight cube of the gmodel backend
// (core/gmodel/gmodels.hpp:9-82).
#include "common.hpp"
#include "moments.hpp"

namespace gbkhip {

// dmodels.hpp:7-46.  One lane per destination voxel, x fastest.
__global__ void downscale_kernel(int sx, int sy, int sz, int ox, int oy,
                                 int oz, int snx, int sny, int dnx, int dny,
                                 int dnz, const float* __restrict__ src,
                                 float* __restrict__ dst, long long bsrc,
                                 long long bdst)
{
    src += blockIdx.z * bsrc;
    dst += blockIdx.z * bdst;
    const long long n = (long long)dnx * dny * dnz;
    const float nf = 1.0f / (float)(sx * sy * sz);
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const int x = (int)(i % dnx);
        const int y = (int)((i / dnx) % dny);
        const int z = (int)(i / ((long long)dnx * dny));
        const int bx = ox + x * sx, by = oy + y * sy, bz = oz + z * sz;
        float sum = 0.0f;
        for (int dz = 0; dz < sz; ++dz)
            for (int dy = 0; dy < sy; ++dy)
                for (int dx = 0; dx < sx; ++dx)
                    sum += src[(bx + dx) + (size_t)(by + dy) * snx +
                               (size_t)(bz + dz) * snx * sny];
        dst[i] = sum * nf;
    }
}

// dmodels.hpp:48-74
__global__ void mask_kernel(float cutoff, int apply, long long n, float* d,
                            float* m)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        float mv = 1.0f;
        if (fabsf(d[i]) <= cutoff) {
            mv = 0.0f;
            if (apply) d[i] = __builtin_nanf("");
        }
        if (m) m[i] = mv;
    }
}

// dmodels.hpp:76-228.  One lane per spaxel; lanes consecutive in x so each
// channel plane is read as coalesced rows (moment sums: moments.hpp).
__global__ void moments_kernel(int nx, int ny, int nz, float step_z,
                               float zero_z, const float* __restrict__ d,
                               const float* __restrict__ w, float cutoff,
                               int norders, const int* __restrict__ orders,
                               float* __restrict__ md, float* __restrict__ mm,
                               float* __restrict__ mw, long long bcube,
                               long long bmap)
{
    const long long plane = (long long)nx * ny;
    const long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (c >= plane) return;
    d += blockIdx.y * bcube;
    if (w) w += blockIdx.y * bcube;
    md += blockIdx.y * bmap * norders;
    mm += blockIdx.y * bmap;
    if (mw) mw += blockIdx.y * bmap * norders;
    int ord[kMaxOrders];
    for (int i = 0; i < kMaxOrders; ++i) ord[i] = i < norders ? orders[i] : 0;
    float mo[kMaxOrders], m0;
    const bool valid = spaxel_moments(d + c, plane, nz, step_z, zero_z,
                                      cutoff, norders, ord, mo, m0);
    mm[c] = valid ? 1.0f : 0.0f;
    for (int i = 0; i < norders; ++i) md[c + i * plane] = mo[i];
    if (w) {
        float ws = 0.0f;
        for (int z = 0; valid && z < nz; ++z)
            ws += w[c + z * plane] * d[c + z * plane] * step_z / m0;
        mw[c] = valid ? ws : __builtin_nanf("");
    }
}

// gmodels.hpp:52-82: spectral weight = mean / max of the spatial column.
__global__ void wcube_kernel(int nx, int ny, int nzs, int nzspec,
                             const float* __restrict__ spat,
                             float* __restrict__ spec)
{
    const long long plane = (long long)nx * ny;
    const long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (c >= plane) return;
    float sum = 0.0f, mx = 0.0f;
    for (int z = 0; z < nzs; ++z) {
        const float v = spat[c + z * plane];
        sum += v;
        mx = fmaxf(mx, v);
    }
    const float mean = sum / nzs;
    const float val = mx > 0 ? mean / mx : 0.0f;
    for (int z = 0; z < nzspec; ++z) spec[c + z * plane] = val;
}

// gmodels.hpp:9-50: zero-padded direct convolution.
__global__ void convolve_kernel(int n1x, int n1y, int n1z, int n2x, int n2y,
                                int n2z, const float* __restrict__ a,
                                const float* __restrict__ k,
                                float* __restrict__ out)
{
    const long long n = (long long)n1x * n1y * n1z;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const int x = (int)(i % n1x);
        const int y = (int)((i / n1x) % n1y);
        const int z = (int)(i / ((long long)n1x * n1y));
        float s = 0.0f;
        for (int z2 = 0; z2 < n2z; ++z2) {
            const int z1 = z - n2z / 2 + z2;
            if (z1 < 0 || z1 >= n1z) continue;
            for (int y2 = 0; y2 < n2y; ++y2) {
                const int y1 = y - n2y / 2 + y2;
                if (y1 < 0 || y1 >= n1y) continue;
                for (int x2 = 0; x2 < n2x; ++x2) {
                    const int x1 = x - n2x / 2 + x2;
                    if (x1 < 0 || x1 >= n1x) continue;
                    s += a[x1 + (size_t)y1 * n1x + (size_t)z1 * n1x * n1y] *
                         k[x2 + (size_t)y2 * n2x + (size_t)z2 * n2x * n2y];
                }
            }
        }
        out[i] = s;
    }
}

static dim3 grid1d(long long n, int block = 256)
{
    long long b = (n + block - 1) / block;
    if (b > 65536LL * 8) b = 65536LL * 8;
    if (b < 1) b = 1;
    return dim3((unsigned)b);
}

void launch_downscale(Context& c, const int* scale, const int* offset,
                      const int* src_size, const int* dst_size,
                      const float* src, float* dst, int batch, long long bsrc,
                      long long bdst)
{
    const long long n = (long long)dst_size[0] * dst_size[1] * dst_size[2];
    if (n <= 0) return;
    dim3 g = grid1d(n);
    g.z = batch;
    hipLaunchKernelGGL(downscale_kernel, g, dim3(256), 0, c.stream, scale[0],
                       scale[1], scale[2], offset[0], offset[1], offset[2],
                       src_size[0], src_size[1], dst_size[0], dst_size[1],
                       dst_size[2], src, dst, bsrc, bdst);
    GBK_HIP(hipGetLastError());
}

void launch_mask(Context& c, float cutoff, int apply, long long n, float* d,
                 float* m)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(mask_kernel, grid1d(n), dim3(256), 0, c.stream, cutoff,
                       apply, n, d, m);
    GBK_HIP(hipGetLastError());
}

void launch_moments(Context& c, const int* size, const float* step,
                    const float* zero, const float* d, const float* w,
                    float cutoff, int norders, const int* orders, float* md,
                    float* mm, float* mw, int batch, long long bcube,
                    long long bmap)
{
    const long long plane = (long long)size[0] * size[1];
    if (plane <= 0) return;
    dim3 g(ceil_div(plane, 256), batch);
    hipLaunchKernelGGL(moments_kernel, g, dim3(256), 0, c.stream, size[0],
                       size[1], size[2], step[2], zero[2], d, w, cutoff,
                       norders, orders, md, mm, mw, bcube, bmap);
    GBK_HIP(hipGetLastError());
}

void launch_wcube(Context& c, const int* spat_size, int spec_size,
                  const float* spat, float* spec)
{
    const long long plane = (long long)spat_size[0] * spat_size[1];
    if (plane <= 0) return;
    hipLaunchKernelGGL(wcube_kernel, dim3(ceil_div(plane, 256)), dim3(256), 0,
                       c.stream, spat_size[0], spat_size[1], spat_size[2],
                       spec_size, spat, spec);
    GBK_HIP(hipGetLastError());
}

void launch_convolve(Context& c, const int* s1, const int* s2,
                     const float* a, const float* k, float* out)
{
    const long long n = (long long)s1[0] * s1[1] * s1[2];
    if (n <= 0) return;
    hipLaunchKernelGGL(convolve_kernel, grid1d(n), dim3(256), 0, c.stream,
                       s1[0], s1[1], s1[2], s2[0], s2[1], s2[2], a, k, out);
    GBK_HIP(hipGetLastError());
}

}  // namespace gbkhip
