// smdisk evaluation kernels (core/gmodel/gmodels.hpp:645-958; host loop
// gbkfit_host/include/gbkfit/host/kernels.hpp:308-423).
//
// The reference runs one thread per spatial voxel and adds every channel of
// every voxel's line into the cube with an atomic.  Here:
//   * thin disks (no height traits, one voxel per column): one lane per
//     spaxel; the wave walks the union of its lanes' channel windows so the
//     read-modify-write of each channel plane is one coalesced 256-byte row.
//   * thick disks: a 256-thread block owns a row segment of 64 columns; its
//     four waves split the line of sight, add their lines into a
//     (channels x 64) spectrum tile in LDS, and the block then adds the tile
//     to the cube with coalesced rows.  No global atomics on the cube.
#include "common.hpp"
#include "disk.hpp"

#include <climits>

namespace gbkhip {

constexpr int kWave = 64;

__device__ __forceinline__ int wave_min(int v)
{
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, kWave));
    return v;
}

__device__ __forceinline__ int wave_max(int v)
{
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, kWave));
    return v;
}

// Per-voxel side outputs (gmodels.hpp:931-957).  idx is the spatial voxel.
__device__ __forceinline__ void side_outputs(const gbkhip_outputs& o,
                                             size_t idx, const VoxelLine& L)
{
    if (o.wdata) o.wdata[idx] += L.w;
    if (o.wdata_cmp) o.wdata_cmp[idx] = L.w;
    if (o.rdata) o.rdata[idx] += L.rvalue;
    if (o.rdata_cmp) o.rdata_cmp[idx] = L.rvalue;
    if (o.ordata) o.ordata[idx] += L.orvalue;
    if (o.ordata_cmp) o.ordata_cmp[idx] = L.orvalue;
    if (o.vdata_cmp) o.vdata_cmp[idx] = L.v;
    if (o.ddata_cmp) o.ddata_cmp[idx] = L.s;
}

// Thin disk: block (64, 4) covers 64 x 4 spaxels.
__global__ __launch_bounds__(256) void smdisk_thin_kernel(
    gbkhip_disk d, gbkhip_grid g, gbkhip_outputs o)
{
    const int lane = threadIdx.x;
    const int x = blockIdx.x * kWave + lane;
    const int y = blockIdx.y * 4 + threadIdx.y;
    const int nx = g.spat_size[0], ny = g.spat_size[1];
    if (y >= ny) return;     // whole wave leaves together
    const bool need_theta = disk_needs_theta(d);
    const NodeGeom ng = node_geom(d, 0);

    VoxelLine L;
    L.on = false;
    if (x < nx) L = smdisk_voxel(d, g, ng, x, y, 0, need_theta, o.wdata);
    const size_t col = (size_t)y * nx + x;
    int k0 = 1, k1 = 0;
    LineEval e;
    if (L.on) {
        if (o.image) o.image[col] += L.orvalue;
        if (o.scube) {
            line_window(g, L.v, L.s, k0, k1);
            e = line_eval(g, L.orvalue, L.v, L.s);
        }
        side_outputs(o, col, L);
    }
    if (!o.scube) return;
    const int kmin = wave_min(k0 <= k1 ? k0 : INT_MAX);
    const int kmax = wave_max(k0 <= k1 ? k1 : INT_MIN);
    const size_t plane = (size_t)nx * ny;
    for (int k = kmin; k <= kmax; ++k)
        if (k >= k0 && k <= k1) o.scube[k * plane + col] += e.at(k);
}

// Thick disk: block of 4 waves on the row segment [x0, x0 + 64) of row y;
// wave w evaluates z = w, w + 4, ...  Channels are processed in chunks of
// `chunk` (the LDS tile height) when the spectrum is taller than the tile.
__global__ __launch_bounds__(256) void smdisk_thick_kernel(
    gbkhip_disk d, gbkhip_grid g, gbkhip_outputs o, int chunk)
{
    extern __shared__ float tile[];        // [chunk][64]
    __shared__ int s_kmin, s_kmax;
    __shared__ float s_img[kWave];
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int x = blockIdx.x * kWave + lane;
    const int y = blockIdx.y;
    const int nx = g.spat_size[0], ny = g.spat_size[1], nz = g.spat_size[2];
    const size_t plane = (size_t)nx * ny;
    const bool need_theta = disk_needs_theta(d);
    const NodeGeom ng = node_geom(d, 0);
    const int nspec = g.spec_size;

    for (int c0 = 0; c0 < (o.scube ? nspec : 1); c0 += chunk) {
        const bool first = c0 == 0;
        const int c1 = min(c0 + chunk, nspec);   // exclusive
        for (int i = threadIdx.x; i < chunk * kWave; i += blockDim.x)
            tile[i] = 0.0f;
        if (threadIdx.x == 0) { s_kmin = INT_MAX; s_kmax = INT_MIN; }
        if (threadIdx.x < kWave) s_img[threadIdx.x] = 0.0f;
        __syncthreads();
        int kmin = INT_MAX, kmax = INT_MIN;
        if (x < nx) {
            for (int z = wave; z < nz; z += 4) {
                const VoxelLine L = smdisk_voxel(d, g, ng, x, y, z,
                                                 need_theta, o.wdata);
                if (!L.on) continue;
                if (first) {
                    if (o.image) atomicAdd(&s_img[lane], L.orvalue);
                    side_outputs(o, (size_t)z * plane + (size_t)y * nx + x, L);
                }
                if (!o.scube) continue;
                int k0, k1;
                line_window(g, L.v, L.s, k0, k1);
                k0 = max(k0, c0);
                k1 = min(k1, c1 - 1);
                if (k0 > k1) continue;
                const LineEval e = line_eval(g, L.orvalue, L.v, L.s);
                for (int k = k0; k <= k1; ++k)
                    atomicAdd(&tile[(k - c0) * kWave + lane], e.at(k));
                kmin = min(kmin, k0);
                kmax = max(kmax, k1);
            }
        }
        kmin = wave_min(kmin);
        kmax = wave_max(kmax);
        if (lane == 0 && kmin <= kmax) {
            atomicMin(&s_kmin, kmin);
            atomicMax(&s_kmax, kmax);
        }
        __syncthreads();
        if (first && o.image && threadIdx.x < kWave && x < nx &&
            s_img[lane] != 0.0f)
            o.image[(size_t)y * nx + x] += s_img[lane];
        if (o.scube && s_kmin <= s_kmax) {
            const int rows = s_kmax - s_kmin + 1;
            for (int i = threadIdx.x; i < rows * kWave; i += blockDim.x) {
                const int k = s_kmin + i / kWave;
                const int l = i % kWave;
                const int xx = blockIdx.x * kWave + l;
                if (xx < nx)
                    o.scube[k * plane + (size_t)y * nx + xx] +=
                        tile[(k - c0) * kWave + l];
            }
        }
        __syncthreads();
    }
}

void launch_smdisk(Context& c, const gbkhip_disk& d, const gbkhip_grid& g,
                   const gbkhip_outputs& o)
{
    const int nx = g.spat_size[0], ny = g.spat_size[1];
    if (nx <= 0 || ny <= 0 || g.spat_size[2] <= 0) return;
    const bool thin = d.rh.uids == nullptr;
    if (thin) {
        dim3 block(kWave, 4);
        dim3 grid(ceil_div(nx, kWave), ceil_div(ny, 4));
        hipLaunchKernelGGL(smdisk_thin_kernel, grid, block, 0, c.stream, d, g,
                           o);
    } else {
        static std::once_flag once;
        std::call_once(once, [] {
            (void)hipFuncSetAttribute(
                (const void*)smdisk_thick_kernel,
                hipFuncAttributeMaxDynamicSharedMemorySize, 96 * 1024);
        });
        const int chunk = o.scube ? std::max(1, std::min(g.spec_size, 256))
                                  : 1;
        const size_t lds = (size_t)chunk * kWave * sizeof(float);
        dim3 grid(ceil_div(nx, kWave), ny);
        hipLaunchKernelGGL(smdisk_thick_kernel, grid, dim3(256), lds,
                           c.stream, d, g, o, chunk);
    }
    GBK_HIP(hipGetLastError());
}

}  // namespace gbkhip
