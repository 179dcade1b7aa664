// smdisk evaluation kernels (core/gmodel/gmodels.hpp:645-958; host loop
// gbkfit_host/include/gbkfit/host/kernels.hpp:308-423).
//
// The reference runs one thread per spatial voxel and adds every channel of
// every voxel's line into the cube with an atomic.  Here:
//   * thin disks (no height traits, one voxel per column): one lane per
//     spaxel; the wave walks the union of its lanes' channel windows so the
//     read-modify-write of each channel plane is one coalesced 256-byte row.
//   * thick disks: a 256-thread block owns a row segment of 64 columns; its
//     four waves split the line of sight, add their lines into per-wave
//     LDS partial spectra (each lane owns its column) and the block adds
//     their fixed-order sum to the cube with coalesced rows (thick.hpp).
//     No atomics: results are bit-identical from run to run.
#include "common.hpp"
#include "disk.hpp"
#include "thick.hpp"

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

// Thick disk: block of 4 waves on the row segment [x0, x0 + 64) of row y,
// the line of sight split over the waves, every sum in a fixed order
// (thick.hpp); lines added to the cube (+=).
template <int J>
__global__ __launch_bounds__(256) void smdisk_thick_kernel(
    gbkhip_disk d, gbkhip_grid g, gbkhip_outputs o)
{
    __shared__ ThickLds S;
    const bool need_theta = disk_needs_theta(d);
    const NodeGeom ng = node_geom(d, 0);
    thick_row_segment<J, true>(d, g, ng, need_theta, blockIdx.y,
                               g.spat_size[0], g.spec_size, o.scube, false,
                               o, S);
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
        GBK_CHECK(g.spec_size <= 65535, GBKHIP_ERR_UNSUPPORTED,
                  "thick smdisk: at most 65535 channels");
        dim3 grid(ceil_div(nx, kWave), ny);
        const int nz = g.spat_size[2];
        if (nz <= 16)
            hipLaunchKernelGGL(smdisk_thick_kernel<4>, grid, dim3(256), 0,
                               c.stream, d, g, o);
        else if (nz <= 32)
            hipLaunchKernelGGL(smdisk_thick_kernel<8>, grid, dim3(256), 0,
                               c.stream, d, g, o);
        else
            hipLaunchKernelGGL(smdisk_thick_kernel<16>, grid, dim3(256), 0,
                               c.stream, d, g, o);
    }
    GBK_HIP(hipGetLastError());
}

}  // namespace gbkhip
