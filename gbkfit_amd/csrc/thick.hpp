// Thick-disk spectral deposit of one row segment of 64 columns, shared by the
// plain smdisk kernel and the batched one (core/gmodel/gmodels.hpp:645-958;
// the host driver loops over every voxel and adds each line with an atomic,
// gbkfit_host/include/gbkfit/host/kernels.hpp:308-423).
//
// A 256-thread block owns the columns [64 bx, 64 bx + 64) of row y.  The
// four waves split each line of sight (wave w takes z = w, w + 4, ...), in
// groups of 4 J voxels:
//   A. every lane evaluates its J voxels of the group once and keeps their
//      lines (amplitude, centre, 1 / 2 sigma^2, channel window) in
//      registers;
//   B. per chunk of kThickC channels, every wave adds its lines into its own
//      LDS partial spectrum, in which each lane owns its column (no
//      atomics: the lane adds its voxels in z order), and the block sums
//      the four partials in a fixed order (w0 + w1 + w2 + w3) into the cube.
// Every sum has a fixed order, so results are bit-identical from run to run
// and between the kernels that use it.
#pragma once

#include "disk.hpp"

#include <climits>

namespace gbkhip {

constexpr int kThickC = 32;     // channels per LDS chunk
constexpr int kThickWaves = 4;

// LDS the block needs: the four partial spectra of a chunk, the four
// partial images and the four waves' channel ranges.
struct ThickLds {
    float tile[kThickWaves][kThickC][64];
    float img[kThickWaves][64];
    int kr[2][kThickWaves];
};

__device__ __forceinline__ int thick_wave_min(int v)
{
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ int thick_wave_max(int v)
{
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

// Deposits the row segment into `cube` (spec_size x ny x nx, x fastest):
//   assign: the block's channels [0, nze) of columns x < nxe are written
//           (zero where no line lands: the cube needs no zero fill), else
//           the lines are added to the cube (+=, like the reference);
//   plain:  also the image and the per-voxel side outputs of `o`.
// Columns x >= nxe and channels k >= nze are neither evaluated nor written.
// The channel window is packed in 16 bits per bound: spec_size <= 65535.
template <int J, bool PLAIN>
__device__ __forceinline__ void thick_row_segment(
    const gbkhip_disk& d, const gbkhip_grid& g, const NodeGeom& ng,
    bool need_theta, int y, int nxe, int nze, float* __restrict__ cube,
    bool assign, const gbkhip_outputs& o, ThickLds& S)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int x = blockIdx.x * 64 + lane;
    const int nx = g.spat_size[0], ny = g.spat_size[1], nz = g.spat_size[2];
    const size_t plane = (size_t)nx * ny;
    const size_t row = (size_t)y * nx;
    float img = 0.0f;
    for (int zg = 0; zg < nz; zg += kThickWaves * J) {
        // A: this lane's voxels of the group, lines in registers
        float amp[J], vc[J], i2s[J];
        unsigned win[J];
        int kmin = INT_MAX, kmax = INT_MIN;
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const int z = zg + wave + kThickWaves * j;
            amp[j] = 0.0f;
            vc[j] = 0.0f;
            i2s[j] = 0.0f;
            win[j] = 1u;                           // empty: k0 = 1 > k1 = 0
            if (x >= nxe || z >= nz) continue;
            const VoxelLine L = smdisk_voxel(d, g, ng, x, y, z, need_theta,
                                             PLAIN && o.wdata);
            if (!L.on) continue;
            if (PLAIN) {
                img += L.orvalue;
                const size_t idx = (size_t)z * plane + row + x;
                if (o.wdata) o.wdata[idx] += L.w;
                if (o.wdata_cmp) o.wdata_cmp[idx] = L.w;
                if (o.rdata) o.rdata[idx] += L.rvalue;
                if (o.rdata_cmp) o.rdata_cmp[idx] = L.rvalue;
                if (o.ordata) o.ordata[idx] += L.orvalue;
                if (o.ordata_cmp) o.ordata_cmp[idx] = L.orvalue;
                if (o.vdata_cmp) o.vdata_cmp[idx] = L.v;
                if (o.ddata_cmp) o.ddata_cmp[idx] = L.s;
            }
            if (!cube) continue;
            int k0, k1;
            line_window(g, L.v, L.s, k0, k1);
            k1 = min(k1, nze - 1);
            if (k0 > k1) continue;
            const LineEval e = line_eval(g, L.orvalue, L.v, L.s);
            amp[j] = e.amp;
            vc[j] = e.v;
            i2s[j] = e.inv2s2;
            win[j] = (unsigned)k0 | ((unsigned)k1 << 16);
            kmin = min(kmin, k0);
            kmax = max(kmax, k1);
        }
        // the block's channel range (uniform over the block)
        kmin = thick_wave_min(kmin);
        kmax = thick_wave_max(kmax);
        if (lane == 0) {
            S.kr[0][wave] = kmin;
            S.kr[1][wave] = kmax;
        }
        __syncthreads();
        int bmin = INT_MAX, bmax = INT_MIN;
#pragma unroll
        for (int w = 0; w < kThickWaves; ++w) {
            bmin = min(bmin, S.kr[0][w]);
            bmax = max(bmax, S.kr[1][w]);
        }
        const bool first = assign && zg == 0;
        const int lo = first ? 0 : bmin;
        const int hi = first ? nze - 1 : bmax;
        if (cube) {
            for (int c0 = lo; c0 <= hi; c0 += kThickC) {
                const int cn = min(kThickC, hi - c0 + 1);
                const bool live = c0 <= bmax && c0 + cn - 1 >= bmin;
                if (!live && !first) continue;
                if (live) {
                    // B: this lane's column of its wave's partial spectrum
                    // (read-modify-write of LDS the lane owns, channels two
                    // at a time so two reads are in flight).
                    float* my = &S.tile[wave][0][lane];
                    for (int k = 0; k < cn; ++k) my[k * 64] = 0.0f;
#pragma unroll
                    for (int j = 0; j < J; ++j) {
                        const int k0 = max((int)(win[j] & 0xffffu), c0);
                        const int k1 = min((int)(win[j] >> 16), c0 + cn - 1);
                        // flux N(zero + k step; v, s): exp(-t) as one
                        // v_exp_f32 of -t log2(e)
                        const float zmv = g.spec_zero - vc[j];
                        const float cc = -i2s[j] * 1.44269504088896341f;
                        const float a = amp[j];
                        float* dst = my + (k0 - c0) * 64;
                        float kf = (float)k0;
                        int k = k0;
                        for (; k < k1; k += 2, kf += 2.0f, dst += 128) {
                            const float m0 = dst[0], m1 = dst[64];
                            const float d0 = fmaf(kf, g.spec_step, zmv);
                            const float d1 = fmaf(kf + 1.0f, g.spec_step,
                                                  zmv);
                            dst[0] = fmaf(a, __builtin_amdgcn_exp2f(
                                                 d0 * d0 * cc), m0);
                            dst[64] = fmaf(a, __builtin_amdgcn_exp2f(
                                                  d1 * d1 * cc), m1);
                        }
                        if (k == k1) {
                            const float d0 = fmaf(kf, g.spec_step, zmv);
                            dst[0] = fmaf(a, __builtin_amdgcn_exp2f(
                                                 d0 * d0 * cc), dst[0]);
                        }
                    }
                    __syncthreads();
                }
                for (int i = threadIdx.x; i < cn * 64; i += blockDim.x) {
                    const int k = i >> 6, l = i & 63;
                    const int xx = blockIdx.x * 64 + l;
                    if (xx >= nxe) continue;
                    float v = 0.0f;
                    if (live)
                        v = ((S.tile[0][k][l] + S.tile[1][k][l]) +
                             S.tile[2][k][l]) + S.tile[3][k][l];
                    float* dst = cube + (size_t)(c0 + k) * plane + row + xx;
                    if (first)
                        *dst = v;
                    else
                        *dst += v;
                }
                if (live) __syncthreads();    // the tile is free again
            }
        }
        __syncthreads();                      // kr is free for the next group
    }
    if (PLAIN && o.image) {
        S.img[wave][lane] = img;
        __syncthreads();
        if (wave == 0 && x < nxe) {
            const float t = ((S.img[0][lane] + S.img[1][lane]) +
                             S.img[2][lane]) + S.img[3][lane];
            if (t != 0.0f) o.image[row + x] += t;
        }
    }
}

}  // namespace gbkhip
