// Register-blocked decimating FIR filter along rows held in LDS, shared by
// the batched direct pipeline (x pass) and the plain path's separable
// convolution.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace gbkhip {

constexpr int kSepThreads = 256;

// outputs per thread of S1 and S2 (fewer for wide taps: registers)
__host__ __device__ constexpr int sep_jx(int taps) { return taps <= 64 ? 8 : 4; }
__host__ __device__ constexpr int sep_jy(int taps) { return taps <= 64 ? 16 : 4; }
// smallest d with stride d = 16 (mod 64) floats (stride a multiple of 4)
__host__ __device__ constexpr int sep_dxb(int stride)
{
    int d = 1;
    while ((stride * d) % 64 != 16) ++d;
    return d;
}

// S1 filter: the block's LDS tile M [zc][mp] holds the hi-res values of
// one row for channels [z0, z0 + zn); a thread filters J consecutive X
// outputs of one channel into out (the row's P1 line of channel 0).
// Outputs (z, X .. X + J - 1) read inputs sx X + [0, sx (J - 1) + TX).
// Lane l of a wave takes channel l % 16; its four 16-lane groups take
// output blocks DXB apart, whose inputs start 16 banks apart (SX J DXB = 16
// mod 64), and the 16 channels of a group, mp (odd) apart, hit 16 distinct
// banks: no bank conflicts with 32 or 64 banks.  The taps w are wave-uniform
// (global memory: scalar loads into SGPRs; or an LDS copy).
template <int SX, int TX>
__device__ __forceinline__ void sep_xfilter(const float* M, const float* taps,
                                            int mp, int zn, int z0, int nxl,
                                            int ny1, float* __restrict__ out)
{
    constexpr int T = kSepThreads, J = sep_jx(TX);
    constexpr int NIN = SX * (J - 1) + TX;
    constexpr int DXB = sep_dxb(SX * J);
    const int nxb = (nxl + J - 1) / J;
    const int lane = threadIdx.x & 63, zz = lane & 15, grp = lane >> 4;
    for (int w = threadIdx.x >> 6;; w += T / 64) {
        const int g = w / DXB, xb = g * 4 * DXB + w % DXB + grp * DXB;
        if (g * 4 * DXB >= nxb) break;           // uniform over the wave
        if (zz >= zn || xb >= nxb) continue;
        const int X0 = xb * J;
        const float* in = M + zz * mp + SX * X0;
        float v[NIN];
#pragma unroll
        for (int i = 0; i < NIN; ++i) v[i] = in[i];
        float o[J] = {};
#pragma unroll
        for (int u = 0; u < TX; ++u) {
            const float w = taps[u];
#pragma unroll
            for (int j = 0; j < J; ++j) o[j] = fmaf(w, v[SX * j + u], o[j]);
        }
        float* dst = out + (size_t)(z0 + zz) * ny1 * nxl + X0;
        if (X0 + J <= nxl &&
            (reinterpret_cast<uintptr_t>(dst) & 15) == 0) {
#pragma unroll
            for (int j = 0; j < J; j += 4)
                *reinterpret_cast<float4*>(dst + j) =
                    make_float4(o[j], o[j + 1], o[j + 2], o[j + 3]);
        } else {
#pragma unroll
            for (int j = 0; j < J; ++j)
                if (X0 + j < nxl) dst[j] = o[j];
        }
    }
}

}  // namespace gbkhip
