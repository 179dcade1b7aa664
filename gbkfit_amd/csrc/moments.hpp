// Moments of one spaxel's spectrum (core/dmodel/dmodels.hpp:76-228), shared
// by the plain moments kernel and the batched moment-map objective so both
// give the same bits.  Sums run over z in the reference's order, in FP32.
#pragma once

#include "common.hpp"

namespace gbkhip {

constexpr int kMaxOrders = 8;

// d points at the spaxel's channel 0; channel z is d[z * plane].  Writes
// the requested moments (orders sorted, each 0-7) to md[0, norders) and
// returns whether the spaxel is valid (|sum I dv| > cutoff); invalid
// spaxels get NaN.  m0 is also returned for the weight map.
__device__ __forceinline__ bool spaxel_moments(
    const float* __restrict__ d, long long plane, int nz, float step_z,
    float zero_z, float cutoff, int norders, const int* ord, float* md,
    float& m0_out)
{
    const float nan = __builtin_nanf("");
    const int max_order = ord[norders - 1];
    int m = 0;
    float m0s = 0.0f;
    for (int z = 0; z < nz; ++z) m0s += d[z * plane] * step_z;
    const bool valid = fabsf(m0s) > cutoff;
    const float m0 = valid ? m0s : nan;
    m0_out = m0;
    if (ord[m] == 0) md[m++] = m0;
    if (max_order == 0) return valid;
    float m1s = 0.0f;
    for (int z = 0; valid && z < nz; ++z)
        m1s += d[z * plane] * (zero_z + z * step_z) * step_z;
    const float m1 = valid ? m1s / m0 : nan;
    if (m < norders && ord[m] == 1) md[m++] = m1;
    if (max_order == 1) return valid;
    float m2s = 0.0f;
    for (int z = 0; valid && z < nz; ++z) {
        const float dv = zero_z + z * step_z - m1;
        m2s += d[z * plane] * dv * dv * step_z;
    }
    const float m2 = valid ? sqrtf(m2s / m0) : nan;
    if (m < norders && ord[m] == 2) md[m++] = m2;
    for (; m < norders; ++m) {
        float s = 0.0f;
        for (int z = 0; valid && z < nz; ++z) {
            const float dv = zero_z + z * step_z - m1;
            float p = 1.0f;
            for (int e = 0; e < ord[m]; ++e) p *= dv;
            s += d[z * plane] * p;
        }
        md[m] = valid ? s / m0 : nan;
    }
    return valid;
}

}  // namespace gbkhip
