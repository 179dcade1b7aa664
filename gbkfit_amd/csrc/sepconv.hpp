// Separable circular convolution of a padded cube (sepconv.hip).
#pragma once

#include "common.hpp"

namespace gbkhip {

// A kernel cube that factors as kx(x) ky(y) kz(z) with short support: its
// half-widths, tap buckets, the x pass's LDS row length and the taps
// (x, then y, then z, each zero-padded to its bucket).
struct SepConv {
    bool ok = false;
    int e[3] = {0, 0, 0};
    int taps[3] = {0, 0, 0};
    int mp = 0;
    DevBuf w;
};

// Analyses the kernel cube (size x, y, z; device pointer) on the host;
// false (P.ok false) when the direct convolution does not apply.
bool sepconv_plan(Context& c, const int size[3], const float* kernel,
                  SepConv& P);
// data <- data (circularly convolved with) the planned kernel, in place.
void sepconv_run(Context& c, const int size[3], const SepConv& P,
                 float* data);

}  // namespace gbkhip
