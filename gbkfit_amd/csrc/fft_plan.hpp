// FFT plan shared by fft.hip (kernels) and api.hip (buffer sizing).
#pragma once

#include "common.hpp"

namespace gbkhip {

constexpr int kMaxStages = 24;

struct Axis {
    int n = 1;                    // transform length
    int nst = 0;                  // number of stages
    int radix[kMaxStages] = {};   // radices in stage order
};

struct FFTPlan {
    int nx, ny, nz, nxh;
    bool even;                    // nx even: half-length row transform
    Axis ax;                      // row transform (nx/2 or nx)
    Axis ay, az;
    DevBuf twx, twy, twz;         // twiddle tables of the three transforms
    DevBuf twpost;                // exp(-2 pi i k / nx), k = 0..nx/2
};

}  // namespace gbkhip
