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

// Stage table of the register-tiled engine (lengths whose prime factors are
// 2, 3 and 5).  ns[s] is the product of the radices before stage s; nsm[s]
// is the magic multiplier of a division by ns[s].
constexpr int kFMax = 16;

struct FAxis {
    int n = 1, nst = 0;
    int R[kFMax] = {};
    int ns[kFMax] = {};
    int nb[kFMax] = {};           // n / R[s], butterflies per line
    int tws[kFMax] = {};          // n / (ns[s] R[s]), twiddle index step
    unsigned nsm[kFMax] = {};
};

// Launch shape of one axis on the register-tiled engine: lshift = log2 of
// the lines per block, threads = block size (256 or 512), lds = bytes of
// the twiddle table plus the [n][L+1] tile.
struct FastShape {
    bool ok = false;
    FAxis a;
    int lshift = 0;
    int threads = 256;
    size_t lds = 0;
};

struct FFTPlan {
    int nx, ny, nz, nxh;
    bool even;                    // nx even: half-length row transform
    Axis ax;                      // row transform (nx/2 or nx)
    Axis ay, az;
    FastShape fx, fy, fz;         // register-tiled engine, when applicable
    DevBuf twx, twy, twz;         // twiddle tables of the three transforms
    DevBuf twsx, twsy, twsz;      // per-stage tables (compile-time plans)
    DevBuf twpost;                // exp(-2 pi i k / nx), k = 0..nx/2
};

}  // namespace gbkhip
