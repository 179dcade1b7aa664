// Separable circular convolution of a padded cube (the plain path's
// fft_convolve_cached, host/include/gbkfit/host/fft.hpp:98-125, when the
// kernel factors).  DCube builds the PSF x LSF cube as psf(y, x) lsf(z)
// (_dcube.py:191-209), and a Gaussian PSF cut at 8 sigma is px(x) py(y), so
// the circular convolution c(q) = sum_t k(t) m(q - t) is three 1-D circular
// convolutions of 2e + 1 taps each.  Three in-place passes over the cube
// (rows, then y and z columns) move about 6 cube sizes of bytes instead of
// the FFTs' passes.  The kernel is analysed once per cached kernel; when it
// does not factor (to 1e-6 of its maximum), or its support is too wide, the
// FFT convolution runs instead.
#include "common.hpp"
#include "sepfilter.hpp"
#include "sepconv.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace gbkhip {

namespace {

constexpr int kCvRows = 16;          // rows per x-pass block (lanes' rows)
constexpr int kCvW = 32;             // columns per y / z block
constexpr int kCvPitch = kCvW + 1;
constexpr int kCvLds = 64 * 1024;

int cv_bucket(int n)
{
    for (int b : {8, 16, 24, 32, 44, 48, 64, 96})
        if (n <= b) return b;
    return 0;
}

// i mod n for i in [-n, ...): the staged halo wraps at most a few times
__device__ __forceinline__ int wrap(int i, int n)
{
    if (i < 0) i += n;
    while (i >= n) i -= n;
    return i;
}

// Rows: block = kCvRows consecutive rows of length nx.  The LDS tile holds
// s[i] = m((i - e) mod nx) per row, so output q reads s[q .. q + 2e]
// (taps W[u] = k(e - u)); the shared row filter writes the row in place.
template <int TAPS>
__global__ __launch_bounds__(256) void cv_rows(float* __restrict__ data,
                                               int nx, long long nrows,
                                               int mp, int e,
                                               const float* __restrict__ w)
{
    extern __shared__ float M[];          // [kCvRows][mp]
    __shared__ float s_w[TAPS];
    constexpr int T = kSepThreads;
    for (int i = threadIdx.x; i < TAPS; i += T) s_w[i] = w[i];
    const long long r0 = (long long)blockIdx.x * kCvRows;
    const int nr = (int)std::min<long long>(kCvRows, nrows - r0);
    float* rows = data + r0 * nx;
    for (int i = threadIdx.x; i < nr * mp; i += T) {
        const int r = i / mp, c = i - r * mp;
        M[i] = rows[(size_t)r * nx + wrap(c - e, nx)];
    }
    __syncthreads();
    sep_xfilter<1, TAPS>(M, s_w, mp, nr, 0, nx, 1, rows);
}

// Columns: block = kCvW consecutive elements of n rows rstride apart (y
// columns of one z plane, or z columns of one y row), staged with the
// circular halo as s[i][c] = m((i - e) mod n); a thread filters J
// consecutive outputs of one column and writes them in place.
template <int TAPS>
__global__ __launch_bounds__(256) void cv_cols(float* __restrict__ data,
                                               int n, long long rstride,
                                               int ncols, long long ostride,
                                               int e,
                                               const float* __restrict__ w)
{
    constexpr int T = kSepThreads, J = sep_jy(TAPS), G = T / kCvW;
    constexpr int NIN = J - 1 + TAPS;
    extern __shared__ float tile[];       // [rows][kCvPitch]
    __shared__ float s_w[TAPS];
    for (int i = threadIdx.x; i < TAPS; i += T) s_w[i] = w[i];
    const int strips = (ncols + kCvW - 1) / kCvW;
    const int X0 = (blockIdx.x % strips) * kCvW;
    const int o = blockIdx.x / strips;
    float* base = data + (long long)o * ostride + X0;
    const int nb = (n + J - 1) / J;
    const int rows = nb * J + TAPS - 1;
    const int cw = std::min(kCvW, ncols - X0);
    for (int i = threadIdx.x; i < rows * kCvW; i += T) {
        const int r = i / kCvW, c = i - r * kCvW;
        tile[r * kCvPitch + c] =
            c < cw ? base[wrap(r - e, n) * rstride + c] : 0.0f;
    }
    __syncthreads();
    const int c = threadIdx.x % kCvW;
    if (c >= cw) return;
    for (int b = threadIdx.x / kCvW; b < nb; b += G) {
        const float* in = tile + b * J * kCvPitch + c;
        float v[NIN];
#pragma unroll
        for (int i = 0; i < NIN; ++i) v[i] = in[i * kCvPitch];
        float acc[J] = {};
#pragma unroll
        for (int u = 0; u < TAPS; ++u) {
            const float wu = s_w[u];
#pragma unroll
            for (int j = 0; j < J; ++j) acc[j] = fmaf(wu, v[j + u], acc[j]);
        }
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const int y = b * J + j;
            if (y < n) base[y * rstride + c] = acc[j];
        }
    }
}

#define GBK_CV_DISPATCH(KERNEL, TAPS, ...)                                    \
    switch (TAPS) {                                                          \
    case 8: hipLaunchKernelGGL(KERNEL<8>, __VA_ARGS__); break;               \
    case 16: hipLaunchKernelGGL(KERNEL<16>, __VA_ARGS__); break;             \
    case 24: hipLaunchKernelGGL(KERNEL<24>, __VA_ARGS__); break;             \
    case 32: hipLaunchKernelGGL(KERNEL<32>, __VA_ARGS__); break;             \
    case 44: hipLaunchKernelGGL(KERNEL<44>, __VA_ARGS__); break;             \
    case 48: hipLaunchKernelGGL(KERNEL<48>, __VA_ARGS__); break;             \
    case 64: hipLaunchKernelGGL(KERNEL<64>, __VA_ARGS__); break;             \
    case 96: hipLaunchKernelGGL(KERNEL<96>, __VA_ARGS__); break;             \
    default:                                                                 \
        throw Error{GBKHIP_ERR_INTERNAL, "no convolution kernel for taps"};  \
    }

// Half-width e of a circular 1-D kernel a (length n): the smallest e with
// at most 1e-9 of the absolute weight at circular distance > e; -1 if none
// leaves room (2e + 1 > n).
int support(const std::vector<double>& a)
{
    const int n = (int)a.size();
    double tot = 0.0;
    for (double v : a) tot += std::fabs(v);
    if (!(tot > 0.0)) return -1;
    std::vector<double> by(n / 2 + 1, 0.0);   // weight at each distance
    for (int i = 0; i < n; ++i) by[std::min(i, n - i)] += std::fabs(a[i]);
    double out = tot;
    for (int e = 0; e <= n / 2; ++e) {
        out -= by[e];
        if (out <= 1e-9 * tot) return 2 * e + 1 <= n ? e : -1;
    }
    return -1;
}

// W[u] = k(e - u) / norm, u in [0, 2e], zero-padded to the bucket.
std::vector<float> taps_of(const std::vector<double>& a, int e, double norm,
                           int bucket)
{
    const int n = (int)a.size();
    std::vector<float> w(bucket, 0.0f);
    for (int u = 0; u <= 2 * e; ++u) {
        const int t = e - u;
        w[u] = (float)(a[((t % n) + n) % n] / norm);
    }
    return w;
}

}  // namespace

bool sepconv_plan(Context& c, const int size[3], const float* kernel,
                  SepConv& P)
{
    P.ok = false;
    const int nx = size[0], ny = size[1], nz = size[2];
    const size_t N = (size_t)nx * ny * nz;
    std::vector<float> k(N);
    GBK_HIP(hipMemcpyAsync(k.data(), kernel, N * sizeof(float),
                           hipMemcpyDeviceToHost, c.stream));
    GBK_HIP(hipStreamSynchronize(c.stream));
    std::vector<double> ax(nx, 0.0), ay(ny, 0.0), az(nz, 0.0);
    double tot = 0.0, kmax = 0.0;
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                const double v = k[((size_t)z * ny + y) * nx + x];
                ax[x] += v;
                ay[y] += v;
                az[z] += v;
                tot += v;
                kmax = std::max(kmax, std::fabs(v));
            }
    if (!(tot > 0.0) || !std::isfinite(tot)) return false;
    // k = ax(x) ay(y) az(z) / tot^2 when it factors
    const double t2 = tot * tot;
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                const double d = k[((size_t)z * ny + y) * nx + x] -
                                 ax[x] * ay[y] * az[z] / t2;
                if (std::fabs(d) > 1e-6 * kmax) return false;
            }
    const int e[3] = {support(ax), support(ay), support(az)};
    const int n[3] = {nx, ny, nz};
    for (int i = 0; i < 3; ++i) {
        if (e[i] < 0) return false;
        P.e[i] = e[i];
        P.taps[i] = cv_bucket(2 * e[i] + 1);
        if (!P.taps[i]) return false;
    }
    // LDS: x tile kCvRows x mp, column tiles (n + taps) x pitch
    const int jx = sep_jx(P.taps[0]);
    P.mp = (((nx + jx - 1) / jx) * jx + P.taps[0]) | 1;
    if ((size_t)kCvRows * P.mp * 4 > kCvLds) return false;
    for (int i = 1; i < 3; ++i) {
        const int j = sep_jy(P.taps[i]);
        const int rows = (n[i] + j - 1) / j * j + P.taps[i] - 1;
        if ((size_t)rows * kCvPitch * 4 > kCvLds) return false;
    }
    const std::vector<float> wx = taps_of(ax, e[0], tot, P.taps[0]);
    const std::vector<float> wy = taps_of(ay, e[1], tot, P.taps[1]);
    const std::vector<float> wz = taps_of(az, e[2], 1.0, P.taps[2]);
    P.w.alloc((size_t)(P.taps[0] + P.taps[1] + P.taps[2]) * sizeof(float));
    float* d = P.w.as<float>();
    GBK_HIP(hipMemcpy(d, wx.data(), wx.size() * 4, hipMemcpyHostToDevice));
    GBK_HIP(hipMemcpy(d + P.taps[0], wy.data(), wy.size() * 4,
                      hipMemcpyHostToDevice));
    GBK_HIP(hipMemcpy(d + P.taps[0] + P.taps[1], wz.data(), wz.size() * 4,
                      hipMemcpyHostToDevice));
    P.ok = true;
    return true;
}

void sepconv_run(Context& c, const int size[3], const SepConv& P,
                 float* data)
{
    const int nx = size[0], ny = size[1], nz = size[2];
    const float* w = P.w.as<float>();
    const long long nrows = (long long)ny * nz;
    {
        dim3 grid((unsigned)((nrows + kCvRows - 1) / kCvRows));
        const size_t lds = (size_t)kCvRows * P.mp * sizeof(float);
        GBK_CV_DISPATCH(cv_rows, P.taps[0], grid, dim3(kSepThreads), lds,
                        c.stream, data, nx, nrows, P.mp, P.e[0], w);
        GBK_HIP(hipGetLastError());
    }
    const int strips = (nx + kCvW - 1) / kCvW;
    {
        const int j = sep_jy(P.taps[1]);
        const int rows = (ny + j - 1) / j * j + P.taps[1] - 1;
        dim3 grid((unsigned)(strips * nz));
        GBK_CV_DISPATCH(cv_cols, P.taps[1], grid, dim3(kSepThreads),
                        (size_t)rows * kCvPitch * sizeof(float), c.stream,
                        data, ny, (long long)nx, nx, (long long)nx * ny,
                        P.e[1], w + P.taps[0]);
        GBK_HIP(hipGetLastError());
    }
    {
        const int j = sep_jy(P.taps[2]);
        const int rows = (nz + j - 1) / j * j + P.taps[2] - 1;
        dim3 grid((unsigned)(strips * ny));
        GBK_CV_DISPATCH(cv_cols, P.taps[2], grid, dim3(kSepThreads),
                        (size_t)rows * kCvPitch * sizeof(float), c.stream,
                        data, nz, (long long)nx * ny, nx, (long long)nx,
                        P.e[2], w + P.taps[0] + P.taps[1]);
        GBK_HIP(hipGetLastError());
    }
}

}  // namespace gbkhip
