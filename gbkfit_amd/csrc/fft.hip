// Hand-written 3-D real FFT convolution for gfx950.
//
// Replaces the FFTW / cuFFT plans of the reference FFT backend
// (host/include/gbkfit/host/fft.hpp:98-207, cuda/include/gbkfit/cuda/
// fft.hpp:97-184).  Layout and scaling follow FFTW's r2c/c2r convention:
// real (nz, ny, nx) <-> complex (nz, ny, nx/2+1), both unnormalised, and the
// convolution multiplies by 1/N (host kernels.hpp:32-53).
//
// Structure (all passes stage whole lines in LDS):
//   * row pass (x): r2c via a complex FFT of half length on (x[2n], x[2n+1])
//     plus the split post-processing; c2r is the mirror image.  Odd lengths
//     use a full-length complex FFT.
//   * line pass (y, z): tiles of `tl` consecutive kx columns, strided lines.
//     The last nontrivial axis of a convolution runs forward FFT, multiply by
//     the cached kernel spectrum times 1/N, and inverse FFT in one kernel, so
//     the spectrum of that axis never goes back to HBM between them.
//   * 1-D transforms are mixed-radix Stockham (autosort) stages ping-ponging
//     between two LDS buffers: radix 8, 4, 2, 3, 5 butterflies and a generic
//     direct DFT for other primes (the reference's padding rounds to
//     multiples of 32, which can leave factors like 17).
//   * twiddles exp(-2 pi i m / n) are generated in FP64 on the host and
//     rounded to FP32 once per plan; the pass stages them in LDS.
#include "fft_plan.hpp"
#include "fft_engine.hpp"

#include <cmath>
#include <vector>

namespace gbkhip {

static Axis make_axis(int n)
{
    Axis a;
    a.n = n;
    int m = n;
    auto take = [&](int r) {
        while (m % r == 0 && m > 1) {
            GBK_CHECK(a.nst < kMaxStages, GBKHIP_ERR_UNSUPPORTED,
                      "FFT length has too many factors");
            a.radix[a.nst++] = r;
            m /= r;
        }
    };
    take(8);
    take(4);
    take(2);
    take(3);
    take(5);
    for (int p = 7; m > 1 && p * p <= m; p += 2) take(p);
    if (m > 1) take(m);
    return a;
}

// Per-stage twiddle table of the compile-time plan of length n:
// [stage][k][t - 1] = exp(-2 pi i t k / (NS R)), FP64 rounded once to FP32.
void upload_stage_twiddles(DevBuf& buf, int n)
{
    const CPlan P = cplan(n);
    std::vector<float> h(2 * (size_t)std::max(P.tw_total, 1), 0.0f);
    for (int s = 0; s < P.nst; ++s)
        for (int k = 0; k < P.NS[s]; ++k)
            for (int t = 1; t < P.R[s]; ++t) {
                const double ang = -2.0 * M_PI * (double)t * k /
                                   ((double)P.NS[s] * P.R[s]);
                const size_t i = P.tw_off[s] + (size_t)k * (P.R[s] - 1) +
                                 (t - 1);
                h[2 * i] = (float)std::cos(ang);
                h[2 * i + 1] = (float)std::sin(ang);
            }
    buf.alloc(h.size() * sizeof(float));
    GBK_HIP(hipMemcpy(buf.ptr, h.data(), h.size() * sizeof(float),
                      hipMemcpyHostToDevice));
}

void upload_twiddles(DevBuf& buf, int n, int count)
{
    std::vector<float> h(2 * (size_t)count);
    for (int m = 0; m < count; ++m) {
        // FP64 generation, one rounding to FP32
        const double ang = -2.0 * M_PI * (double)m / (double)n;
        h[2 * m] = (float)std::cos(ang);
        h[2 * m + 1] = (float)std::sin(ang);
    }
    buf.alloc(h.size() * sizeof(float));
    GBK_HIP(hipMemcpy(buf.ptr, h.data(), h.size() * sizeof(float),
                      hipMemcpyHostToDevice));
}

// Radices of the register-tiled engine: the power of two is split evenly
// into stages of radix <= 16, then one stage per factor 3, 5, 7, 11 and 13
// (the reference's padding to multiples of 32 leaves such factors, e.g.
// 224 = 32 x 7); larger primes take the generic kernels.
static bool fast_radices(int n, std::vector<int>& R)
{
    int m = n, a2 = 0;
    while (m % 2 == 0) { m /= 2; ++a2; }
    int odd[5] = {0, 0, 0, 0, 0};
    const int primes[5] = {3, 5, 7, 11, 13};
    for (int i = 0; i < 5; ++i)
        while (m % primes[i] == 0) { m /= primes[i]; ++odd[i]; }
    if (m != 1) return false;
    const int s2 = (a2 + 3) / 4;
    for (int s = 0; s < s2; ++s) {
        const int bits = a2 / s2 + (s < a2 % s2 ? 1 : 0);
        R.push_back(1 << bits);
    }
    for (int i = 0; i < 5; ++i)
        for (int j = 0; j < odd[i]; ++j) R.push_back(primes[i]);
    return (int)R.size() <= kFMax;
}

// Magic multiplier m with floor(x / d) = umulhi(x, m) whenever x * d < 2^32
// (m = ceil(2^32 / d); d = 1 is flagged by m = 0).
unsigned div_magic(unsigned d)
{
    return d <= 1 ? 0u
                  : (unsigned)((((unsigned long long)1 << 32) + d - 1) / d);
}

constexpr size_t kFastLdsTarget = 48 * 1024;

// Engine shape for n-point lines, 2^lshift lines per block: 256 threads when
// the tile holds at most 16 elements per thread, else 512.
FastShape fast_shape_lines(int n, int lshift)
{
    FastShape f;
    std::vector<int> R;
    if (n < 2 || lshift < 0 || lshift > 6 || !fast_radices(n, R)) return f;
    f.a.n = n;
    f.a.nst = (int)R.size();
    int ns = 1;
    for (int s = 0; s < f.a.nst; ++s) {
        f.a.R[s] = R[s];
        f.a.ns[s] = ns;
        f.a.nb[s] = n / R[s];
        f.a.tws[s] = n / (ns * R[s]);
        f.a.nsm[s] = div_magic((unsigned)ns);
        ns *= R[s];
    }
    const long long tile = (long long)n << lshift;
    const size_t lds = ((size_t)std::max(n, cplan(n).tw_total) +
                        (size_t)n * ((1 << lshift) + 1)) * sizeof(float2);
    if (lds > 96 * 1024) return f;
    if (tile <= (long long)kFastE * 256) f.threads = 256;
    else if (tile <= (long long)kFastE * 512 && lshift >= 2) f.threads = 512;
    else return f;
    f.ok = true;
    f.lshift = lshift;
    f.lds = lds;
    return f;
}

// Strided lines: the most lines per block (the block reads 8 L contiguous
// bytes per element) whose tile fits the LDS target.
FastShape fast_shape(int n)
{
    for (int ls = 5; ls >= 2; --ls) {
        const size_t bytes = (size_t)n * ((1 << ls) + 2) * sizeof(float2);
        if (bytes <= kFastLdsTarget || (ls == 2 && bytes <= 96 * 1024)) {
            FastShape f = fast_shape_lines(n, ls);
            if (f.ok) return f;
        }
    }
    return FastShape();
}

// Contiguous rows: the rows per block do not affect coalescing, so take the
// most rows that keep 256 threads at 16 elements each.
FastShape fast_shape_rows(int n)
{
    for (int ls = 5; ls >= 0; --ls) {
        const size_t bytes = (size_t)n * ((1 << ls) + 2) * sizeof(float2);
        if (((long long)n << ls) <= kFastE * 256 && bytes <= kFastLdsTarget) {
            FastShape f = fast_shape_lines(n, ls);
            if (f.ok) return f;
        }
    }
    return fast_shape(n);
}

std::shared_ptr<FFTPlan> get_plan(Context& c, int nx, int ny, int nz)
{
    GBK_CHECK(nx > 0 && ny > 0 && nz > 0, GBKHIP_ERR_ARG,
              "FFT size must be positive");
    const auto key = std::make_tuple(nx, ny, nz);
    auto it = c.plans.find(key);
    if (it != c.plans.end()) return it->second;
    auto p = std::make_shared<FFTPlan>();
    p->nx = nx;
    p->ny = ny;
    p->nz = nz;
    p->nxh = nx / 2 + 1;
    p->even = nx % 2 == 0;
    const int rn = p->even ? nx / 2 : nx;
    GBK_CHECK(rn <= 4096 && ny <= 4096 && nz <= 4096, GBKHIP_ERR_UNSUPPORTED,
              "FFT axis longer than 4096 (x) / 4096 (y, z) is not supported");
    p->ax = make_axis(rn);
    p->ay = make_axis(ny);
    p->az = make_axis(nz);
    upload_twiddles(p->twx, rn, rn);
    upload_twiddles(p->twy, ny, ny);
    upload_twiddles(p->twz, nz, nz);
    upload_twiddles(p->twpost, nx, nx / 2 + 1);
    if (p->even) p->fx = fast_shape_rows(rn);
    p->fy = fast_shape(ny);
    p->fz = fast_shape(nz);
    if (p->fx.ok) upload_stage_twiddles(p->twsx, rn);
    if (p->fy.ok) upload_stage_twiddles(p->twsy, ny);
    if (p->fz.ok) upload_stage_twiddles(p->twsz, nz);
    c.plans[key] = p;
    return p;
}

// ---------------------------------------------------------------------------
// Device side
// ---------------------------------------------------------------------------
struct AxisArg {
    int n, nst;
    int radix[kMaxStages];
};

static AxisArg axis_arg(const Axis& a)
{
    AxisArg r;
    r.n = a.n;
    r.nst = a.nst;
    for (int i = 0; i < kMaxStages; ++i) r.radix[i] = a.radix[i];
    return r;
}

// One Stockham stage of radix r over `lines` lines of length n stored with
// leading dimension ld in `src`, results to `dst`.  ns = product of the
// radices of the previous stages.
__device__ void stockham_stage(const float2* src, float2* dst, int lines,
                               int ld, int n, int r, int ns,
                               const float2* tw, float sign)
{
    const int nb = n / r;                 // butterflies per line
    const int span = ns * r;
    const int twstep = n / span;          // twiddle index scale
    if (r == 2 || r == 3 || r == 4 || r == 5 || r == 8) {
        for (int it = threadIdx.x; it < lines * nb; it += blockDim.x) {
            const int l = it / nb;
            const int j = it - l * nb;
            const int k = j % ns;
            const float2* in = src + (size_t)l * ld;
            float2* out = dst + (size_t)l * ld;
            float2 v[8];
            for (int t = 0; t < r; ++t) {
                v[t] = in[j + t * nb];
                if (t && k) v[t] = cmul(v[t], twid(tw, (t * k * twstep) % n,
                                                   sign));
            }
            if (r == 2) {
                const float2 a = v[0], b = v[1];
                v[0] = cadd(a, b);
                v[1] = csub(a, b);
            } else if (r == 3) {
                const float2 s = cadd(v[1], v[2]);
                const float2 d = mul_si(csub(v[1], v[2]),
                                        sign * 0.866025403784438647f);
                const float2 t = make_float2(v[0].x - 0.5f * s.x,
                                             v[0].y - 0.5f * s.y);
                v[0] = cadd(v[0], s);
                v[1] = cadd(t, d);
                v[2] = csub(t, d);
            } else if (r == 4) {
                dft4(v[0], v[1], v[2], v[3], sign);
            } else if (r == 5) {
                const float c1 = 0.309016994374947424f;
                const float c2 = -0.809016994374947424f;
                const float s1 = 0.951056516295153572f * sign;
                const float s2 = 0.587785252292473129f * sign;
                const float2 b1 = cadd(v[1], v[4]), b2 = cadd(v[2], v[3]);
                const float2 d1 = csub(v[1], v[4]), d2 = csub(v[2], v[3]);
                const float2 t1 = make_float2(v[0].x + c1 * b1.x + c2 * b2.x,
                                              v[0].y + c1 * b1.y + c2 * b2.y);
                const float2 t2 = make_float2(v[0].x + c2 * b1.x + c1 * b2.x,
                                              v[0].y + c2 * b1.y + c1 * b2.y);
                const float2 u1 = make_float2(s1 * d1.x + s2 * d2.x,
                                              s1 * d1.y + s2 * d2.y);
                const float2 u2 = make_float2(s2 * d1.x - s1 * d2.x,
                                              s2 * d1.y - s1 * d2.y);
                v[0] = cadd(v[0], cadd(b1, b2));
                v[1] = cadd(t1, mul_si(u1, 1.0f));
                v[4] = csub(t1, mul_si(u1, 1.0f));
                v[2] = cadd(t2, mul_si(u2, 1.0f));
                v[3] = csub(t2, mul_si(u2, 1.0f));
            } else {  // r == 8: two radix-4 halves and the W8 twiddles
                float2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
                float2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
                dft4(e0, e1, e2, e3, sign);
                dft4(o0, o1, o2, o3, sign);
                const float h = 0.707106781186547524f;
                // W8^1 = h (1 + sign i), W8^2 = sign i, W8^3 = h (-1 + sign i)
                o1 = make_float2(h * (o1.x - sign * o1.y),
                                 h * (o1.y + sign * o1.x));
                o2 = mul_si(o2, sign);
                o3 = make_float2(h * (-o3.x - sign * o3.y),
                                 h * (-o3.y + sign * o3.x));
                v[0] = cadd(e0, o0); v[4] = csub(e0, o0);
                v[1] = cadd(e1, o1); v[5] = csub(e1, o1);
                v[2] = cadd(e2, o2); v[6] = csub(e2, o2);
                v[3] = cadd(e3, o3); v[7] = csub(e3, o3);
            }
            const int base = (j / ns) * span + k;
            for (int q = 0; q < r; ++q) out[base + q * ns] = v[q];
        }
    } else {
        // generic prime radix: one output per work item, direct DFT
        const int rstep = n / r;
        for (int it = threadIdx.x; it < lines * n; it += blockDim.x) {
            const int l = it / n;
            const int rem = it - l * n;
            const int j = rem % nb;
            const int q = rem / nb;
            const int k = j % ns;
            const float2* in = src + (size_t)l * ld;
            float2 acc = make_float2(0.0f, 0.0f);
            for (int t = 0; t < r; ++t) {
                float2 x = in[j + t * nb];
                if (t && k) x = cmul(x, twid(tw, (t * k * twstep) % n, sign));
                const int m = (t * q) % r;
                acc = cadd(acc, m ? cmul(x, twid(tw, m * rstep, sign)) : x);
            }
            dst[(size_t)l * ld + (j / ns) * span + k + q * ns] = acc;
        }
    }
}

// Full 1-D transform of `lines` lines in buf0 (result ends up in buf0 or
// buf1; returns the one holding it).
__device__ float2* lds_fft(float2* buf0, float2* buf1, int lines, int ld,
                           const AxisArg& a, const float2* tw, float sign)
{
    float2* src = buf0;
    float2* dst = buf1;
    int ns = 1;
    for (int s = 0; s < a.nst; ++s) {
        stockham_stage(src, dst, lines, ld, a.n, a.radix[s], ns, tw, sign);
        __syncthreads();
        ns *= a.radix[s];
        float2* t = src;
        src = dst;
        dst = t;
    }
    return src;
}

__device__ __forceinline__ void load_twiddles(float2* s_tw, const float2* tw,
                                              int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) s_tw[i] = tw[i];
}

// ---- row pass: real <-> half complex ---------------------------------------
// in: rows of nx reals (row stride nx); out: rows of nxh complex.
__global__ __launch_bounds__(256) void fft_rows_r2c(
    const float* __restrict__ in, float2* __restrict__ out, int nrows, int nx,
    AxisArg a, const float2* tw, const float2* twpost, int lines, int ld,
    long long bstride_r, long long bstride_c)
{
    extern __shared__ float2 lds[];
    float2* s_tw = lds;
    float2* b0 = s_tw + a.n;
    float2* b1 = b0 + (size_t)lines * ld;
    const int row0 = blockIdx.x * lines;
    const int nl = min(lines, nrows - row0);
    const int nxh = nx / 2 + 1;
    in += blockIdx.y * bstride_r;
    out += blockIdx.y * bstride_c;
    load_twiddles(s_tw, tw, a.n);
    const bool even = (nx & 1) == 0;
    const int n = a.n;
    for (int it = threadIdx.x; it < nl * n; it += blockDim.x) {
        const int l = it / n, m = it - l * n;
        const float* row = in + (size_t)(row0 + l) * nx;
        b0[(size_t)l * ld + m] = even ? make_float2(row[2 * m], row[2 * m + 1])
                                      : make_float2(row[m], 0.0f);
    }
    __syncthreads();
    const float2* Z = lds_fft(b0, b1, nl, ld, a, s_tw, -1.0f);
    for (int it = threadIdx.x; it < nl * nxh; it += blockDim.x) {
        const int l = it / nxh, k = it - l * nxh;
        const float2* z = Z + (size_t)l * ld;
        float2 X;
        if (even) {
            const float2 zk = z[k == n ? 0 : k];
            const float2 zm = conjf2(z[k == 0 ? 0 : n - k]);
            const float2 fe = make_float2(0.5f * (zk.x + zm.x),
                                          0.5f * (zk.y + zm.y));
            // fo = (zk - zm) / (2 i)
            const float2 dd = csub(zk, zm);
            const float2 fo = make_float2(0.5f * dd.y, -0.5f * dd.x);
            X = cadd(fe, cmul(twpost[k], fo));
        } else {
            X = z[k];
        }
        out[(size_t)(row0 + l) * nxh + k] = X;
    }
}

__global__ __launch_bounds__(256) void fft_rows_c2r(
    const float2* __restrict__ in, float* __restrict__ out, int nrows, int nx,
    AxisArg a, const float2* tw, const float2* twpost, int lines, int ld,
    long long bstride_c, long long bstride_r)
{
    extern __shared__ float2 lds[];
    float2* s_tw = lds;
    float2* b0 = s_tw + a.n;
    float2* b1 = b0 + (size_t)lines * ld;
    const int row0 = blockIdx.x * lines;
    const int nl = min(lines, nrows - row0);
    const int nxh = nx / 2 + 1;
    in += blockIdx.y * bstride_c;
    out += blockIdx.y * bstride_r;
    load_twiddles(s_tw, tw, a.n);
    const bool even = (nx & 1) == 0;
    const int n = a.n;
    for (int it = threadIdx.x; it < nl * n; it += blockDim.x) {
        const int l = it / n, k = it - l * n;
        const float2* row = in + (size_t)(row0 + l) * nxh;
        float2 Z;
        if (even) {
            // Z' = (X[k] + conj X[M-k]) + i (X[k] - conj X[M-k]) W^-k
            const float2 xk = row[k];
            const float2 xm = conjf2(row[n - k]);
            const float2 fe = cadd(xk, xm);
            const float2 w = conjf2(twpost[k]);
            const float2 fo = cmul(csub(xk, xm), w);
            Z = cadd(fe, make_float2(-fo.y, fo.x));
        } else {
            Z = k < nxh ? row[k] : conjf2(row[n - k]);
        }
        b0[(size_t)l * ld + k] = Z;
    }
    __syncthreads();
    const float2* z = lds_fft(b0, b1, nl, ld, a, s_tw, 1.0f);
    for (int it = threadIdx.x; it < nl * n; it += blockDim.x) {
        const int l = it / n, m = it - l * n;
        float* row = out + (size_t)(row0 + l) * nx;
        const float2 v = z[(size_t)l * ld + m];
        if (even) {
            row[2 * m] = v.x;
            row[2 * m + 1] = v.y;
        } else {
            row[m] = v.x;
        }
    }
}

// ---- line pass along y or z --------------------------------------------------
// Element n of line (outer, kx) is at outer*ostride + n*stride + kx.
// mode 0: forward, 1: inverse, 2: forward * K * scale then inverse.
__global__ __launch_bounds__(256) void fft_lines(
    const float2* __restrict__ src, float2* __restrict__ dst,
    const float2* __restrict__ K, int nxh, long long stride, long long ostride,
    AxisArg a, const float2* tw, int mode, float scale, int tl, int ld,
    long long bstride, long long kbstride)
{
    extern __shared__ float2 lds[];
    float2* s_tw = lds;
    float2* b0 = s_tw + a.n;
    float2* b1 = b0 + (size_t)tl * ld;
    const int kx0 = blockIdx.x * tl;
    const int nl = min(tl, nxh - kx0);
    const int n = a.n;
    const long long base = (long long)blockIdx.y * ostride + kx0;
    src += blockIdx.z * bstride + base;
    dst += blockIdx.z * bstride + base;
    if (K) K += blockIdx.z * kbstride + base;
    load_twiddles(s_tw, tw, n);
    for (int it = threadIdx.x; it < n * tl; it += blockDim.x) {
        const int m = it / tl, l = it - m * tl;
        if (l < nl) b0[(size_t)l * ld + m] = src[m * stride + l];
    }
    __syncthreads();
    float2* res = lds_fft(b0, b1, nl, ld, a, s_tw, mode == 1 ? 1.0f : -1.0f);
    if (mode == 2) {
        for (int it = threadIdx.x; it < n * tl; it += blockDim.x) {
            const int m = it / tl, l = it - m * tl;
            if (l < nl) {
                float2& v = res[(size_t)l * ld + m];
                const float2 kv = K[m * stride + l];
                v = make_float2((v.x * kv.x - v.y * kv.y) * scale,
                                (v.x * kv.y + v.y * kv.x) * scale);
            }
        }
        __syncthreads();
        float2* other = res == b0 ? b1 : b0;
        res = lds_fft(res, other, nl, ld, a, s_tw, 1.0f);
    }
    for (int it = threadIdx.x; it < n * tl; it += blockDim.x) {
        const int m = it / tl, l = it - m * tl;
        if (l < nl) dst[m * stride + l] = res[(size_t)l * ld + m];
    }
}

// Strided line pass.  Line q (of nlines) starts at
// base0 + (q / inner) * ostride + q % inner; its elements are `stride` apart.
// mode 0: forward, 1: inverse, 2: forward * K * scale then inverse.  Only
// elements m in [st_lo, st_hi) are stored.  Thread (l, m0) = (tid % L,
// tid / L) moves elements m0 + c T/L of line l.
template <int LSH, int T, int N = 0>
__global__ __launch_bounds__(T) void fx_lines(
    const float2* __restrict__ src, float2* __restrict__ dst,
    const float2* __restrict__ K, int nlines, int inner, unsigned inner_m,
    long long base0, long long ostride, long long stride, FAxis a,
    const float2* tw, const float2* tws, int mode, float scale, int st_lo, int st_hi,
    long long bstride, long long kbstride)
{
    extern __shared__ float2 lds[];
    constexpr int L = 1 << LSH, LS = L + 1, JS = T >> LSH;
    const int n = N ? N : a.n;
    float2* s_tw = lds;
    float2* buf = lds + tw_entries<N>(n);
    src += blockIdx.y * bstride;
    dst += blockIdx.y * bstride;
    if (K) K += blockIdx.y * kbstride;
    for (int i = threadIdx.x; i < tw_entries<N>(n); i += T)
        s_tw[i] = N ? tws[i] : tw[i];
    const int l = threadIdx.x & (L - 1);
    const int m0 = threadIdx.x >> LSH;
    const int q = (blockIdx.x << LSH) + l;
    const bool valid = q < nlines;
    const int qo = inner >= nlines ? 0 : (int)fdiv((unsigned)q, inner_m);
    const long long base = base0 + (long long)qo * ostride + (q - qo * inner) +
                           (long long)m0 * stride;
    const long long cstride = (long long)JS * stride;
    const int la = m0 * LS + l;
    // all global loads of a thread are issued before the first LDS write
    float2 r[kFastE];
#pragma unroll
    for (int c = 0; c < kFastE; ++c)
        r[c] = valid && m0 + c * JS < n ? src[base + c * cstride]
                                        : make_float2(0, 0);
#pragma unroll
    for (int c = 0; c < kFastE; ++c)
        if (m0 + c * JS < n) buf[la + c * JS * LS] = r[c];
    __syncthreads();
    tile_fft<LSH, T, N>(buf, a, s_tw, mode == 1 ? 1.0f : -1.0f);
    if (mode == 2) {
#pragma unroll
        for (int c = 0; c < kFastE; ++c)
            r[c] = valid && m0 + c * JS < n ? K[base + c * cstride]
                                            : make_float2(0, 0);
#pragma unroll
        for (int c = 0; c < kFastE; ++c)
            if (m0 + c * JS < n) {
                const float2 v = buf[la + c * JS * LS], kv = r[c];
                buf[la + c * JS * LS] =
                    make_float2((v.x * kv.x - v.y * kv.y) * scale,
                                (v.x * kv.y + v.y * kv.x) * scale);
            }
        __syncthreads();
        tile_fft<LSH, T, N>(buf, a, s_tw, 1.0f);
    }
    if (valid) {
#pragma unroll
        for (int c = 0; c < kFastE; ++c) {
            const int m = m0 + c * JS;
            if (m < n && m >= st_lo && m < st_hi)
                dst[base + c * cstride] = buf[la + c * JS * LS];
        }
    }
}

// Row pass, real -> half spectrum, nx even: a complex FFT of length n = nx/2
// over z[m] = x[2m] + i x[2m+1], then
// X[k] = (Z[k] + conj Z[n-k]) / 2 + W^k (Z[k] - conj Z[n-k]) / (2i).
// Element i of the block's contiguous tile is (row i / n, m = i % n).
template <int LSH, int T, int N = 0>
__global__ __launch_bounds__(T) void fx_rows_r2c(
    const float* __restrict__ in, float2* __restrict__ out, int nrows,
    FAxis a, unsigned n_m, unsigned nxh_m, const float2* tw, const float2* tws,
    const float2* __restrict__ twpost, long long bstride_r,
    long long bstride_c)
{
    extern __shared__ float2 lds[];
    constexpr int L = 1 << LSH, LS = L + 1;
    const int n = N ? N : a.n, nxh = n + 1;
    float2* s_tw = lds;
    float2* buf = lds + tw_entries<N>(n);
    float2* s_post = buf + n * LS;      // twpost staged in LDS
    const int r0 = blockIdx.x << LSH;
    const int nl = min(L, nrows - r0);
    const float2* in2 = reinterpret_cast<const float2*>(
                            in + blockIdx.y * bstride_r) + (size_t)r0 * n;
    out += blockIdx.y * bstride_c + (size_t)r0 * nxh;
    for (int i = threadIdx.x; i < tw_entries<N>(n); i += T)
        s_tw[i] = N ? tws[i] : tw[i];
    for (int i = threadIdx.x; i < nxh; i += T) s_post[i] = twpost[i];
    const int tile = n << LSH;
    const int valid = n * nl;
    float2 r[kFastE];
#pragma unroll
    for (int c = 0; c < kFastE; ++c) {
        const int i = threadIdx.x + c * T;
        r[c] = i < valid ? in2[i] : make_float2(0, 0);
    }
#pragma unroll
    for (int c = 0; c < kFastE; ++c) {
        const int i = threadIdx.x + c * T;
        if (i < tile) {
            const int l = (int)fdiv((unsigned)i, n_m), m = i - l * n;
            buf[m * LS + l] = r[c];
        }
    }
    __syncthreads();
    tile_fft<LSH, T, N>(buf, a, s_tw, -1.0f);
    const int cnt = nl * nxh;
    for (int i = threadIdx.x; i < cnt; i += T) {
        const int l = (int)fdiv((unsigned)i, nxh_m), k = i - l * nxh;
        const float2 zk = buf[(k == n ? 0 : k) * LS + l];
        const float2 zm = conjf2(buf[(k == 0 ? 0 : n - k) * LS + l]);
        const float2 fe = make_float2(0.5f * (zk.x + zm.x),
                                      0.5f * (zk.y + zm.y));
        const float2 dd = csub(zk, zm);
        const float2 fo = make_float2(0.5f * dd.y, -0.5f * dd.x);
        out[i] = cadd(fe, cmul(s_post[k], fo));
    }
}

// Row pass, half spectrum -> real, nx even (mirror of fx_rows_r2c):
// Z[k] = (X[k] + conj X[n-k]) + i W^-k (X[k] - conj X[n-k]).
template <int LSH, int T, int N = 0>
__global__ __launch_bounds__(T) void fx_rows_c2r(
    const float2* __restrict__ in, float* __restrict__ out, int nrows,
    FAxis a, unsigned n_m, const float2* tw, const float2* tws,
    const float2* __restrict__ twpost, long long bstride_c,
    long long bstride_r)
{
    extern __shared__ float2 lds[];
    constexpr int L = 1 << LSH, LS = L + 1;
    const int n = N ? N : a.n, nxh = n + 1;
    float2* s_tw = lds;
    float2* buf = lds + tw_entries<N>(n);
    float2* s_post = buf + n * LS;      // twpost staged in LDS
    const int r0 = blockIdx.x << LSH;
    const int nl = min(L, nrows - r0);
    in += blockIdx.y * bstride_c + (size_t)r0 * nxh;
    float2* out2 = reinterpret_cast<float2*>(out + blockIdx.y * bstride_r) +
                   (size_t)r0 * n;
    for (int i = threadIdx.x; i < tw_entries<N>(n); i += T)
        s_tw[i] = N ? tws[i] : tw[i];
    for (int i = threadIdx.x; i < nxh; i += T) s_post[i] = twpost[i];
    __syncthreads();
    const int tile = n << LSH;
    float2 xa[kFastE], xb[kFastE];
#pragma unroll
    for (int c = 0; c < kFastE; ++c) {
        const int i = threadIdx.x + c * T;
        const int l = (int)fdiv((unsigned)i, n_m), k = i - l * n;
        xa[c] = xb[c] = make_float2(0, 0);
        if (i < tile && l < nl) {
            const float2* row = in + (size_t)l * nxh;
            xa[c] = row[k];
            xb[c] = row[n - k];
        }
    }
#pragma unroll
    for (int c = 0; c < kFastE; ++c) {
        const int i = threadIdx.x + c * T;
        if (i < tile) {
            const int l = (int)fdiv((unsigned)i, n_m), k = i - l * n;
            const float2 xk = xa[c];
            const float2 xm = conjf2(xb[c]);
            const float2 fe = cadd(xk, xm);
            const float2 fo = cmul(csub(xk, xm), conjf2(s_post[k]));
            buf[k * LS + l] = cadd(fe, make_float2(-fo.y, fo.x));
        }
    }
    __syncthreads();
    tile_fft<LSH, T, N>(buf, a, s_tw, 1.0f);
    const int cnt = nl * n;
    for (int i = threadIdx.x; i < cnt; i += T) {
        const int l = (int)fdiv((unsigned)i, n_m), m = i - l * n;
        out2[i] = buf[m * LS + l];
    }
}

// Pointwise a <- a * K * scale (convolution of arrays with ny = nz = 1).
__global__ void spectrum_multiply(float2* a, const float2* K, long long n,
                                  float scale, long long bstride,
                                  long long kbstride)
{
    a += blockIdx.y * bstride;
    K += blockIdx.y * kbstride;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const float2 u = a[i], v = K[i];
        a[i] = make_float2((u.x * v.x - u.y * v.y) * scale,
                           (u.x * v.y + u.y * v.x) * scale);
    }
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
constexpr size_t kLdsBudget = 64 * 1024;

static int pad_ld(int n) { return n + (n % 2 == 0 ? 1 : 0); }

void launch_fx_lines(Context& c, const FastShape& f, const float2* tw,
                     const float2* tws, const float2* src, float2* dst,
                     const float2* K,
                     int nlines, int inner, long long base0,
                     long long ostride, long long stride, int st_lo,
                     int st_hi, int mode, float scale, int batch,
                     long long bstride, long long kbstride)
{
    GBK_CHECK(f.ok && nlines > 0 && inner > 0 &&
              (long long)nlines * inner < (1LL << 32) && st_lo >= 0 &&
              st_hi <= f.a.n, GBKHIP_ERR_INTERNAL, "bad line pass shape");
    const unsigned inner_m = inner >= nlines ? 0u : div_magic((unsigned)inner);
    dim3 grid(ceil_div(nlines, 1 << f.lshift), batch);
    GBK_DISPATCH(GBK_SPEC_LINES, f, fx_lines, grid, f.lds, c.stream, src,
                 dst, K, nlines, inner, inner_m, base0, ostride, stride, f.a,
                 tw, tws, mode, scale, st_lo, st_hi, bstride, kbstride);
    GBK_HIP(hipGetLastError());
}

static void launch_rows_fast(Context& c, const FFTPlan& p, bool forward,
                             const void* in, void* out, int batch,
                             long long br, long long bc)
{
    const FastShape& f = p.fx;
    const int nrows = p.ny * p.nz;
    const int n = f.a.n;
    const unsigned n_m = div_magic((unsigned)n);
    const unsigned nxh_m = div_magic((unsigned)(n + 1));
    const size_t lds = f.lds + (size_t)(n + 1) * sizeof(float2);
    dim3 grid(ceil_div(nrows, 1 << f.lshift), batch);
    if (forward) {
        GBK_DISPATCH(GBK_SPEC_ROWS, f, fx_rows_r2c, grid, lds, c.stream,
                     (const float*)in, (float2*)out, nrows, f.a, n_m, nxh_m,
                     p.twx.as<float2>(), p.twsx.as<float2>(),
                     p.twpost.as<float2>(), br, bc);
    } else {
        GBK_DISPATCH(GBK_SPEC_ROWS, f, fx_rows_c2r, grid, lds, c.stream,
                     (const float2*)in, (float*)out, nrows, f.a, n_m,
                     p.twx.as<float2>(), p.twsx.as<float2>(),
                     p.twpost.as<float2>(), bc, br);
    }
    GBK_HIP(hipGetLastError());
}

static void launch_rows(Context& c, const FFTPlan& p, bool forward,
                        const void* in, void* out, int batch,
                        long long bstride_real, long long bstride_cplx)
{
    if (p.fx.ok) {
        launch_rows_fast(c, p, forward, in, out, batch, bstride_real,
                         bstride_cplx);
        return;
    }
    const int nrows = p.ny * p.nz;
    const int n = p.ax.n;
    const int ld = pad_ld(n);
    int lines = (int)((kLdsBudget / sizeof(float2) - n) / (2 * (size_t)ld));
    lines = std::max(1, std::min(lines, 16));
    const size_t lds = ((size_t)n + 2 * (size_t)lines * ld) * sizeof(float2);
    GBK_CHECK(lds <= 160 * 1024, GBKHIP_ERR_UNSUPPORTED,
              "FFT row too long for LDS");
    dim3 grid(ceil_div(nrows, lines), batch);
    if (forward)
        hipLaunchKernelGGL(fft_rows_r2c, grid, dim3(256), lds, c.stream,
                           (const float*)in, (float2*)out, nrows, p.nx,
                           axis_arg(p.ax), p.twx.as<float2>(),
                           p.twpost.as<float2>(), lines, ld, bstride_real,
                           bstride_cplx);
    else
        hipLaunchKernelGGL(fft_rows_c2r, grid, dim3(256), lds, c.stream,
                           (const float2*)in, (float*)out, nrows, p.nx,
                           axis_arg(p.ax), p.twx.as<float2>(),
                           p.twpost.as<float2>(), lines, ld, bstride_cplx,
                           bstride_real);
    GBK_HIP(hipGetLastError());
}

// axis 1 = y, 2 = z
static void launch_lines(Context& c, const FFTPlan& p, int axis, int mode,
                         const float2* src, float2* dst, const float2* K,
                         float scale, int batch, long long bstride,
                         long long kbstride)
{
    const Axis& A = axis == 1 ? p.ay : p.az;
    const float2* tw = axis == 1 ? p.twy.as<float2>() : p.twz.as<float2>();
    const long long stride = axis == 1 ? p.nxh : (long long)p.nxh * p.ny;
    const long long ostride = axis == 1 ? (long long)p.nxh * p.ny : p.nxh;
    const int outer = axis == 1 ? p.nz : p.ny;
    const FastShape& f = axis == 1 ? p.fy : p.fz;
    // line q of the flattened (outer, kx) / (y, kx) index; the y pass needs
    // q / nxh, the z pass has all lines contiguous
    const long long nlines = (long long)p.nxh * outer;
    const int inner = axis == 1 ? p.nxh : (int)nlines;
    if (f.ok && nlines * (long long)inner < (1LL << 32)) {
        const unsigned inner_m = axis == 1 ? div_magic((unsigned)inner) : 0u;
        const long long os = axis == 1 ? ostride : 0;
        dim3 grid(ceil_div(nlines, 1 << f.lshift), batch);
        const float2* tws = axis == 1 ? p.twsy.as<float2>()
                                      : p.twsz.as<float2>();
        launch_fx_lines(c, f, tw, tws, src, dst, K, (int)nlines, inner, 0,
                        os, stride, 0, f.a.n, mode, scale, batch, bstride,
                        kbstride);
        return;
    }
    const int n = A.n;
    const int ld = pad_ld(n);
    int tl = (int)((kLdsBudget / sizeof(float2) - n) / (2 * (size_t)ld));
    tl = std::max(1, std::min(tl, 16));
    const size_t lds = ((size_t)n + 2 * (size_t)tl * ld) * sizeof(float2);
    GBK_CHECK(lds <= 160 * 1024, GBKHIP_ERR_UNSUPPORTED,
              "FFT line too long for LDS");
    dim3 grid(ceil_div(p.nxh, tl), outer, batch);
    hipLaunchKernelGGL(fft_lines, grid, dim3(256), lds, c.stream, src, dst, K,
                       p.nxh, stride, ostride, axis_arg(A), tw, mode, scale,
                       tl, ld, bstride, kbstride);
    GBK_HIP(hipGetLastError());
}

void fft_prepare_attributes()
{
    static std::once_flag once;
    std::call_once(once, [] {
        (void)hipFuncSetAttribute((const void*)fft_rows_r2c,
                                  hipFuncAttributeMaxDynamicSharedMemorySize,
                                  160 * 1024);
        (void)hipFuncSetAttribute((const void*)fft_rows_c2r,
                                  hipFuncAttributeMaxDynamicSharedMemorySize,
                                  160 * 1024);
        (void)hipFuncSetAttribute((const void*)fft_lines,
                                  hipFuncAttributeMaxDynamicSharedMemorySize,
                                  160 * 1024);
        const void* fast[] = {GBK_FAST_KERNELS(fx_lines),
                              GBK_FAST_KERNELS(fx_rows_r2c),
                              GBK_FAST_KERNELS(fx_rows_c2r)};
        for (const void* k : fast)
            (void)hipFuncSetAttribute(k,
                                      hipFuncAttributeMaxDynamicSharedMemorySize,
                                      96 * 1024);
    });
}

// Forward r2c of `batch` cubes (real batch stride br, complex stride bc).
void fft_forward(Context& c, const FFTPlan& p, const float* in, float2* out,
                 int batch, long long br, long long bc)
{
    fft_prepare_attributes();
    launch_rows(c, p, true, in, out, batch, br, bc);
    if (p.ny > 1) launch_lines(c, p, 1, 0, out, out, nullptr, 1, batch, bc, 0);
    if (p.nz > 1) launch_lines(c, p, 2, 0, out, out, nullptr, 1, batch, bc, 0);
}

// Inverse c2r; `in` is read (not modified) when work != in.
void fft_inverse(Context& c, const FFTPlan& p, const float2* in, float2* work,
                 float* out, int batch, long long bc, long long br)
{
    fft_prepare_attributes();
    const float2* cur = in;
    if (p.nz > 1) {
        launch_lines(c, p, 2, 1, cur, work, nullptr, 1, batch, bc, 0);
        cur = work;
    }
    if (p.ny > 1) {
        launch_lines(c, p, 1, 1, cur, work, nullptr, 1, batch, bc, 0);
        cur = work;
    }
    launch_rows(c, p, false, cur, out, batch, br, bc);
}

// data_r <- irfft(rfft(data_r) * K) / N using `spec` as the work spectrum.
void fft_convolve_with(Context& c, const FFTPlan& p, float* data_r,
                       float2* spec, const float2* K, int batch, long long br,
                       long long bc, long long kbc)
{
    fft_prepare_attributes();
    const float scale = 1.0f / ((float)p.nx * (float)p.ny * (float)p.nz);
    launch_rows(c, p, true, data_r, spec, batch, br, bc);
    if (p.nz > 1) {
        if (p.ny > 1)
            launch_lines(c, p, 1, 0, spec, spec, nullptr, 1, batch, bc, 0);
        launch_lines(c, p, 2, 2, spec, spec, K, scale, batch, bc, kbc);
        if (p.ny > 1)
            launch_lines(c, p, 1, 1, spec, spec, nullptr, 1, batch, bc, 0);
    } else if (p.ny > 1) {
        launch_lines(c, p, 1, 2, spec, spec, K, scale, batch, bc, kbc);
    } else {
        const long long n = (long long)p.nxh;
        dim3 grid(ceil_div(n, 256), batch);
        hipLaunchKernelGGL(spectrum_multiply, grid, dim3(256), 0, c.stream,
                           spec, K, n, scale, bc, kbc);
        GBK_HIP(hipGetLastError());
    }
    launch_rows(c, p, false, spec, data_r, batch, br, bc);
}

}  // namespace gbkhip
