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

static void upload_twiddles(DevBuf& buf, int n, int count)
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
// into stages of radix <= 16, then one stage per factor 3 and 5.
static bool fast_radices(int n, std::vector<int>& R)
{
    int m = n, a2 = 0, a3 = 0, a5 = 0;
    while (m % 2 == 0) { m /= 2; ++a2; }
    while (m % 3 == 0) { m /= 3; ++a3; }
    while (m % 5 == 0) { m /= 5; ++a5; }
    if (m != 1) return false;
    const int s2 = (a2 + 3) / 4;
    for (int s = 0; s < s2; ++s) {
        const int bits = a2 / s2 + (s < a2 % s2 ? 1 : 0);
        R.push_back(1 << bits);
    }
    for (int i = 0; i < a3; ++i) R.push_back(3);
    for (int i = 0; i < a5; ++i) R.push_back(5);
    return (int)R.size() <= kFMax;
}

// Magic multiplier m with floor(x / d) = umulhi(x, m) whenever x * d < 2^32
// (m = ceil(2^32 / d); d = 1 is flagged by m = 0).
static unsigned div_magic(unsigned d)
{
    return d <= 1 ? 0u
                  : (unsigned)((((unsigned long long)1 << 32) + d - 1) / d);
}

constexpr int kFastThreads = 256;
constexpr size_t kFastLdsTarget = 48 * 1024;

static FastShape fast_shape(int n)
{
    FastShape f;
    std::vector<int> R;
    if (n < 2 || !fast_radices(n, R)) return f;
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
    // most lines per block (best coalescing) whose tile fits the LDS target
    // and whose elements per thread stay within the register buckets (32
    // complex values per thread keeps the kernels near 150 VGPRs, no spills)
    int lshift = -1;
    for (int ls = 5; ls >= 2; --ls) {
        const size_t bytes = (size_t)n * ((1 << ls) + 2) * sizeof(float2);
        const int e = ((n << ls) + kFastThreads - 1) / kFastThreads;
        if ((bytes <= kFastLdsTarget || (ls == 2 && bytes <= 96 * 1024)) &&
            e <= 32) {
            lshift = ls;
            break;
        }
    }
    if (lshift < 0) return f;
    const size_t lds = (size_t)n * ((1 << lshift) + 2) * sizeof(float2);
    const int e = ((n << lshift) + kFastThreads - 1) / kFastThreads;
    const int emax = e <= 8 ? 8 : e <= 16 ? 16 : e <= 24 ? 24 : 32;
    f.ok = true;
    f.lshift = lshift;
    f.emax = emax;
    f.lds = lds;
    return f;
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
    if (p->even) p->fx = fast_shape(rn);
    p->fy = fast_shape(ny);
    p->fz = fast_shape(nz);
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

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b)
{ return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b)
{ return make_float2(a.x - b.x, a.y - b.y); }
// multiply by (sign * i)
__device__ __forceinline__ float2 mul_si(float2 a, float sign)
{ return make_float2(-sign * a.y, sign * a.x); }
__device__ __forceinline__ float2 conjf2(float2 a)
{ return make_float2(a.x, -a.y); }

// twiddle exp(sign * 2 pi i m / n) from the forward table
__device__ __forceinline__ float2 twid(const float2* tw, int m, float sign)
{
    const float2 t = tw[m];
    return sign < 0 ? t : make_float2(t.x, -t.y);
}

__device__ __forceinline__ void dft4(float2& a0, float2& a1, float2& a2,
                                     float2& a3, float sign)
{
    const float2 s02 = cadd(a0, a2), d02 = csub(a0, a2);
    const float2 s13 = cadd(a1, a3), d13 = mul_si(csub(a1, a3), sign);
    a0 = cadd(s02, s13);
    a2 = csub(s02, s13);
    a1 = cadd(d02, d13);
    a3 = csub(d02, d13);
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

// ---------------------------------------------------------------------------
// Register-tiled engine (2/3/5-smooth lengths)
//
// A block transforms L = 2^lshift lines held in one LDS tile laid out
// [m][L+1] (element m of line l at m*(L+1) + l).  Threads walk butterflies
// with the line index fastest, so the lanes of a wave touch consecutive LDS
// words and share one twiddle.  Each Stockham stage reads its butterflies
// into registers (compile-time radix, statically indexed arrays), syncs, and
// writes them back in autosort order into the same tile.
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned fdiv(unsigned x, unsigned magic)
{
    return magic ? __umulhi(x, magic) : x;
}

template <int R> struct Bfly;

template <> struct Bfly<2> {
    __device__ static constexpr int perm(int q) { return q; }
    __device__ static void run(float2* v, float)
    {
        const float2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};
template <> struct Bfly<3> {
    __device__ static constexpr int perm(int q) { return q; }
    __device__ static void run(float2* v, float sign)
    {
        const float2 s = cadd(v[1], v[2]);
        const float2 d = mul_si(csub(v[1], v[2]),
                                sign * 0.866025403784438647f);
        const float2 t = make_float2(v[0].x - 0.5f * s.x, v[0].y - 0.5f * s.y);
        v[0] = cadd(v[0], s);
        v[1] = cadd(t, d);
        v[2] = csub(t, d);
    }
};
template <> struct Bfly<4> {
    __device__ static constexpr int perm(int q) { return q; }
    __device__ static void run(float2* v, float sign)
    {
        dft4(v[0], v[1], v[2], v[3], sign);
    }
};
template <> struct Bfly<5> {
    __device__ static constexpr int perm(int q) { return q; }
    __device__ static void run(float2* v, float sign)
    {
        // cos/sin of 2 pi / 5 and 4 pi / 5
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
        // u1 = s1 d1 + s2 d2, u2 = s2 d1 - s1 d2, each times i
        const float2 u1 = make_float2(-(s1 * d1.y + s2 * d2.y),
                                      s1 * d1.x + s2 * d2.x);
        const float2 u2 = make_float2(-(s2 * d1.y - s1 * d2.y),
                                      s2 * d1.x - s1 * d2.x);
        v[0] = cadd(v[0], cadd(b1, b2));
        v[1] = cadd(t1, u1);
        v[4] = csub(t1, u1);
        v[2] = cadd(t2, u2);
        v[3] = csub(t2, u2);
    }
};
template <> struct Bfly<8> {
    __device__ static constexpr int perm(int q) { return q; }
    __device__ static void run(float2* v, float sign)
    {
        float2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
        float2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
        dft4(e0, e1, e2, e3, sign);
        dft4(o0, o1, o2, o3, sign);
        const float h = 0.707106781186547524f;
        // W8^1 = h (1 + sign i), W8^2 = sign i, W8^3 = h (-1 + sign i)
        o1 = make_float2(h * (o1.x - sign * o1.y), h * (o1.y + sign * o1.x));
        o2 = mul_si(o2, sign);
        o3 = make_float2(h * (-o3.x - sign * o3.y), h * (-o3.y + sign * o3.x));
        v[0] = cadd(e0, o0); v[4] = csub(e0, o0);
        v[1] = cadd(e1, o1); v[5] = csub(e1, o1);
        v[2] = cadd(e2, o2); v[6] = csub(e2, o2);
        v[3] = cadd(e3, o3); v[7] = csub(e3, o3);
    }
};
template <> struct Bfly<16> {
    // 16 = 4 x 4: index n = n1 + 4 n2 in, k = 4 k1 + k2 out;
    // X[4 k1 + k2] = sum_n1 W4^(n1 k1) W16^(n1 k2) sum_n2 x[n1 + 4 n2] W4^(n2 k2)
    // computed in place: X[4 k1 + k2] ends in v[4 k2 + k1].
    __device__ static constexpr int perm(int q) { return 4 * (q % 4) + q / 4; }
    __device__ static void run(float2* v, float sign)
    {
        // cos and sin of 2 pi m / 16, m = 0..9
        constexpr float C[10] = {1.0f, 0.923879532511286756f,
                                 0.707106781186547524f, 0.382683432365089772f,
                                 0.0f, -0.382683432365089772f,
                                 -0.707106781186547524f, -0.923879532511286756f,
                                 -1.0f, -0.923879532511286756f};
        constexpr float S[10] = {0.0f, 0.382683432365089772f,
                                 0.707106781186547524f, 0.923879532511286756f,
                                 1.0f, 0.923879532511286756f,
                                 0.707106781186547524f, 0.382683432365089772f,
                                 0.0f, -0.382683432365089772f};
#pragma unroll
        for (int n1 = 0; n1 < 4; ++n1)
            dft4(v[n1], v[n1 + 4], v[n1 + 8], v[n1 + 12], sign);
        // v[n1 + 4 k2] now holds Y[n1][k2]
#pragma unroll
        for (int n1 = 1; n1 < 4; ++n1)
#pragma unroll
            for (int k2 = 1; k2 < 4; ++k2) {
                const int m = n1 * k2;
                v[n1 + 4 * k2] = cmul(v[n1 + 4 * k2],
                                      make_float2(C[m], sign * S[m]));
            }
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2)
            dft4(v[4 * k2], v[4 * k2 + 1], v[4 * k2 + 2], v[4 * k2 + 3], sign);
    }
};

// One Stockham stage of radix R over the tile; CAP = butterflies per thread.
// The per-stage constants come from the stage table (not from n and R), so
// the compiler cannot hoist one radix case's addresses out of the stage loop
// and keep all cases' live at once.
template <int R, int CAP>
__device__ __forceinline__ void fstage(float2* buf, int lshift, int nb, int ns,
                                       int twstep, unsigned nsm,
                                       const float2* tw, float sign)
{
    constexpr int T = kFastThreads;
    const int LS = (1 << lshift) + 1;
    const int lmask = (1 << lshift) - 1;
    const int total = nb << lshift;
    float2 v[CAP][R];
#pragma unroll
    for (int c = 0; c < CAP; ++c) {
        const int b = threadIdx.x + c * T;
        if (b < total) {
            const int l = b & lmask, j = b >> lshift;
            const int k = j - ns * (int)fdiv((unsigned)j, nsm);
#pragma unroll
            for (int t = 0; t < R; ++t) v[c][t] = buf[(j + t * nb) * LS + l];
            if (k) {
#pragma unroll
                for (int t = 1; t < R; ++t)
                    v[c][t] = cmul(v[c][t], twid(tw, t * k * twstep, sign));
            }
            Bfly<R>::run(v[c], sign);
        }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < CAP; ++c) {
        const int b = threadIdx.x + c * T;
        if (b < total) {
            const int l = b & lmask, j = b >> lshift;
            const int k = j - ns * (int)fdiv((unsigned)j, nsm);
            const int base = (j - k) * R + k;
#pragma unroll
            for (int q = 0; q < R; ++q)
                buf[(base + q * ns) * LS + l] = v[c][Bfly<R>::perm(q)];
        }
    }
    __syncthreads();
}

// Whole transform of the tile (the tile must be complete on entry; it is
// complete again on return).
template <int EMAX>
__device__ void fast_fft(float2* buf, const FAxis& a, int lshift,
                         const float2* tw, float sign)
{
#pragma unroll 1
    for (int s = 0; s < a.nst; ++s) {
        const int nb = a.nb[s], ns = a.ns[s], tws = a.tws[s];
        const unsigned nsm = a.nsm[s];
        switch (a.R[s]) {
        case 2: fstage<2, (EMAX + 1) / 2>(buf, lshift, nb, ns, tws, nsm, tw, sign); break;
        case 3: fstage<3, (EMAX + 2) / 3>(buf, lshift, nb, ns, tws, nsm, tw, sign); break;
        case 4: fstage<4, (EMAX + 3) / 4>(buf, lshift, nb, ns, tws, nsm, tw, sign); break;
        case 5: fstage<5, (EMAX + 4) / 5>(buf, lshift, nb, ns, tws, nsm, tw, sign); break;
        case 8: fstage<8, (EMAX + 7) / 8>(buf, lshift, nb, ns, tws, nsm, tw, sign); break;
        default: fstage<16, (EMAX + 15) / 16>(buf, lshift, nb, ns, tws, nsm, tw, sign); break;
        }
    }
}

// Strided line pass along y or z.  Line q (of nlines) starts at
// (q / inner) * ostride + q % inner; its elements are `stride` apart.
// mode 0: forward, 1: inverse, 2: forward * K * scale then inverse.
template <int EMAX>
__global__ __launch_bounds__(256) void fx_lines(
    const float2* __restrict__ src, float2* __restrict__ dst,
    const float2* __restrict__ K, int nlines, int inner, unsigned inner_m,
    long long ostride, long long stride, FAxis a, const float2* tw,
    int lshift, int mode, float scale, long long bstride, long long kbstride)
{
    extern __shared__ float2 lds[];
    constexpr int T = kFastThreads;
    const int n = a.n;
    const int LS = (1 << lshift) + 1;
    float2* s_tw = lds;
    float2* buf = lds + n;
    src += blockIdx.y * bstride;
    dst += blockIdx.y * bstride;
    if (K) K += blockIdx.y * kbstride;
    for (int i = threadIdx.x; i < n; i += T) s_tw[i] = tw[i];
    const int l = threadIdx.x & ((1 << lshift) - 1);
    const int q = (blockIdx.x << lshift) + l;
    const bool valid = q < nlines;
    const int qo = inner >= nlines ? 0 : (int)fdiv((unsigned)q, inner_m);
    const long long base = (long long)qo * ostride + (q - qo * inner);
    // a thread owns elements m0 + c * mstep (c < EMAX) of line l; all its
    // global loads are issued before the first LDS write so that EMAX loads
    // per lane are in flight
    const int m0 = threadIdx.x >> lshift, mstep = T >> lshift;
    float2 r[EMAX];
#pragma unroll
    for (int c = 0; c < EMAX; ++c) {
        const int m = m0 + c * mstep;
        r[c] = valid && m < n ? src[base + (long long)m * stride]
                              : make_float2(0, 0);
    }
#pragma unroll
    for (int c = 0; c < EMAX; ++c) {
        const int m = m0 + c * mstep;
        if (m < n) buf[m * LS + l] = r[c];
    }
    __syncthreads();
    fast_fft<EMAX>(buf, a, lshift, s_tw, mode == 1 ? 1.0f : -1.0f);
    if (mode == 2) {
#pragma unroll
        for (int c = 0; c < EMAX; ++c) {
            const int m = m0 + c * mstep;
            r[c] = valid && m < n ? K[base + (long long)m * stride]
                                  : make_float2(0, 0);
        }
#pragma unroll
        for (int c = 0; c < EMAX; ++c) {
            const int m = m0 + c * mstep;
            if (m < n) {
                const float2 v = buf[m * LS + l], kv = r[c];
                buf[m * LS + l] = make_float2((v.x * kv.x - v.y * kv.y) * scale,
                                              (v.x * kv.y + v.y * kv.x) * scale);
            }
        }
        __syncthreads();
        fast_fft<EMAX>(buf, a, lshift, s_tw, 1.0f);
    }
    if (valid) {
#pragma unroll
        for (int c = 0; c < EMAX; ++c) {
            const int m = m0 + c * mstep;
            if (m < n) dst[base + (long long)m * stride] = buf[m * LS + l];
        }
    }
}

// Row pass, real -> half spectrum, nx even: a complex FFT of length n = nx/2
// over z[m] = x[2m] + i x[2m+1], then
// X[k] = (Z[k] + conj Z[n-k]) / 2 + W^k (Z[k] - conj Z[n-k]) / (2i).
template <int EMAX>
__global__ __launch_bounds__(256) void fx_rows_r2c(
    const float* __restrict__ in, float2* __restrict__ out, int nrows,
    FAxis a, unsigned n_m, unsigned nxh_m, const float2* tw,
    const float2* __restrict__ twpost, int lshift, long long bstride_r,
    long long bstride_c)
{
    extern __shared__ float2 lds[];
    constexpr int T = kFastThreads;
    const int n = a.n, nxh = n + 1;
    const int LS = (1 << lshift) + 1;
    float2* s_tw = lds;
    float2* buf = lds + n;
    const int r0 = blockIdx.x << lshift;
    const int nl = min(1 << lshift, nrows - r0);
    const float2* in2 = reinterpret_cast<const float2*>(
                            in + blockIdx.y * bstride_r) + (size_t)r0 * n;
    out += blockIdx.y * bstride_c + (size_t)r0 * nxh;
    for (int i = threadIdx.x; i < n; i += T) s_tw[i] = tw[i];
    const int tile = n << lshift;
    const int valid = n * nl;
    float2 r[EMAX];
#pragma unroll
    for (int c = 0; c < EMAX; ++c) {
        const int i = threadIdx.x + c * T;
        r[c] = i < valid ? in2[i] : make_float2(0, 0);
    }
#pragma unroll
    for (int c = 0; c < EMAX; ++c) {
        const int i = threadIdx.x + c * T;
        if (i < tile) {
            const int l = (int)fdiv((unsigned)i, n_m), m = i - l * n;
            buf[m * LS + l] = r[c];
        }
    }
    __syncthreads();
    fast_fft<EMAX>(buf, a, lshift, s_tw, -1.0f);
    const int cnt = nl * nxh;
    for (int i = threadIdx.x; i < cnt; i += T) {
        const int l = (int)fdiv((unsigned)i, nxh_m), k = i - l * nxh;
        const float2 zk = buf[(k == n ? 0 : k) * LS + l];
        const float2 zm = conjf2(buf[(k == 0 ? 0 : n - k) * LS + l]);
        const float2 fe = make_float2(0.5f * (zk.x + zm.x),
                                      0.5f * (zk.y + zm.y));
        const float2 dd = csub(zk, zm);
        const float2 fo = make_float2(0.5f * dd.y, -0.5f * dd.x);
        out[i] = cadd(fe, cmul(twpost[k], fo));
    }
}

// Row pass, half spectrum -> real, nx even (mirror of fx_rows_r2c):
// Z[k] = (X[k] + conj X[n-k]) + i W^-k (X[k] - conj X[n-k]).
template <int EMAX>
__global__ __launch_bounds__(256) void fx_rows_c2r(
    const float2* __restrict__ in, float* __restrict__ out, int nrows,
    FAxis a, unsigned n_m, const float2* tw,
    const float2* __restrict__ twpost, int lshift, long long bstride_c,
    long long bstride_r)
{
    extern __shared__ float2 lds[];
    constexpr int T = kFastThreads;
    const int n = a.n, nxh = n + 1;
    const int LS = (1 << lshift) + 1;
    float2* s_tw = lds;
    float2* buf = lds + n;
    const int r0 = blockIdx.x << lshift;
    const int nl = min(1 << lshift, nrows - r0);
    in += blockIdx.y * bstride_c + (size_t)r0 * nxh;
    float2* out2 = reinterpret_cast<float2*>(out + blockIdx.y * bstride_r) +
                   (size_t)r0 * n;
    for (int i = threadIdx.x; i < n; i += T) s_tw[i] = tw[i];
    const int tile = n << lshift;
    float2 xa[EMAX], xb[EMAX];
#pragma unroll
    for (int c = 0; c < EMAX; ++c) {
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
    for (int c = 0; c < EMAX; ++c) {
        const int i = threadIdx.x + c * T;
        if (i < tile) {
            const int l = (int)fdiv((unsigned)i, n_m), k = i - l * n;
            const float2 xk = xa[c];
            const float2 xm = conjf2(xb[c]);
            const float2 fe = cadd(xk, xm);
            const float2 fo = cmul(csub(xk, xm), conjf2(twpost[k]));
            buf[k * LS + l] = cadd(fe, make_float2(-fo.y, fo.x));
        }
    }
    __syncthreads();
    fast_fft<EMAX>(buf, a, lshift, s_tw, 1.0f);
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

#define GBK_FAST_DISPATCH(EMAX, KERNEL, ...)                                 \
    switch (EMAX) {                                                          \
    case 8: hipLaunchKernelGGL(KERNEL<8>, __VA_ARGS__); break;               \
    case 16: hipLaunchKernelGGL(KERNEL<16>, __VA_ARGS__); break;             \
    case 24: hipLaunchKernelGGL(KERNEL<24>, __VA_ARGS__); break;             \
    default: hipLaunchKernelGGL(KERNEL<32>, __VA_ARGS__); break;             \
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
    const size_t lds = f.lds;
    dim3 grid(ceil_div(nrows, 1 << f.lshift), batch);
    if (forward) {
        GBK_FAST_DISPATCH(f.emax, fx_rows_r2c, grid, dim3(kFastThreads), lds,
                          c.stream, (const float*)in, (float2*)out, nrows,
                          f.a, n_m, nxh_m, p.twx.as<float2>(),
                          p.twpost.as<float2>(), f.lshift, br, bc);
    } else {
        GBK_FAST_DISPATCH(f.emax, fx_rows_c2r, grid, dim3(kFastThreads), lds,
                          c.stream, (const float2*)in, (float*)out, nrows,
                          f.a, n_m, p.twx.as<float2>(),
                          p.twpost.as<float2>(), f.lshift, bc, br);
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
        GBK_FAST_DISPATCH(f.emax, fx_lines, grid, dim3(kFastThreads), f.lds,
                          c.stream, src, dst, K, (int)nlines, inner, inner_m,
                          os, stride, f.a, tw, f.lshift, mode, scale, bstride,
                          kbstride);
        GBK_HIP(hipGetLastError());
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
        const void* fast[] = {
            (const void*)fx_lines<8>, (const void*)fx_lines<16>,
            (const void*)fx_lines<24>, (const void*)fx_lines<32>,
            (const void*)fx_rows_r2c<8>, (const void*)fx_rows_r2c<16>,
            (const void*)fx_rows_r2c<24>, (const void*)fx_rows_r2c<32>,
            (const void*)fx_rows_c2r<8>, (const void*)fx_rows_c2r<16>,
            (const void*)fx_rows_c2r<24>, (const void*)fx_rows_c2r<32>};
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
