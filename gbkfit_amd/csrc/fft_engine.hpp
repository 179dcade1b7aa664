// Device side of the register-tiled FFT engine, shared by the FFT passes
// and the fused batched pipeline.
#pragma once

#include "fft_plan.hpp"

namespace gbkhip {

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
// Odd prime radices 7, 11, 13: X[k] = sum_n v[n] W^(nk), W = exp(i sign
// 2 pi / P), with the inputs paired as (v[n] + v[P-n], v[n] - v[P-n]):
//   X[k], X[P-k] = v0 + sum_n cos(2 pi n k / P) b_n
//                    +- i sign sum_n sin(2 pi n k / P) d_n.
// cos / sin of 2 pi m / P for m = 1 .. (P-1)/2, in double, rounded once.
template <int P> struct OddPrime;
template <> struct OddPrime<7> {
    static constexpr float c[3] = {0.6234898018587336f, -0.22252093395631434f,
                                   -0.900968867902419f};
    static constexpr float s[3] = {0.7818314824680298f, 0.9749279121818236f,
                                   0.43388373911755823f};
};
template <> struct OddPrime<11> {
    static constexpr float c[5] = {0.8412535328311812f, 0.41541501300188644f,
                                   -0.142314838273285f, -0.654860733945285f,
                                   -0.9594929736144974f};
    static constexpr float s[5] = {0.5406408174555976f, 0.9096319953545183f,
                                   0.9898214418809328f, 0.7557495743542583f,
                                   0.28173255684142967f};
};
template <> struct OddPrime<13> {
    static constexpr float c[6] = {0.8854560256532099f, 0.5680647467311559f,
                                   0.120536680255323f, -0.35460488704253545f,
                                   -0.7485107481711012f, -0.970941817426052f};
    static constexpr float s[6] = {0.4647231720437685f, 0.8229838658936564f,
                                   0.992708874098054f, 0.9350162426854148f,
                                   0.6631226582407952f, 0.23931566428755768f};
};

template <int P> struct BflyOdd {
    __device__ static constexpr int perm(int q) { return q; }
    __device__ static void run(float2* v, float sign)
    {
        constexpr int H = (P - 1) / 2;
        float2 b[H], d[H];
        float2 s0 = v[0];
#pragma unroll
        for (int n = 1; n <= H; ++n) {
            b[n - 1] = cadd(v[n], v[P - n]);
            d[n - 1] = csub(v[n], v[P - n]);
            s0 = cadd(s0, b[n - 1]);
        }
        float2 out[P];
        out[0] = s0;
#pragma unroll
        for (int k = 1; k <= H; ++k) {
            float2 t = v[0];
            float ux = 0.0f, uy = 0.0f;
#pragma unroll
            for (int n = 1; n <= H; ++n) {
                const int m = (n * k) % P;
                const int mm = m <= H ? m : P - m;
                const float c = OddPrime<P>::c[mm - 1];
                const float sn = m <= H ? OddPrime<P>::s[mm - 1]
                                        : -OddPrime<P>::s[mm - 1];
                t.x = fmaf(c, b[n - 1].x, t.x);
                t.y = fmaf(c, b[n - 1].y, t.y);
                ux = fmaf(sn, d[n - 1].x, ux);
                uy = fmaf(sn, d[n - 1].y, uy);
            }
            const float2 u = make_float2(-sign * uy, sign * ux);
            out[k] = cadd(t, u);
            out[P - k] = csub(t, u);
        }
#pragma unroll
        for (int q = 0; q < P; ++q) v[q] = out[q];
    }
};
template <> struct Bfly<7> : BflyOdd<7> {};
template <> struct Bfly<11> : BflyOdd<11> {};
template <> struct Bfly<13> : BflyOdd<13> {};

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

// Elements per thread: a tile of n x L complex values is spread over the T
// threads of a block with n L <= kFastE T.
constexpr int kFastE = 16;

// ---- composite radices for compile-time plans --------------------------
// cos and sin of 2 pi m / R at compile time: reduce to the nearest quarter
// turn, Taylor series on |theta| <= pi / 4 (terms to x^25: error far below
// FP64 rounding), then rotate by the quarter turns.
constexpr double kTwoPi = 6.283185307179586476925286766559;

constexpr double cx_sin_small(double x)
{
    double term = x, sum = x;
    for (int i = 1; i <= 12; ++i) {
        term *= -x * x / ((2.0 * i) * (2.0 * i + 1.0));
        sum += term;
    }
    return sum;
}

constexpr double cx_cos_small(double x)
{
    double term = 1.0, sum = 1.0;
    for (int i = 1; i <= 12; ++i) {
        term *= -x * x / ((2.0 * i - 1.0) * (2.0 * i));
        sum += term;
    }
    return sum;
}

// (cos, sin)(2 pi m / R); m taken mod R
constexpr double cx_cos_turn(int m, int R)
{
    m = ((m % R) + R) % R;
    const int q = (4 * m + R / 2) / R;              // nearest quarter turn
    const double theta = kTwoPi * ((double)m / R - 0.25 * q);
    const double c = cx_cos_small(theta), s = cx_sin_small(theta);
    switch (q & 3) {
    case 0: return c;
    case 1: return -s;
    case 2: return -c;
    default: return s;
    }
}

constexpr double cx_sin_turn(int m, int R)
{
    m = ((m % R) + R) % R;
    const int q = (4 * m + R / 2) / R;
    const double theta = kTwoPi * ((double)m / R - 0.25 * q);
    const double c = cx_cos_small(theta), s = cx_sin_small(theta);
    switch (q & 3) {
    case 0: return s;
    case 1: return c;
    case 2: return -s;
    default: return -c;
    }
}

// v * exp(sign 2 pi i m / R) for constant m, R: exact quarter turns are
// swaps and negations, other angles one complex multiply.
template <int M, int R>
__device__ __forceinline__ float2 rot_const(float2 v, float sign)
{
    constexpr int m = ((M % R) + R) % R;
    if constexpr (m == 0) {
        return v;
    } else if constexpr (4 * m == R) {
        return mul_si(v, sign);
    } else if constexpr (2 * m == R) {
        return make_float2(-v.x, -v.y);
    } else if constexpr (4 * m == 3 * R) {
        return mul_si(v, -sign);
    } else {
        constexpr float c = (float)cx_cos_turn(m, R);
        constexpr float s = (float)cx_sin_turn(m, R);
        return cmul(v, make_float2(c, sign * s));
    }
}

// R = A B point DFT in registers: with n = n1 + A n2 and k = B k1 + k2,
// X[k] = sum_n1 W_A^(n1 k1) W_R^(n1 k2) sum_n2 v[n1 + A n2] W_B^(n2 k2).
// B-point DFTs over n2, twiddles, A-point DFTs over n1; X[B k1 + k2] ends
// in v[k1 + A k2].
template <int A, int B> struct Bfly2 {
    static constexpr int R = A * B;
    __device__ static constexpr int perm(int q) { return q / B + A * (q % B); }
    template <int N1, int K2>
    __device__ static __forceinline__ void twiddle(float2* v, float sign)
    {
        if constexpr (K2 < B) {
            v[N1 + A * K2] = rot_const<N1 * K2, R>(v[N1 + A * K2], sign);
            twiddle<N1, K2 + 1>(v, sign);
        }
    }
    template <int N1>
    __device__ static __forceinline__ void twiddles(float2* v, float sign)
    {
        if constexpr (N1 < A) {
            twiddle<N1, 1>(v, sign);
            twiddles<N1 + 1>(v, sign);
        }
    }
    __device__ static void run(float2* v, float sign);
};

template <int A, int B>
__device__ void Bfly2<A, B>::run(float2* v, float sign)
{
    float2 t[B > A ? B : A];
#pragma unroll
    for (int n1 = 0; n1 < A; ++n1) {
#pragma unroll
        for (int n2 = 0; n2 < B; ++n2) t[n2] = v[n1 + A * n2];
        Bfly<B>::run(t, sign);
#pragma unroll
        for (int k2 = 0; k2 < B; ++k2)
            v[n1 + A * k2] = t[Bfly<B>::perm(k2)];
    }
    twiddles<1>(v, sign);
#pragma unroll
    for (int k2 = 0; k2 < B; ++k2) {
#pragma unroll
        for (int n1 = 0; n1 < A; ++n1) t[n1] = v[n1 + A * k2];
        Bfly<A>::run(t, sign);
#pragma unroll
        for (int k1 = 0; k1 < A; ++k1)
            v[k1 + A * k2] = t[Bfly<A>::perm(k1)];
    }
}

#define GBK_COMPOSITE(R_, A_, B_)                                            \
    template <> struct Bfly<R_> : Bfly2<A_, B_> {};
GBK_COMPOSITE(6, 2, 3)
GBK_COMPOSITE(9, 3, 3)
GBK_COMPOSITE(10, 2, 5)
GBK_COMPOSITE(12, 4, 3)
GBK_COMPOSITE(15, 3, 5)
GBK_COMPOSITE(18, 2, 9)
GBK_COMPOSITE(20, 4, 5)
GBK_COMPOSITE(24, 8, 3)
GBK_COMPOSITE(25, 5, 5)
#undef GBK_COMPOSITE

// One Stockham stage of radix R.  Thread (l, j0) = (tid % L, tid / L)
// handles butterflies j = j0 + c JS (JS = T / L) of line l, so its LDS
// addresses differ from slot to slot by compile-time offsets and from
// element to element by uniform strides; k = j mod ns is carried from slot
// to slot instead of divided out.
template <int R, int LSH, int T>
__device__ __forceinline__ void fstage(float2* buf, int a0, int l, int j0,
                                       int nb, int ns, int twstep,
                                       unsigned nsm, const float2* tw,
                                       float sign)
{
    constexpr int LS = (1 << LSH) + 1;
    constexpr int JS = T >> LSH;
    constexpr int CAP = (kFastE + R - 1) / R;
    const int rstride = nb * LS;              // uniform
    const int wstride = ns * LS;              // uniform
    const int jm = JS - ns * (int)fdiv((unsigned)JS, nsm);   // JS mod ns
    const int jmt = jm * twstep, wrapt = ns * twstep;
    int k = j0 - ns * (int)fdiv((unsigned)j0, nsm);
    int kt = k * twstep;
    float2 v[CAP][R];
    int ks[CAP];
#pragma unroll
    for (int c = 0; c < CAP; ++c) {
        ks[c] = k;
        if (j0 + c * JS < nb) {
            const int ra = a0 + c * JS * LS;
#pragma unroll
            for (int t = 0; t < R; ++t) v[c][t] = buf[ra + t * rstride];
            if (k) {
#pragma unroll
                for (int t = 1; t < R; ++t)
                    v[c][t] = cmul(v[c][t], twid(tw, t * kt, sign));
            }
            Bfly<R>::run(v[c], sign);
        }
        k += jm;
        kt += jmt;
        if (k >= ns) {
            k -= ns;
            kt -= wrapt;
        }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < CAP; ++c) {
        const int j = j0 + c * JS;
        if (j < nb) {
            const int wa = ((j - ks[c]) * R + ks[c]) * LS + l;
#pragma unroll
            for (int q = 0; q < R; ++q)
                buf[wa + q * wstride] = v[c][Bfly<R>::perm(q)];
        }
    }
    __syncthreads();
}

// Whole transform of the tile [m][L+1] (complete on entry and on return).
template <int LSH, int T>
__device__ void fast_fft(float2* buf, const FAxis& a, const float2* tw,
                         float sign)
{
    constexpr int LS = (1 << LSH) + 1;
    const int l = threadIdx.x & ((1 << LSH) - 1);
    const int j0 = threadIdx.x >> LSH;
    const int a0 = j0 * LS + l;
#pragma unroll 1
    for (int s = 0; s < a.nst; ++s) {
        const int nb = a.nb[s], ns = a.ns[s], tws = a.tws[s];
        const unsigned nsm = a.nsm[s];
        switch (a.R[s]) {
        case 2: fstage<2, LSH, T>(buf, a0, l, j0, nb, ns, tws, nsm, tw, sign); break;
        case 3: fstage<3, LSH, T>(buf, a0, l, j0, nb, ns, tws, nsm, tw, sign); break;
        case 4: fstage<4, LSH, T>(buf, a0, l, j0, nb, ns, tws, nsm, tw, sign); break;
        case 5: fstage<5, LSH, T>(buf, a0, l, j0, nb, ns, tws, nsm, tw, sign); break;
        case 8: fstage<8, LSH, T>(buf, a0, l, j0, nb, ns, tws, nsm, tw, sign); break;
        case 7: fstage<7, LSH, T>(buf, a0, l, j0, nb, ns, tws, nsm, tw, sign); break;
        case 11: fstage<11, LSH, T>(buf, a0, l, j0, nb, ns, tws, nsm, tw, sign); break;
        case 13: fstage<13, LSH, T>(buf, a0, l, j0, nb, ns, tws, nsm, tw, sign); break;
        default: fstage<16, LSH, T>(buf, a0, l, j0, nb, ns, tws, nsm, tw, sign); break;
        }
    }
}

// ---------------------------------------------------------------------------
// Compile-time plans: for the lengths in the specialisation lists below the
// radices, strides and slot counts are constants, LDS offsets become
// instruction immediates and each stage reads its twiddles from a table
// laid out [k][t - 1], tw_s[k][t - 1] = exp(-2 pi i t k / (NS_s R_s)).
// ---------------------------------------------------------------------------
struct CPlan {
    int nst = 0;
    int R[kFMax] = {};
    int NS[kFMax] = {};
    int tw_off[kFMax] = {};
    int tw_total = 0;
};

// Radices of compile-time plans: the fewest stages with radices from
// kCRadices (composites run as two-level butterflies in registers), ties
// broken by the smallest largest radix; stages in descending radix order.
constexpr int kCRadices[] = {25, 24, 20, 18, 16, 15, 12, 10, 9, 8, 6, 5, 4,
                             3, 2};
constexpr int kCPlanMax = 4096;

__host__ __device__ constexpr CPlan cplan(int n)
{
    CPlan P;
    if (n < 2 || n > kCPlanMax) return P;
    int st[kCPlanMax + 1] = {}, mx[kCPlanMax + 1] = {}, ch[kCPlanMax + 1] = {};
    for (int m = 2; m <= n; ++m) st[m] = 1 << 20;
    for (int m = 2; m <= n; ++m) {
        if (n % m) continue;
        for (int r : kCRadices) {
            if (m % r || st[m / r] >= (1 << 20)) continue;
            const int sc = st[m / r] + 1;
            const int xc = mx[m / r] > r ? mx[m / r] : r;
            if (sc < st[m] || (sc == st[m] && xc < mx[m])) {
                st[m] = sc;
                mx[m] = xc;
                ch[m] = r;
            }
        }
    }
    if (st[n] >= (1 << 20) || st[n] > kFMax) return P;
    for (int m = n; m > 1; m /= ch[m]) P.R[P.nst++] = ch[m];
    for (int i = 0; i < P.nst; ++i)          // descending radix order
        for (int j = i + 1; j < P.nst; ++j)
            if (P.R[j] > P.R[i]) {
                const int t = P.R[i];
                P.R[i] = P.R[j];
                P.R[j] = t;
            }
    int ns = 1, off = 0;
    for (int s = 0; s < P.nst; ++s) {
        P.NS[s] = ns;
        P.tw_off[s] = off;
        off += ns * (P.R[s] - 1);
        ns *= P.R[s];
    }
    P.tw_total = off;
    return P;
}

// The plan of length N as a compile-time constant (no object at run time).
template <int N> struct CP {
    static constexpr CPlan P = cplan(N);
};

template <int N, int LSH, int T, int S>
__device__ __forceinline__ void cstage(float2* buf, int a0, int j0,
                                       const float2* twst, float sign)
{
    constexpr int R = CP<N>::P.R[S], NS = CP<N>::P.NS[S], NB = N / R;
    constexpr int LS = (1 << LSH) + 1, JS = T >> LSH;
    constexpr int CAP = (NB + JS - 1) / JS;
    constexpr int JM = JS % NS;
    constexpr int TO = CP<N>::P.tw_off[S];
    int k = NS == 1 ? 0 : j0 % NS;
    float2 v[CAP][R];
    int ks[CAP];
#pragma unroll
    for (int c = 0; c < CAP; ++c) {
        ks[c] = k;
        if ((c + 1) * JS <= NB || j0 + c * JS < NB) {
#pragma unroll
            for (int t = 0; t < R; ++t)
                v[c][t] = buf[a0 + (c * JS + t * NB) * LS];
            if (NS > 1) {
                const float2* w = twst + TO + k * (R - 1);
#pragma unroll
                for (int t = 1; t < R; ++t) {
                    const float2 wt = w[t - 1];
                    v[c][t] = cmul(v[c][t], sign < 0 ? wt : conjf2(wt));
                }
            }
            Bfly<R>::run(v[c], sign);
        }
        if (NS > 1) {
            k += JM;
            if (k >= NS) k -= NS;
        }
    }
    __syncthreads();
    // write index (j - k) R + k with j = j0 + c JS
    const int w0 = a0 + j0 * (R - 1) * LS;
#pragma unroll
    for (int c = 0; c < CAP; ++c) {
        if ((c + 1) * JS <= NB || j0 + c * JS < NB) {
            const int wa = w0 + c * JS * R * LS - ks[c] * (R - 1) * LS;
#pragma unroll
            for (int q = 0; q < R; ++q)
                buf[wa + q * NS * LS] = v[c][Bfly<R>::perm(q)];
        }
    }
    __syncthreads();
}

template <int N, int LSH, int T, int S = 0>
__device__ __forceinline__ void cfft_stages(float2* buf, int a0, int j0,
                                            const float2* twst, float sign)
{
    if constexpr (S < CP<N>::P.nst) {
        cstage<N, LSH, T, S>(buf, a0, j0, twst, sign);
        cfft_stages<N, LSH, T, S + 1>(buf, a0, j0, twst, sign);
    }
}

// Whole transform of the tile: the runtime engine (N = 0, twiddles `tw` =
// exp(-2 pi i m / n)) or the compile-time one (twiddles = stage table).
template <int LSH, int T, int N>
__device__ __forceinline__ void tile_fft(float2* buf, const FAxis& a,
                                         const float2* tw, float sign)
{
    if constexpr (N == 0) {
        fast_fft<LSH, T>(buf, a, tw, sign);
    } else {
        constexpr int LS = (1 << LSH) + 1;
        const int l = threadIdx.x & ((1 << LSH) - 1);
        const int j0 = threadIdx.x >> LSH;
        cfft_stages<N, LSH, T>(buf, j0 * LS + l, j0, tw, sign);
    }
}

// Twiddle entries a kernel stages in LDS (n = N when N > 0).
template <int N>
__device__ __forceinline__ constexpr int tw_entries(int n)
{
    if constexpr (N == 0) return n;
    else return CP<N>::P.tw_total > N ? CP<N>::P.tw_total : N;
}

// Specialised shapes (n, lshift, threads) per kernel family: the sizes of
// the benchmark workloads (fused pipeline 300^2 x 120, padded cubes
// 320^2 x 128 and 192^2 x 96).
#define GBK_SPEC_LINES(X, ...)                                               \
    X(300, 4, 512, __VA_ARGS__) X(320, 4, 512, __VA_ARGS__)                  \
    X(128, 5, 256, __VA_ARGS__) X(192, 4, 256, __VA_ARGS__)                  \
    X(96, 5, 256, __VA_ARGS__)
#define GBK_SPEC_ROWS(X, ...)                                                \
    X(160, 4, 256, __VA_ARGS__) X(96, 5, 256, __VA_ARGS__)                   \
    X(300, 3, 256, __VA_ARGS__)
#define GBK_SPEC_Z(X, ...) X(60, 6, 256, __VA_ARGS__)

#define GBK_SPEC_LAUNCH(N_, L_, T_, SHAPE, KERNEL, GRID, LDS, STREAM, ...)   \
    if ((SHAPE).a.n == (N_) && (SHAPE).lshift == (L_) &&                     \
        (SHAPE).threads == (T_))                                             \
        hipLaunchKernelGGL((KERNEL<L_, T_, N_>), GRID, dim3(T_), LDS,        \
                           STREAM, __VA_ARGS__);                             \
    else
#define GBK_SPEC_MATCH(N_, L_, T_, SHAPE)                                    \
    || ((SHAPE).a.n == (N_) && (SHAPE).lshift == (L_) &&                     \
        (SHAPE).threads == (T_))

// Launch of KERNEL<LSH, T> for the shape's runtime (lshift, threads).
#define GBK_FAST_CASE(KERNEL, L_, T_, ...)                                   \
    case (L_) * 2 + ((T_) == 512):                                           \
        hipLaunchKernelGGL((KERNEL<L_, T_, 0>), __VA_ARGS__);                \
        break;
#define GBK_FAST_DISPATCH(SHAPE, KERNEL, GRID, LDS, STREAM, ...)             \
    switch ((SHAPE).lshift * 2 + ((SHAPE).threads == 512)) {                 \
    GBK_FAST_CASE(KERNEL, 0, 256, GRID, dim3(256), LDS, STREAM, __VA_ARGS__) \
    GBK_FAST_CASE(KERNEL, 1, 256, GRID, dim3(256), LDS, STREAM, __VA_ARGS__) \
    GBK_FAST_CASE(KERNEL, 2, 256, GRID, dim3(256), LDS, STREAM, __VA_ARGS__) \
    GBK_FAST_CASE(KERNEL, 3, 256, GRID, dim3(256), LDS, STREAM, __VA_ARGS__) \
    GBK_FAST_CASE(KERNEL, 4, 256, GRID, dim3(256), LDS, STREAM, __VA_ARGS__) \
    GBK_FAST_CASE(KERNEL, 5, 256, GRID, dim3(256), LDS, STREAM, __VA_ARGS__) \
    GBK_FAST_CASE(KERNEL, 6, 256, GRID, dim3(256), LDS, STREAM, __VA_ARGS__) \
    GBK_FAST_CASE(KERNEL, 2, 512, GRID, dim3(512), LDS, STREAM, __VA_ARGS__) \
    GBK_FAST_CASE(KERNEL, 3, 512, GRID, dim3(512), LDS, STREAM, __VA_ARGS__) \
    GBK_FAST_CASE(KERNEL, 4, 512, GRID, dim3(512), LDS, STREAM, __VA_ARGS__) \
    GBK_FAST_CASE(KERNEL, 5, 512, GRID, dim3(512), LDS, STREAM, __VA_ARGS__) \
    GBK_FAST_CASE(KERNEL, 6, 512, GRID, dim3(512), LDS, STREAM, __VA_ARGS__) \
    default:                                                                 \
        throw Error{GBKHIP_ERR_INTERNAL, "no kernel for this FFT shape"};    \
    }

// Every (lshift, threads) pair the dispatch above instantiates.
#define GBK_FAST_KERNELS(KERNEL)                                             \
    (const void*)KERNEL<0, 256, 0>, (const void*)KERNEL<1, 256, 0>,          \
    (const void*)KERNEL<2, 256, 0>, (const void*)KERNEL<3, 256, 0>,          \
    (const void*)KERNEL<4, 256, 0>, (const void*)KERNEL<5, 256, 0>,          \
    (const void*)KERNEL<6, 256, 0>, (const void*)KERNEL<2, 512, 0>,          \
    (const void*)KERNEL<3, 512, 0>, (const void*)KERNEL<4, 512, 0>,          \
    (const void*)KERNEL<5, 512, 0>, (const void*)KERNEL<6, 512, 0>

// Specialised launch when (n, lshift, threads) is on LIST, else the runtime
// engine.
#define GBK_DISPATCH(LIST, SHAPE, KERNEL, GRID, LDS, STREAM, ...)            \
    LIST(GBK_SPEC_LAUNCH, SHAPE, KERNEL, GRID, LDS, STREAM, __VA_ARGS__)     \
    { GBK_FAST_DISPATCH(SHAPE, KERNEL, GRID, LDS, STREAM, __VA_ARGS__) }

}  // namespace gbkhip
