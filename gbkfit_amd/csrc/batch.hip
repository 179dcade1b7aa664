// Batched evaluation: many parameter sets of one galaxy model per call.
//
// The reference evaluates one parameter set per Objective.log_likelihood
// call (src/gbkfit/fitting/fitters/emcee/sampler.py:34-57 ->
// objective.py:208-226).  A batch here shares the geometry, the PSF x LSF
// spectrum and the observed cube, and runs for each chunk of sets:
//   1. smdisk into the padded hi-res cube of every set (columns written in
//      full, which fuses the cube's zero fill);
//   2. the FFT convolution of all cubes against the one cached spectrum;
//   3. one fused pass: block-mean downscale, residual against the data and
//      FP64 partial sums of r^2, reduced per set in a fixed order (chi^2
//      bit-identical from run to run and independent of batch composition);
// optionally also writing the low-res model cubes.  The launch sequence of
// a given (chunk size, pointers) is captured once in a HIP graph and
// replayed.
#include "fft_plan.hpp"
#include "fft_engine.hpp"
#include "launch.hpp"
#include "disk.hpp"
#include "mcdisk.hpp"
#include "thick.hpp"
#include "sepfilter.hpp"
#include "moments.hpp"

#include <climits>
#include <cmath>
#include <cstring>

namespace gbkhip {

// Offsets of the per-set real parameters inside one set's parameter vector.
struct SetLayout {
    int vsys, xpos, ypos, posa, incl;      // -1: absent (vsys only)
    int cvalues[9], pvalues[9];            // per trait slot, -1: absent
};

__device__ __forceinline__ gbkhip_disk set_disk(const gbkhip_disk& t,
                                                const SetLayout& L,
                                                const float* p)
{
    gbkhip_disk d = t;
    d.vsys = L.vsys >= 0 ? p + L.vsys : nullptr;
    d.xpos = p + L.xpos;
    d.ypos = p + L.ypos;
    d.posa = p + L.posa;
    d.incl = p + L.incl;
    // one statement per slot: no member addresses in a runtime-indexed
    // array, so the struct stays in registers
    auto bind = [&](gbkhip_traits& t, int k) {
        if (t.uids) {
            t.cvalues = p + L.cvalues[k];
            t.pvalues = p + L.pvalues[k];
        }
    };
    bind(d.rp, 0);
    bind(d.rh, 1);
    bind(d.vp, 2);
    bind(d.vh, 3);
    bind(d.dp, 4);
    bind(d.dh, 5);
    bind(d.zp, 6);
    bind(d.sp, 7);
    bind(d.wp, 8);
    return d;
}

// Thin disk, whole columns: every channel of every spaxel of the padded
// cube is written (zero outside the line window).
__global__ __launch_bounds__(256) void batch_smdisk_thin(
    gbkhip_disk tmpl, SetLayout lay, const float* __restrict__ params,
    int pstride, gbkhip_grid g, float* __restrict__ cubes, long long cstride)
{
    const int lane = threadIdx.x;
    const int x = blockIdx.x * 64 + lane;
    const int y = blockIdx.y * 4 + threadIdx.y;
    const int set = blockIdx.z;
    const int nx = g.spat_size[0], ny = g.spat_size[1];
    if (y >= ny || x >= nx) return;
    const gbkhip_disk d = set_disk(tmpl, lay, params + (size_t)set * pstride);
    const bool need_theta = disk_needs_theta(d);
    const NodeGeom ng = node_geom(d, 0);
    const VoxelLine L = smdisk_voxel(d, g, ng, x, y, 0, need_theta, false);
    int k0 = 1, k1 = 0;
    LineEval e;
    if (L.on) {
        line_window(g, L.v, L.s, k0, k1);
        e = line_eval(g, L.orvalue, L.v, L.s);
    }
    float* col = cubes + set * cstride + (size_t)y * nx + x;
    const size_t plane = (size_t)nx * ny;
    for (int k = 0; k < g.spec_size; ++k)
        col[k * plane] = (k >= k0 && k <= k1) ? e.at(k) : 0.0f;
}

// Thick disk: 4 waves split the line of sight of 64 columns, every sum in
// a fixed order (thick.hpp).  Only columns x < nxe and channels k < nze
// are evaluated and written (zero where no line lands), on the rows the
// grid launches (the direct pipeline reads nothing beyond its inputs; the
// padded-cube pipeline passes the whole grid).
template <int J>
__global__ __launch_bounds__(256) void batch_smdisk_thick(
    gbkhip_disk tmpl, SetLayout lay, const float* __restrict__ params,
    int pstride, gbkhip_grid g, float* __restrict__ cubes, long long cstride,
    int nxe, int nze, int assign)
{
    __shared__ ThickLds S;
    const int set = blockIdx.z;
    const gbkhip_disk d = set_disk(tmpl, lay, params + (size_t)set * pstride);
    const bool need_theta = disk_needs_theta(d);
    const NodeGeom ng = node_geom(d, 0);
    gbkhip_outputs o;
    memset(&o, 0, sizeof(o));
    thick_row_segment<J, false>(d, g, ng, need_theta, blockIdx.y, nxe, nze,
                                cubes + set * cstride, assign != 0, o, S);
}

struct DataArgs {
    // (nz, ny, nx) low-res observed cube; nullptr: no cube residual (the
    // moment-map objective takes the model cubes instead)
    const float* obs;
    const float* err;     // same shape or nullptr
    const float* mask;    // same shape or nullptr
    float err_scalar;     // used when err == nullptr (1 = none)
    float inv_err_scalar; // 1 / err_scalar
    float weight;
};

// Downscale + residual + FP64 partial sums of r^2 (one partial per block).
__global__ __launch_bounds__(256) void batch_downscale_chi2(
    const float* __restrict__ hi, long long hstride, int hnx, int hny,
    int sx, int sy, int sz, int ox, int oy, int oz, int nx, int ny, int nz,
    DataArgs da, float* __restrict__ lo_out, long long lstride,
    double* __restrict__ partials, int nparts)
{
    __shared__ double sh[4];
    const int set = blockIdx.y;
    const float* src = hi + set * hstride;
    const long long n = (long long)nx * ny * nz;
    const float nf = 1.0f / (float)(sx * sy * sz);
    double acc = 0.0;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const int x = (int)(i % nx);
        const int y = (int)((i / nx) % ny);
        const int z = (int)(i / ((long long)nx * ny));
        const int bx = ox + x * sx, by = oy + y * sy, bz = oz + z * sz;
        float sum = 0.0f;
        for (int dz = 0; dz < sz; ++dz)
            for (int dy = 0; dy < sy; ++dy)
                for (int dx = 0; dx < sx; ++dx)
                    sum += src[(bx + dx) + (size_t)(by + dy) * hnx +
                               (size_t)(bz + dz) * hnx * hny];
        const float m = sum * nf;
        if (lo_out) lo_out[set * lstride + i] = m;
        if (!da.obs) continue;
        // objective.hpp:7-33 with mdl_w = mdl_m = none
        float r = m - da.obs[i];
        r /= da.err ? da.err[i] : da.err_scalar;
        if (da.mask) r *= da.mask[i];
        r *= da.weight;
        acc += (double)r * (double)r;
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0)
        partials[(size_t)set * nparts + blockIdx.x] =
            sh[0] + sh[1] + sh[2] + sh[3];
}

__global__ void batch_chi2_final(const double* __restrict__ partials,
                                 int nparts, double* __restrict__ out,
                                 int set0)
{
    __shared__ double sh[4];
    const int set = blockIdx.x;
    double acc = 0.0;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x)
        acc += partials[(size_t)set * nparts + i];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) out[set0 + set] = sh[0] + sh[1] + sh[2] + sh[3];
}

// Moment-map objective (mmaps.py:136-201 through objective.py:279-308):
// per spaxel the requested moments of the set's low-res model cube
// (moments.hpp, the plain kernel's arithmetic) and, per order, the residual
// against the data map (objective.hpp:7-33 with the model's mask = the
// spaxel's validity and unit model weights), FP64 partial sums of r^2, one
// per block.  A zero mask entry (data or model) gives a zero residual, so
// spaxels without model flux, whose moments are NaN, add nothing.
struct MMapArgs {
    int norders;
    int ord[kMaxOrders];
    float cutoff, zero_z, step_z;
    const float* obs;      // (norders, ny, nx) data maps
    const float* err;      // same shape or nullptr
    const float* mask;     // same shape or nullptr
    float err_scalar[kMaxOrders];
    float weight[kMaxOrders];
};

__global__ __launch_bounds__(256) void batch_mmaps_chi2(
    const float* __restrict__ lo, long long lstride, int nx, int ny, int nz,
    MMapArgs ma, double* __restrict__ partials, int nparts)
{
    __shared__ double sh[4];
    const int set = blockIdx.y;
    const long long plane = (long long)nx * ny;
    const long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    double acc = 0.0;
    if (c < plane) {
        float md[kMaxOrders], m0;
        const bool valid = spaxel_moments(lo + set * lstride + c, plane, nz,
                                          ma.step_z, ma.zero_z, ma.cutoff,
                                          ma.norders, ma.ord, md, m0);
        for (int i = 0; i < ma.norders; ++i) {
            const long long gi = i * plane + c;
            float r = md[i] - ma.obs[gi];
            r /= ma.err ? ma.err[gi] : ma.err_scalar[i];
            if (ma.mask) r = ma.mask[gi] != 0.0f ? r * ma.mask[gi] : 0.0f;
            r = valid ? r : 0.0f;
            r *= ma.weight[i];
            acc += (double)r * (double)r;
        }
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0)
        partials[(size_t)set * nparts + blockIdx.x] =
            sh[0] + sh[1] + sh[2] + sh[3];
}

// ---------------------------------------------------------------------------
// Spectral-first fused pipeline (thin disk, separable PSF x LSF kernel)
//
// The convolution is a product of three 1-D transforms, so their order is
// free.  Taking the spectral axis first lets each spaxel's line go straight
// into the z transform, and the inverse z transform
// hand the convolved line straight to downscale + chi^2: the hi-res cube
// never reaches HBM.  Per chunk of sets, with S the (nz/2+1, ny, nx)
// complex work spectrum of each set:
//   K0 batch_lines_thin  (flux, v, sigma) of each spaxel's line
//   K1 fz_deposit_r2c    lines of 2^lshift spaxels into an LDS tile, real
//                        FFT along z, half spectrum -> S
//   K2 fx_lines (y)      forward along y
//   K3 fx_rows_conv      forward along x, times Lhat[kz] Phat[ky, kx] (the
//                        1/N scale folded in), inverse along x; only the x
//                        range read by the downscale is stored
//   K4 fx_lines (y)      inverse along y over those x; only the needed y
//                        range is stored
//   K5 fz_c2r_down_chi2  inverse real FFT along z of a tile of spaxel
//                        columns, block-mean downscale, residual, FP64
//                        partial sum of r^2 per tile
// Inputs beyond S_i + 2 edge_i along each axis (S_i the supersampled data
// size) reach no stored output through a kernel of half-width edge_i, so
// they are zero and the FFT sizes are the smallest 5-smooth ones above it.
// ---------------------------------------------------------------------------
__global__ void sep_zero_masks(unsigned long long* m, long long n,
                               unsigned long long* m2)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
         i < n; i += (long long)gridDim.x * blockDim.x) {
        m[i] = 0ull;
        if (m2) m2[i] = 0ull;
    }
}

// Line of sight of each spaxel (thin disk): the line it deposits, flux 0
// when it deposits none or lies beyond the needed region.  Form 0: (flux, v,
// sigma, 0).  Form 1, ready for evaluation: (amp, v, 1 / (2 sigma^2),
// k0 | k1 << 16) with the LineEval amplitude and the rint channel window
// (amp 0 and an empty window when there is no line).
__global__ __launch_bounds__(256) void batch_lines_thin(
    gbkhip_disk tmpl, SetLayout lay, const float* __restrict__ params,
    int pstride, gbkhip_grid g, int need_x, int need_y, int nx, int ny,
    float4* __restrict__ lines, long long lstride, int form,
    unsigned long long* __restrict__ rowmask, int zc, int need_z,
    unsigned long long* __restrict__ segmask)
{
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int y = blockIdx.y * 4 + threadIdx.y;
    const int set = blockIdx.z;
    const gbkhip_disk d = set_disk(tmpl, lay, params + (size_t)set * pstride);
    // the set's node-0 geometry (two sincos) once per block, and whether
    // the block's 64 x 4 spaxels all lie beyond the disk edge: a spaxel's
    // deprojected radius is at least its sky distance from the centre
    // (the thin deprojection only stretches y), so a block whose nearest
    // point is farther than the last node (with a margin far above FP32
    // rounding) holds no line (flat disks; loose / tilted ones evaluate
    // every spaxel)
    __shared__ NodeGeom s_ng;
    __shared__ int s_cull;
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        const NodeGeom ng = node_geom(d, 0);
        s_ng = ng;
        int cull = 0;
        if (!d.loose && !d.tilted) {
            const float rmax = d.rnodes[d.nrnodes - 1] * 1.0001f + 1e-6f;
            auto gap = [](float a, float b, float c) {
                const float lo = fminf(a, b), hi = fmaxf(a, b);
                return fmaxf(fmaxf(lo - c, c - hi), 0.0f);
            };
            const int xa = blockIdx.x * 64, ya = blockIdx.y * 4;
            const float dx = gap(g.spat_zero[0] + xa * g.spat_step[0],
                                 g.spat_zero[0] + (xa + 63) * g.spat_step[0],
                                 ng.xpos);
            const float dy = gap(g.spat_zero[1] + ya * g.spat_step[1],
                                 g.spat_zero[1] + (ya + 3) * g.spat_step[1],
                                 ng.ypos);
            cull = dx * dx + dy * dy > rmax * rmax;
        }
        s_cull = cull;
    }
    __syncthreads();
    if (y >= ny) return;               // whole waves: a wave is one row
    if (s_cull && form == 1 && segmask && blockIdx.x < 63) return;
    float4 out = make_float4(0, 0, 1, 0);
    if (!s_cull && x < nx && x < need_x && y < need_y) {
        const bool need_theta = disk_needs_theta(d);
        const NodeGeom ng = s_ng;
        const VoxelLine V = smdisk_voxel(d, g, ng, x, y, 0, need_theta, false);
        if (V.on) out = make_float4(V.orvalue, V.v, V.s, 0);
    }
    if (form == 1) {
        int k0 = 1, k1 = 0;
        float amp = 0.0f, inv2s2 = 0.0f;
        if (out.x != 0.0f) {
            line_window(g, out.y, out.z, k0, k1);
            const LineEval e = line_eval(g, out.x, out.y, out.z);
            amp = e.amp;
            inv2s2 = e.inv2s2;
            if (k0 > k1) {          // empty window
                k0 = 1;
                k1 = 0;
            }
        }
        out = make_float4(amp, out.y, inv2s2,
                          __uint_as_float((unsigned)k0 |
                                          ((unsigned)k1 << 16)));
        // channel chunks of the x pass this spaxel's line reaches, OR-ed
        // over the row (the wave) into the row's mask
        unsigned long long bits = 0;
        const int k1c = min(k1, need_z - 1);
        if (amp != 0.0f && k0 <= k1c) {
            const int c0 = k0 / zc, c1 = k1c / zc;
            bits = (c1 - c0 + 1 >= 64 ? ~0ull
                                      : ((1ull << (c1 - c0 + 1)) - 1ull))
                   << c0;
        }
        for (int o = 32; o > 0; o >>= 1) bits |= __shfl_xor(bits, o, 64);
        if (threadIdx.x == 0 && bits) {
            atomicOr(&rowmask[(size_t)set * ny + y], bits);
            // the row's 64-spaxel segment (this wave) carries a line
            if (segmask)
                atomicOr(&segmask[(size_t)set * ny + y],
                         1ull << min((int)blockIdx.x, 63));
        }
        // a segment below the 63rd with no line reaching the x pass's
        // channels is not written: the x pass reads only live segments
        // (and every segment from the 63rd on) and takes zeros for the rest
        if (segmask && !bits && blockIdx.x < 63) return;
    }
    if (x < nx) lines[set * lstride + (size_t)y * nx + x] = out;
}

// Spectral lines of L spaxels into an LDS tile, real FFT along z, half
// spectrum -> S.
template <int LSH, int T, int N = 0>
__global__ __launch_bounds__(T) void fz_deposit_r2c(
    const float4* __restrict__ lines, long long lstride, gbkhip_grid g,
    FAxis a, const float2* __restrict__ tw,
    const float2* __restrict__ tws, const float2* __restrict__ tpz,
    int need_z, int nx, int ny, int px, int shift, float2* __restrict__ S,
    long long sstride)
{
    extern __shared__ float2 lds[];
    constexpr int L = 1 << LSH, LS = L + 1;
    const int n = N ? N : a.n;
    float2* s_tw = lds;
    float2* buf = lds + tw_entries<N>(n);
    float2* s_post = buf + n * LS;      // tpz staged in LDS
    const int x0 = blockIdx.x << LSH, y = blockIdx.y, set = blockIdx.z;
    for (int i = threadIdx.x; i < tw_entries<N>(n); i += T)
        s_tw[i] = N ? tws[i] : tw[i];
    for (int i = threadIdx.x; i <= n; i += T) s_post[i] = tpz[i];
    for (int i = threadIdx.x; i < n * LS; i += T) buf[i] = make_float2(0, 0);
    __syncthreads();
    // spaxel l of the tile; the T / L threads of a spaxel split its channels
    const int l = threadIdx.x & (L - 1);
    const int part = threadIdx.x >> LSH;
    constexpr int nparts = T >> LSH;
    const int c = x0 + l;                // FFT column
    if (c < nx) {
        const int x = c >= shift ? c - shift : c - shift + nx;   // spaxel
        const float4 V = lines[set * lstride + (size_t)y * nx + x];
        if (V.x != 0.0f) {
            int k0, k1;
            line_window(g, V.y, V.z, k0, k1);
            k1 = min(k1, need_z - 1);
            const LineEval e = line_eval(g, V.x, V.y, V.z);
            float* fb = reinterpret_cast<float*>(buf);
            for (int k = k0 + part; k <= k1; k += nparts)
                fb[2 * ((k >> 1) * LS + l) + (k & 1)] = e.at(k);
        }
    }
    __syncthreads();
    // z[m] = x[2m] + i x[2m+1]; complex FFT of length n = nz/2, then
    // X[k] = (Z[k] + conj Z[n-k]) / 2 + W^k (Z[k] - conj Z[n-k]) / (2i)
    tile_fft<LSH, T, N>(buf, a, s_tw, -1.0f);
    float2* out = S + set * sstride + (size_t)y * px + x0;
    const size_t plane = (size_t)px * ny;
    const int cnt = (n + 1) << LSH;
    for (int i = threadIdx.x; i < cnt; i += T) {
        const int ll = i & (L - 1), k = i >> LSH;
        if (x0 + ll >= nx) continue;
        const float2 zk = buf[(k == n ? 0 : k) * LS + ll];
        const float2 zm = conjf2(buf[(k == 0 ? 0 : n - k) * LS + ll]);
        const float2 fe = make_float2(0.5f * (zk.x + zm.x),
                                      0.5f * (zk.y + zm.y));
        const float2 dd = csub(zk, zm);
        const float2 fo = make_float2(0.5f * dd.y, -0.5f * dd.x);
        out[k * plane + ll] = cadd(fe, cmul(s_post[k], fo));
    }
}

// Rows of S along x: forward FFT, kernel spectrum, inverse FFT.  Row r is
// plane kz = r / ny, line ky = r % ny; element i of the block's contiguous
// tile is (row i / n, x = i % n).
template <int LSH, int T, int N = 0>
__global__ __launch_bounds__(T) void fx_rows_conv(
    float2* __restrict__ S, long long sstride, int nrows, int ny,
    unsigned ny_m, FAxis a, unsigned n_m, const float2* __restrict__ tw,
    const float2* __restrict__ tws,
    const float2* __restrict__ phat, const float2* __restrict__ lhat,
    int x_lo, int x_cnt, unsigned xcnt_m, int px)
{
    extern __shared__ float2 lds[];
    constexpr int L = 1 << LSH, LS = L + 1;
    const int n = N ? N : a.n;
    float2* s_tw = lds;
    float2* buf = lds + tw_entries<N>(n);
    const int r0 = blockIdx.x << LSH;
    const int nl = min(L, nrows - r0);
    float2* rows = S + blockIdx.y * sstride + (size_t)r0 * px;
    for (int i = threadIdx.x; i < tw_entries<N>(n); i += T)
        s_tw[i] = N ? tws[i] : tw[i];
    const int tile = n << LSH, valid = n * nl;
    float2 r[kFastE];
#pragma unroll
    for (int c = 0; c < kFastE; ++c) {
        const int i = threadIdx.x + c * T;
        const int l = (int)fdiv((unsigned)i, n_m), m = i - l * n;
        r[c] = i < valid ? rows[(size_t)l * px + m] : make_float2(0, 0);
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
#pragma unroll
    for (int c = 0; c < kFastE; ++c) {
        const int i = threadIdx.x + c * T;
        r[c] = make_float2(0, 0);
        if (i < valid) {
            const int l = (int)fdiv((unsigned)i, n_m), m = i - l * n;
            const int row = r0 + l;
            const int kz = (int)fdiv((unsigned)row, ny_m);
            r[c] = cmul(lhat[kz], phat[(size_t)(row - kz * ny) * n + m]);
        }
    }
#pragma unroll
    for (int c = 0; c < kFastE; ++c) {
        const int i = threadIdx.x + c * T;
        if (i < valid) {
            const int l = (int)fdiv((unsigned)i, n_m), m = i - l * n;
            buf[m * LS + l] = cmul(buf[m * LS + l], r[c]);
        }
    }
    __syncthreads();
    tile_fft<LSH, T, N>(buf, a, s_tw, 1.0f);
    const int cnt = nl * x_cnt;
    for (int i = threadIdx.x; i < cnt; i += T) {
        const int l = (int)fdiv((unsigned)i, xcnt_m), m = x_lo + i - l * x_cnt;
        rows[(size_t)l * px + m] = buf[m * LS + l];
    }
}

// Inverse real FFT along z of a tile of spaxel columns, then block-mean
// downscale, residual and FP64 partial sum of r^2.  The tile is kx x ky
// low-res spaxels = tx x ty hi-res columns (tx = kx sx, ty = ky sy), column
// (lx, ly) on line l = ly * tx + lx.
template <int LSH, int T, int N = 0>
__global__ __launch_bounds__(T) void fz_c2r_down_chi2(
    const float2* __restrict__ S, long long sstride, int px, int ny, FAxis a,
    const float2* __restrict__ tw,
    const float2* __restrict__ tws, const float2* __restrict__ tpz, int x_lo,
    int y_lo, int kx, int ky, int sx, int sy, int sz, int ez, int nxl,
    int nyl, int nzl, int tiles_x, unsigned pz_m, unsigned kx_m, DataArgs da,
    float* __restrict__ lo_out, long long lstride,
    double* __restrict__ partials, int nparts)
{
    extern __shared__ float2 lds[];
    __shared__ double red[T / 64];
    constexpr int L = 1 << LSH, LS = L + 1;
    const int n = N ? N : a.n;
    const int tx = kx * sx, ty = ky * sy;
    float2* s_tw = lds;
    float2* buf = lds + tw_entries<N>(n);
    float2* s_post = buf + n * LS;      // tpz staged in LDS
    const int set = blockIdx.y;
    const int X0 = (blockIdx.x % tiles_x) * kx, Y0 = (blockIdx.x / tiles_x) * ky;
    const float2* src = S + set * sstride;
    const size_t plane = (size_t)px * ny;
    for (int i = threadIdx.x; i < tw_entries<N>(n); i += T)
        s_tw[i] = N ? tws[i] : tw[i];
    for (int i = threadIdx.x; i <= n; i += T) s_post[i] = tpz[i];
    __syncthreads();
    // pairs (k, n - k), k <= n/2, each plane read once:
    //   Z[k] = (X[k] + conj X[n-k]) + i W^-k (X[k] - conj X[n-k]), k < n
    const int h = n >> 1;
    float2 xa[kFastE], xb[kFastE];
#pragma unroll
    for (int c = 0; c < kFastE; ++c) {
        const int i = threadIdx.x + c * T;
        const int l = i & (L - 1), k = i >> LSH;
        const int lx = l % tx, ly = l / tx;
        xa[c] = xb[c] = make_float2(0, 0);
        if (k <= h && ly < ty && X0 + lx / sx < nxl && Y0 + ly / sy < nyl) {
            const size_t col = (size_t)(y_lo + Y0 * sy + ly) * px +
                               (x_lo + X0 * sx + lx);
            xa[c] = src[k * plane + col];
            xb[c] = src[(n - k) * plane + col];
        }
    }
#pragma unroll
    for (int c = 0; c < kFastE; ++c) {
        const int i = threadIdx.x + c * T;
        const int l = i & (L - 1), k = i >> LSH;
        if (k <= h) {
            const float2 A = xa[c], B = xb[c];
            const float2 ca = conjf2(A), cb = conjf2(B);
            const float2 fo = cmul(csub(A, cb), conjf2(s_post[k]));
            buf[k * LS + l] = cadd(cadd(A, cb), make_float2(-fo.y, fo.x));
            if (k > 0 && n - k != k) {
                const float2 go = cmul(csub(B, ca), conjf2(s_post[n - k]));
                buf[(n - k) * LS + l] =
                    cadd(cadd(B, ca), make_float2(-go.y, go.x));
            }
        }
    }
    __syncthreads();
    tile_fft<LSH, T, N>(buf, a, s_tw, 1.0f);
    const float* fb = reinterpret_cast<const float*>(buf);
    const float nf = 1.0f / (float)(sx * sy * sz);
    const int per_z = kx * ky;
    const int nvox = per_z * nzl;
    double acc = 0.0;
    // voxels v = threadIdx.x + j T, kV at a time: their data loads are all
    // issued before the first is used
    constexpr int kV = 8;
    for (int v0 = 0; v0 < nvox; v0 += kV * T) {
        long long gi[kV];
        int lb[kV], zb[kV];
        float ob[kV], er[kV], mk[kV];
#pragma unroll
        for (int j = 0; j < kV; ++j) {
            const int v = v0 + threadIdx.x + j * T;
            gi[j] = -1;
            lb[j] = zb[j] = 0;
            if (v < nvox) {
                // v = (zl * ky + lyo) * kx + lxo
                const int zl = (int)fdiv((unsigned)v, pz_m);
                const int rem = v - zl * per_z;
                const int lyo = (int)fdiv((unsigned)rem, kx_m);
                const int lxo = rem - lyo * kx;
                const int X = X0 + lxo, Y = Y0 + lyo;
                if (X < nxl && Y < nyl) {
                    gi[j] = ((long long)zl * nyl + Y) * nxl + X;
                    lb[j] = lyo * sy * tx + lxo * sx;
                    zb[j] = ez + zl * sz;
                }
            }
            ob[j] = gi[j] >= 0 && da.obs ? da.obs[gi[j]] : 0.0f;
            er[j] = gi[j] >= 0 && da.err ? da.err[gi[j]] : da.err_scalar;
            mk[j] = gi[j] >= 0 && da.mask ? da.mask[gi[j]] : 1.0f;
        }
#pragma unroll
        for (int j = 0; j < kV; ++j) {
            if (gi[j] < 0) continue;
            float sum = 0.0f;
            for (int dz = 0; dz < sz; ++dz) {
                const int z = zb[j] + dz;
                const int zo = 2 * ((z >> 1) * LS) + (z & 1);
                for (int dy = 0; dy < sy; ++dy)
                    for (int dx = 0; dx < sx; ++dx)
                        sum += fb[zo + 2 * (lb[j] + dy * tx + dx)];
            }
            const float m = sum * nf;
            if (lo_out) lo_out[set * lstride + gi[j]] = m;
            if (!da.obs) continue;
            // objective.hpp:7-33 with mdl_w = mdl_m = none
            float r = m - ob[j];
            r /= er[j];
            if (da.mask) r *= mk[j];
            r *= da.weight;
            acc += (double)r * (double)r;
        }
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < T / 64; ++w) t += red[w];
        partials[(size_t)set * nparts + blockIdx.x] = t;
    }
}

// Batched mcdisk: blockIdx.y = set.  Each set builds its disk from its
// parameter vector and runs the clouds of its own budget into its own int64
// cube, with its own fixed-point scale; the per-cloud code is the plain
// kernel's (mcdisk.hpp), so a set's cube equals a one-model evaluation.
__global__ __launch_bounds__(256) void batch_mcdisk(
    gbkhip_disk tmpl, SetLayout lay, const float* __restrict__ params,
    int pstride, gbkhip_grid g, float cflux, const int* __restrict__ csum,
    int csum_len, const unsigned char* hoi, uint32_t key0, uint32_t key1,
    const double* __restrict__ qscale, unsigned long long* icube,
    long long n_hi)
{
    extern __shared__ int s_csum[];
    __shared__ float s_nodes[kMcNodesLds];
    const int set = blockIdx.y;
    gbkhip_disk d = set_disk(tmpl, lay, params + (size_t)set * pstride);
    d.rnodes = mc_stage_nodes(d.rnodes, d.nrnodes, s_nodes);
    McParams mp;
    mp.cflux = cflux;
    mp.csum = mc_stage_csum(csum + (size_t)set * csum_len, csum_len, s_csum);
    mp.csum_len = csum_len;
    mp.nclouds = mp.csum[csum_len - 1];
    mp.hoi = hoi;
    mp.key0 = key0;
    mp.key1 = key1;
    mp.qscale = qscale[set];
    gbkhip_outputs o;
    memset(&o, 0, sizeof(o));
    unsigned long long* ic = icube + (size_t)set * n_hi;
    for (long long ci = blockIdx.x * (long long)blockDim.x + threadIdx.x;
         ci < mp.nclouds; ci += (long long)gridDim.x * blockDim.x)
        mc_cloud(ci, mp, d, g, o, ic);
}

// Batched binned mcdisk: blockIdx.y = set.  The same cloud body writes a
// record per cloud (at most cap per set) and ranks it in its tile, with the
// wave-aggregated LDS ranks and per-block reservations of the plain path.
__global__ __launch_bounds__(256) void batch_mcdisk_bin(
    gbkhip_disk tmpl, SetLayout lay, const float* __restrict__ params,
    int pstride, gbkhip_grid g, float cflux, const int* __restrict__ csum,
    int csum_len, const unsigned char* hoi, uint32_t key0, uint32_t key1,
    McBin bin, long long cap, int ntiles, int* __restrict__ overflow)
{
    extern __shared__ int s_dyn[];            // mc_bin_lds layout
    const McBinLds L = mc_bin_lds(csum_len, tmpl.nrnodes, ntiles);
    float* s_nodes = reinterpret_cast<float*>(s_dyn + L.csum);
    int* s_cnt = s_dyn + L.csum + L.nodes;
    const int set = blockIdx.y;
    gbkhip_disk d = set_disk(tmpl, lay, params + (size_t)set * pstride);
    d.rnodes = mc_stage_nodes(d.rnodes, d.nrnodes, s_nodes);
    McParams mp;
    mp.cflux = cflux;
    mp.csum = mc_stage_csum(csum + (size_t)set * csum_len, csum_len, s_dyn);
    mp.csum_len = csum_len;
    mp.nclouds = (int)min((long long)mp.csum[csum_len - 1], cap);
    // a budget beyond the reserved capacity would lose clouds: flag it
    // (the host reports it; the batch's results are not valid)
    if (mp.csum[csum_len - 1] > cap && blockIdx.x == 0 && threadIdx.x == 0)
        *overflow = 1;
    mp.hoi = hoi;
    mp.key0 = key0;
    mp.key1 = key1;
    mp.qscale = 1.0;                 // unused by the record path
    bin.rec += set * cap;
    bin.tr += set * cap;
    bin.tcount += (size_t)set * ntiles;
    const bool local = L.cnt > 0;
    if (local) {
        for (int i = threadIdx.x; i < ntiles; i += blockDim.x) s_cnt[i] = 0;
        __syncthreads();
    }
    gbkhip_outputs o;
    memset(&o, 0, sizeof(o));
    __shared__ int s_wscr[4][128];
    int* wscr = s_wscr[threadIdx.x >> 6];
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long b0 = (long long)blockIdx.x * blockDim.x; b0 < mp.nclouds;
         b0 += stride) {
        const long long ci = b0 + threadIdx.x;
        const int tile = ci < mp.nclouds
                             ? mc_cloud(ci, mp, d, g, o, nullptr, &bin) : -1;
        const int rank = wave_rank(tile, local ? s_cnt : bin.tcount, wscr);
        if (ci < mp.nclouds) bin.tr[ci] = make_int2(tile, rank);
    }
    if (!local) return;
    __syncthreads();
    for (int i = threadIdx.x; i < ntiles; i += blockDim.x) {
        const int n = s_cnt[i];
        s_cnt[i] = n ? atomicAdd(&bin.tcount[i], n) : 0;
    }
    __syncthreads();
    for (long long ci = blockIdx.x * (long long)blockDim.x + threadIdx.x;
         ci < mp.nclouds; ci += stride) {
        const int2 t = bin.tr[ci];
        if (t.x >= 0) bin.tr[ci] = make_int2(t.x, t.y + s_cnt[t.x]);
    }
}

__global__ void batch_zero_f32(float* p, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
         i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = 0.0f;
}

__global__ void batch_zero_i32(int* p, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
         i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = 0;
}

// hi = icube / 2^S per set (0 where nothing landed), icube <- 0.
__global__ void batch_fixed_to_hi(unsigned long long* icube, float* hi,
                                  long long n, const double* qscale)
{
    const int set = blockIdx.y;
    const double inv = 1.0 / qscale[set];
    unsigned long long* ic = icube + (size_t)set * n;
    float* h = hi + (size_t)set * n;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const long long q = (long long)ic[i];
        float v = 0.0f;
        if (q) {
            v = (float)((double)q * inv);
            ic[i] = 0ull;
        }
        h[i] = v;
    }
}

__global__ void batch_zero_u64(unsigned long long* p, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

// ---------------------------------------------------------------------------
// Direct separable pipeline (thin disk, separable PSF x LSF kernel)
//
// With K = Wz(z) Wy(y) Wx(x) of half-widths e, and the block mean of the
// downscale folded into the taps, the low-resolution cube is
//   lo(X, Y, Z) = sum_u,v,w Wx[u] Wy[v] Wz[w] m(sx X + u, sy Y + v, sz Z + w)
// with W_a[u] = (1/s_a) sum_{d < s_a} K_a[d - u + e_a], u in [0, s_a + 2 e_a)
// (indices relative to the first input of the block).  Three passes:
//   S1 sep_xpass   lines of one hi-res row deposited in LDS for a chunk of
//                  channels, x taps -> P1 (z, y, X)
//   S2 sep_ypass   y taps over LDS column tiles -> P2 (z, Y, X)
//   S3 sep_zpass   z taps, residual, FP64 partial sums of r^2 per tile
// Every output is a fixed-order FP32 dot product of at most TX (TY, TZ)
// taps; templates fix the tap buckets and downscale factors so the inner
// loops are register-blocked with constant offsets.
// ---------------------------------------------------------------------------
constexpr int kGroupBranches = 4;   // parallel streams of a ragged group
constexpr int kSepZC = 16;          // most channels per S1 block
constexpr int kSepJZ = 16;          // outputs per thread, S3
constexpr int kSepXW = 32;          // X columns per S2 block
constexpr int kSepYPitch = kSepXW + 1;
constexpr int kSepMaxRows = (48 * 1024) / (kSepYPitch * 4);   // S2 LDS rows
// channels of one x-pass chunk an S2 block filters in turn (the chunk's
// zero rows and live groups are found once per block)
constexpr int kSepYChannels = 16;
// lines per thread an S1 row block keeps in registers (wider rows read the
// rest again per chunk)
constexpr int kSepXLines = 2;

// S1: block = (row y, set); the row's channel chunks [z0, z0 + zn) in
// turn.  The hi-res values of the row are deposited into an LDS tile
// [zc][mp] (zero outside each spaxel's rint window, gmodels.hpp:91-116),
// then filtered along x.  A chunk no line of the row reaches filters to
// zeros: it is skipped (its rows are not written; the y pass reads the same
// mask and uses zeros), and a row no line reaches returns at once.  The
// row's first T UX lines are loaded once for all its chunks.
template <int SX, int TX>
__global__ __launch_bounds__(256) void sep_xpass(
    const float4* __restrict__ lines, long long lstride, int lpitch,
    gbkhip_grid g, const float* __restrict__ wx, int ox, int need_x,
    int need_z, int nxl, int mp, int zc, float* __restrict__ P1,
    long long p1stride, int ny1,
    const unsigned long long* __restrict__ rowmask,
    const unsigned long long* __restrict__ segmask)
{
    extern __shared__ float4 M4[];        // [zc][mp] floats
    float* M = reinterpret_cast<float*>(M4);
    constexpr int T = kSepThreads, UX = kSepXLines;
    const int y = blockIdx.x, set = blockIdx.z;
    unsigned long long live = rowmask[(size_t)set * ny1 + y];
    if (!live) return;
    // 64-spaxel segments with a line (batch_lines_thin wrote only those
    // below the 63rd; the rest are zero)
    const unsigned long long seg = segmask[(size_t)set * ny1 + y];
    auto written = [&](int x) {
        const int sg = x >> 6;
        return sg >= 63 || ((seg >> sg) & 1ull);
    };
    // lines in form 1 (batch_lines_thin): amplitude and window precomputed
    const float4* lrow = lines + set * lstride + (size_t)y * lpitch;
    float4 V[UX];
#pragma unroll
    for (int u = 0; u < UX; ++u) {
        const int x = threadIdx.x + u * T;
        V[u] = x < need_x && written(x) ? lrow[x] : make_float4(0, 0, 0, 0);
    }
    const int n4 = (zc * mp) >> 2;
    float* out = P1 + set * p1stride + (size_t)y * nxl;
    while (live) {                            // uniform over the block
        const int ch = __ffsll((long long)live) - 1;
        live &= live - 1ull;
        const int z0 = ch * zc, zn = min(zc, need_z - z0);
        for (int i = threadIdx.x; i < n4; i += T)
            M4[i] = make_float4(0, 0, 0, 0);
        for (int i = (n4 << 2) + threadIdx.x; i < zc * mp; i += T)
            M[i] = 0.0f;
        __syncthreads();
        auto deposit = [&](const float4& L, int x) {
            const unsigned kw = __float_as_uint(L.w);
            const int k0 = max((int)(kw & 0xffffu), z0);
            const int k1 = min((int)(kw >> 16), z0 + zn - 1);
            const LineEval e = line_from(g, L.x, L.y, L.z);
            for (int k = k0; k <= k1; ++k) M[(k - z0) * mp + x] = e.at(k);
        };
#pragma unroll
        for (int u = 0; u < UX; ++u)
            if (V[u].x != 0.0f) deposit(V[u], threadIdx.x + u * T);
        for (int x = threadIdx.x + UX * T; x < need_x; x += T) {
            if (!written(x)) continue;
            const float4 L = lrow[x];
            if (L.x != 0.0f) deposit(L, x);
        }
        __syncthreads();
        sep_xfilter<SX, TX>(M + ox, wx, mp, zn, z0, nxl, ny1, out);
        __syncthreads();            // the tile is free for the next chunk
    }
}

// S1 from a deposited hi-res cube (thick disks, mcdisk): block = (row y,
// channels [z0, z0 + zn), set).  The row's need_x values of each channel
// are read from the cube (a wave per channel row, 8 loads in flight per
// lane) into the LDS tile, zero beyond need_x.  A tile that is zero
// everywhere filters to zeros: it returns without writing P1 and leaves its
// bit of the row mask clear (the mask starts at zero; the y pass reads it
// and uses zeros for the row); otherwise it sets the bit and filters.
template <int SX, int TX>
__global__ __launch_bounds__(256) void sep_xpass_cube(
    const float* __restrict__ hi, long long hstride, int hnx, int hny,
    const float* __restrict__ wx, int ox, int need_x, int need_z, int nxl,
    int mp, int zc, float* __restrict__ P1, long long p1stride, int ny1,
    unsigned long long* __restrict__ rowmask)
{
    extern __shared__ float4 M4[];        // [zc][mp] floats
    float* M = reinterpret_cast<float*>(M4);
    __shared__ float s_w[TX];
    constexpr int T = kSepThreads, U = 8;
    const int y = blockIdx.x, z0 = blockIdx.y * zc, set = blockIdx.z;
    const int zn = min(zc, need_z - z0);
    for (int i = threadIdx.x; i < TX; i += T) s_w[i] = wx[i];
    const int lane = threadIdx.x & 63;
    const size_t plane = (size_t)hny * hnx;
    int live = 0;
    for (int zz = threadIdx.x >> 6; zz < zn; zz += T / 64) {
        const float* src = hi + set * hstride + (size_t)(z0 + zz) * plane +
                           (size_t)y * hnx;
        float* dst = M + zz * mp;
        for (int x0 = lane; x0 < mp; x0 += 64 * U) {
            float t[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int x = x0 + 64 * u;
                t[u] = x < need_x ? src[x] : 0.0f;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int x = x0 + 64 * u;
                if (x < mp) dst[x] = t[u];
                live |= t[u] != 0.0f;
            }
        }
    }
    if (!__syncthreads_or(live)) return;
    if (threadIdx.x == 0)
        atomicOr(&rowmask[(size_t)set * ny1 + y], 1ull << blockIdx.y);
    sep_xfilter<SX, TX>(M + ox, s_w, mp, zn, z0, nxl, ny1,
                        P1 + set * p1stride + (size_t)y * nxl);
}

// S2: P2[z][Y][X] = sum_v Wy[v] P1[z][sy Y + v][X].  Block = (32 X
// columns, Y outputs [Y0, Y0 + yblk), one x-pass channel chunk, set), in
// groups of J consecutive Y.  The x pass leaves a row of a chunk unwritten
// when no line reaches it, so every channel of the chunk has the same zero
// rows: a group whose input rows are all zero is zero for the whole chunk,
// neither computed nor written, and its flag tells the z pass to read
// zeros.  The block lists the chunk's live groups once, then per channel
// stages only the rows they read in LDS (pitch 33) and hands each live
// group to one wave, in list order: lanes 0-31 filter the group's first
// J / 2 outputs of their column, lanes 32-63 the last J / 2, so no lane of a
// working wave idles on a dead group.
template <int SY, int TY>
__global__ __launch_bounds__(256, 4) void sep_ypass(
    const float* __restrict__ P1, long long p1stride,
    const float* __restrict__ wy, int oy, int nxl, int nyl, int ny1,
    int need_z, int yblk, float* __restrict__ P2, long long p2stride,
    const unsigned long long* __restrict__ rowmask, int zc, int zpb,
    unsigned char* __restrict__ gflag, int ngy, const float4* __restrict__ zq)
{
    constexpr int T = kSepThreads, J = sep_jy(TY), XW = kSepXW;
    constexpr int PITCH = kSepYPitch;
    constexpr int NIN = SY * (J - 1) + TY;
    static_assert(2 * XW == 64 && J % 2 == 0, "a wave is two column halves");
    extern __shared__ float tile[];       // [rows][PITCH]
    __shared__ unsigned char s_flag[kSepMaxRows];
    // one word per 64 rows, written by every wave of each 256-row pass
    __shared__ unsigned long long s_bits[(kSepMaxRows + 255) / 256 * 4];
    __shared__ unsigned char s_gl[T];     // live groups, in order
    __shared__ int s_wcnt[T / 64];
    __shared__ float4 s_junk[T / 64];     // staging stores past the rows
    const int X0 = blockIdx.x * XW, Y0 = blockIdx.y * yblk;
    const int R0 = SY * Y0 + oy;          // input row of staged row 0
    // blockIdx.z = (set, chunk, sub-block of zpb channels of the chunk)
    const int nsub = (zc + zpb - 1) / zpb;
    const int nzb = (need_z + zc - 1) / zc * nsub;
    const int zb = blockIdx.z % nzb, set = blockIdx.z / nzb;
    const int chunk = zb / nsub;
    const int z0 = chunk * zc + (zb % nsub) * zpb;
    const int z1 = min(min(z0 + zpb, chunk * zc + zc), need_z);
    const int ngrp = (min(yblk, nyl - Y0) + J - 1) / J;
    const int nrows = SY * (ngrp * J - 1) + TY;
    // rows whose x-pass tile no line reaches are zero (never written): a
    // byte per row for the staging and a bit per row for the group test
    const unsigned long long* rm = rowmask + (size_t)set * ny1;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int r0 = 0; r0 < nrows; r0 += T) {          // uniform trip count
        const int r = r0 + threadIdx.x;
        const int y = R0 + r;
        const bool f = r < nrows && y < ny1 && ((rm[y] >> chunk) & 1ull);
        if (r < nrows) s_flag[r] = f;
        const unsigned long long b = __ballot(f);
        if (lane == 0) s_bits[(r0 >> 6) + wv] = b;
    }
    __syncthreads();
    // group gi reads rows [SY J gi, SY J gi + NIN): live if any is live
    const int gi = threadIdx.x;
    bool live = false;
    if (gi < ngrp) {
        const int a = SY * J * gi, e = a + NIN;
        for (int w = a >> 6; w <= (e - 1) >> 6; ++w) {
            const int lo = max(a, w << 6) - (w << 6);
            const int hi = min(e, (w << 6) + 64) - (w << 6);
            const unsigned long long m =
                (hi - lo == 64 ? ~0ull : ((1ull << (hi - lo)) - 1ull)) << lo;
            live |= (s_bits[w] & m) != 0ull;
        }
        for (int z = z0; z < z1; ++z)
            gflag[((size_t)set * need_z + z) * ngy + Y0 / J + gi] = live;
    }
    const unsigned long long gm = __ballot(live);
    if (lane == 0) s_wcnt[wv] = __popcll(gm);
    __syncthreads();
    int base = 0, nl = 0;
#pragma unroll
    for (int w = 0; w < T / 64; ++w) {
        base += w < wv ? s_wcnt[w] : 0;
        nl += s_wcnt[w];
    }
    if (nl == 0) return;                 // every group of the chunk is zero
    if (live) s_gl[base + __popcll(gm & ((1ull << lane) - 1ull))] = gi;
    __syncthreads();
    // the rows the live groups read
    const int rlo = SY * J * s_gl[0];
    const int rhi = SY * J * s_gl[nl - 1] + NIN;
    constexpr int JH = J / 2, NH = SY * (JH - 1) + TY;
    const int c = lane & (XW - 1), h = lane / XW, X = X0 + c;
    // staging: U independent loads in flight per thread before the stores
    constexpr int U = 8;
    constexpr int Q = XW / 4;                 // float4s per staged row
    constexpr int RS = T / Q;                 // rows between a thread's quads
    static_assert(T % Q == 0, "a thread keeps one column quad");
    const bool quads = (nxl & 3) == 0 && X0 + XW <= nxl;
    // this thread stages quad q4 of rows rt + e RS, e < ne: its addresses
    // are a base plus a uniform stride.  The rows do not change from
    // channel to channel: bit e of lb says whether row rt + e RS was
    // written (the others read a zero quad, so every load issues
    // unconditionally and no flag is read per channel)
    const int q4 = threadIdx.x % Q, rt = rlo + threadIdx.x / Q;
    const int ne = rt < rhi ? (rhi - rt + RS - 1) / RS : 0;
    const int nemax = (rhi - rlo + RS - 1) / RS;          // uniform
    static_assert((kSepMaxRows + RS - 1) / RS + U <= 32, "bits of lb");
    unsigned lb = 0u;
    if (quads)
        for (int e = 0; e < ne; ++e)
            lb |= (s_flag[rt + e * RS] ? 1u : 0u) << e;
    float* const lt = tile + rt * PITCH + 4 * q4;
    const int rstep = RS * nxl;               // floats between a thread's rows
    for (int z = z0; z < z1; ++z) {
        const float* src = P1 + set * p1stride + (size_t)z * ny1 * nxl;
        if (quads) {
            // whole strip, rows 16-byte aligned: float4 loads
            const float* st = src + (size_t)(R0 + min(rt, rhi - 1)) * nxl +
                              X0 + 4 * q4;
            for (int e0 = 0; e0 < nemax; e0 += U) {
                float4 t[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int e = e0 + u;
                    const float4* p = ((lb >> e) & 1u)
                        ? reinterpret_cast<const float4*>(st + e * rstep)
                        : zq;
                    t[u] = *p;
                }
                // every store issues (no branch the loads could sink
                // into): quads past this thread's rows go to a junk slot
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int e = e0 + u;
                    float* d = e < ne ? lt + e * (RS * PITCH)
                                      : reinterpret_cast<float*>(s_junk + wv);
                    d[0] = t[u].x;
                    d[1] = t[u].y;
                    d[2] = t[u].z;
                    d[3] = t[u].w;
                }
            }
        } else {
            const int n1 = (rhi - rlo) * XW;
            for (int i0 = threadIdx.x; i0 < n1; i0 += T * U) {
                float t[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int i = i0 + u * T, r = rlo + i / XW, c = i % XW;
                    const int X = X0 + c;
                    t[u] = (i < n1 && s_flag[r] && X < nxl)
                               ? src[(size_t)(R0 + r) * nxl + X] : 0.0f;
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int i = i0 + u * T, r = rlo + i / XW, c = i % XW;
                    if (i < n1) tile[r * PITCH + c] = t[u];
                }
            }
        }
        __syncthreads();
        float* dst = P2 + set * p2stride + (size_t)z * nyl * nxl + X;
        for (int k = wv; k < nl; k += T / 64) {
            const int g = s_gl[k];
            const float* in = tile + SY * (J * g + JH * h) * PITCH + c;
            float v[NH];
#pragma unroll
            for (int i = 0; i < NH; ++i) v[i] = in[i * PITCH];
            float o[JH] = {};
#pragma unroll
            for (int u = 0; u < TY; ++u) {
                const float w = wy[u];        // wave-uniform: held in SGPRs
#pragma unroll
                for (int j = 0; j < JH; ++j)
                    o[j] = fmaf(w, v[SY * j + u], o[j]);
            }
            // full halves (all but the last rows' group) store without a
            // bounds check per output
            const int Yb = Y0 + g * J + JH * h;
            float* d = dst + (size_t)Yb * nxl;
            if (X < nxl && Yb + JH <= nyl) {
#pragma unroll
                for (int j = 0; j < JH; ++j) d[(size_t)j * nxl] = o[j];
            } else {
#pragma unroll
                for (int j = 0; j < JH; ++j)
                    if (X < nxl && Yb + j < nyl) d[(size_t)j * nxl] = o[j];
            }
        }
        __syncthreads();                      // the tile is free again
    }
}

// lo[Z][Y][X] = sum_w Wz[w] P2[sz Z + w][Y][X], then the residual against
// the data and one FP64 partial sum of r^2 per block (fixed order).  A block
// covers 64 X x 4 Y columns and kSepJZ channels: wave w takes row Y0 + w, so
// whether an input row was written by the y pass is uniform over the wave
// (one ballot per wave).  Every input load is issued: rows the y pass left
// unwritten read a 64-float zero row instead (cached, no HBM traffic), so
// the loads go out back to back with a scalar base and a lane offset.  The
// residual divides by the error through its reciprocal (v_rcp_f32, within
// an ulp of the quotient).
// Raw buffer loads from a wave-uniform base: the base goes into a buffer
// resource (scalar registers, full 4 GiB range, so no load is clipped), a
// wave-uniform byte offset into the instruction's scalar offset and the
// lane's byte offset is the only vector operand.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t lane_rsrc(const float* base)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(base), 0,
                                             (int)0xffffffffu, 0x00020000);
}

__device__ __forceinline__ float load_at(__amdgpu_buffer_rsrc_t r,
                                         unsigned lane_off, unsigned uni_off)
{
    return __builtin_bit_cast(
        float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)lane_off,
                                                    (int)uni_off, 0));
}

template <int SZ, int TZ>
__global__ __launch_bounds__(256) void sep_zpass(
    const float* __restrict__ P2, long long p2stride, const float* wz,
    int oz, int nxl, int nyl, int nzl, int need_z, int zblocks, DataArgs da,
    float* __restrict__ lo_out, long long lstride,
    double* __restrict__ partials, int nparts,
    const unsigned char* __restrict__ gflag, int ngy, int jy,
    const double* __restrict__ zero_part, const float* __restrict__ zrow)
{
    constexpr int T = kSepThreads, J = kSepJZ, TXB = 64;
    constexpr int NIN = SZ * (J - 1) + TZ;
    constexpr int NW = (NIN + 63) / 64;         // ballot words
    __shared__ float s_w[TZ];
    __shared__ double red[T / 64];
    const int set = blockIdx.z / zblocks, zb = blockIdx.z % zblocks;
    const int lane = threadIdx.x & 63;
    // the wave index in a scalar register: row bases are wave-uniform
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int X0 = blockIdx.x * TXB;
    const int Y = blockIdx.y * (T / TXB) + wv;          // uniform per wave
    const int Z0 = zb * J;
    for (int i = threadIdx.x; i < TZ; i += T) s_w[i] = wz[i];
    // which of the wave's input rows (channel SZ Z0 + oz + i of row Y) the
    // y pass wrote: the others are zero
    unsigned long long live[NW];
    unsigned long long any = 0ull;
#pragma unroll
    for (int k = 0; k < NW; ++k) {
        const int i = 64 * k + lane, z = SZ * Z0 + oz + i;
        const bool f = i < NIN && z < need_z && Y < nyl &&
            gflag[((size_t)set * need_z + z) * ngy + Y / jy] != 0;
        live[k] = __ballot(f);
        any |= live[k];
    }
    const int part = (blockIdx.y * gridDim.x + blockIdx.x) * zblocks + zb;
    const int plane = nyl * nxl;
    // a block whose model is zero everywhere has the data's own partial
    // sum, computed once at set-up by this kernel with no live input (same
    // arithmetic, so the same bits): the data are not read again
    if (!__syncthreads_or(any != 0ull) && zero_part) {
        if (threadIdx.x == 0)
            partials[(size_t)set * nparts + part] = zero_part[part];
        const int X = X0 + lane;
        if (lo_out && X < nxl && Y < nyl) {
            float* lo = lo_out + set * lstride + Y * nxl + X;
            for (int j = 0; j < J && Z0 + j < nzl; ++j)
                lo[(size_t)(Z0 + j) * plane] = 0.0f;
        }
        return;
    }
    double acc = 0.0;
    if (Y < nyl) {
        // lanes past the last column read the row's last value (in bounds)
        const unsigned xo = min(lane, nxl - 1 - X0);
        const float* row = P2 + set * p2stride + (size_t)Y * nxl + X0;
        // two resources (the column's P2 slab and the zero row) and the
        // channel's byte offset in the scalar operand: per load only a
        // scalar select and add (the set's P2 fits 32-bit offsets: the
        // host checks)
        const unsigned xb = xo * (unsigned)sizeof(float);
        const unsigned pb = (unsigned)plane * (unsigned)sizeof(float);
        const __amdgpu_buffer_rsrc_t rp = lane_rsrc(row), rz = lane_rsrc(zrow);
        const unsigned zb0 = (unsigned)(SZ * Z0 + oz) * pb;
        float v[NIN];
#pragma unroll
        for (int i = 0; i < NIN; ++i) {
            const bool lv = (live[i >> 6] >> (i & 63)) & 1ull;
            v[i] = load_at(lv ? rp : rz, xb, lv ? zb0 + (unsigned)i * pb : 0u);
        }
        float o[J] = {};
#pragma unroll
        for (int w = 0; w < TZ; ++w) {
            const float ww = s_w[w];
#pragma unroll
            for (int j = 0; j < J; ++j) o[j] = fmaf(ww, v[SZ * j + w], o[j]);
        }
        const int X = X0 + lane;
        if (X < nxl) {
            // column offsets: a wave-uniform part and the lane's column
            const int col0 = Y * nxl + X0;
            const int jn = min(J, nzl - Z0);      // outputs of this block
            if (lo_out) {
                float* lo = lo_out + set * lstride + col0;
#pragma unroll
                for (int j = 0; j < J; ++j)
                    if (j < jn) lo[(size_t)(Z0 + j) * plane + xo] = o[j];
            }
            if (da.obs) {
                // every data load in flight before the first use (channels
                // past the cube read the last one and are not summed)
                unsigned zo[J];
                float ob[J], ie[J], mk[J];
                const __amdgpu_buffer_rsrc_t ro = lane_rsrc(da.obs + col0);
#pragma unroll
                for (int j = 0; j < J; ++j) {
                    zo[j] = (unsigned)min(Z0 + j, nzl - 1) * pb;
                    ob[j] = load_at(ro, xb, zo[j]);
                }
                if (da.err) {
                    const __amdgpu_buffer_rsrc_t re = lane_rsrc(da.err + col0);
#pragma unroll
                    for (int j = 0; j < J; ++j) ie[j] = load_at(re, xb, zo[j]);
                }
                if (da.mask) {
                    const __amdgpu_buffer_rsrc_t rm = lane_rsrc(da.mask + col0);
#pragma unroll
                    for (int j = 0; j < J; ++j) mk[j] = load_at(rm, xb, zo[j]);
                }
                // objective.hpp:7-33 with mdl_w = mdl_m = none; the
                // uniform choices are branches around whole loops (the same
                // products in the same order per output; a weight of 1 is
                // an exact no-op and skipped)
                float r[J];
#pragma unroll
                for (int j = 0; j < J; ++j) r[j] = o[j] - ob[j];
                if (da.err) {
#pragma unroll
                    for (int j = 0; j < J; ++j)
                        r[j] *= __builtin_amdgcn_rcpf(ie[j]);
                } else {
#pragma unroll
                    for (int j = 0; j < J; ++j) r[j] *= da.inv_err_scalar;
                }
                if (da.mask) {
#pragma unroll
                    for (int j = 0; j < J; ++j) r[j] *= mk[j];
                }
                if (da.weight != 1.0f) {
#pragma unroll
                    for (int j = 0; j < J; ++j) r[j] *= da.weight;
                }
                if (jn == J) {
#pragma unroll
                    for (int j = 0; j < J; ++j)
                        acc += (double)r[j] * (double)r[j];
                } else {
#pragma unroll
                    for (int j = 0; j < J; ++j)
                        if (j < jn) acc += (double)r[j] * (double)r[j];
                }
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if (lane == 0) red[wv] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < T / 64; ++w) t += red[w];
        partials[(size_t)set * nparts + part] = t;
    }
}

struct Sep {
    bool on = false;
    bool cube = false;            // S1 reads a deposited hi-res cube
    int need[3] = {0, 0, 0};      // inputs that reach the outputs
    int tx = 0, ty = 0, tz = 0;   // tap buckets (zero-padded)
    int ox = 0, oy = 0, oz = 0;   // inputs before an output's first tap
    int kept[3] = {0, 0, 0};      // taps an output uses, per axis
    int mp = 0;                   // S1 LDS row length (odd)
    int zc = 0;                   // channels per S1 block (tile <= 64 KB)
    int yblk = 0;                 // Y outputs per S2 block
    int yblk_full = 0;            // ... when one block covers every Y
    DevBuf wx, wy, wz;            // taps
    DevBuf P1, P2;                // work: per set (z, y, X) and (z, Y, X)
    DevBuf masks;                 // work: per set and row, live chunks
    DevBuf segmask;               // work: per set and row, live segments
    DevBuf gflags;                // work: per set, channel, y-pass group
    DevBuf zero_part;             // z-pass partials of a zero model (data)
    DevBuf zrow;                  // 64 zeros: the z pass's unwritten rows
    int ngy = 0;                  // y-pass output groups per channel
    int nchunks = 0;              // S1 channel chunks
    int last_sets = 0;            // sets of the last chunk run on lane 0
    long long p1stride = 0, p2stride = 0;
    int zblocks = 1;
};

struct Fused {
    bool on = false;
    int nx = 0, ny = 0, nz = 0, nzh = 0;   // FFT sizes (nz even)
    int need[3] = {0, 0, 0};               // S_i + 2 edge_i
    int xlo = 0, xcnt = 0, ylo = 0, ycnt = 0;
    // S rows are px >= nx complex apart (a multiple of 16: 128-byte aligned
    // column tiles) and column c holds model spaxel (c - shift) mod nx, the
    // shift putting the needed x range at a multiple of 16 (the convolution
    // is circular, so a cyclic shift of its input shifts its output)
    int px = 0, shift = 0;
    FastShape fz, fy, fx;                  // z columns, y lines, x rows
    DevBuf twz, twy, twx, tpz;
    DevBuf twsz, twsy, twsx;               // per-stage tables
    DevBuf phat, lhat;
    DevBuf S;
    long long sstride = 0;
    DevBuf lines;                          // per spaxel (flux, v, sigma)
    long long lstride = 0;
    int kx = 1, ky = 1, tiles_x = 1, tiles_y = 1;
};

// A disk of the batch: the template (device trait arrays and sub-ring
// nodes) and where each set's parameters of it are in the set vector.
struct DiskLayout {
    gbkhip_disk tmpl;
    SetLayout lay;
    DevBuf ints, rnodes;
    std::vector<int> h_uids;
};

struct Batch {
    Context* ctx;
    int device = 0;
    gbkhip_batch_desc desc;
    gbkhip_disk tmpl;
    SetLayout lay;
    std::vector<int> h_uids, h_ccounts, h_pcounts;   // packed int arrays
    DevBuf ints;                                      // device copy of them
    DevBuf rnodes;
    std::shared_ptr<FFTPlan> plan;
    Fused fused;            // spectral-first pipeline, when it applies
    Sep sep;                // direct separable pipeline, when it applies
    // separable kernel found by the analysis of the PSF x LSF cube: sums of
    // pcube over z (psf, ny x nx) and over y, x (lsf); empty when absent
    std::vector<double> ka_psf, ka_lsf;
    double ka_total = 0.0;
    DevBuf spectrum;        // kernel spectrum (nz, ny, nxh) complex
    DevBuf hi;              // chunk x padded cube
    DevBuf spec;            // chunk x complex work spectrum
    DevBuf partials;
    // moment-map objective (gbkhip_batch_set_mmaps): the model cubes of a
    // chunk (when the caller asks for none) and the per-block partials
    bool mm_on = false;
    MMapArgs mm;
    int mm_nparts = 0;
    DevBuf mm_lo, mm_partials;
    DevBuf icube;           // mcdisk: chunk x int64 fixed-point cube
    // binned mcdisk deposits (when a tile spectrum fits LDS): per set of a
    // chunk, mc_cap records, ranks and sorted records, tile counts, offsets
    int mc_tw = 0, mc_ntx = 0, mc_ntiles = 0;
    long long mc_cap = 0;
    DevBuf mcbin;
    DevBuf overflow;        // int: set when a set's clouds exceed mc_cap
    int max_grid_z = 65535; // device limit on gridDim.z (and y)
    unsigned long long work_gen = 0;    // bumped whenever work is allocated
    // components after the first (smdisk), added into the hi-res cube
    std::vector<std::unique_ptr<DiskLayout>> extras;
    // caller-owned work arena (gbkhip_batch_bind_work): the work buffers
    // are views into it, or allocations of their own when it is null
    char* arena = nullptr;
    size_t arena_bytes = 0;
    const int* csum = nullptr;       // mcdisk inputs of the current call
    const double* qscale = nullptr;
    int nparts = 0;
    long long n_hi = 0, n_c = 0, n_lo = 0;
    bool thin = true;
    bool work = false;      // work buffers allocated
    hipStream_t stream = nullptr;   // evaluate() stream; null: the context's
    // graph cache: (params, nsets, chi2, cubes) -> executable graph
    std::map<std::tuple<uintptr_t, int, uintptr_t, uintptr_t, uintptr_t,
                        uintptr_t, uintptr_t>,
             hipGraphExec_t> graphs;
    hipStream_t cap_stream = nullptr;
    // chunks in flight: with 2 lanes, odd chunks run on the aux stream with
    // their own work buffers, so one chunk's compute-bound x pass overlaps
    // the memory-bound y and z passes of the previous one
    int lanes = 1;
    hipStream_t aux = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // stage timing: events around smdisk / convolution / chi^2 of every
    // chunk (recorded into graphs too), read back after the work completes
    bool timing = false;
    std::vector<hipEvent_t> events;      // chunk-major, 4 per chunk
    int timed_chunks = 0;
    ~Batch()
    {
        for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
        for (auto e : events) (void)hipEventDestroy(e);
        if (cap_stream) (void)hipStreamDestroy(cap_stream);
        if (aux) (void)hipStreamDestroy(aux);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
    }
};

static void build_disk(const gbkhip_batch_desc& D, DiskLayout& L)
{
    const int nn = D.nrnodes;
    GBK_CHECK(nn >= 2 && D.rnodes, GBKHIP_ERR_ARG,
              "a disk needs at least two rnodes");
    L.lay.vsys = D.vsys_count > 0 ? D.off_vsys : -1;
    L.lay.xpos = D.off_xpos;
    L.lay.ypos = D.off_ypos;
    L.lay.posa = D.off_posa;
    L.lay.incl = D.off_incl;
    std::memset(&L.tmpl, 0, sizeof(L.tmpl));
    L.tmpl.loose = D.loose;
    L.tmpl.tilted = D.tilted;
    L.tmpl.nrnodes = nn;
    int total = 0;
    for (int k = 0; k < 9; ++k) total += 3 * D.traits[k].n;
    L.h_uids.assign(std::max(total, 1), 0);
    int pos = 0;
    std::vector<std::pair<int, int>> where(9, {-1, 0});
    for (int k = 0; k < 9; ++k) {
        const gbkhip_batch_traits& t = D.traits[k];
        L.lay.cvalues[k] = t.n ? t.off_cvalues : -1;
        L.lay.pvalues[k] = t.n ? t.off_pvalues : -1;
        if (!t.n) continue;
        GBK_CHECK(t.n <= 4, GBKHIP_ERR_ARG, "at most 4 traits per kind");
        where[k] = {pos, t.n};
        for (int i = 0; i < t.n; ++i) {
            L.h_uids[pos + i] = t.uids[i];
            L.h_uids[pos + t.n + i] = t.ccounts[i];
            L.h_uids[pos + 2 * t.n + i] = t.pcounts[i];
        }
        pos += 3 * t.n;
    }
    L.ints.alloc(L.h_uids.size() * sizeof(int));
    GBK_HIP(hipMemcpy(L.ints.ptr, L.h_uids.data(),
                      L.h_uids.size() * sizeof(int), hipMemcpyHostToDevice));
    gbkhip_traits* tr[9] = {&L.tmpl.rp, &L.tmpl.rh, &L.tmpl.vp, &L.tmpl.vh,
                            &L.tmpl.dp, &L.tmpl.dh, &L.tmpl.zp, &L.tmpl.sp,
                            &L.tmpl.wp};
    const int* base = L.ints.as<int>();
    for (int k = 0; k < 9; ++k) {
        if (where[k].first < 0) continue;
        const int p0 = where[k].first, n = where[k].second;
        tr[k]->n = n;
        tr[k]->uids = base + p0;
        tr[k]->ccounts = base + p0 + n;
        tr[k]->pcounts = base + p0 + 2 * n;
    }
    L.rnodes.alloc(nn * sizeof(float));
    GBK_HIP(hipMemcpy(L.rnodes.ptr, D.rnodes, nn * sizeof(float),
                      hipMemcpyHostToDevice));
    L.tmpl.rnodes = L.rnodes.as<float>();
    L.tmpl.opacity = nullptr;
}

static void build_layout(Batch& b)
{
    DiskLayout L;
    build_disk(b.desc, L);
    b.tmpl = L.tmpl;
    b.lay = L.lay;
    b.h_uids.swap(L.h_uids);
    std::swap(b.ints.ptr, L.ints.ptr);
    std::swap(b.ints.bytes, L.ints.bytes);
    std::swap(b.rnodes.ptr, L.rnodes.ptr);
    std::swap(b.rnodes.bytes, L.rnodes.bytes);
    b.thin = b.desc.traits[1].n == 0;
    // further smdisk components of the same gmodel, summed into the cube
    // after the first (gmodels summing components, _detail.py:405-422)
    for (int i = 0; i < b.desc.nextra; ++i) {
        const gbkhip_batch_desc& E = b.desc.extra[i];
        GBK_CHECK(!E.mcdisk, GBKHIP_ERR_UNSUPPORTED,
                  "further components must be smdisk disks");
        GBK_CHECK((E.traits[1].n == 0) == b.thin || b.desc.mcdisk,
                  GBKHIP_ERR_ARG,
                  "components of one gmodel are all thin or all thick");
        b.extras.emplace_back(new DiskLayout);
        build_disk(E, *b.extras.back());
    }
}

// Work buffers: everything sized by max_batch (per-set spectra or padded
// cubes, line parameters, chi^2 partials).  They can be released between
// evaluations and are re-allocated by the next one; the resident state
// (layout, kernel spectra, plans, twiddles) stays.
// Binned mcdisk deposits of one lane (max_batch sets): records, ranks,
// sorted records, tile counts, offsets and split-tile accumulators.
static size_t mc_lane_bytes(const Batch& B)
{
    return mc_bufs(nullptr, B.desc.max_batch, B.mc_cap, B.mc_ntiles,
                   B.mc_tw, B.desc.grid.spec_size, nullptr);
}

static size_t work_bytes(const Batch& B)
{
    const size_t mb = (size_t)B.desc.max_batch * B.lanes;
    size_t n = mb * (size_t)B.nparts * sizeof(double);
    if (B.mm_on)
        n += mb * ((size_t)B.n_lo * sizeof(float) +
                   (size_t)B.mm_nparts * sizeof(double));
    if (B.sep.on) {
        if (B.sep.cube)
            n += mb * (size_t)B.n_hi * sizeof(float);
        else
            n += mb * (size_t)B.sep.need[1] * B.sep.need[0] * sizeof(float4);
        n += mb * (size_t)(B.sep.p1stride + B.sep.p2stride) * sizeof(float);
        n += mb * (size_t)B.sep.need[1] * sizeof(uint64_t) *
             (B.sep.cube ? 1 : 2);
        n += mb * (size_t)B.sep.need[2] * B.sep.ngy;
    } else if (B.fused.on) {
        n += mb * (size_t)B.fused.sstride * sizeof(float2);
        n += mb * (size_t)B.fused.lstride * sizeof(float4);
    } else {
        n += mb * (size_t)B.n_hi * sizeof(float);
        if (B.desc.has_conv) n += mb * (size_t)B.n_c * sizeof(float2);
    }
    if (B.desc.mcdisk && B.mc_tw)
        n += (size_t)B.lanes * mc_lane_bytes(B);
    else if (B.desc.mcdisk)
        n += mb * (size_t)B.n_hi * sizeof(uint64_t);
    return n;
}

static void drop_graphs(Batch& B)
{
    for (auto& kv : B.graphs) (void)hipGraphExecDestroy(kv.second);
    B.graphs.clear();
}

// Places the work buffers, in a fixed order: views into the bound arena
// (256-byte aligned, so one arena always gets the same layout and graphs
// captured with it stay valid) or allocations of their own.  dry: only the
// arena bytes the layout needs.  The int64 cube of unbinned mcdisk relies on
// staying zero between evaluations, so it is always the batch's own.
static size_t place_work(Batch& B, bool dry)
{
    const size_t mb = (size_t)B.desc.max_batch * B.lanes;
    size_t off = 0;
    auto take = [&](DevBuf& d, size_t n) {
        off = (off + 255) & ~(size_t)255;
        if (!dry) {
            if (B.arena) d.view(B.arena + off, n);
            else d.alloc(n);
        }
        off += n;
    };
    take(B.partials, mb * B.nparts * sizeof(double));
    if (B.mm_on) {
        take(B.mm_lo, mb * (size_t)B.n_lo * sizeof(float));
        take(B.mm_partials, mb * (size_t)B.mm_nparts * sizeof(double));
    }
    if (B.sep.on) {
        if (B.sep.cube)
            take(B.hi, mb * B.n_hi * sizeof(float));
        else
            take(B.fused.lines, mb * (size_t)B.sep.need[1] *
                                    B.sep.need[0] * sizeof(float4));
        take(B.sep.P1, mb * (size_t)B.sep.p1stride * sizeof(float));
        take(B.sep.P2, mb * (size_t)B.sep.p2stride * sizeof(float));
        take(B.sep.masks, mb * (size_t)B.sep.need[1] * sizeof(uint64_t));
        if (!B.sep.cube)
            take(B.sep.segmask, mb * (size_t)B.sep.need[1] *
                                    sizeof(uint64_t));
        take(B.sep.gflags, mb * (size_t)B.sep.need[2] * B.sep.ngy);
    } else if (B.fused.on) {
        take(B.fused.S, mb * B.fused.sstride * sizeof(float2));
        take(B.fused.lines, mb * B.fused.lstride * sizeof(float4));
    } else {
        take(B.hi, mb * B.n_hi * sizeof(float));
        if (B.desc.has_conv) take(B.spec, mb * B.n_c * sizeof(float2));
    }
    if (B.desc.mcdisk && B.mc_tw) {
        take(B.mcbin, (size_t)B.lanes * mc_lane_bytes(B));
    } else if (B.desc.mcdisk && !dry) {
        // zeroed once; the conversion re-zeroes what it reads
        B.icube.alloc(mb * B.n_hi * sizeof(uint64_t));
        hipLaunchKernelGGL(batch_zero_u64, dim3(4096), dim3(256), 0,
                           B.ctx->stream, B.icube.as<unsigned long long>(),
                           (long long)(mb * B.n_hi));
        GBK_HIP(hipGetLastError());
    }
    return off;
}

static void alloc_work(Batch& B)
{
    if (B.work) return;
    if (B.arena)
        GBK_CHECK(place_work(B, true) <= B.arena_bytes, GBKHIP_ERR_ARG,
                  "the bound work arena is smaller than the batch needs "
                  "(gbkhip_batch_memory)");
    place_work(B, false);
    B.work = true;
    ++B.work_gen;
}

// Lets go of the work buffers (frees the batch's own; views of an arena
// just end) without dropping graphs.
static void drop_work_buffers(Batch& B)
{
    B.partials.release();
    B.mm_lo.release();
    B.mm_partials.release();
    B.fused.S.release();
    B.fused.lines.release();
    B.sep.P1.release();
    B.sep.P2.release();
    B.sep.masks.release();
    B.sep.segmask.release();
    B.sep.gflags.release();
    B.hi.release();
    B.spec.release();
    B.icube.release();
    B.mcbin.release();
    B.work = false;
}

// Frees the work buffers for the current binding: graphs captured with the
// batch's own buffers go with them, graphs of arenas stay (their memory is
// the caller's and keeps its layout).
static void release_binding(Batch& B)
{
    for (auto it = B.graphs.begin(); it != B.graphs.end();) {
        if (std::get<6>(it->first) == 0) {
            (void)hipGraphExecDestroy(it->second);
            it = B.graphs.erase(it);
        } else {
            ++it;
        }
    }
    drop_work_buffers(B);
}

// The layout changes (lanes, capacity, objective): every graph goes.
static void release_work(Batch& B)
{
    drop_graphs(B);          // graphs hold the buffer addresses
    B.partials.release();
    B.mm_lo.release();
    B.mm_partials.release();
    B.fused.S.release();
    B.fused.lines.release();
    B.sep.P1.release();
    B.sep.P2.release();
    B.sep.masks.release();
    B.sep.segmask.release();
    B.sep.gflags.release();
    B.hi.release();
    B.spec.release();
    B.icube.release();
    B.mcbin.release();
    B.work = false;
}

// Timing marks per chunk: the padded path records 4 (start, after smdisk,
// after convolution, after chi^2), the fused one 8 (one after each kernel).
constexpr int kMarks = 8;

static int n_marks(const Batch& b)
{
    return b.sep.on ? 6 : b.fused.on ? 8 : 4;
}

static void ensure_events(Batch& b, int chunks)
{
    while ((int)b.events.size() < kMarks * chunks) {
        hipEvent_t e;
        GBK_HIP(hipEventCreate(&e));
        b.events.push_back(e);
    }
}

static void mark(Batch& b, int chunk, int i)
{
    if (b.timing)
        GBK_HIP(hipEventRecord(b.events[kMarks * chunk + i], b.ctx->stream));
}

// Smallest 2^a 3^b 5^c >= n (even when asked).
static int smooth5_at_least(int n, bool even)
{
    for (int m = std::max(n, 2);; ++m) {
        if (even && (m & 1)) continue;
        int r = m;
        for (int p : {2, 3, 5})
            while (r % p == 0) r /= p;
        if (r == 1) return m;
    }
}

// Decides whether the spectral-first pipeline applies and prepares it:
// thin disk, convolution present, and a PSF x LSF cube that factors as
// lsf(z) psf(y, x) with (up to FP32 rounding of the product) and whose
// taps beyond the DCube edge carry no weight.  Kernel spectra are computed
// on the host in FP64 from the taps within the edge:
//   Phat[ky][kx] = sum psf(dy, dx) e^(-2 pi i (ky dy / ny + kx dx / nx))
//   Lhat[kz]     = sum lsf(dz) e^(-2 pi i kz dz / nz)
// with 1 / (sum pcube * nx ny nz) folded into Phat.
static void plan_fused(Context& c, Batch& B)
{
    const gbkhip_batch_desc& D = B.desc;
    Fused& F = B.fused;
    F.on = false;
    if (!D.has_conv || D.path == 1) return;
    const int Nx = D.size_hi[0], Ny = D.size_hi[1], Nz = D.size_hi[2];
    const size_t N = (size_t)Nx * Ny * Nz;
    std::vector<float> pc(N);
    GBK_HIP(hipMemcpy(pc.data(), D.pcube, N * sizeof(float),
                      hipMemcpyDeviceToHost));
    std::vector<double> psf((size_t)Nx * Ny, 0.0), lsf(Nz, 0.0);
    double total = 0.0, amax = 0.0;
    for (int z = 0; z < Nz; ++z)
        for (size_t i = 0; i < (size_t)Nx * Ny; ++i) {
            const double v = pc[(size_t)z * Nx * Ny + i];
            psf[i] += v;
            lsf[z] += v;
            total += v;
            amax = std::max(amax, std::fabs(v));
        }
    if (!(total > 0.0) || !std::isfinite(total)) return;
    double worst = 0.0;
    for (int z = 0; z < Nz; ++z)
        for (size_t i = 0; i < (size_t)Nx * Ny; ++i)
            worst = std::max(worst,
                             std::fabs(pc[(size_t)z * Nx * Ny + i] -
                                       lsf[z] * psf[i] / total));
    if (worst > 1e-6 * amax) return;
    // taps beyond the edge: at most 1e-9 of the absolute weight
    auto dist = [](int i, int n) { return std::min(i, n - i); };
    const int ex = D.edge_hi[0], ey = D.edge_hi[1], ez = D.edge_hi[2];
    double in_p = 0, out_p = 0, in_l = 0, out_l = 0;
    for (int y = 0; y < Ny; ++y)
        for (int x = 0; x < Nx; ++x) {
            const double a = std::fabs(psf[(size_t)y * Nx + x]);
            (dist(x, Nx) <= ex && dist(y, Ny) <= ey ? in_p : out_p) += a;
        }
    for (int z = 0; z < Nz; ++z)
        (dist(z, Nz) <= ez ? in_l : out_l) += std::fabs(lsf[z]);
    if (out_p > 1e-9 * in_p || out_l > 1e-9 * in_l) return;
    B.ka_psf = psf;
    B.ka_lsf = lsf;
    B.ka_total = total;
    // the kernel analysis above also serves the direct pipeline of thick
    // disks and mcdisk; the spectral-first deposit is for thin disks only
    if (!B.thin || D.path == 3) return;
    for (int i = 0; i < 3; ++i)
        F.need[i] = D.size_lo[i] * D.scale[i] + 2 * D.edge_hi[i];
    // FFT sizes with engine shapes: x rows, y lines, z columns of 64 or 32
    int nx = smooth5_at_least(F.need[0], false);
    int ny = smooth5_at_least(F.need[1], false);
    int nz = smooth5_at_least(F.need[2], true);
    FastShape fx = fast_shape_rows(nx), fy = fast_shape(ny);
    FastShape fz = fast_shape_lines(nz / 2, 6);
    if (!fz.ok) fz = fast_shape_lines(nz / 2, 5);
    if (!fx.ok || !fy.ok || !fz.ok) return;
    const int L = 1 << fz.lshift;
    const int sxy = D.scale[0] * D.scale[1];
    if (sxy > L) return;
    const int nzh = nz / 2 + 1;
    if ((long long)nzh * ny * ny >= (1LL << 32) ||
        (long long)nzh * nx * nx >= (1LL << 32))
        return;
    F.nx = nx;
    F.ny = ny;
    F.nz = nz;
    F.nzh = nzh;
    F.fx = fx;
    F.fy = fy;
    F.fz = fz;
    F.xcnt = D.size_lo[0] * D.scale[0];
    F.px = (nx + 15) / 16 * 16;
    F.shift = (16 - ex % 16) % 16;
    if (ex + F.shift + F.xcnt > nx) F.shift = 0;
    F.xlo = ex + F.shift;
    F.ylo = ey;
    F.ycnt = D.size_lo[1] * D.scale[1];
    F.kx = std::max(1, std::min(L / sxy, D.size_lo[0]));
    F.ky = 1;
    F.tiles_x = (D.size_lo[0] + F.kx - 1) / F.kx;
    F.tiles_y = (D.size_lo[1] + F.ky - 1) / F.ky;
    // kernel spectra
    const double scale = 1.0 / (total * (double)nx * ny * nz);
    const int wx = 2 * ex + 1, wy = 2 * ey + 1;
    std::vector<double> A((size_t)wy * nx * 2, 0.0);
    for (int dy = -ey; dy <= ey; ++dy) {
        const int iy = (dy % Ny + Ny) % Ny;
        for (int kx = 0; kx < nx; ++kx) {
            double re = 0, im = 0;
            for (int dx = -ex; dx <= ex; ++dx) {
                const int ix = (dx % Nx + Nx) % Nx;
                const double v = psf[(size_t)iy * Nx + ix];
                if (v == 0.0) continue;
                const long long m = (((long long)kx * dx) % nx + nx) % nx;
                const double ang = -2.0 * M_PI * (double)m / nx;
                re += v * std::cos(ang);
                im += v * std::sin(ang);
            }
            A[2 * ((size_t)(dy + ey) * nx + kx)] = re;
            A[2 * ((size_t)(dy + ey) * nx + kx) + 1] = im;
        }
    }
    std::vector<float> ph((size_t)ny * nx * 2);
    std::vector<double> cy(wy), sy(wy);
    for (int ky = 0; ky < ny; ++ky) {
        for (int dy = -ey; dy <= ey; ++dy) {
            const long long m = (((long long)ky * dy) % ny + ny) % ny;
            const double ang = -2.0 * M_PI * (double)m / ny;
            cy[dy + ey] = std::cos(ang);
            sy[dy + ey] = std::sin(ang);
        }
        for (int kx = 0; kx < nx; ++kx) {
            double re = 0, im = 0;
            for (int j = 0; j < wy; ++j) {
                const double ar = A[2 * ((size_t)j * nx + kx)];
                const double ai = A[2 * ((size_t)j * nx + kx) + 1];
                re += ar * cy[j] - ai * sy[j];
                im += ar * sy[j] + ai * cy[j];
            }
            ph[2 * ((size_t)ky * nx + kx)] = (float)(re * scale);
            ph[2 * ((size_t)ky * nx + kx) + 1] = (float)(im * scale);
        }
    }
    std::vector<float> lh((size_t)nzh * 2);
    for (int kz = 0; kz < nzh; ++kz) {
        double re = 0, im = 0;
        for (int dz = -ez; dz <= ez; ++dz) {
            const double v = lsf[(dz % Nz + Nz) % Nz];
            const long long m = (((long long)kz * dz) % nz + nz) % nz;
            const double ang = -2.0 * M_PI * (double)m / nz;
            re += v * std::cos(ang);
            im += v * std::sin(ang);
        }
        lh[2 * kz] = (float)re;
        lh[2 * kz + 1] = (float)im;
    }
    (void)wx;
    F.phat.alloc(ph.size() * sizeof(float));
    GBK_HIP(hipMemcpy(F.phat.ptr, ph.data(), ph.size() * sizeof(float),
                      hipMemcpyHostToDevice));
    F.lhat.alloc(lh.size() * sizeof(float));
    GBK_HIP(hipMemcpy(F.lhat.ptr, lh.data(), lh.size() * sizeof(float),
                      hipMemcpyHostToDevice));
    upload_twiddles(F.twz, nz / 2, nz / 2);
    upload_twiddles(F.tpz, nz, nzh);
    upload_twiddles(F.twy, ny, ny);
    upload_twiddles(F.twx, nx, nx);
    upload_stage_twiddles(F.twsz, nz / 2);
    upload_stage_twiddles(F.twsy, ny);
    upload_stage_twiddles(F.twsx, nx);
    F.sstride = (long long)nzh * ny * F.px;
    F.lstride = (long long)ny * nx;
    static std::once_flag once;
    std::call_once(once, [] {
        const void* ks[] = {GBK_FAST_KERNELS(fz_deposit_r2c),
                            GBK_FAST_KERNELS(fx_rows_conv),
                            GBK_FAST_KERNELS(fz_c2r_down_chi2)};
        for (const void* k : ks)
            (void)hipFuncSetAttribute(
                k, hipFuncAttributeMaxDynamicSharedMemorySize, 96 * 1024);
    });
    fft_prepare_attributes();
    (void)c;
    F.on = true;
}

// cubes: the chunk's model cubes (set set0 first), or nullptr
static void run_chunk_fused(Batch& b, const float* params, int set0,
                            int nsets, double* chi2, float* cubes, int ci)
{
    Context& c = *b.ctx;
    const gbkhip_batch_desc& D = b.desc;
    const Fused& F = b.fused;
    const float* P = params + (size_t)set0 * D.pstride;
    float2* S = F.S.as<float2>();
    const long long plane = (long long)F.ny * F.px;
    {
        float4* lines = F.lines.as<float4>();
        dim3 g0(ceil_div(F.nx, 64), ceil_div(F.ny, 4), nsets);
        hipLaunchKernelGGL(batch_lines_thin, g0, dim3(64, 4), 0, c.stream,
                           b.tmpl, b.lay, P, D.pstride, D.grid, F.need[0],
                           F.need[1], F.nx, F.ny, lines, F.lstride, 0,
                           nullptr, 1, 0, nullptr);
        GBK_HIP(hipGetLastError());
        mark(b, ci, 1);
        dim3 grid(ceil_div(F.nx, 1 << F.fz.lshift), F.ny, nsets);
        GBK_DISPATCH(GBK_SPEC_Z, F.fz, fz_deposit_r2c, grid,
                     F.fz.lds + (F.fz.a.n + 1) * sizeof(float2), c.stream,
                     lines, F.lstride, D.grid, F.fz.a, F.twz.as<float2>(),
                     F.twsz.as<float2>(), F.tpz.as<float2>(), F.need[2],
                     F.nx, F.ny, F.px, F.shift, S, F.sstride);
        GBK_HIP(hipGetLastError());
    }
    mark(b, ci, 2);
    launch_fx_lines(c, F.fy, F.twy.as<float2>(), F.twsy.as<float2>(), S, S,
                    nullptr,
                    F.nzh * F.nx, F.nx, 0, plane, F.px, 0, F.ny, 0, 1.0f,
                    nsets, F.sstride, 0);
    mark(b, ci, 3);
    {
        const int nrows = F.nzh * F.ny;
        dim3 grid(ceil_div(nrows, 1 << F.fx.lshift), nsets);
        GBK_DISPATCH(GBK_SPEC_ROWS, F.fx, fx_rows_conv, grid, F.fx.lds,
                     c.stream, S, F.sstride, nrows, F.ny,
                     div_magic((unsigned)F.ny), F.fx.a,
                     div_magic((unsigned)F.nx), F.twx.as<float2>(),
                     F.twsx.as<float2>(), F.phat.as<float2>(),
                          F.lhat.as<float2>(), F.xlo, F.xcnt,
                          div_magic((unsigned)F.xcnt), F.px);
        GBK_HIP(hipGetLastError());
    }
    mark(b, ci, 4);
    launch_fx_lines(c, F.fy, F.twy.as<float2>(), F.twsy.as<float2>(), S, S,
                    nullptr,
                    F.nzh * F.xcnt, F.xcnt, F.xlo, plane, F.px, F.ylo,
                    F.ylo + F.ycnt, 1, 1.0f, nsets, F.sstride, 0);
    mark(b, ci, 5);
    DataArgs da;
    da.obs = D.obs;
    da.err = D.err;
    da.mask = D.mask;
    da.err_scalar = D.err_scalar;
    da.inv_err_scalar = 1.0f / D.err_scalar;
    da.weight = D.weight;
    {
        dim3 grid(F.tiles_x * F.tiles_y, nsets);
        GBK_DISPATCH(GBK_SPEC_Z, F.fz, fz_c2r_down_chi2, grid,
                     F.fz.lds + (F.fz.a.n + 1) * sizeof(float2), c.stream, S,
                     F.sstride, F.px, F.ny, F.fz.a, F.twz.as<float2>(),
                     F.twsz.as<float2>(), F.tpz.as<float2>(), F.xlo,
                          F.ylo, F.kx, F.ky, D.scale[0], D.scale[1],
                          D.scale[2], D.edge_hi[2], D.size_lo[0],
                          D.size_lo[1], D.size_lo[2], F.tiles_x,
                          div_magic((unsigned)(F.kx * F.ky)),
                          div_magic((unsigned)F.kx), da, cubes,
                          b.n_lo, b.partials.as<double>(), b.nparts);
        GBK_HIP(hipGetLastError());
    }
    mark(b, ci, 6);
    hipLaunchKernelGGL(batch_chi2_final, dim3(nsets), dim3(256), 0, c.stream,
                       b.partials.as<double>(), b.nparts, chi2, set0);
    GBK_HIP(hipGetLastError());
    mark(b, ci, 7);
}

static int tap_bucket(int n)
{
    for (int b : {8, 12, 16, 24, 32, 44, 48, 64, 96, 128})
        if (n <= b) return b;
    return 0;
}

// Output d of a block sits at padded index s X + d + e and sees input
// s X + u through the kernel offset t = d + e - u (c(q) = sum_t k(t)
// m(q - t), k wrapped so k(t) = k[t mod N]):
//   W[u] = (1/s) sum_{d < s, |d + e - u| <= e} k[(d + e - u) mod N] / norm,
// u in [0, s + 2e).
static std::vector<double> block_taps(const std::vector<double>& k, int N,
                                      int e, int s, double norm)
{
    std::vector<double> w(s + 2 * e, 0.0);
    for (int u = 0; u < s + 2 * e; ++u) {
        double acc = 0.0;
        for (int d = 0; d < s; ++d) {
            const int t = d + e - u;
            if (t < -e || t > e) continue;
            acc += k[(t % N + N) % N] / norm;
        }
        w[u] = acc / s;
    }
    return w;
}

// The taps an output needs: W without its negligible ends.  Leading and
// trailing taps are dropped while each end's sum of |W| stays within 2^-31
// of the total, so the dropped terms of any output add up to at most 2^-30
// of the largest input in its window: 2^6 below FP32's unit roundoff, far
// below the rounding of the FP32 sum itself and of an FFT convolution
// (whose error is spread over the whole cube at ~1e-7 of its maximum).  A
// Gaussian cut at 8 sigma keeps about +-5.7 sigma.  Returns the taps
// zero-padded to their bucket and the count dropped in front (the offset
// of an output's first tap); bucket 0 when none fits.
static std::vector<float> trimmed_taps(const std::vector<double>& w,
                                       int& bucket, int& lead, int& kept)
{
    const int n = (int)w.size();
    double tot = 0.0;
    for (double v : w) tot += std::fabs(v);
    const double cut = std::ldexp(tot, -31);
    int a = 0, b = n;
    for (double acc = 0.0; a < n - 1; ++a) {
        acc += std::fabs(w[a]);
        if (acc > cut) break;
    }
    for (double acc = 0.0; b > a + 1; --b) {
        acc += std::fabs(w[b - 1]);
        if (acc > cut) break;
    }
    lead = a;
    kept = b - a;
    bucket = tap_bucket(b - a);
    std::vector<float> out(std::max(bucket, 0), 0.0f);
    for (int u = a; u < b && bucket; ++u) out[u - a] = (float)w[u];
    return out;
}

// Direct separable pipeline: needs the kernel analysis of plan_fused and a
// PSF that factors as px(x) py(y) (a Gaussian cut at 8 sigma does, to
// ~1e-14).  Returns false when it does not apply.
static bool plan_sep(Batch& B)
{
    const gbkhip_batch_desc& D = B.desc;
    Sep& P = B.sep;
    P.on = false;
    if (B.ka_psf.empty()) return false;
    const int Nx = D.size_hi[0], Ny = D.size_hi[1], Nz = D.size_hi[2];
    const std::vector<double>& psf = B.ka_psf;
    std::vector<double> px(Nx, 0.0), py(Ny, 0.0);
    double ptot = 0.0, pmax = 0.0;
    for (int y = 0; y < Ny; ++y)
        for (int x = 0; x < Nx; ++x) {
            const double v = psf[(size_t)y * Nx + x];
            px[x] += v;
            py[y] += v;
            ptot += v;
            pmax = std::max(pmax, std::fabs(v));
        }
    if (!(ptot > 0.0)) return false;
    double worst = 0.0;
    for (int y = 0; y < Ny; ++y)
        for (int x = 0; x < Nx; ++x)
            worst = std::max(worst, std::fabs(psf[(size_t)y * Nx + x] -
                                              px[x] * py[y] / ptot));
    if (worst > 1e-6 * pmax) return false;
    const int ex = D.edge_hi[0], ey = D.edge_hi[1], ez = D.edge_hi[2];
    const int sx = D.scale[0], sy = D.scale[1], sz = D.scale[2];
    if (sx < 1 || sx > 3 || sy < 1 || sy > 3 || sz < 1 || sz > 3) return false;
    if (D.grid.spec_size > 65535) return false;    // 16-bit channel windows
    // the taps lose their negligible ends (the passes read from offsets
    // ox, oy, oz)
    const std::vector<float> wx =
        trimmed_taps(block_taps(px, Nx, ex, sx, ptot), P.tx, P.ox, P.kept[0]);
    const std::vector<float> wy =
        trimmed_taps(block_taps(py, Ny, ey, sy, ptot), P.ty, P.oy, P.kept[1]);
    const std::vector<float> wz =
        trimmed_taps(block_taps(B.ka_lsf, Nz, ez, sz, B.ka_total), P.tz,
                     P.oz, P.kept[2]);
    if (!P.tx || !P.ty || !P.tz) return false;
    for (int i = 0; i < 3; ++i)
        P.need[i] = D.size_lo[i] * D.scale[i] + 2 * D.edge_hi[i];
    auto up = [](DevBuf& b, const std::vector<float>& h) {
        b.alloc(h.size() * sizeof(float));
        GBK_HIP(hipMemcpy(b.ptr, h.data(), h.size() * sizeof(float),
                          hipMemcpyHostToDevice));
    };
    up(P.wx, wx);
    up(P.wy, wy);
    up(P.wz, wz);
    const int nxl = D.size_lo[0], nyl = D.size_lo[1], nzl = D.size_lo[2];
    const int jx = sep_jx(P.tx), jy = sep_jy(P.ty);
    const int nxb = (nxl + jx - 1) / jx;
    // the tile holds every deposited column and every filter input
    P.mp = std::max(sx * nxb * jx + P.ox + P.tx, P.need[0]) | 1;
    // channels per S1 block: at most 16 (the lanes of a group) in a tile of
    // at most 64 KB, spread evenly over the chunks
    const int zmax = std::min(kSepZC, (int)((64 * 1024) / (P.mp * 4)));
    if (zmax < 1) return false;
    const int nchunks = (P.need[2] + zmax - 1) / zmax;
    P.zc = (P.need[2] + nchunks - 1) / nchunks;
    P.nchunks = (P.need[2] + P.zc - 1) / P.zc;
    if (P.nchunks > 64) return false;      // 64-bit row masks
    // Y outputs per S2 block: all of them when the rows fit 48 KB
    const int max_rows = kSepMaxRows;
    const int gmax = ((max_rows - P.ty) / sy + 1) / jy;
    if (gmax < 1) return false;
    P.yblk = std::min((nyl + jy - 1) / jy, gmax) * jy;
    P.yblk_full = P.yblk;
    P.ngy = (nyl + jy - 1) / jy;
    P.p1stride = (long long)P.need[2] * P.need[1] * nxl;
    P.p2stride = (long long)P.need[2] * nyl * nxl;
    // the z pass addresses a set's P2 and the data cube with 32-bit byte
    // offsets
    if ((double)P.p2stride * 4 >= 4294967295.0 ||
        (double)nzl * nyl * nxl * 4 >= 4294967295.0)
        return false;
    P.zblocks = (nzl + kSepJZ - 1) / kSepJZ;
    return true;
}

// Multiply-adds per set of the direct pipeline and (complex butterfly
// work, 6 flops per element per radix-2 level) of the fused FFT one.
static bool sep_cheaper(const Batch& B)
{
    const gbkhip_batch_desc& D = B.desc;
    const Sep& P = B.sep;
    const double nxl = D.size_lo[0], nyl = D.size_lo[1], nzl = D.size_lo[2];
    const double direct = (double)P.need[2] * P.need[1] * nxl * P.tx +
                          (double)P.need[2] * nyl * nxl * P.ty +
                          nzl * nyl * nxl * P.tz;
    if (!B.fused.on) return true;
    const Fused& F = B.fused;
    const double nc = (double)F.nzh * F.ny * F.nx;
    const double fft = 6.0 * nc * (2 * std::log2((double)F.nx) +
                                   2 * std::log2((double)F.ny) +
                                   std::log2((double)F.nz));
    return direct < fft;
}

#define GBK_SEP_CASE(KERNEL, S_, T_, SEL, ...)                               \
    case (S_) * 256 + (T_):                                                  \
        hipLaunchKernelGGL((KERNEL<S_, T_>), __VA_ARGS__);                   \
        break;
#define GBK_SEP_DISPATCH(KERNEL, S, TAPS, ...)                               \
    switch ((S) * 256 + (TAPS)) {                                            \
    GBK_SEP_CASE(KERNEL, 1, 8, 0, __VA_ARGS__)                               \
    GBK_SEP_CASE(KERNEL, 1, 12, 0, __VA_ARGS__)                              \
    GBK_SEP_CASE(KERNEL, 1, 16, 0, __VA_ARGS__)                              \
    GBK_SEP_CASE(KERNEL, 1, 24, 0, __VA_ARGS__)                              \
    GBK_SEP_CASE(KERNEL, 1, 32, 0, __VA_ARGS__)                              \
    GBK_SEP_CASE(KERNEL, 1, 44, 0, __VA_ARGS__)                              \
    GBK_SEP_CASE(KERNEL, 1, 48, 0, __VA_ARGS__)                              \
    GBK_SEP_CASE(KERNEL, 1, 64, 0, __VA_ARGS__)                              \
    GBK_SEP_CASE(KERNEL, 1, 96, 0, __VA_ARGS__)                              \
    GBK_SEP_CASE(KERNEL, 1, 128, 0, __VA_ARGS__)                             \
    GBK_SEP_CASE(KERNEL, 2, 8, 0, __VA_ARGS__)                               \
    GBK_SEP_CASE(KERNEL, 2, 12, 0, __VA_ARGS__)                              \
    GBK_SEP_CASE(KERNEL, 2, 16, 0, __VA_ARGS__)                              \
    GBK_SEP_CASE(KERNEL, 2, 24, 0, __VA_ARGS__)                              \
    GBK_SEP_CASE(KERNEL, 2, 32, 0, __VA_ARGS__)                              \
    GBK_SEP_CASE(KERNEL, 2, 44, 0, __VA_ARGS__)                              \
    GBK_SEP_CASE(KERNEL, 2, 48, 0, __VA_ARGS__)                              \
    GBK_SEP_CASE(KERNEL, 2, 64, 0, __VA_ARGS__)                              \
    GBK_SEP_CASE(KERNEL, 2, 96, 0, __VA_ARGS__)                              \
    GBK_SEP_CASE(KERNEL, 2, 128, 0, __VA_ARGS__)                             \
    GBK_SEP_CASE(KERNEL, 3, 8, 0, __VA_ARGS__)                               \
    GBK_SEP_CASE(KERNEL, 3, 12, 0, __VA_ARGS__)                              \
    GBK_SEP_CASE(KERNEL, 3, 16, 0, __VA_ARGS__)                              \
    GBK_SEP_CASE(KERNEL, 3, 24, 0, __VA_ARGS__)                              \
    GBK_SEP_CASE(KERNEL, 3, 32, 0, __VA_ARGS__)                              \
    GBK_SEP_CASE(KERNEL, 3, 44, 0, __VA_ARGS__)                              \
    GBK_SEP_CASE(KERNEL, 3, 48, 0, __VA_ARGS__)                              \
    GBK_SEP_CASE(KERNEL, 3, 64, 0, __VA_ARGS__)                              \
    GBK_SEP_CASE(KERNEL, 3, 96, 0, __VA_ARGS__)                              \
    GBK_SEP_CASE(KERNEL, 3, 128, 0, __VA_ARGS__)                             \
    default:                                                                 \
        throw Error{GBKHIP_ERR_INTERNAL, "no separable kernel for shape"};   \
    }

// Hi-res model cubes of sets [set0, set0 + nsets) of a chunk on `lane` into
// hi (n_hi per set): mcdisk (binned deposits when a tile spectrum fits LDS,
// else int64 atomics into a fixed-point cube) or a thick smdisk.
// smdisk voxels of one disk into the sets' hi-res cubes (assign: write,
// zero where no line lands; else add).  The direct pipeline reads only its
// inputs [0, need) of each axis; the padded-cube pipeline the whole grid.
static void launch_thick(Batch& b, const gbkhip_disk& tmpl,
                         const SetLayout& lay, const float* P, int nsets,
                         float* hi, int assign)
{
    Context& c = *b.ctx;
    const gbkhip_batch_desc& D = b.desc;
    const gbkhip_grid& g = D.grid;
    const bool part = b.sep.on && b.sep.cube;
    const int nxe = part ? b.sep.need[0] : g.spat_size[0];
    const int nye = part ? b.sep.need[1] : g.spat_size[1];
    const int nze = part ? b.sep.need[2] : g.spec_size;
    dim3 grid(ceil_div(nxe, 64), nye, nsets);
    const int nz = g.spat_size[2];
    if (nz <= 16)
        hipLaunchKernelGGL(batch_smdisk_thick<4>, grid, dim3(256), 0,
                           c.stream, tmpl, lay, P, D.pstride, g, hi, b.n_hi,
                           nxe, nze, assign);
    else if (nz <= 32)
        hipLaunchKernelGGL(batch_smdisk_thick<8>, grid, dim3(256), 0,
                           c.stream, tmpl, lay, P, D.pstride, g, hi, b.n_hi,
                           nxe, nze, assign);
    else
        hipLaunchKernelGGL(batch_smdisk_thick<16>, grid, dim3(256), 0,
                           c.stream, tmpl, lay, P, D.pstride, g, hi, b.n_hi,
                           nxe, nze, assign);
    GBK_HIP(hipGetLastError());
}

// The components after the first, added into the cubes in order.
static void deposit_extras(Batch& b, const float* P, int nsets, float* hi)
{
    for (const auto& e : b.extras)
        launch_thick(b, e->tmpl, e->lay, P, nsets, hi, 0);
}

static void deposit_hi(Batch& b, const float* P, int set0, int nsets,
                       int lane, float* hi)
{
    Context& c = *b.ctx;
    const gbkhip_batch_desc& D = b.desc;
    const gbkhip_grid& g = D.grid;
    if (D.mcdisk && b.mc_tw) {
        GBK_CHECK(b.mc_cap > 0, GBKHIP_ERR_ARG,
                  "mcdisk batch: reserve cloud capacity first");
        const long long cap = b.mc_cap;
        const int nt = b.mc_ntiles;
        char* base = static_cast<char*>(b.mcbin.ptr) +
                     (size_t)lane * mc_lane_bytes(b);
        McBufs bb;
        mc_bufs(base, D.max_batch, cap, nt, b.mc_tw, g.spec_size, &bb);
        McBin bin;
        bin.rec = bb.rec;
        bin.tr = bb.tr;
        bin.tcount = bb.tcount;
        bin.tw = bin.th = b.mc_tw;
        bin.ntx = b.mc_ntx;
        hipLaunchKernelGGL(batch_zero_i32, dim3(ceil_div(nsets * nt, 256)),
                           dim3(256), 0, c.stream, bin.tcount,
                           (long long)nsets * nt);
        hipLaunchKernelGGL(batch_zero_f32, dim3(4096), dim3(256), 0,
                           c.stream, hi, (long long)nsets * b.n_hi);
        const size_t lds =
            mc_bin_lds(D.csum_len, D.nrnodes, nt).total() * sizeof(int);
        const int bx = std::max(1, std::min(ceil_div(cap, 256),
                                            8192 / std::max(nsets, 1)));
        hipLaunchKernelGGL(batch_mcdisk_bin, dim3(bx, nsets), dim3(256), lds,
                           c.stream, b.tmpl, b.lay, P, D.pstride, g, D.cflux,
                           b.csum + (size_t)set0 * D.csum_len, D.csum_len,
                           D.hoi, (uint32_t)D.seed, (uint32_t)(D.seed >> 32),
                           bin, cap, nt, b.overflow.as<int>());
        GBK_HIP(hipGetLastError());
        launch_mc_binned_tail(c, bin, bb, nsets, 0,
                              b.csum + (size_t)set0 * D.csum_len, D.csum_len,
                              g, b.qscale + set0, 1.0, hi, b.n_hi);
    } else if (D.mcdisk) {
        // the same 256-block x 256-thread cloud loop for every set
        unsigned long long* icube = b.icube.as<unsigned long long>() +
                                    (size_t)lane * D.max_batch * b.n_hi;
        const size_t lds = D.csum_len <= kMcCsumLds
                               ? D.csum_len * sizeof(int) : 0;
        hipLaunchKernelGGL(batch_mcdisk, dim3(256, nsets), dim3(256), lds,
                           c.stream, b.tmpl, b.lay, P, D.pstride, g, D.cflux,
                           b.csum + (size_t)set0 * D.csum_len, D.csum_len,
                           D.hoi, (uint32_t)D.seed, (uint32_t)(D.seed >> 32),
                           b.qscale + set0, icube, b.n_hi);
        GBK_HIP(hipGetLastError());
        hipLaunchKernelGGL(batch_fixed_to_hi, dim3(1024, nsets), dim3(256), 0,
                           c.stream, icube, hi, b.n_hi, b.qscale + set0);
    } else {
        launch_thick(b, b.tmpl, b.lay, P, nsets, hi, 1);
    }
    GBK_HIP(hipGetLastError());
}

// cubes: the chunk's model cubes (set set0 first), or nullptr
static void run_chunk_sep(Batch& b, const float* params, int set0,
                          int nsets, double* chi2, float* cubes, int ci,
                          int lane)
{
    Context& c = *b.ctx;
    const gbkhip_batch_desc& D = b.desc;
    const Sep& P = b.sep;
    const float* par = params + (size_t)set0 * D.pstride;
    const int nxl = D.size_lo[0], nyl = D.size_lo[1], nzl = D.size_lo[2];
    const int nx = P.need[0], ny = P.need[1], nz = P.need[2];
    const size_t mb = (size_t)D.max_batch * lane;    // sets before this lane
    const long long lstride = (long long)ny * nx;
    float4* lines = P.cube ? nullptr
                           : b.fused.lines.as<float4>() + mb * lstride;
    float* P1 = P.P1.as<float>() + mb * P.p1stride;
    float* P2 = P.P2.as<float>() + mb * P.p2stride;
    double* partials = b.partials.as<double>() + mb * b.nparts;
    unsigned long long* masks = P.masks.as<unsigned long long>() +
                                mb * (size_t)ny;
    unsigned char* gflags = P.gflags.as<unsigned char>() +
                            mb * (size_t)nz * P.ngy;
    if (lane == 0) b.sep.last_sets = nsets;
    if (P.cube) {
        // model deposited into the hi-res cube; the x pass sets the masks
        float* hi = b.hi.as<float>() + mb * b.n_hi;
        hipLaunchKernelGGL(sep_zero_masks, dim3(ceil_div(nsets * ny, 256)),
                           dim3(256), 0, c.stream, masks,
                           (long long)nsets * ny, nullptr);
        GBK_HIP(hipGetLastError());
        deposit_hi(b, par, set0, nsets, lane, hi);
        deposit_extras(b, par, nsets, hi);
        mark(b, ci, 1);
        dim3 grid(ny, ceil_div(nz, P.zc), nsets);
        const size_t lds = (size_t)P.zc * P.mp * sizeof(float);
        GBK_SEP_DISPATCH(sep_xpass_cube, D.scale[0], P.tx, grid,
                         dim3(kSepThreads), lds, c.stream, hi, b.n_hi,
                         D.size_hi[0], D.size_hi[1], P.wx.as<float>(), P.ox,
                         nx, nz, nxl, P.mp, P.zc, P1, P.p1stride, ny, masks);
        GBK_HIP(hipGetLastError());
    } else {
        unsigned long long* segs = P.segmask.as<unsigned long long>() +
                                   mb * (size_t)ny;
        hipLaunchKernelGGL(sep_zero_masks, dim3(ceil_div(nsets * ny, 256)),
                           dim3(256), 0, c.stream, masks,
                           (long long)nsets * ny, segs);
        GBK_HIP(hipGetLastError());
        dim3 g0(ceil_div(nx, 64), ceil_div(ny, 4), nsets);
        hipLaunchKernelGGL(batch_lines_thin, g0, dim3(64, 4), 0, c.stream,
                           b.tmpl, b.lay, par, D.pstride, D.grid, nx, ny, nx,
                           ny, lines, lstride, 1, masks, P.zc, nz, segs);
        GBK_HIP(hipGetLastError());
        mark(b, ci, 1);
        dim3 grid(ny, 1, nsets);
        const size_t lds = (size_t)P.zc * P.mp * sizeof(float);
        GBK_SEP_DISPATCH(sep_xpass, D.scale[0], P.tx, grid, dim3(kSepThreads),
                         lds, c.stream, lines, lstride, nx, D.grid,
                         P.wx.as<float>(), P.ox, nx, nz, nxl, P.mp, P.zc,
                         P1, P.p1stride, ny, masks, segs);
        GBK_HIP(hipGetLastError());
    }
    mark(b, ci, 2);
    // the y and z passes put (channel or channel block, set) on grid z:
    // launched over slices of sets that keep it within the device limit
    auto slices = [&](int per_set, auto&& launch) {
        const int step = std::max(1, b.max_grid_z / per_set);
        for (int s0 = 0; s0 < nsets; s0 += step)
            launch(s0, std::min(step, nsets - s0));
    };
    // y-pass blocks per set: per x-pass chunk, its channels kSepYChannels
    // at a time
    const int zpb = std::min(kSepYChannels, P.zc);
    const int nzb = P.nchunks * ceil_div(P.zc, zpb);
    slices(nzb, [&](int s0, int ns) {
        dim3 grid(ceil_div(nxl, kSepXW), ceil_div(nyl, P.yblk), nzb * ns);
        const int rows = D.scale[1] * (P.yblk - 1) + P.ty;
        const size_t lds = (size_t)rows * kSepYPitch * sizeof(float);
        GBK_SEP_DISPATCH(sep_ypass, D.scale[1], P.ty, grid, dim3(kSepThreads),
                         lds, c.stream, P1 + s0 * P.p1stride, P.p1stride,
                         P.wy.as<float>(), P.oy, nxl, nyl, ny, nz, P.yblk,
                         P2 + s0 * P.p2stride, P.p2stride,
                         masks + (size_t)s0 * ny, P.zc, zpb,
                         gflags + (size_t)s0 * nz * P.ngy, P.ngy,
                         P.zrow.as<float4>());
        GBK_HIP(hipGetLastError());
    });
    mark(b, ci, 3);
    DataArgs da;
    da.obs = D.obs;
    da.err = D.err;
    da.mask = D.mask;
    da.err_scalar = D.err_scalar;
    da.inv_err_scalar = 1.0f / D.err_scalar;
    da.weight = D.weight;
    slices(P.zblocks, [&](int s0, int ns) {
        dim3 grid(ceil_div(nxl, 64), ceil_div(nyl, kSepThreads / 64),
                  P.zblocks * ns);
        GBK_SEP_DISPATCH(sep_zpass, D.scale[2], P.tz, grid, dim3(kSepThreads),
                         0, c.stream, P2 + s0 * P.p2stride, P.p2stride,
                         P.wz.as<float>(), P.oz, nxl, nyl, nzl, nz,
                         P.zblocks, da,
                         cubes ? cubes + (size_t)s0 * b.n_lo : nullptr,
                         b.n_lo, partials + (size_t)s0 * b.nparts, b.nparts,
                         gflags + (size_t)s0 * nz * P.ngy, P.ngy,
                         sep_jy(P.ty), P.zero_part.as<double>(),
                         P.zrow.as<float>());
        GBK_HIP(hipGetLastError());
    });
    mark(b, ci, 4);
    hipLaunchKernelGGL(batch_chi2_final, dim3(nsets), dim3(256), 0, c.stream,
                       partials, b.nparts, chi2, set0);
    GBK_HIP(hipGetLastError());
    mark(b, ci, 5);
}

// The z pass's per-block partial sums of r^2 for a model that is zero
// everywhere (data only): blocks with no live input take them instead of
// reading the data again.  Computed once by the z pass itself with every
// input flagged dead, so the bits are those it would produce.
static void compute_zero_partials(Batch& b)
{
    Context& c = *b.ctx;
    const gbkhip_batch_desc& D = b.desc;
    Sep& P = b.sep;
    const int nxl = D.size_lo[0], nyl = D.size_lo[1], nzl = D.size_lo[2];
    const int nz = P.need[2];
    P.zero_part.alloc((size_t)b.nparts * sizeof(double));
    P.zrow.alloc(64 * sizeof(float));
    GBK_HIP(hipMemsetAsync(P.zrow.ptr, 0, 64 * sizeof(float), c.stream));
    DevBuf dead((size_t)nz * P.ngy);
    GBK_HIP(hipMemsetAsync(dead.ptr, 0, (size_t)nz * P.ngy, c.stream));
    DataArgs da;
    da.obs = D.obs;
    da.err = D.err;
    da.mask = D.mask;
    da.err_scalar = D.err_scalar;
    da.inv_err_scalar = 1.0f / D.err_scalar;
    da.weight = D.weight;
    dim3 grid(ceil_div(nxl, 64), ceil_div(nyl, kSepThreads / 64), P.zblocks);
    GBK_SEP_DISPATCH(sep_zpass, D.scale[2], P.tz, grid, dim3(kSepThreads), 0,
                     c.stream, P.P2.as<float>(), P.p2stride,
                     P.wz.as<float>(), P.oz, nxl, nyl, nzl, nz, P.zblocks,
                     da, nullptr, b.n_lo, P.zero_part.as<double>(), b.nparts,
                     dead.as<unsigned char>(), P.ngy, sep_jy(P.ty), nullptr,
                     P.zrow.as<float>());
    GBK_HIP(hipGetLastError());
    GBK_HIP(hipStreamSynchronize(c.stream));
}

// The padded-cube pipeline; cubes: the chunk's model cubes or nullptr.
static void run_chunk_padded(Batch& b, const float* params, int set0,
                             int nsets, double* chi2, float* cubes, int ci)
{
    Context& c = *b.ctx;
    const gbkhip_batch_desc& D = b.desc;
    const gbkhip_grid& g = D.grid;
    const float* P = params + (size_t)set0 * D.pstride;
    float* hi = b.hi.as<float>();
    if (D.mcdisk || !b.thin) {
        deposit_hi(b, P, set0, nsets, 0, hi);
    } else {
        dim3 grid(ceil_div(g.spat_size[0], 64), ceil_div(g.spat_size[1], 4),
                  nsets);
        hipLaunchKernelGGL(batch_smdisk_thin, grid, dim3(64, 4), 0, c.stream,
                           b.tmpl, b.lay, P, D.pstride, g, hi, b.n_hi);
        GBK_HIP(hipGetLastError());
    }
    deposit_extras(b, P, nsets, hi);
    mark(b, ci, 1);
    if (D.has_conv)
        fft_convolve_with(c, *b.plan, hi, b.spec.as<float2>(),
                          b.spectrum.as<float2>(), nsets, b.n_hi, b.n_c, 0);
    mark(b, ci, 2);
    DataArgs da;
    da.obs = D.obs;
    da.err = D.err;
    da.mask = D.mask;
    da.err_scalar = D.err_scalar;
    da.inv_err_scalar = 1.0f / D.err_scalar;
    da.weight = D.weight;
    dim3 grid(b.nparts, nsets);
    hipLaunchKernelGGL(batch_downscale_chi2, grid, dim3(256), 0, c.stream, hi,
                       b.n_hi, D.size_hi[0], D.size_hi[1], D.scale[0],
                       D.scale[1], D.scale[2], D.edge_hi[0], D.edge_hi[1],
                       D.edge_hi[2], D.size_lo[0], D.size_lo[1],
                       D.size_lo[2], da, cubes,
                       b.n_lo, b.partials.as<double>(), b.nparts);
    GBK_HIP(hipGetLastError());
    hipLaunchKernelGGL(batch_chi2_final, dim3(nsets), dim3(256), 0, c.stream,
                       b.partials.as<double>(), b.nparts, chi2, set0);
    GBK_HIP(hipGetLastError());
    mark(b, ci, 3);
}

static void run_chunk(Batch& b, const float* params, int set0, int nsets,
                      double* chi2, float* cubes, int lane)
{
    Context& c = *b.ctx;
    const int ci = set0 / b.desc.max_batch;
    const size_t mb = (size_t)b.desc.max_batch * lane;   // sets before lane
    // the chunk's model cubes: the caller's, or (moment-map objective) the
    // lane's own buffer
    float* lo = cubes ? cubes + (size_t)set0 * b.n_lo
                      : b.mm_on ? b.mm_lo.as<float>() + mb * b.n_lo
                                : nullptr;
    mark(b, ci, 0);
    if (b.sep.on)
        run_chunk_sep(b, params, set0, nsets, chi2, lo, ci, lane);
    else if (b.fused.on)
        run_chunk_fused(b, params, set0, nsets, chi2, lo, ci);
    else
        run_chunk_padded(b, params, set0, nsets, chi2, lo, ci);
    if (b.mm_on) {
        // the moment-map chi^2 replaces the (empty) cube residual's
        double* part = b.mm_partials.as<double>() + mb * b.mm_nparts;
        const gbkhip_batch_desc& D = b.desc;
        hipLaunchKernelGGL(batch_mmaps_chi2, dim3(b.mm_nparts, nsets),
                           dim3(256), 0, c.stream, lo, b.n_lo, D.size_lo[0],
                           D.size_lo[1], D.size_lo[2], b.mm, part,
                           b.mm_nparts);
        GBK_HIP(hipGetLastError());
        hipLaunchKernelGGL(batch_chi2_final, dim3(nsets), dim3(256), 0,
                           c.stream, part, b.mm_nparts, chi2, set0);
        GBK_HIP(hipGetLastError());
    }
}

static void ensure_lane_streams(Batch& B)
{
    if (B.aux) return;
    GBK_HIP(hipStreamCreateWithFlags(&B.aux, hipStreamNonBlocking));
    GBK_HIP(hipEventCreateWithFlags(&B.ev_fork, hipEventDisableTiming));
    GBK_HIP(hipEventCreateWithFlags(&B.ev_join, hipEventDisableTiming));
}

static void run_all(Batch& b, const float* params, int nsets, double* chi2,
                    float* cubes)
{
    Context& c = *b.ctx;
    const int chunks = (nsets + b.desc.max_batch - 1) / b.desc.max_batch;
    if (b.timing) ensure_events(b, chunks);
    b.timed_chunks = b.timing ? chunks : 0;
    // per-kernel timing measures each kernel alone: one lane then
    const bool two = b.lanes > 1 && chunks > 1 && !b.timing;
    const hipStream_t main = c.stream;
    if (two) {
        GBK_HIP(hipEventRecord(b.ev_fork, main));
        GBK_HIP(hipStreamWaitEvent(b.aux, b.ev_fork, 0));
    }
    struct Restore {
        Context& c;
        hipStream_t s;
        ~Restore() { c.stream = s; }
    } restore{c, main};
    for (int s0 = 0, ci = 0; s0 < nsets; s0 += b.desc.max_batch, ++ci) {
        const int lane = two ? ci & 1 : 0;
        c.stream = lane ? b.aux : main;
        run_chunk(b, params, s0, std::min(b.desc.max_batch, nsets - s0), chi2,
                  cubes, lane);
    }
    c.stream = main;
    if (two) {
        GBK_HIP(hipEventRecord(b.ev_join, b.aux));
        GBK_HIP(hipStreamWaitEvent(main, b.ev_join, 0));
    }
}

}  // namespace gbkhip

struct gbkhip_batch {
    gbkhip::Batch impl;
};

// A ragged group: member batches (galaxies of different geometry, model
// and data) launched together, on up to kGroupBranches parallel streams,
// captured in one HIP graph per set of call pointers.
struct gbkhip_group {
    gbkhip::Context* ctx = nullptr;
    std::vector<gbkhip_batch*> members;
    std::vector<hipStream_t> branch;
    std::vector<hipEvent_t> ev_join;
    hipEvent_t ev_fork = nullptr;
    hipStream_t cap = nullptr;
    std::map<std::vector<uintptr_t>, hipGraphExec_t> graphs;
    ~gbkhip_group()
    {
        for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
        for (auto s : branch) (void)hipStreamDestroy(s);
        for (auto e : ev_join) (void)hipEventDestroy(e);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (cap) (void)hipStreamDestroy(cap);
    }
};

using namespace gbkhip;

extern "C" {

int gbkhip_batch_create(gbkhip_ctx* ctx, const gbkhip_batch_desc* desc,
                        gbkhip_batch** out)
{
    return guarded([&] {
        Context* c = ctx_of(ctx);
        GBK_CHECK(desc && out, GBKHIP_ERR_ARG, "null argument");
        const gbkhip_batch_desc& D = *desc;
        GBK_CHECK(D.nrnodes >= 2 && D.rnodes, GBKHIP_ERR_ARG,
                  "need at least two rnodes");
        GBK_CHECK(D.max_batch >= 1, GBKHIP_ERR_ARG, "max_batch must be >= 1");
        GBK_CHECK(D.pstride >= 1, GBKHIP_ERR_ARG, "pstride must be >= 1");
        for (int i = 0; i < 3; ++i) {
            GBK_CHECK(D.size_lo[i] > 0 && D.size_hi[i] > 0 && D.scale[i] > 0,
                      GBKHIP_ERR_ARG, "sizes must be positive");
            GBK_CHECK(D.edge_hi[i] >= 0 && D.edge_hi[i] + D.size_lo[i] *
                      D.scale[i] <= D.size_hi[i], GBKHIP_ERR_ARG,
                      "downscale window exceeds the padded cube");
        }
        GBK_CHECK(D.mcdisk || D.traits[1].n == 0 ||
                  D.grid.spec_size <= 65535,
                  GBKHIP_ERR_UNSUPPORTED,
                  "thick disks: at most 65535 channels");
        GBK_CHECK(D.grid.spat_size[0] == D.size_hi[0] &&
                  D.grid.spat_size[1] == D.size_hi[1] &&
                  D.grid.spec_size == D.size_hi[2], GBKHIP_ERR_ARG,
                  "gmodel grid must cover the padded cube");
        int gy = 0, gz = 0;
        GBK_HIP(hipDeviceGetAttribute(&gy, hipDeviceAttributeMaxGridDimY,
                                      c->device));
        GBK_HIP(hipDeviceGetAttribute(&gz, hipDeviceAttributeMaxGridDimZ,
                                      c->device));
        // one block row per set on grid y or z of the per-set kernels
        GBK_CHECK(D.max_batch <= std::min(gy, gz), GBKHIP_ERR_ARG,
                  "max_batch exceeds the device's grid y/z limit");
        auto* b = new gbkhip_batch();
        Batch& B = b->impl;
        B.ctx = c;
        B.device = c->device;
        B.desc = D;
        B.max_grid_z = gz;
        try {
            B.overflow.alloc(sizeof(int));
            GBK_HIP(hipMemset(B.overflow.ptr, 0, sizeof(int)));
            build_layout(B);
            const int nx = D.size_hi[0], ny = D.size_hi[1], nz = D.size_hi[2];
            B.n_hi = (long long)nx * ny * nz;
            B.n_lo = (long long)D.size_lo[0] * D.size_lo[1] * D.size_lo[2];
            B.nparts = (int)std::max<long long>(
                1, std::min<long long>(512, (B.n_lo + 4095) / 4096));
            GBK_CHECK(D.path >= 0 && D.path <= 3, GBKHIP_ERR_ARG,
                      "path must be 0 (automatic), 1 (padded-cube FFT), "
                      "2 (fused FFT) or 3 (direct separable)");
            if (D.mcdisk) {
                GBK_CHECK(D.path != 2, GBKHIP_ERR_UNSUPPORTED,
                          "mcdisk batches run the direct or padded-cube "
                          "pipeline");
                GBK_CHECK(D.hoi != nullptr && D.csum_len > 0, GBKHIP_ERR_ARG,
                          "mcdisk batch needs hoi and csum_len");
                // cloud i of trait t samples height trait t
                // (gmodels.hpp:406-416)
                GBK_CHECK(D.traits[0].n > 0 && D.traits[1].n >= D.traits[0].n,
                          GBKHIP_ERR_ARG,
                          "mcdisk needs one density height trait per polar "
                          "trait");
                B.thin = false;
                B.mc_tw = mc_tile_width(D.grid.spec_size);
                if (B.mc_tw) {
                    B.mc_ntx = ceil_div(D.grid.spat_size[0], B.mc_tw);
                    B.mc_ntiles = B.mc_ntx *
                                  ceil_div(D.grid.spat_size[1], B.mc_tw);
                }
            }
            plan_fused(*c, B);
            if (D.nextra > 0) {
                // further components: the model is a deposited cube
                GBK_CHECK(D.path != 2, GBKHIP_ERR_UNSUPPORTED,
                          "the fused FFT pipeline takes one thin disk");
                B.fused.on = false;
            }
            if ((D.path == 0 || D.path == 3) && plan_sep(B) &&
                (D.path == 3 || sep_cheaper(B))) {
                B.fused.on = false;
                B.nparts = (int)(ceil_div(D.size_lo[0], 64) *
                                 ceil_div(D.size_lo[1], kSepThreads / 64) *
                                 B.sep.zblocks);
                B.sep.on = true;
                // thick disks, mcdisk and several components deposit a
                // hi-res cube first
                B.sep.cube = D.mcdisk || !B.thin || D.nextra > 0;
                B.ka_psf.clear();
                B.ka_lsf.clear();
                // two lanes overlap one chunk's compute-bound x pass with
                // the memory-bound passes of the previous one; a deposited
                // cube per lane costs memory, so those start with one
                B.lanes = B.sep.cube ? 1 : 2;
                if (B.lanes == 2) ensure_lane_streams(B);
                alloc_work(B);
                compute_zero_partials(B);
                *out = b;
                return;
            }
            GBK_CHECK(D.path != 3, GBKHIP_ERR_UNSUPPORTED,
                      "direct separable path unavailable for this model "
                      "(a PSF x LSF that factors as lsf(z) px(x) py(y) "
                      "within the edge needed)");
            GBK_CHECK(D.path != 2 || B.fused.on, GBKHIP_ERR_UNSUPPORTED,
                      "fused FFT path unavailable for this model");
            B.ka_psf.clear();
            B.ka_lsf.clear();
            if (B.fused.on) {
                B.nparts = B.fused.tiles_x * B.fused.tiles_y;
                alloc_work(B);
                GBK_HIP(hipStreamSynchronize(c->stream));
                *out = b;
                return;
            }
            if (D.has_conv) {
                GBK_CHECK(D.pcube != nullptr, GBKHIP_ERR_ARG,
                          "pcube required with convolution");
                std::lock_guard<std::mutex> lk(c->mu);
                B.plan = get_plan(*c, nx, ny, nz);
                B.n_c = (long long)B.plan->nxh * ny * nz;
                B.spectrum.alloc(B.n_c * sizeof(float2));
                fft_forward(*c, *B.plan, D.pcube, B.spectrum.as<float2>(), 1,
                            0, 0);
            }
            alloc_work(B);
            fft_prepare_attributes();
            GBK_HIP(hipStreamSynchronize(c->stream));
        } catch (...) {
            delete b;
            throw;
        }
        *out = b;
    });
}

int gbkhip_batch_destroy(gbkhip_batch* b)
{
    // Touches neither the context (it may already be gone) nor its stream:
    // the graph executables and buffers are released on the batch's device,
    // and hipFree waits for outstanding work itself.
    return guarded([&] {
        if (!b) return;
        (void)hipSetDevice(b->impl.device);
        delete b;
    });
}

static void evaluate_impl(gbkhip_batch* b, const float* params,
                          const int* csum, const double* qscale, int nsets,
                          double* chi2, float* cubes, int use_graph)
{
    GBK_CHECK(b != nullptr, GBKHIP_ERR_ARG, "null batch");
    Batch& B = b->impl;
    Context& c = *B.ctx;
    GBK_HIP(hipSetDevice(c.device));
    GBK_CHECK(nsets >= 0, GBKHIP_ERR_ARG, "negative set count");
    if (nsets == 0) return;
    GBK_CHECK(params && chi2, GBKHIP_ERR_ARG, "null params or chi2");
    GBK_CHECK(B.desc.obs || B.mm_on, GBKHIP_ERR_ARG,
              "no data: an observed cube at creation or moment maps "
              "(gbkhip_batch_set_mmaps)");
    // the launches go to the context's stream member: point it at this
    // batch's stream for the duration of the call (under the context lock,
    // so batches of one context can be driven from several threads)
    std::lock_guard<std::mutex> lk(c.mu);
    const hipStream_t ctx_stream = c.stream;
    const hipStream_t target = B.stream ? B.stream : ctx_stream;
    struct Restore {
        Context& c;
        hipStream_t s;
        ~Restore() { c.stream = s; }
    } restore{c, ctx_stream};
    // work buffers (and the zero fill of a fresh int64 cube) on the
    // evaluation's own stream, ahead of the launches that use them
    c.stream = target;
    alloc_work(B);
    B.csum = csum;
    B.qscale = qscale;
    if (!use_graph || B.timing) {   // stage events need direct launches
        run_all(B, params, nsets, chi2, cubes);
        return;
    }
    const auto key = std::make_tuple((uintptr_t)params, nsets,
                                     (uintptr_t)chi2, (uintptr_t)cubes,
                                     (uintptr_t)csum, (uintptr_t)qscale,
                                     (uintptr_t)B.arena);
    auto it = B.graphs.find(key);
    if (it == B.graphs.end()) {
        if (!B.cap_stream)
            GBK_HIP(hipStreamCreateWithFlags(&B.cap_stream,
                                             hipStreamNonBlocking));
        // capture on a private stream, ordered after the target stream
        GBK_HIP(hipStreamSynchronize(target));
        c.stream = B.cap_stream;
        hipGraph_t graph = nullptr;
        GBK_HIP(hipStreamBeginCapture(B.cap_stream,
                                      hipStreamCaptureModeThreadLocal));
        try {
            run_all(B, params, nsets, chi2, cubes);
        } catch (...) {
            (void)hipStreamEndCapture(B.cap_stream, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        GBK_HIP(hipStreamEndCapture(B.cap_stream, &graph));
        c.stream = target;
        hipGraphExec_t exec = nullptr;
        GBK_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        GBK_HIP(hipGraphDestroy(graph));
        it = B.graphs.emplace(key, exec).first;
    }
    GBK_HIP(hipGraphLaunch(it->second, target));
}

int gbkhip_batch_evaluate(gbkhip_batch* b, const float* params, int nsets,
                          double* chi2, float* cubes, int use_graph)
{
    return guarded([&] {
        GBK_CHECK(b != nullptr && !b->impl.desc.mcdisk, GBKHIP_ERR_ARG,
                  "mcdisk batches evaluate with gbkhip_batch_evaluate_clouds");
        evaluate_impl(b, params, nullptr, nullptr, nsets, chi2, cubes,
                      use_graph);
    });
}

int gbkhip_batch_evaluate_clouds(gbkhip_batch* b, const float* params,
                                 const int* csum, const double* qscale,
                                 int nsets, double* chi2, float* cubes,
                                 int use_graph)
{
    return guarded([&] {
        GBK_CHECK(b != nullptr && b->impl.desc.mcdisk, GBKHIP_ERR_ARG,
                  "not an mcdisk batch");
        GBK_CHECK(nsets == 0 || (csum && qscale), GBKHIP_ERR_ARG,
                  "null csum or qscale");
        evaluate_impl(b, params, csum, qscale, nsets, chi2, cubes,
                      use_graph);
    });
}

int gbkhip_mcdisk_fixed_scale(float cflux, int nclouds,
                              const gbkhip_grid* grid, double* out)
{
    return guarded([&] {
        GBK_CHECK(grid && out, GBKHIP_ERR_ARG, "null grid or out");
        *out = mcdisk_fixed_scale(cflux, nclouds, *grid);
    });
}

int gbkhip_batch_set_mmaps(gbkhip_batch* b, int norders, const int* orders,
                           float cutoff, float zero_z, float step_z,
                           const float* obs, const float* err,
                           const float* err_scalars, const float* mask,
                           const float* weights)
{
    return guarded([&] {
        GBK_CHECK(b != nullptr, GBKHIP_ERR_ARG, "null batch");
        Batch& B = b->impl;
        GBK_HIP(hipSetDevice(B.device));
        if (norders == 0) {
            GBK_CHECK(B.desc.obs != nullptr, GBKHIP_ERR_ARG,
                      "no observed cube to return to");
            release_work(B);
            B.mm_on = false;
            return;
        }
        GBK_CHECK(norders > 0 && norders <= kMaxOrders && orders && obs,
                  GBKHIP_ERR_ARG, "1-8 orders and the data maps required");
        GBK_CHECK(err || err_scalars, GBKHIP_ERR_ARG,
                  "error maps or one error per order required");
        MMapArgs m{};
        m.norders = norders;
        for (int i = 0; i < kMaxOrders; ++i) {
            m.ord[i] = i < norders ? orders[i] : 0;
            m.err_scalar[i] = i < norders && err_scalars ? err_scalars[i]
                                                          : 1.0f;
            m.weight[i] = i < norders && weights ? weights[i] : 1.0f;
        }
        for (int i = 0; i < norders; ++i) {
            GBK_CHECK(m.ord[i] >= 0 && m.ord[i] <= 7 &&
                      (i == 0 || m.ord[i] > m.ord[i - 1]), GBKHIP_ERR_ARG,
                      "orders must be sorted, distinct and within 0-7");
        }
        m.cutoff = cutoff;
        m.zero_z = zero_z;
        m.step_z = step_z;
        m.obs = obs;
        m.err = err;
        m.mask = mask;
        release_work(B);      // the moment buffers join the work buffers
        B.mm = m;
        B.mm_on = true;
        B.mm_nparts = (int)ceil_div((long long)B.desc.size_lo[0] *
                                    B.desc.size_lo[1], 256);
    });
}

int gbkhip_batch_set_stream(gbkhip_batch* b, void* stream)
{
    return guarded([&] {
        GBK_CHECK(b != nullptr, GBKHIP_ERR_ARG, "null batch");
        b->impl.stream = static_cast<hipStream_t>(stream);
    });
}

int gbkhip_batch_release(gbkhip_batch* b)
{
    return guarded([&] {
        GBK_CHECK(b != nullptr, GBKHIP_ERR_ARG, "null batch");
        Batch& B = b->impl;
        GBK_HIP(hipSetDevice(B.device));
        // hipFree waits for work still using the buffers
        release_binding(B);
    });
}

int gbkhip_batch_memory(gbkhip_batch* b, unsigned long long* resident,
                        unsigned long long* work, int* allocated)
{
    return guarded([&] {
        GBK_CHECK(b && resident && work && allocated, GBKHIP_ERR_ARG,
                  "null argument");
        const Batch& B = b->impl;
        const Fused& F = B.fused;
        size_t r = B.ints.bytes + B.rnodes.bytes + B.spectrum.bytes +
                   F.phat.bytes + F.lhat.bytes + F.twz.bytes + F.twy.bytes +
                   F.twx.bytes + F.tpz.bytes + F.twsz.bytes + F.twsy.bytes +
                   F.twsx.bytes + B.sep.wx.bytes + B.sep.wy.bytes +
                   B.sep.wz.bytes + B.sep.zero_part.bytes;
        *resident = r;
        *work = work_bytes(B);
        *allocated = B.work ? 1 : 0;
    });
}

int gbkhip_batch_arena_bytes(gbkhip_batch* b, unsigned long long* bytes)
{
    return guarded([&] {
        GBK_CHECK(b && bytes, GBKHIP_ERR_ARG, "null argument");
        *bytes = place_work(b->impl, true);
    });
}

int gbkhip_batch_bind_work(gbkhip_batch* b, void* arena,
                           unsigned long long bytes)
{
    return guarded([&] {
        GBK_CHECK(b != nullptr, GBKHIP_ERR_ARG, "null batch");
        Batch& B = b->impl;
        char* a = static_cast<char*>(arena);
        if (a == B.arena && (size_t)bytes == B.arena_bytes) return;
        GBK_CHECK(!a || place_work(B, true) <= bytes, GBKHIP_ERR_ARG,
                  "the work arena is smaller than the batch needs "
                  "(gbkhip_batch_arena_bytes)");
        GBK_HIP(hipSetDevice(B.device));
        // own buffers are freed with their graphs; an arena's graphs stay
        // valid for it (keyed by its address)
        release_binding(B);
        B.arena = a;
        B.arena_bytes = a ? (size_t)bytes : 0;
    });
}

int gbkhip_batch_set_tiles(gbkhip_batch* b, int lanes, int ysplit)
{
    return guarded([&] {
        GBK_CHECK(b != nullptr, GBKHIP_ERR_ARG, "null batch");
        GBK_CHECK(lanes == 1 || lanes == 2, GBKHIP_ERR_ARG,
                  "lanes must be 1 or 2");
        GBK_CHECK(ysplit == 1 || ysplit == 2 || ysplit == 4, GBKHIP_ERR_ARG,
                  "ysplit must be 1, 2 or 4");
        Batch& B = b->impl;
        GBK_HIP(hipSetDevice(B.device));
        if (!B.sep.on) {
            GBK_CHECK(lanes == 1 && ysplit == 1, GBKHIP_ERR_UNSUPPORTED,
                      "tile settings apply to the direct pipeline");
            return;
        }
        std::lock_guard<std::mutex> lk(B.ctx->mu);
        const int jy = sep_jy(B.sep.ty);
        const int groups = B.sep.yblk_full / jy;
        B.sep.yblk = ceil_div(groups, ysplit) * jy;
        if (lanes != B.lanes) {
            release_work(B);          // re-allocated for the new lane count
            B.lanes = lanes;
            if (lanes == 2) ensure_lane_streams(B);
        }
        drop_graphs(B);
    });
}

int gbkhip_group_create(gbkhip_ctx* ctx, gbkhip_batch* const* members, int n,
                        gbkhip_group** out)
{
    return guarded([&] {
        Context* c = ctx_of(ctx);
        GBK_CHECK(out && members && n >= 1, GBKHIP_ERR_ARG,
                  "need at least one member");
        GBK_HIP(hipSetDevice(c->device));
        auto* g = new gbkhip_group();
        try {
            g->ctx = c;
            for (int i = 0; i < n; ++i) {
                GBK_CHECK(members[i] != nullptr, GBKHIP_ERR_ARG,
                          "null member");
                GBK_CHECK(members[i]->impl.ctx == c, GBKHIP_ERR_ARG,
                          "members must share the group's context");
                g->members.push_back(members[i]);
            }
            const int k = std::min(n, kGroupBranches);
            for (int i = 0; i < k; ++i) {
                hipStream_t s;
                GBK_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
                g->branch.push_back(s);
                hipEvent_t e;
                GBK_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
                g->ev_join.push_back(e);
            }
            GBK_HIP(hipEventCreateWithFlags(&g->ev_fork,
                                            hipEventDisableTiming));
            GBK_HIP(hipStreamCreateWithFlags(&g->cap, hipStreamNonBlocking));
        } catch (...) {
            delete g;
            throw;
        }
        *out = g;
    });
}

int gbkhip_group_destroy(gbkhip_group* g)
{
    return guarded([&] {
        if (!g) return;
        GBK_HIP(hipSetDevice(g->ctx->device));
        for (auto s : g->branch) GBK_HIP(hipStreamSynchronize(s));
        delete g;
    });
}

int gbkhip_group_evaluate(gbkhip_group* g, const float* const* params,
                          const int* const* csum,
                          const double* const* qscale, const int* nsets,
                          double* const* chi2, float* const* cubes,
                          int use_graph)
{
    return guarded([&] {
        GBK_CHECK(g && params && nsets && chi2, GBKHIP_ERR_ARG,
                  "null argument");
        Context& c = *g->ctx;
        GBK_HIP(hipSetDevice(c.device));
        const int n = (int)g->members.size();
        for (int i = 0; i < n; ++i) {
            const Batch& B = g->members[i]->impl;
            GBK_CHECK(nsets[i] >= 0, GBKHIP_ERR_ARG, "negative set count");
            GBK_CHECK(nsets[i] == 0 || (params[i] && chi2[i]),
                      GBKHIP_ERR_ARG, "null params or chi2");
            GBK_CHECK(!B.desc.mcdisk || nsets[i] == 0 ||
                          (csum && qscale && csum[i] && qscale[i]),
                      GBKHIP_ERR_ARG, "mcdisk member without cloud budgets");
            GBK_CHECK(!B.timing, GBKHIP_ERR_ARG,
                      "member has stage timing on");
        }
        std::lock_guard<std::mutex> lk(c.mu);
        const hipStream_t target = c.stream;
        struct Restore {
            Context& c;
            hipStream_t s;
            ~Restore() { c.stream = s; }
        } restore{c, target};
        for (auto* m : g->members) alloc_work(m->impl);
        const int k = (int)g->branch.size();
        auto launch_all = [&](hipStream_t main) {
            GBK_HIP(hipEventRecord(g->ev_fork, main));
            for (int j = 0; j < k; ++j)
                GBK_HIP(hipStreamWaitEvent(g->branch[j], g->ev_fork, 0));
            for (int i = 0; i < n; ++i) {
                if (nsets[i] == 0) continue;
                Batch& B = g->members[i]->impl;
                c.stream = g->branch[i % k];
                B.csum = csum ? csum[i] : nullptr;
                B.qscale = qscale ? qscale[i] : nullptr;
                run_all(B, params[i], nsets[i], chi2[i],
                        cubes ? cubes[i] : nullptr);
            }
            c.stream = main;
            for (int j = 0; j < k; ++j) {
                GBK_HIP(hipEventRecord(g->ev_join[j], g->branch[j]));
                GBK_HIP(hipStreamWaitEvent(main, g->ev_join[j], 0));
            }
        };
        if (!use_graph) {
            launch_all(target);
            return;
        }
        std::vector<uintptr_t> key;
        for (int i = 0; i < n; ++i) {
            // a member whose work buffers were re-allocated needs a new graph
            key.push_back((uintptr_t)g->members[i]->impl.work_gen);
            key.push_back((uintptr_t)params[i]);
            key.push_back((uintptr_t)nsets[i]);
            key.push_back((uintptr_t)chi2[i]);
            key.push_back(cubes ? (uintptr_t)cubes[i] : 0);
            key.push_back(csum ? (uintptr_t)csum[i] : 0);
            key.push_back(qscale ? (uintptr_t)qscale[i] : 0);
        }
        auto it = g->graphs.find(key);
        if (it == g->graphs.end()) {
            GBK_HIP(hipStreamSynchronize(target));
            hipGraph_t graph = nullptr;
            GBK_HIP(hipStreamBeginCapture(g->cap,
                                          hipStreamCaptureModeThreadLocal));
            try {
                launch_all(g->cap);
            } catch (...) {
                (void)hipStreamEndCapture(g->cap, &graph);
                if (graph) (void)hipGraphDestroy(graph);
                throw;
            }
            GBK_HIP(hipStreamEndCapture(g->cap, &graph));
            hipGraphExec_t exec = nullptr;
            GBK_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
            GBK_HIP(hipGraphDestroy(graph));
            it = g->graphs.emplace(key, exec).first;
        }
        GBK_HIP(hipGraphLaunch(it->second, target));
    });
}

int gbkhip_batch_reserve_clouds(gbkhip_batch* b, long long clouds)
{
    return guarded([&] {
        GBK_CHECK(b != nullptr, GBKHIP_ERR_ARG, "null batch");
        Batch& B = b->impl;
        GBK_CHECK(B.desc.mcdisk, GBKHIP_ERR_ARG, "not an mcdisk batch");
        GBK_CHECK(clouds >= 0 && clouds < (1LL << 31), GBKHIP_ERR_ARG,
                  "cloud capacity out of range");
        if (!B.mc_tw || clouds <= B.mc_cap) return;
        GBK_HIP(hipSetDevice(B.device));
        std::lock_guard<std::mutex> lk(B.ctx->mu);
        release_work(B);             // re-allocated for the new capacity
        B.mc_cap = (clouds + 1023) / 1024 * 1024;
    });
}

int gbkhip_batch_tile_stats(gbkhip_batch* b, long long* counts)
{
    return guarded([&] {
        GBK_CHECK(b && counts, GBKHIP_ERR_ARG, "null argument");
        Batch& B = b->impl;
        for (int i = 0; i < 6; ++i) counts[i] = 0;
        if (!B.sep.on || !B.work || B.sep.last_sets == 0) return;
        GBK_HIP(hipSetDevice(B.device));
        const hipStream_t s = B.stream ? B.stream : B.ctx->stream;
        GBK_HIP(hipStreamSynchronize(s));
        const size_t n = (size_t)B.sep.last_sets * B.sep.need[1];
        std::vector<unsigned long long> h(n);
        GBK_HIP(hipMemcpy(h.data(), B.sep.masks.ptr, n * sizeof(uint64_t),
                          hipMemcpyDeviceToHost));
        for (unsigned long long v : h) counts[0] += __builtin_popcountll(v);
        counts[1] = (long long)n * B.sep.nchunks;
        const size_t ng = (size_t)B.sep.last_sets * B.sep.need[2] *
                          B.sep.ngy;
        std::vector<unsigned char> f(ng);
        GBK_HIP(hipMemcpy(f.data(), B.sep.gflags.ptr, ng,
                          hipMemcpyDeviceToHost));
        for (unsigned char v : f) counts[2] += v != 0;
        counts[3] = (long long)ng;
        // z-pass blocks (per set, 16-channel output block and 4-row block)
        // with a live input; the others take the zero-model partials
        const gbkhip_batch_desc& D = B.desc;
        const int nz = B.sep.need[2], nyl = D.size_lo[1];
        const int sz = D.scale[2], jy = sep_jy(B.sep.ty);
        const int nin = sz * (kSepJZ - 1) + B.sep.tz;
        const int yb = ceil_div(nyl, kSepThreads / 64);
        for (int st = 0; st < B.sep.last_sets; ++st)
            for (int zb = 0; zb < B.sep.zblocks; ++zb)
                for (int by = 0; by < yb; ++by) {
                    bool live = false;
                    for (int yy = 0; yy < kSepThreads / 64 && !live; ++yy) {
                        const int Y = by * (kSepThreads / 64) + yy;
                        if (Y >= nyl) break;
                        for (int k = 0; k < nin && !live; ++k) {
                            const int z = sz * zb * kSepJZ + B.sep.oz + k;
                            live = z < nz &&
                                   f[((size_t)st * nz + z) * B.sep.ngy +
                                     Y / jy] != 0;
                        }
                    }
                    counts[4] += live;
                    counts[5] += 1;
                }
    });
}

int gbkhip_batch_line_stats(gbkhip_batch* b, long long* counts)
{
    return guarded([&] {
        GBK_CHECK(b && counts, GBKHIP_ERR_ARG, "null argument");
        Batch& B = b->impl;
        counts[0] = counts[1] = 0;
        if (!B.sep.on || B.sep.cube || !B.work || B.sep.last_sets == 0)
            return;
        GBK_HIP(hipSetDevice(B.device));
        const hipStream_t s = B.stream ? B.stream : B.ctx->stream;
        GBK_HIP(hipStreamSynchronize(s));
        const int nx = B.sep.need[0];
        const size_t rows = (size_t)B.sep.last_sets * B.sep.need[1];
        std::vector<unsigned long long> h(rows);
        GBK_HIP(hipMemcpy(h.data(), B.sep.segmask.ptr,
                          rows * sizeof(uint64_t), hipMemcpyDeviceToHost));
        const int nseg = ceil_div(nx, 64);
        for (unsigned long long m : h)
            for (int sg = 0; sg < nseg; ++sg)
                if (sg >= 63 || ((m >> sg) & 1ull))
                    counts[0] += std::min(64, nx - 64 * sg);
        counts[1] = (long long)rows * nx;
    });
}

int gbkhip_batch_check(gbkhip_batch* b, int* overflow)
{
    return guarded([&] {
        GBK_CHECK(b && overflow, GBKHIP_ERR_ARG, "null argument");
        Batch& B = b->impl;
        GBK_HIP(hipSetDevice(B.device));
        const hipStream_t s = B.stream ? B.stream : B.ctx->stream;
        GBK_HIP(hipStreamSynchronize(s));
        int h = 0;
        GBK_HIP(hipMemcpy(&h, B.overflow.ptr, sizeof(int),
                          hipMemcpyDeviceToHost));
        if (h) GBK_HIP(hipMemset(B.overflow.ptr, 0, sizeof(int)));
        *overflow = h;
    });
}

int gbkhip_batch_info(gbkhip_batch* b, int* info)
{
    return guarded([&] {
        GBK_CHECK(b && info, GBKHIP_ERR_ARG, "null argument");
        const Batch& B = b->impl;
        const Fused& F = B.fused;
        const Sep& P = B.sep;
        info[0] = P.on ? 2 : F.on ? 1 : 0;
        info[1] = P.on ? P.need[0] : F.on ? F.nx : B.desc.size_hi[0];
        info[2] = P.on ? P.need[1] : F.on ? F.ny : B.desc.size_hi[1];
        info[3] = P.on ? P.need[2] : F.on ? F.nz : B.desc.size_hi[2];
        info[4] = B.nparts;
    });
}

int gbkhip_batch_taps(gbkhip_batch* b, int* taps)
{
    return guarded([&] {
        GBK_CHECK(b && taps, GBKHIP_ERR_ARG, "null argument");
        const Sep& P = b->impl.sep;
        GBK_CHECK(P.on, GBKHIP_ERR_UNSUPPORTED,
                  "taps apply to the direct pipeline");
        for (int i = 0; i < 3; ++i) taps[i] = P.kept[i];
        taps[3] = P.ox;
        taps[4] = P.oy;
        taps[5] = P.oz;
    });
}

int gbkhip_batch_set_timing(gbkhip_batch* b, int enable)
{
    return guarded([&] {
        GBK_CHECK(b != nullptr, GBKHIP_ERR_ARG, "null batch");
        b->impl.timing = enable != 0;
    });
}

int gbkhip_batch_stage_times(gbkhip_batch* b, double* ms, int* chunks)
{
    return guarded([&] {
        GBK_CHECK(b && ms && chunks, GBKHIP_ERR_ARG, "null argument");
        Batch& B = b->impl;
        ms[0] = ms[1] = ms[2] = 0.0;
        *chunks = B.timed_chunks;
        // stage boundaries among the marks
        static const int kFusedBnd[4] = {0, 2, 5, 7};
        static const int kSepBnd[4] = {0, 1, 3, 5};
        static const int kPaddedBnd[4] = {0, 1, 2, 3};
        const int* bnd = B.sep.on ? kSepBnd
                         : B.fused.on ? kFusedBnd : kPaddedBnd;
        for (int c = 0; c < B.timed_chunks; ++c) {
            GBK_HIP(hipEventSynchronize(
                B.events[kMarks * c + n_marks(B) - 1]));
            for (int s = 0; s < 3; ++s) {
                float t = 0.0f;
                GBK_HIP(hipEventElapsedTime(&t,
                                            B.events[kMarks * c + bnd[s]],
                                            B.events[kMarks * c + bnd[s + 1]]));
                ms[s] += t;
            }
        }
    });
}

int gbkhip_batch_kernel_times(gbkhip_batch* b, double* ms, int cap, int* n)
{
    return guarded([&] {
        GBK_CHECK(b && ms && n, GBKHIP_ERR_ARG, "null argument");
        Batch& B = b->impl;
        const int k = n_marks(B) - 1;
        GBK_CHECK(cap >= k, GBKHIP_ERR_ARG, "kernel_times: array too short");
        *n = k;
        for (int i = 0; i < k; ++i) ms[i] = 0.0;
        for (int c = 0; c < B.timed_chunks; ++c) {
            GBK_HIP(hipEventSynchronize(B.events[kMarks * c + k]));
            for (int i = 0; i < k; ++i) {
                float t = 0.0f;
                GBK_HIP(hipEventElapsedTime(&t, B.events[kMarks * c + i],
                                            B.events[kMarks * c + i + 1]));
                ms[i] += t;
            }
        }
    });
}

}  // extern "C"
