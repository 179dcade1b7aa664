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
#include "launch.hpp"
#include "disk.hpp"

#include <climits>
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
    gbkhip_traits* tr[9] = {&d.rp, &d.rh, &d.vp, &d.vh, &d.dp, &d.dh, &d.zp,
                            &d.sp, &d.wp};
    for (int k = 0; k < 9; ++k) {
        if (!tr[k]->uids) continue;
        tr[k]->cvalues = p + L.cvalues[k];
        tr[k]->pvalues = p + L.pvalues[k];
    }
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

// Thick disk: 4 waves split the line of sight of 64 columns, spectrum tile
// in LDS (chunked over channels), tile written in full.
__global__ __launch_bounds__(256) void batch_smdisk_thick(
    gbkhip_disk tmpl, SetLayout lay, const float* __restrict__ params,
    int pstride, gbkhip_grid g, float* __restrict__ cubes, long long cstride,
    int chunk)
{
    extern __shared__ float tile[];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int x = blockIdx.x * 64 + lane;
    const int y = blockIdx.y;
    const int set = blockIdx.z;
    const int nx = g.spat_size[0], ny = g.spat_size[1], nz = g.spat_size[2];
    const size_t plane = (size_t)nx * ny;
    const gbkhip_disk d = set_disk(tmpl, lay, params + (size_t)set * pstride);
    const bool need_theta = disk_needs_theta(d);
    const NodeGeom ng = node_geom(d, 0);
    float* cube = cubes + set * cstride;
    for (int c0 = 0; c0 < g.spec_size; c0 += chunk) {
        const int c1 = min(c0 + chunk, g.spec_size);
        for (int i = threadIdx.x; i < chunk * 64; i += blockDim.x)
            tile[i] = 0.0f;
        __syncthreads();
        if (x < nx) {
            for (int z = wave; z < nz; z += 4) {
                const VoxelLine L = smdisk_voxel(d, g, ng, x, y, z,
                                                 need_theta, false);
                if (!L.on) continue;
                int k0, k1;
                line_window(g, L.v, L.s, k0, k1);
                k0 = max(k0, c0);
                k1 = min(k1, c1 - 1);
                if (k0 > k1) continue;
                const LineEval e = line_eval(g, L.orvalue, L.v, L.s);
                for (int k = k0; k <= k1; ++k)
                    atomicAdd(&tile[(k - c0) * 64 + lane], e.at(k));
            }
        }
        __syncthreads();
        for (int i = threadIdx.x; i < (c1 - c0) * 64; i += blockDim.x) {
            const int k = c0 + i / 64, l = i % 64;
            const int xx = blockIdx.x * 64 + l;
            if (xx < nx)
                cube[k * plane + (size_t)y * nx + xx] = tile[i];
        }
        __syncthreads();
    }
}

struct DataArgs {
    const float* obs;     // (nz, ny, nx) low-res observed cube
    const float* err;     // same shape or nullptr
    const float* mask;    // same shape or nullptr
    float err_scalar;     // used when err == nullptr (1 = none)
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
    DevBuf spectrum;        // kernel spectrum (nz, ny, nxh) complex
    DevBuf hi;              // chunk x padded cube
    DevBuf spec;            // chunk x complex work spectrum
    DevBuf partials;
    int nparts = 0;
    long long n_hi = 0, n_c = 0, n_lo = 0;
    bool thin = true;
    // graph cache: (params, nsets, chi2, cubes) -> executable graph
    std::map<std::tuple<uintptr_t, int, uintptr_t, uintptr_t>,
             hipGraphExec_t> graphs;
    hipStream_t cap_stream = nullptr;
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
    }
};

static void build_layout(Batch& b)
{
    const gbkhip_batch_desc& D = b.desc;
    const int nn = D.nrnodes;
    b.lay.vsys = D.vsys_count > 0 ? D.off_vsys : -1;
    b.lay.xpos = D.off_xpos;
    b.lay.ypos = D.off_ypos;
    b.lay.posa = D.off_posa;
    b.lay.incl = D.off_incl;
    std::memset(&b.tmpl, 0, sizeof(b.tmpl));
    b.tmpl.loose = D.loose;
    b.tmpl.tilted = D.tilted;
    b.tmpl.nrnodes = nn;
    int total = 0;
    for (int k = 0; k < 9; ++k) total += 3 * D.traits[k].n;
    b.h_uids.assign(std::max(total, 1), 0);
    int pos = 0;
    std::vector<std::pair<int, int>> where(9, {-1, 0});
    for (int k = 0; k < 9; ++k) {
        const gbkhip_batch_traits& t = D.traits[k];
        b.lay.cvalues[k] = t.n ? t.off_cvalues : -1;
        b.lay.pvalues[k] = t.n ? t.off_pvalues : -1;
        if (!t.n) continue;
        GBK_CHECK(t.n <= 4, GBKHIP_ERR_ARG, "at most 4 traits per kind");
        where[k] = {pos, t.n};
        for (int i = 0; i < t.n; ++i) {
            b.h_uids[pos + i] = t.uids[i];
            b.h_uids[pos + t.n + i] = t.ccounts[i];
            b.h_uids[pos + 2 * t.n + i] = t.pcounts[i];
        }
        pos += 3 * t.n;
    }
    b.ints.alloc(b.h_uids.size() * sizeof(int));
    GBK_HIP(hipMemcpy(b.ints.ptr, b.h_uids.data(),
                      b.h_uids.size() * sizeof(int), hipMemcpyHostToDevice));
    gbkhip_traits* tr[9] = {&b.tmpl.rp, &b.tmpl.rh, &b.tmpl.vp, &b.tmpl.vh,
                            &b.tmpl.dp, &b.tmpl.dh, &b.tmpl.zp, &b.tmpl.sp,
                            &b.tmpl.wp};
    const int* base = b.ints.as<int>();
    for (int k = 0; k < 9; ++k) {
        if (where[k].first < 0) continue;
        const int p0 = where[k].first, n = where[k].second;
        tr[k]->n = n;
        tr[k]->uids = base + p0;
        tr[k]->ccounts = base + p0 + n;
        tr[k]->pcounts = base + p0 + 2 * n;
    }
    b.rnodes.alloc(nn * sizeof(float));
    GBK_HIP(hipMemcpy(b.rnodes.ptr, D.rnodes, nn * sizeof(float),
                      hipMemcpyHostToDevice));
    b.tmpl.rnodes = b.rnodes.as<float>();
    b.tmpl.opacity = nullptr;
    b.thin = D.traits[1].n == 0;
}

static void ensure_events(Batch& b, int chunks)
{
    while ((int)b.events.size() < 4 * chunks) {
        hipEvent_t e;
        GBK_HIP(hipEventCreate(&e));
        b.events.push_back(e);
    }
}

static void mark(Batch& b, int chunk, int i)
{
    if (b.timing)
        GBK_HIP(hipEventRecord(b.events[4 * chunk + i], b.ctx->stream));
}

static void run_chunk(Batch& b, const float* params, int set0, int nsets,
                      double* chi2, float* cubes)
{
    Context& c = *b.ctx;
    const int ci = set0 / b.desc.max_batch;
    mark(b, ci, 0);
    const gbkhip_batch_desc& D = b.desc;
    const gbkhip_grid& g = D.grid;
    const float* P = params + (size_t)set0 * D.pstride;
    float* hi = b.hi.as<float>();
    if (b.thin) {
        dim3 grid(ceil_div(g.spat_size[0], 64), ceil_div(g.spat_size[1], 4),
                  nsets);
        hipLaunchKernelGGL(batch_smdisk_thin, grid, dim3(64, 4), 0, c.stream,
                           b.tmpl, b.lay, P, D.pstride, g, hi, b.n_hi);
    } else {
        const int chunk = std::max(1, std::min(g.spec_size, 240));
        dim3 grid(ceil_div(g.spat_size[0], 64), g.spat_size[1], nsets);
        hipLaunchKernelGGL(batch_smdisk_thick, grid, dim3(256),
                           (size_t)chunk * 64 * sizeof(float), c.stream,
                           b.tmpl, b.lay, P, D.pstride, g, hi, b.n_hi, chunk);
    }
    GBK_HIP(hipGetLastError());
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
    da.weight = D.weight;
    dim3 grid(b.nparts, nsets);
    hipLaunchKernelGGL(batch_downscale_chi2, grid, dim3(256), 0, c.stream, hi,
                       b.n_hi, D.size_hi[0], D.size_hi[1], D.scale[0],
                       D.scale[1], D.scale[2], D.edge_hi[0], D.edge_hi[1],
                       D.edge_hi[2], D.size_lo[0], D.size_lo[1],
                       D.size_lo[2], da,
                       cubes ? cubes + (size_t)set0 * b.n_lo : nullptr,
                       b.n_lo, b.partials.as<double>(), b.nparts);
    GBK_HIP(hipGetLastError());
    hipLaunchKernelGGL(batch_chi2_final, dim3(nsets), dim3(256), 0, c.stream,
                       b.partials.as<double>(), b.nparts, chi2, set0);
    GBK_HIP(hipGetLastError());
    mark(b, ci, 3);
}

static void run_all(Batch& b, const float* params, int nsets, double* chi2,
                    float* cubes)
{
    const int chunks = (nsets + b.desc.max_batch - 1) / b.desc.max_batch;
    if (b.timing) ensure_events(b, chunks);
    b.timed_chunks = b.timing ? chunks : 0;
    for (int s0 = 0; s0 < nsets; s0 += b.desc.max_batch)
        run_chunk(b, params, s0, std::min(b.desc.max_batch, nsets - s0), chi2,
                  cubes);
}

}  // namespace gbkhip

struct gbkhip_batch {
    gbkhip::Batch impl;
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
        GBK_CHECK(D.obs != nullptr, GBKHIP_ERR_ARG, "observed cube required");
        for (int i = 0; i < 3; ++i) {
            GBK_CHECK(D.size_lo[i] > 0 && D.size_hi[i] > 0 && D.scale[i] > 0,
                      GBKHIP_ERR_ARG, "sizes must be positive");
            GBK_CHECK(D.edge_hi[i] >= 0 && D.edge_hi[i] + D.size_lo[i] *
                      D.scale[i] <= D.size_hi[i], GBKHIP_ERR_ARG,
                      "downscale window exceeds the padded cube");
        }
        GBK_CHECK(D.grid.spat_size[0] == D.size_hi[0] &&
                  D.grid.spat_size[1] == D.size_hi[1] &&
                  D.grid.spec_size == D.size_hi[2], GBKHIP_ERR_ARG,
                  "gmodel grid must cover the padded cube");
        auto* b = new gbkhip_batch();
        Batch& B = b->impl;
        B.ctx = c;
        B.device = c->device;
        B.desc = D;
        try {
            build_layout(B);
            const int nx = D.size_hi[0], ny = D.size_hi[1], nz = D.size_hi[2];
            B.n_hi = (long long)nx * ny * nz;
            B.n_lo = (long long)D.size_lo[0] * D.size_lo[1] * D.size_lo[2];
            B.nparts = (int)std::max<long long>(
                1, std::min<long long>(512, (B.n_lo + 4095) / 4096));
            B.hi.alloc((size_t)D.max_batch * B.n_hi * sizeof(float));
            B.partials.alloc((size_t)D.max_batch * B.nparts * sizeof(double));
            if (D.has_conv) {
                GBK_CHECK(D.pcube != nullptr, GBKHIP_ERR_ARG,
                          "pcube required with convolution");
                std::lock_guard<std::mutex> lk(c->mu);
                B.plan = get_plan(*c, nx, ny, nz);
                B.n_c = (long long)B.plan->nxh * ny * nz;
                B.spectrum.alloc(B.n_c * sizeof(float2));
                fft_forward(*c, *B.plan, D.pcube, B.spectrum.as<float2>(), 1,
                            0, 0);
                B.spec.alloc((size_t)D.max_batch * B.n_c * sizeof(float2));
            }
            fft_prepare_attributes();
            if (!B.thin) {
                (void)hipFuncSetAttribute(
                    (const void*)batch_smdisk_thick,
                    hipFuncAttributeMaxDynamicSharedMemorySize, 96 * 1024);
            }
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

int gbkhip_batch_evaluate(gbkhip_batch* b, const float* params, int nsets,
                          double* chi2, float* cubes, int use_graph)
{
    return guarded([&] {
        GBK_CHECK(b != nullptr, GBKHIP_ERR_ARG, "null batch");
        Batch& B = b->impl;
        Context& c = *B.ctx;
        GBK_HIP(hipSetDevice(c.device));
        GBK_CHECK(nsets >= 0, GBKHIP_ERR_ARG, "negative set count");
        if (nsets == 0) return;
        GBK_CHECK(params && chi2, GBKHIP_ERR_ARG, "null params or chi2");
        if (!use_graph || B.timing) {   // stage events need direct launches
            run_all(B, params, nsets, chi2, cubes);
            return;
        }
        const auto key = std::make_tuple((uintptr_t)params, nsets,
                                         (uintptr_t)chi2, (uintptr_t)cubes);
        auto it = B.graphs.find(key);
        if (it == B.graphs.end()) {
            if (!B.cap_stream)
                GBK_HIP(hipStreamCreateWithFlags(&B.cap_stream,
                                                 hipStreamNonBlocking));
            // capture on a private stream, ordered after the context stream
            GBK_HIP(hipStreamSynchronize(c.stream));
            hipStream_t saved = c.stream;
            c.stream = B.cap_stream;
            hipGraph_t graph = nullptr;
            GBK_HIP(hipStreamBeginCapture(B.cap_stream,
                                          hipStreamCaptureModeThreadLocal));
            try {
                run_all(B, params, nsets, chi2, cubes);
            } catch (...) {
                (void)hipStreamEndCapture(B.cap_stream, &graph);
                if (graph) (void)hipGraphDestroy(graph);
                c.stream = saved;
                throw;
            }
            c.stream = saved;
            GBK_HIP(hipStreamEndCapture(B.cap_stream, &graph));
            hipGraphExec_t exec = nullptr;
            GBK_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
            GBK_HIP(hipGraphDestroy(graph));
            it = B.graphs.emplace(key, exec).first;
        }
        GBK_HIP(hipGraphLaunch(it->second, c.stream));
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
        for (int c = 0; c < B.timed_chunks; ++c) {
            GBK_HIP(hipEventSynchronize(B.events[4 * c + 3]));
            for (int s = 0; s < 3; ++s) {
                float t = 0.0f;
                GBK_HIP(hipEventElapsedTime(&t, B.events[4 * c + s],
                                            B.events[4 * c + s + 1]));
                ms[s] += t;
            }
        }
    });
}

}  // extern "C"
