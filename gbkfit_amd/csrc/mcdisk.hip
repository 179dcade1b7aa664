// mcdisk: Monte Carlo cloud disk (core/gmodel/gmodels.hpp:325-643; host
// loop gbkfit_host/include/gbkfit/host/kernels.hpp:183-306).  The per-cloud
// body (Philox stream per (seed, cloud id), samplers, projection, int64
// fixed-point channel deposits) is in mcdisk.hpp.  The reference seeds every
// OpenMP thread with 42 (kernels.hpp:239-242), so its output depends on the
// thread count; this stream depends only on (seed, cloud id).  The int64
// cube is converted and added to the FP32 cube afterwards: integer addition
// is associative, so reruns are bit-identical.
#include "launch.hpp"
#include "mcdisk.hpp"

namespace gbkhip {

__global__ __launch_bounds__(256) void mcdisk_kernel(
    McParams mp, gbkhip_disk d, gbkhip_grid g, gbkhip_outputs o,
    unsigned long long* icube)
{
    extern __shared__ int s_csum[];
    __shared__ float s_nodes[kMcNodesLds];
    mp.csum = mc_stage_csum(mp.csum, mp.csum_len, s_csum);
    d.rnodes = mc_stage_nodes(d.rnodes, d.nrnodes, s_nodes);
    for (long long ci = blockIdx.x * (long long)blockDim.x + threadIdx.x;
         ci < mp.nclouds; ci += (long long)gridDim.x * blockDim.x)
        mc_cloud(ci, mp, d, g, o, icube);
}

// ---- Binned deposits (scube only): clouds -> records ranked per spatial
// tile -> tile offsets -> records sorted by tile -> per-tile LDS sums.
__global__ void mc_zero_i32(int* p, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += gridDim.x * blockDim.x)
        p[i] = 0;
}

// With ntiles <= kMcBinLds the block counts its clouds per tile in LDS and
// reserves one range per tile with a single global atomic, instead of one
// global atomic per cloud: the central tiles of a disk take most clouds,
// and same-address atomics serialise.
constexpr int kMcBinBlocks = 8192;

__global__ __launch_bounds__(256) void mcdisk_bin_kernel(
    McParams mp, gbkhip_disk d, gbkhip_grid g, McBin bin, int ntiles)
{
    extern __shared__ int s_dyn[];            // mc_bin_lds layout
    const McBinLds L = mc_bin_lds(mp.csum_len, d.nrnodes, ntiles);
    float* s_nodes = reinterpret_cast<float*>(s_dyn + L.csum);
    int* s_cnt = s_dyn + L.csum + L.nodes;
    mp.csum = mc_stage_csum(mp.csum, mp.csum_len, s_dyn);
    d.rnodes = mc_stage_nodes(d.rnodes, d.nrnodes, s_nodes);
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
    // every lane of the block runs every round (the rank is a wave-wide
    // operation); lanes past the last cloud carry no key
    for (long long b0 = (long long)blockIdx.x * blockDim.x; b0 < mp.nclouds;
         b0 += stride) {
        const long long ci = b0 + threadIdx.x;
        const int tile = ci < mp.nclouds
                             ? mc_cloud(ci, mp, d, g, o, nullptr, &bin) : -1;
        // rank within the block's clouds of the tile (the block reserves
        // its ranges afterwards), or directly global
        const int rank = wave_rank(tile, local ? s_cnt : bin.tcount, wscr);
        if (ci < mp.nclouds) bin.tr[ci] = make_int2(tile, rank);
    }
    if (!local) return;
    __syncthreads();
    // one reservation per tile this block reached: s_cnt <- its base
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

// Per set (blockIdx.x; one block of 1024 threads): off[t] = records of
// tiles [0, t) (off[n] = all), soff[t] = slices of tiles [0, t) with
// ceil(count / kMcSlice) slices per tile, and the tiles of more than one
// slice numbered in order: sidx[t] (-1 for the others), stile[k], nsplit.
__global__ __launch_bounds__(1024) void mc_scan_tiles(
    const int* __restrict__ count, int* __restrict__ off,
    int* __restrict__ soff, int* __restrict__ sidx, int* __restrict__ stile,
    int* __restrict__ nsplit, int n, int maxsplit)
{
    __shared__ int s_r[1024], s_s[1024], s_p[1024];
    const size_t set = blockIdx.x;
    count += set * n;
    off += set * (n + 1);
    soff += set * (n + 1);
    sidx += set * n;
    stile += set * maxsplit;
    const int t = threadIdx.x, per = (n + 1023) / 1024;
    const int lo = min(t * per, n), hi = min(lo + per, n);
    int r = 0, sl = 0, sp = 0;
    for (int i = lo; i < hi; ++i) {
        const int c = count[i], ns = (c + kMcSlice - 1) / kMcSlice;
        r += c;
        sl += ns;
        sp += ns > 1;
    }
    s_r[t] = r;
    s_s[t] = sl;
    s_p[t] = sp;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {      // inclusive scans
        const int x = t >= d ? s_r[t - d] : 0;
        const int y = t >= d ? s_s[t - d] : 0;
        const int z = t >= d ? s_p[t - d] : 0;
        __syncthreads();
        s_r[t] += x;
        s_s[t] += y;
        s_p[t] += z;
        __syncthreads();
    }
    if (t == 1023) {
        off[n] = s_r[t];
        soff[n] = s_s[t];
        nsplit[set] = s_p[t];
    }
    r = s_r[t] - r;                           // exclusive
    sl = s_s[t] - sl;
    sp = s_p[t] - sp;
    for (int i = lo; i < hi; ++i) {
        const int c = count[i], ns = (c + kMcSlice - 1) / kMcSlice;
        off[i] = r;
        soff[i] = sl;
        if (ns > 1) {
            sidx[i] = sp;
            stile[sp] = i;
            ++sp;
        } else {
            sidx[i] = -1;
        }
        r += c;
        sl += ns;
    }
}

// Records to tile order; blockIdx.y = set.  A set's cloud count is its last
// cumulative count (csum, per set) or n, clamped to the capacity.
__global__ void mc_scatter(const McDep* __restrict__ rec,
                           const int2* __restrict__ tr,
                           const int* __restrict__ off,
                           McDep* __restrict__ sorted, long long n,
                           long long cap, const int* __restrict__ csum,
                           int csum_len, int ntiles)
{
    const int set = blockIdx.y;
    if (csum) n = min((long long)csum[(size_t)set * csum_len + csum_len - 1],
                      cap);
    rec += set * cap;
    tr += set * cap;
    sorted += set * cap;
    off += (size_t)set * (ntiles + 1);
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
         i < n; i += (long long)gridDim.x * blockDim.x) {
        const int2 t = tr[i];
        if (t.x >= 0) sorted[off[t.x] + t.y] = rec[i];
    }
}

// Split tiles' int64 accumulators start at zero (blockIdx.x = split index,
// blockIdx.y = set).
__global__ __launch_bounds__(256) void mc_zero_split(
    unsigned long long* __restrict__ gacc, const int* __restrict__ nsplit,
    int maxsplit, int tile_elems)
{
    const int k = blockIdx.x, set = blockIdx.y;
    if (k >= nsplit[set]) return;
    unsigned long long* a = gacc + ((size_t)set * maxsplit + k) * tile_elems;
    for (int i = threadIdx.x; i < tile_elems; i += blockDim.x) a[i] = 0ull;
}

// One block per (slice, set): the deposits of a slice of a tile's records
// summed in LDS with int64 atomics (the same quantised values the global
// path adds).  A tile of one slice then does cube += sum / 2^S as
// fixed_to_float does; a split tile adds its LDS sums into the tile's int64
// accumulator, converted by mc_split_finish.
__global__ __launch_bounds__(256) void mc_tile_deposit(
    const McDep* __restrict__ sorted, const int* __restrict__ off,
    const int* __restrict__ soff, const int* __restrict__ sidx,
    unsigned long long* __restrict__ gacc, int maxsplit, gbkhip_grid g,
    int tw, int th, int ntx, int ntiles, const double* __restrict__ qs,
    double qscale, float* __restrict__ cube, long long cap,
    long long cstride)
{
    extern __shared__ unsigned long long acc[];      // [nz][tw * th]
    const int slice = blockIdx.x, set = blockIdx.y;
    off += (size_t)set * (ntiles + 1);
    soff += (size_t)set * (ntiles + 1);
    if (slice >= soff[ntiles]) return;
    // the tile t with soff[t] <= slice < soff[t + 1] (the last such t;
    // empty tiles have no slices)
    int a = 0, b = ntiles - 1;
    while (a < b) {
        const int m = (a + b + 1) >> 1;
        if (soff[m] <= slice) a = m; else b = m - 1;
    }
    const int tile = a;
    const int ns = soff[tile + 1] - soff[tile];
    const int lo = off[tile] + (slice - soff[tile]) * kMcSlice;
    const int hi = min(lo + kMcSlice, off[tile + 1]);
    sorted += set * cap;
    cube += set * cstride;
    if (qs) qscale = qs[set];
    const int nx = g.spat_size[0], ny = g.spat_size[1], nz = g.spec_size;
    const int tc = tw * th;
    const size_t plane = (size_t)nx * ny;
    for (int i = threadIdx.x; i < nz * tc; i += blockDim.x) acc[i] = 0ull;
    __syncthreads();
    const int tx0 = (tile % ntx) * tw, ty0 = (tile / ntx) * th;
    for (int i = lo + threadIdx.x; i < hi; i += blockDim.x) {
        const McDep r = sorted[i];
        const int y = r.col / nx, x = r.col - y * nx;
        const int lc = (y - ty0) * tw + (x - tx0);
        int k0, k1;
        line_window(g, r.v, r.s, k0, k1);
        if (k0 > k1) continue;
        const LineEval e = line_eval(g, r.orvalue, r.v, r.s);
        for (int k = k0; k <= k1; ++k) {
            const long long q = __double2ll_rn((double)e.at(k) * qscale);
            if (q) atomicAdd(&acc[k * tc + lc], (unsigned long long)q);
        }
    }
    __syncthreads();
    if (ns > 1) {
        unsigned long long* ga =
            gacc + ((size_t)set * maxsplit + sidx[(size_t)set * ntiles +
                                                   tile]) * (nz * tc);
        for (int i = threadIdx.x; i < nz * tc; i += blockDim.x) {
            const unsigned long long q = acc[i];
            if (q) atomicAdd(&ga[i], q);
        }
        return;
    }
    const double inv = 1.0 / qscale;
    for (int i = threadIdx.x; i < nz * tc; i += blockDim.x) {
        const long long q = (long long)acc[i];
        if (!q) continue;
        const int k = i / tc, lc = i - k * tc;
        const int x = tx0 + lc % tw, y = ty0 + lc / tw;
        if (x < nx && y < ny)
            cube[k * plane + (size_t)y * nx + x] += (float)((double)q * inv);
    }
}

// cube += accumulator / 2^S for the split tiles (blockIdx.x = split index,
// blockIdx.y = set).
__global__ __launch_bounds__(256) void mc_split_finish(
    const unsigned long long* __restrict__ gacc,
    const int* __restrict__ stile, const int* __restrict__ nsplit,
    int maxsplit, gbkhip_grid g, int tw, int th, int ntx,
    const double* __restrict__ qs, double qscale, float* __restrict__ cube,
    long long cstride)
{
    const int k = blockIdx.x, set = blockIdx.y;
    if (k >= nsplit[set]) return;
    const int tile = stile[(size_t)set * maxsplit + k];
    const int nx = g.spat_size[0], ny = g.spat_size[1], nz = g.spec_size;
    const int tc = tw * th;
    const size_t plane = (size_t)nx * ny;
    const unsigned long long* ga =
        gacc + ((size_t)set * maxsplit + k) * (nz * tc);
    cube += set * cstride;
    if (qs) qscale = qs[set];
    const double inv = 1.0 / qscale;
    const int tx0 = (tile % ntx) * tw, ty0 = (tile / ntx) * th;
    for (int i = threadIdx.x; i < nz * tc; i += blockDim.x) {
        const long long q = (long long)ga[i];
        if (!q) continue;
        const int kk = i / tc, lc = i - kk * tc;
        const int x = tx0 + lc % tw, y = ty0 + lc / tw;
        if (x < nx && y < ny)
            cube[kk * plane + (size_t)y * nx + x] +=
                (float)((double)q * inv);
    }
}

int mc_tile_width(int nz)
{
    int tw = 8;
    while (tw > 1 && (size_t)tw * tw * nz * 8 > 64 * 1024) tw /= 2;
    return (size_t)tw * tw * nz * 8 <= 64 * 1024 ? tw : 0;
}

size_t mc_bufs(char* base, int nsets, long long cap, int ntiles, int tw,
               int nz, McBufs* out)
{
    // tile numbers are the wave ranks' sort keys
    GBK_CHECK(ntiles < kMcNoKey, GBKHIP_ERR_UNSUPPORTED,
              "binned mcdisk deposit: " + std::to_string(ntiles) +
              " spatial tiles, beyond its limit of " +
              std::to_string(kMcNoKey - 1));
    McBufs b;
    b.cap = cap;
    b.ntiles = ntiles;
    // a split tile holds more than kMcSlice records; slices beyond one per
    // tile number at most cap / kMcSlice
    const long long extra = (cap + kMcSlice - 1) / kMcSlice;
    b.maxsplit = (int)std::min<long long>(ntiles, extra);
    b.maxslices = (int)std::min<long long>((long long)ntiles + extra,
                                           (1LL << 31) - 1);
    b.tile_elems = tw * tw * nz;
    const size_t ns = (size_t)nsets;
    size_t at = 0;
    auto take = [&](size_t bytes, size_t align) {
        at = (at + align - 1) / align * align;
        char* p = base ? base + at : nullptr;
        at += bytes;
        return p;
    };
    b.rec = reinterpret_cast<McDep*>(take(ns * cap * sizeof(McDep), 256));
    b.sorted = reinterpret_cast<McDep*>(take(ns * cap * sizeof(McDep), 256));
    b.tr = reinterpret_cast<int2*>(take(ns * cap * sizeof(int2), 256));
    b.tcount = reinterpret_cast<int*>(take(ns * ntiles * sizeof(int), 256));
    b.toff = reinterpret_cast<int*>(take(ns * (ntiles + 1) * sizeof(int),
                                         256));
    b.soff = reinterpret_cast<int*>(take(ns * (ntiles + 1) * sizeof(int),
                                         256));
    b.sidx = reinterpret_cast<int*>(take(ns * ntiles * sizeof(int), 256));
    b.stile = reinterpret_cast<int*>(take(ns * b.maxsplit * sizeof(int),
                                          256));
    b.nsplit = reinterpret_cast<int*>(take(ns * sizeof(int), 256));
    b.gacc = reinterpret_cast<unsigned long long*>(
        take(ns * b.maxsplit * (size_t)b.tile_elems * 8, 256));
    at = (at + 255) / 256 * 256;
    if (out) *out = b;
    return at;
}

void launch_mc_binned_tail(Context& c, const McBin& bin, const McBufs& bb,
                           int nsets, long long n, const int* csum,
                           int csum_len, const gbkhip_grid& g,
                           const double* qs, double qscale, float* cube,
                           long long cstride)
{
    const int nt = bb.ntiles;
    hipLaunchKernelGGL(mc_scan_tiles, dim3(nsets), dim3(1024), 0, c.stream,
                       bin.tcount, bb.toff, bb.soff, bb.sidx, bb.stile,
                       bb.nsplit, nt, bb.maxsplit);
    const int sb = std::max(1, std::min(ceil_div(bb.cap, 256), 8192 / nsets));
    hipLaunchKernelGGL(mc_scatter, dim3(sb, nsets), dim3(256), 0, c.stream,
                       bin.rec, bin.tr, bb.toff, bb.sorted, n, bb.cap, csum,
                       csum_len, nt);
    if (bb.maxsplit > 0)
        hipLaunchKernelGGL(mc_zero_split, dim3(bb.maxsplit, nsets), dim3(256),
                           0, c.stream, bb.gacc, bb.nsplit, bb.maxsplit,
                           bb.tile_elems);
    hipLaunchKernelGGL(mc_tile_deposit, dim3(bb.maxslices, nsets), dim3(256),
                       (size_t)bb.tile_elems * 8, c.stream, bb.sorted,
                       bb.toff, bb.soff, bb.sidx, bb.gacc, bb.maxsplit, g,
                       bin.tw, bin.th, bin.ntx, nt, qs, qscale, cube, bb.cap,
                       cstride);
    if (bb.maxsplit > 0)
        hipLaunchKernelGGL(mc_split_finish, dim3(bb.maxsplit, nsets),
                           dim3(256), 0, c.stream, bb.gacc, bb.stile,
                           bb.nsplit, bb.maxsplit, g, bin.tw, bin.th, bin.ntx,
                           qs, qscale, cube, cstride);
    GBK_HIP(hipGetLastError());
}

// cube += icube / 2^S, and icube <- 0 for the next call.
__global__ void fixed_to_float(unsigned long long* icube, float* cube,
                               long long n, double inv_scale)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const long long q = (long long)icube[i];
        if (q) {
            cube[i] += (float)((double)q * inv_scale);
            icube[i] = 0ull;
        }
    }
}

__global__ void zero_u64(unsigned long long* p, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

// Fixed-point exponent: every deposit is |flux| * N(.; v, s) <= |flux| /
// (s_min sqrt(2 pi)) with s_min = 1e-3 channel widths, and at most nclouds
// deposits can land on one voxel, so nclouds * that * 2^S < 2^62.
double mcdisk_fixed_scale(float cflux, int nclouds, const gbkhip_grid& g)
{
    const double flux = std::fabs((double)cflux) /
        ((double)g.spat_step[0] * (double)g.spat_step[1]);
    const double smin = 1e-3 * std::fabs((double)g.spec_step);
    const double peak = flux / (smin * 2.5066282746310002);
    const double bound = std::max(1.0, (double)nclouds) * std::max(peak,
                                                                    1e-30);
    const int e = (int)std::floor(62.0 - std::log2(bound));
    return std::ldexp(1.0, std::min(e, 900));
}

// Fixed scratch of gbkhip_mcdisk_evaluate_noalloc: the binned deposit
// buffers (records, ranks, sorted records, tile offsets, split-tile
// accumulators) of one evaluation, for callers that cannot allocate.  Every
// device holds its own copy until the process ends, at the address
// hipGetSymbolAddress returns on that device; host code uses that address,
// never the symbol.  Calls on one device share it.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

static std::mutex g_scratch_mu;
static std::vector<char*> g_scratch_ptr;          // per device
static std::vector<std::string> g_scratch_err;    // per device, on failure

void noalloc_scratch_setup(int device)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= device) {
        g_scratch_ptr.resize(device + 1, nullptr);
        g_scratch_err.resize(device + 1);
    }
    if (g_scratch_ptr[device]) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        g_scratch_err[device] = std::string("hipGetSymbolAddress(scratch): ") +
                                hipGetErrorString(e);
        return;
    }
    g_scratch_ptr[device] = static_cast<char*>(p);
}

char* noalloc_scratch(int device)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    GBK_CHECK(device >= 0 && device < (int)g_scratch_ptr.size(),
              GBKHIP_ERR_ARG, "no context was created on this device");
    GBK_CHECK(g_scratch_ptr[device] != nullptr, GBKHIP_ERR_HIP,
              g_scratch_err[device].empty()
                  ? std::string("fixed scratch not set up on this device")
                  : g_scratch_err[device]);
    return g_scratch_ptr[device];
}

void launch_mcdisk(Context& c, float cflux, int nclouds, const int* csum,
                   int csum_len, const unsigned char* hoi,
                   const gbkhip_disk& d, const gbkhip_grid& g,
                   const gbkhip_outputs& o, uint64_t seed,
                   char* fixed_scratch, size_t fixed_bytes)
{
    if (nclouds <= 0) return;
    GBK_CHECK(d.rp.uids && d.rh.uids, GBKHIP_ERR_ARG,
              "mcdisk needs density polar and height traits");
    const long long ncube = (long long)g.spat_size[0] * g.spat_size[1] *
                            g.spec_size;
    // binned deposits when the spectral cube is the only output and a tile
    // column of the spectrum fits LDS
    const int nz = g.spec_size;
    const int tw = mc_tile_width(nz);
    const bool only_scube = o.scube && !o.image && !o.wdata &&
        !o.wdata_cmp && !o.rdata && !o.rdata_cmp && !o.ordata &&
        !o.ordata_cmp && !o.vdata_cmp && !o.ddata_cmp;
    const bool binned = only_scube && ncube > 0 && tw > 0;
    if (fixed_scratch) {
        // checked before anything is enqueued: a rejected call does nothing
        GBK_CHECK(binned, GBKHIP_ERR_UNSUPPORTED,
                  "the no-alloc mcdisk evaluation needs the binned deposit: "
                  "the spectral cube as the only output and at most 8192 "
                  "channels");
        const int ntiles = ceil_div(g.spat_size[0], tw) *
                           ceil_div(g.spat_size[1], tw);
        const size_t need = mc_bufs(nullptr, 1, nclouds, ntiles, tw, nz,
                                    nullptr);
        GBK_CHECK(need <= fixed_bytes, GBKHIP_ERR_ARG,
                  "the no-alloc mcdisk evaluation needs " +
                  std::to_string(need) + " bytes of scratch for " +
                  std::to_string(nclouds) + " clouds, beyond its fixed " +
                  std::to_string(fixed_bytes) + " bytes");
    }
    unsigned long long* icube = nullptr;
    if (o.scube && ncube > 0 && !binned) {
        const size_t bytes = (size_t)ncube * sizeof(unsigned long long);
        if (c.int_scratch.bytes < bytes) {
            c.int_scratch.alloc(bytes);
            hipLaunchKernelGGL(zero_u64, dim3(2048), dim3(256), 0, c.stream,
                               c.int_scratch.as<unsigned long long>(), ncube);
            GBK_HIP(hipGetLastError());
        }
        icube = c.int_scratch.as<unsigned long long>();
    }
    McParams mp;
    mp.cflux = cflux;
    mp.nclouds = nclouds;
    const size_t lds = csum_len <= kMcCsumLds ? csum_len * sizeof(int) : 0;
    mp.csum = csum;
    mp.csum_len = csum_len;
    mp.hoi = hoi;
    mp.key0 = (uint32_t)seed;
    mp.key1 = (uint32_t)(seed >> 32);
    mp.qscale = mcdisk_fixed_scale(cflux, nclouds, g);
    const int blocks = std::max(1, std::min(ceil_div(nclouds, 256), 8192));
    if (binned) {
        const int nx = g.spat_size[0], ny = g.spat_size[1];
        const int ntx = ceil_div(nx, tw), nty = ceil_div(ny, tw);
        const int ntiles = ntx * nty;
        const long long nc = nclouds;
        char* base = fixed_scratch;
        if (!base) {
            const size_t need = mc_bufs(nullptr, 1, nc, ntiles, tw, nz,
                                        nullptr);
            if (c.mc_scratch.bytes < need) c.mc_scratch.alloc(need);
            base = static_cast<char*>(c.mc_scratch.ptr);
        }
        McBufs bb;
        mc_bufs(base, 1, nc, ntiles, tw, nz, &bb);
        McBin bin;
        bin.rec = bb.rec;
        bin.tr = bb.tr;
        bin.tcount = bb.tcount;
        bin.tw = bin.th = tw;
        bin.ntx = ntx;
        hipLaunchKernelGGL(mc_zero_i32, dim3(ceil_div(ntiles, 256)),
                           dim3(256), 0, c.stream, bin.tcount, ntiles);
        // each block makes one global reservation per tile it reached (the
        // central tiles are reached by every block); 8192 blocks balance
        // that against the parallelism of the cloud loop (1024 measured 2x
        // slower)
        const int bin_blocks =
            std::max(1, std::min(ceil_div(nclouds, 256), kMcBinBlocks));
        const size_t lds_bin =
            mc_bin_lds(csum_len, d.nrnodes, ntiles).total() * sizeof(int);
        hipLaunchKernelGGL(mcdisk_bin_kernel, dim3(bin_blocks), dim3(256),
                           lds_bin, c.stream, mp, d, g, bin, ntiles);
        launch_mc_binned_tail(c, bin, bb, 1, nc, nullptr, 0, g, nullptr,
                              mp.qscale, o.scube, 0);
        GBK_HIP(hipGetLastError());
        return;
    }
    hipLaunchKernelGGL(mcdisk_kernel, dim3(blocks), dim3(256), lds, c.stream,
                       mp, d, g, o, icube);
    GBK_HIP(hipGetLastError());
    if (icube) {
        hipLaunchKernelGGL(fixed_to_float, dim3(4096), dim3(256), 0, c.stream,
                           icube, o.scube, ncube, 1.0 / mp.qscale);
        GBK_HIP(hipGetLastError());
    }
}

}  // namespace gbkhip
