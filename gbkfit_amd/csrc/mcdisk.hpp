// Per-cloud body of the mcdisk kernels (core/gmodel/gmodels.hpp:325-643):
// the plain one-model kernel and the batched one run the same code.
//
// * Random numbers: Philox4x32-10 (Salmon et al., SC'11) with counter
//   (cloud id, block, 0, 0) and key (seed low, seed high).  Draw i of a cloud
//   is word i % 4 of block i / 4, mapped to ((w >> 9) + 0.5) * 2^-23 in (0, 1)
//   (exact in FP32).  The stream depends only on (seed, cloud id), not on the
//   launch shape.
// * Channel deposits go to an int64 fixed-point cube (scale 2^S) with integer
//   atomics: integer addition is associative, so reruns are bit-identical.
#pragma once
#include "common.hpp"
#include "disk.hpp"

namespace gbkhip {

struct CloudRng {
    uint32_t key0, key1, cid;
    uint32_t block;
    uint32_t next;        // index of the next draw
    uint32_t w0, w1, w2, w3;   // words of the current block (registers)
};

__device__ __forceinline__ void philox(uint32_t c0, uint32_t c1, uint32_t c2,
                                       uint32_t c3, uint32_t k0, uint32_t k1,
                                       uint32_t* out)
{
    for (int r = 0; r < 10; ++r) {
        const uint32_t hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
        const uint32_t hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

__device__ __forceinline__ float rng_next(CloudRng& s)
{
    const uint32_t i = s.next++;
    if ((i >> 2) != s.block) {
        s.block = i >> 2;
        uint32_t w[4];
        philox(s.cid, s.block, 0u, 0u, s.key0, s.key1, w);
        s.w0 = w[0];
        s.w1 = w[1];
        s.w2 = w[2];
        s.w3 = w[3];
    }
    // word i % 4 by selects, not an indexed array (which would live in
    // scratch memory)
    const uint32_t j = i & 3u;
    const uint32_t w = j == 0 ? s.w0 : j == 1 ? s.w1 : j == 2 ? s.w2 : s.w3;
    return ((float)(w >> 9) + 0.5f) * (1.0f / 8388608.0f);
}

// Samplers (math.hpp:152-498) in the reference's draw order.
__device__ __forceinline__ float rnd_uniform(CloudRng& s, float b, float c)
{ return b + rng_next(s) * (c - b); }

__device__ inline float rnd_reject(CloudRng& s, int u, float trunc, float c, float d)
{   // rejection_sampling (math.hpp:206-217) for ggauss (4) / moffat (6)
    float x;
    int guard = 0;
    for (;;) {
        x = rnd_uniform(s, -trunc, trunc);
        const float t0 = u == 4 ? p_ggauss(0, d / (2 * c * tgammaf(1 / d)), 0,
                                           c, d) : 0.0f;
        const float yv = t0 * rng_next(s);
        const float zv = u == 4 ? p_ggauss(x, d / (2 * c * tgammaf(1 / d)), 0,
                                           c, d) : 0.0f;
        if (!(yv > zv) || ++guard >= 100000) break;
    }
    return x;
}

// rh_trait_rnd (traits.hpp:1516-1590); truncated samplers never reject
// (math.hpp:116-124).
__device__ inline float rh_trait_rnd(int uid, const float* c, const float* q,
                              const PPoint& p, CloudRng& s)
{
    const int off = uid >= 1000 ? 1 : 0;
    const int u = uid >= 1000 ? uid - 1000 : uid;
    const bool use_nodes = c[1] != 0.0f;
    const int n = p.nrnodes;
    const float p1 = use_nodes ? lerp_nw(p.r, p.rnidx, p.rnodes, q, off * n)
                               : q[off];
    float p2 = 0.0f;
    if (u == 4 || u == 6)
        p2 = use_nodes ? lerp_nw(p.r, p.rnidx, p.rnodes, q, (off + 1) * n)
                       : q[off + 1];
    switch (u) {
    case 1: return rnd_uniform(s, -p1, p1);
    case 2: { const float v = rng_next(s) * 2 - 1;
              return -p1 * fsgn(v) * logf(1 - 2 * fabsf(v)); }
    case 3: { const float u1 = rng_next(s), u2 = rng_next(s);
              return p1 * sqrtf(-2 * logf(u1)) * cosf(2 * kPi * u2); }
    case 4: return rnd_reject(s, 4, 5 * p1, p1, p2);
    case 5: { const float v = rng_next(s) * 2 - 1;
              return p1 * tanf(kPi * 0.5f * v); }
    case 6: return rnd_reject(s, 6, 5 * p1, p1, p2);
    case 7: { const float v = rng_next(s) * 2 - 1; return p1 * atanhf(v); }
    default: return __builtin_nanf("");
    }
}

// Binned deposits: instead of adding a cloud's channels to the int64 cube
// with global atomics, the cloud writes a record and takes a rank inside
// its spatial tile (tw x th spaxels); a per-tile pass then sums the tile's
// deposits in LDS.  Integer sums: the cube is the same, bit for bit.
struct McDep {
    int col;                 // spaxel y * nx + x
    float orvalue, v, s;     // line flux, velocity, dispersion
};

struct McBin {
    McDep* rec;              // per cloud id
    int2* tr;                // per cloud id: (tile or -1, rank in tile)
    int* tcount;             // per tile: clouds so far
    int tw, th, ntx;         // tile width / height (spaxels), tiles along x
};

// Rank of each lane's key among the lanes of its wave with the same key,
// offset by a reservation of as many slots in counters[key]: the wave sorts
// its 64 keys (bitonic, across lanes), each run of equal keys reserves with
// one atomic at its last lane, and the ranks go back to their lanes through
// 2 x 64 ints of LDS per wave (wscr).  Lanes with key < 0 get 0.  All lanes
// of the wave must call it (inactive lanes pass key = -1).  Same-address
// atomics serialise; this takes one per distinct key and wave, all in one
// instruction.
// Keys must be below kMcNoKey (the host checks the tile count): a key and
// its lane are packed into one int, key * 64 + lane, so the sort moves one
// value per compare-exchange and every packed value is distinct.
constexpr int kMcNoKey = (1 << 25) - 1;

__device__ __forceinline__ int wave_rank(int key, int* counters, int* wscr)
{
    const int lane = __lane_id();
    int kk = ((key < 0 ? kMcNoKey : key) << 6) | lane;
#pragma unroll
    for (int size = 2; size <= 64; size <<= 1) {
#pragma unroll
        for (int j = size >> 1; j > 0; j >>= 1) {
            const int pk = __shfl_xor(kk, j, 64);
            const bool up = (lane & size) == 0, lower = (lane & j) == 0;
            kk = lower == up ? min(kk, pk) : max(kk, pk);
        }
    }
    const int k = kk >> 6, idx = kk & 63;
    // runs of equal keys
    const int prev = __shfl_up(k, 1, 64), next = __shfl_down(k, 1, 64);
    const bool head = lane == 0 || prev != k;
    const bool tail = lane == 63 || next != k;
    // the run's head: the highest head lane at or below this one
    const unsigned long long heads = __ballot(head);
    const unsigned long long upto = lane == 63 ? ~0ull
                                               : (2ull << lane) - 1ull;
    const int start = 63 - __clzll((long long)(heads & upto));
    if (tail && k != kMcNoKey)
        wscr[start] = atomicAdd(&counters[k], lane - start + 1);
    __builtin_amdgcn_wave_barrier();
    const int rank = k != kMcNoKey ? wscr[start] + lane - start : 0;
    __builtin_amdgcn_wave_barrier();
    wscr[64 + idx] = rank;
    __builtin_amdgcn_wave_barrier();
    return wscr[64 + lane];
}

struct Context;
// Tile width (spaxels) of binned deposits for nz channels: the largest
// power of two whose tile spectrum (tw^2 nz int64) fits 64 KB of LDS, or 0.
int mc_tile_width(int nz);

// Records per tile-deposit block: a tile with more records is split into
// slices of at most kMcSlice, one block each, whose LDS sums meet in an int64
// accumulator of the tile (integer sums: the result does not depend on the
// split), converted afterwards.  The central tiles of a disk hold most
// clouds; one block per tile serialised on them.
constexpr int kMcSlice = 8192;

// Binned deposit buffers of nsets sets, set-major within each array.
struct McBufs {
    McDep* rec = nullptr;        // cap per set, by cloud id
    McDep* sorted = nullptr;     // cap per set, by tile
    int2* tr = nullptr;          // cap per set: (tile or -1, rank in tile)
    int* tcount = nullptr;       // ntiles per set
    int* toff = nullptr;         // ntiles + 1 per set: record offsets
    int* soff = nullptr;         // ntiles + 1 per set: slice offsets
    int* sidx = nullptr;         // ntiles per set: split index or -1
    int* stile = nullptr;        // maxsplit per set: tile of a split index
    int* nsplit = nullptr;       // 1 per set
    unsigned long long* gacc = nullptr;   // maxsplit x tile spectrum per set
    long long cap = 0;
    int ntiles = 0, maxsplit = 0, maxslices = 0;
    int tile_elems = 0;          // tw^2 nz
};
// Carves the buffers for (nsets, cap, ntiles, tw, nz) out of base (null:
// sizes only) and returns the bytes they take (256-byte aligned parts).
size_t mc_bufs(char* base, int nsets, long long cap, int ntiles, int tw,
               int nz, McBufs* out);
// Tile offsets, tile-ordered records and per-tile LDS deposits for nsets
// sets whose records, ranks and tile counts were written; cube += the
// deposits.  A set's cloud count is its last cumulative count (csum) or n.
void launch_mc_binned_tail(Context& c, const McBin& bin, const McBufs& bb,
                           int nsets, long long n, const int* csum,
                           int csum_len, const gbkhip_grid& g,
                           const double* qs, double qscale, float* cube,
                           long long cstride);

struct McParams {
    float cflux;
    int nclouds;
    const int* csum;
    int csum_len;
    const unsigned char* hoi;
    uint32_t key0, key1;
    double qscale;          // fixed-point scale 2^S
};

// The cumulative cloud counts are read by every cloud's ring search (a
// linear scan, as gmodels.hpp:401-416): a block stages them in LDS once
// when they fit.  Every thread of the block calls this (it syncs).
constexpr int kMcCsumLds = 8192;

__device__ __forceinline__ const int* mc_stage_csum(const int* csum, int len,
                                                    int* lds)
{
    if (len > kMcCsumLds) return csum;
    for (int i = threadIdx.x; i < len; i += blockDim.x) lds[i] = csum[i];
    __syncthreads();
    return lds;
}

// Sub-ring nodes, read several times per cloud (ring bounds, disk_info,
// interpolation): staged in LDS as well when they fit.
constexpr int kMcNodesLds = 1024;

__device__ __forceinline__ const float* mc_stage_nodes(const float* nodes,
                                                      int n, float* lds)
{
    if (n > kMcNodesLds) return nodes;
    for (int i = threadIdx.x; i < n; i += blockDim.x) lds[i] = nodes[i];
    __syncthreads();
    return lds;
}

// Binned cloud kernels: per-block tile counts in LDS when there are at most
// this many tiles (else ranks come from global atomics).
constexpr int kMcBinLds = 4096;

// Their dynamic LDS, in ints: staged cumulative counts, staged sub-ring
// nodes and per-tile counts, each present only when it fits, so a block
// takes no more LDS than its cube needs (more blocks per CU).
struct McBinLds {
    int csum, nodes, cnt;
    __host__ __device__ int total() const { return csum + nodes + cnt; }
};
__host__ __device__ inline McBinLds mc_bin_lds(int csum_len, int nrnodes,
                                               int ntiles)
{
    McBinLds l;
    l.csum = csum_len <= kMcCsumLds ? csum_len : 0;
    l.nodes = nrnodes <= kMcNodesLds ? nrnodes : 0;
    l.cnt = ntiles <= kMcBinLds ? ntiles : 0;
    return l;
}

// Returns the cloud's tile when `bin` is given and the cloud lands in the
// cube (its record written), else -1.
__device__ __forceinline__ int mc_cloud(long long ci, const McParams& mp,
                                        const gbkhip_disk& d,
                                        const gbkhip_grid& g,
                                        const gbkhip_outputs& o,
                                        unsigned long long* icube,
                                        const McBin* bin = nullptr)
{
    const int nx = g.spat_size[0], ny = g.spat_size[1], nz = g.spat_size[2];
    const size_t plane = (size_t)nx * ny;
    CloudRng S;
    S.key0 = mp.key0;
    S.key1 = mp.key1;
    S.cid = (uint32_t)ci;
    S.block = 0xffffffffu;
    S.next = 0;
    // ring of the cloud (gmodels.hpp:401-416): the first ring whose
    // cumulative count exceeds the cloud id.  The reference scans the
    // counts linearly; they are non-decreasing, so the count of entries
    // <= ci, found by a branch-free binary search (the same trip count on
    // every lane), is the same index.
    int rnidx = 0, tidx;
    {
        const int len = mp.csum_len;
        int step = len > 0 ? 1 << (31 - __clz(len)) : 0;
        for (; step > 0; step >>= 1)
            if (rnidx + step <= len && mp.csum[rnidx + step - 1] <= ci)
                rnidx += step;
    }
    if (rnidx >= mp.csum_len) return -1;     // cloud beyond the budget
    const float* rpc = d.rp.cvalues;
    const float* rpp = d.rp.pvalues;
    const float* rhc = d.rh.cvalues;
    const float* rhp = d.rh.pvalues;
    for (tidx = 0; tidx < d.rp.n; ++tidx) {
        const int size = mp.hoi[tidx] ? 1 : d.nrnodes - 2;
        if (rnidx < size) break;
        rnidx -= size;
        rpc += d.rp.ccounts[tidx];
        rpp += d.rp.pcounts[tidx];
        rhc += d.rh.ccounts[tidx];
        rhp += d.rh.pcounts[tidx];
    }
    if (tidx >= d.rp.n) return -1;           // budget longer than traits
    // rp_trait_rnd (traits.hpp:310-326,1333-1435)
    const int uid = d.rp.uids[tidx];
    const int u = uid >= 1000 ? uid - 1000 : uid;
    float sign = 1.0f, rd = 0.0f, theta = 0.0f;
    {
        const int idx = rnidx + 1;
        float rmin = 0.0f, rmax = 0.0f;
        bool sample = true;
        if (u == 1) {
            rmin = d.rnodes[0];
            rmax = d.rnodes[d.nrnodes - 1];
        } else if ((u >= 2 && u <= 7) || u == 101 || u == 102) {
            const float hw = d.rnodes[1] - d.rnodes[0];
            rmin = d.rnodes[idx] - hw;
            rmax = d.rnodes[idx] + hw;
        } else {
            sample = false;  // mixtures / distortion: not implemented
        }
        if (sample) {
            rd = sqrtf(rmax * rmax + (rmin * rmin - rmax * rmax) *
                       rng_next(S));
            theta = 2 * kPi * rng_next(S);
            if (u == 102) {
                const int k = (int)rintf(rpc[0]);
                const float a = rpp[idx];
                const float ph = rpp[d.nrnodes + idx] * kD2R;
                const float vr = a * (rng_next(S) * 2 - 1);
                const float vf = a * cosf(k * (theta - ph));
                sign = vr > vf ? 1.0f : -1.0f;
            }
        }
    }
    float rvalue = mp.cflux / g.spat_step[2];
    if (sign < 0) rvalue = -rvalue;
    rvalue *= g.spat_step[2];
    rvalue /= g.spat_step[0] * g.spat_step[1];
    float st, ct;
    sincosf(theta, &st, &ct);
    const float xd = rd * ct, yd = rd * st;
    if (!disk_info(rnidx, rd, d.nrnodes, d.rnodes)) return -1;

    PPoint P;
    P.rnidx = rnidx;
    P.nrnodes = d.nrnodes;
    P.rnodes = d.rnodes;
    P.x = xd;
    P.y = yd;
    P.r = rd;
    P.theta = theta;
    P.ct = ct;
    P.st = st;
    float pt[kTraitMax] = {0, 0, 0, 0}, ht[kTraitMax] = {0, 0, 0, 0};
    if (d.sp.uids) {
        float sv = 0;
        p_traits<TK_SP>(d.sp, P, pt);
        for (int i = 0; i < d.sp.n && i < kTraitMax; ++i) sv += pt[i];
        if (!sv) return -1;
    }
    float zd = rh_trait_rnd(d.rh.uids[tidx], rhc, rhp, P, S);
    if (d.zp.uids) {
        float zv = 0;
        p_traits<TK_ZP>(d.zp, P, pt);
        for (int i = 0; i < d.zp.n && i < kTraitMax; ++i) zv += pt[i];
        zd += zv;
    }
    const float vsysi = d.vsys ? (d.loose ? lerp_nw(rd, rnidx, d.rnodes,
                                                    d.vsys, 0)
                                          : d.vsys[0]) : 0.0f;
    const float xposi = d.loose ? lerp_nw(rd, rnidx, d.rnodes, d.xpos, 0)
                                : d.xpos[0];
    const float yposi = d.loose ? lerp_nw(rd, rnidx, d.rnodes, d.ypos, 0)
                                : d.ypos[0];
    const float posai = (d.tilted ? lerp_nw(rd, rnidx, d.rnodes, d.posa, 0)
                                  : d.posa[0]) * kD2R;
    const float incli = (d.tilted ? lerp_nw(rd, rnidx, d.rnodes, d.incl, 0)
                                  : d.incl[0]) * kD2R;
    float sp, cp, si, ci_;
    sincosf(posai, &sp, &cp);
    sincosf(incli, &si, &ci_);
    P.si = si;
    P.ci = ci_;
    // disk -> sky (gmodels.hpp:159-165,496): rotations by -incl, -posa
    float xn = xd, yn, zn;
    rot_lh_x(yn, zn, yd, zd, -si, ci_);
    float xr, yr;
    rot_lh_z(xr, yr, xn, yn, -sp, cp);
    xr += xposi;
    yr += yposi;
    // A NaN position (the exponential height sampler of math.hpp:252-255
    // returns NaN for |u| > 1/2) lands nowhere: on the host driver
    // rint(NaN) converts to INT_MIN, outside the cube, while the GPU's
    // conversion would give 0.
    if (isnan(xr) || isnan(yr) || isnan(zn)) return -1;
    const int x = (int)rintf((xr - g.spat_zero[0]) / g.spat_step[0]);
    const int y = (int)rintf((yr - g.spat_zero[1]) / g.spat_step[1]);
    const int z = (int)rintf((zn - g.spat_zero[2]) / g.spat_step[2]);
    if (x < 0 || x >= nx || y < 0 || y >= ny || z < 0 || z >= nz)
        return -1;

    float vvalue = 0, dvalue = 0, wvalue = 1;
    if (d.vp.uids) p_traits<TK_VP>(d.vp, P, pt);
    if (d.vh.uids) one_traits(d.vh, ht);
    for (int i = 0; i < d.vp.n && i < kTraitMax; ++i) vvalue += pt[i] * ht[i];
    vvalue += vsysi;
    if (d.dp.uids) p_traits<TK_DP>(d.dp, P, pt);
    if (d.dh.uids) one_traits(d.dh, ht);
    for (int i = 0; i < d.dp.n && i < kTraitMax; ++i) dvalue += pt[i] * ht[i];
    dvalue = fabsf(dvalue);
    if (d.wp.uids && o.wdata)
        for (int i = 0; i < d.wp.n && i < kTraitMax; ++i)
            wvalue *= d.wp.uids[i] == 1 ? 1.0f : __builtin_nanf("");

    float orvalue = rvalue;
    if (d.opacity)
        for (int oz = z + 1; oz < nz; ++oz)
            orvalue -= orvalue *
                d.opacity[x + (size_t)y * nx + (size_t)oz * plane];

    const size_t col = (size_t)y * nx + x;
    if (bin) {               // record; the caller ranks it in its tile
        McDep r;
        r.col = (int)col;
        r.orvalue = orvalue;
        r.v = vvalue;
        r.s = dvalue;
        bin->rec[ci] = r;
        return (y / bin->th) * bin->ntx + x / bin->tw;
    }
    if (o.image) atomicAdd(&o.image[col], orvalue);
    if (icube) {
        int k0, k1;
        line_window(g, vvalue, dvalue, k0, k1);
        if (k0 <= k1) {
            const LineEval e = line_eval(g, orvalue, vvalue, dvalue);
            for (int k = k0; k <= k1; ++k) {
                const long long q = __double2ll_rn((double)e.at(k) *
                                                   mp.qscale);
                if (q)
                    atomicAdd(&icube[k * plane + col],
                              (unsigned long long)q);
            }
        }
    }
    const size_t idx = col + (size_t)z * plane;
    if (o.wdata) o.wdata[idx] = wvalue;
    if (o.wdata_cmp) o.wdata_cmp[idx] = wvalue;
    if (o.rdata) atomicAdd(&o.rdata[idx], rvalue);
    if (o.rdata_cmp) atomicAdd(&o.rdata_cmp[idx], rvalue);
    if (o.ordata) atomicAdd(&o.ordata[idx], orvalue);
    if (o.ordata_cmp) atomicAdd(&o.ordata_cmp[idx], orvalue);
    if (o.vdata_cmp) o.vdata_cmp[idx] = vvalue;
    if (o.ddata_cmp) o.ddata_cmp[idx] = dvalue;
    return -1;
}

}  // namespace gbkhip
