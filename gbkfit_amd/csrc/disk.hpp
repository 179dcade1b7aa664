// Device-side (FP32) evaluation of gbkfit disk traits and geometry.
//
// The arithmetic follows the reference's header-only core:
//   profiles            core/math/math.hpp:126-498
//   polar/height traits core/gmodel/traits.hpp:137-1919
//   geometry            core/gmodel/gmodels.hpp:118-323
// Trait UIDs are the reference's (traits.hpp:7-135) because they are the
// wire format the Python layer packs (src/gbkfit/model/gmodels/traits.py).
//
// Departures that do not change results beyond FP32 rounding:
//   * cos(theta), sin(theta) of the in-plane angle are x/r, y/r (atan2 is
//     only evaluated for traits that need theta itself, and at r == 0);
//   * per-model sin/cos of the position angle and inclination are computed
//     once per thread instead of once per voxel.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/gbkhip.h"

namespace gbkhip {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kD2R = 3.14159265358979323846f / 180.0f;
constexpr float kLineWidth = 5.0f;           // core/constants.hpp:5-6
constexpr int kTraitMax = 4;                 // core/constants.hpp:8

// ---- 1-D profiles ------------------------------------------------------
__device__ __forceinline__ float fsgn(float x)
{
    return (float)((0.0f < x) - (x < 0.0f));
}

__device__ __forceinline__ float wrap_angle(float a)
{
    while (a < -kPi) a += 2 * kPi;
    while (a > kPi) a -= 2 * kPi;
    return a;
}

__device__ __forceinline__ float p_uniform(float x, float a, float b, float c)
{ return (x >= b && x <= c) ? a : 0.0f; }
__device__ __forceinline__ float p_expon(float x, float a, float b, float c)
{ return a * expf(-fabsf(x - b) / c); }
__device__ __forceinline__ float p_gauss(float x, float a, float b, float c)
{ return a * expf(-(x - b) * (x - b) / (2 * c * c)); }
__device__ __forceinline__ float p_ggauss(float x, float a, float b, float c,
                                          float d)
{ return a * expf(-powf(fabsf(x - b) / c, d)); }
__device__ __forceinline__ float p_lorentz(float x, float a, float b, float c)
{ return a * c * c / ((x - b) * (x - b) + c * c); }
__device__ __forceinline__ float p_moffat(float x, float a, float b, float c,
                                          float d)
{ float t = (x - b) / c; return a / powf(1 + t * t, d); }
__device__ __forceinline__ float p_sech2(float x, float a, float b, float c)
{ float s = 1.0f / coshf((x - b) / c); return a * s * s; }

__device__ __forceinline__ float c_uniform(float x, float b, float c)
{ return x < b ? 0.0f : (x > c ? 1.0f : (x - b) / (c - b)); }
__device__ __forceinline__ float c_expon(float x, float b, float c)
{ return 0.5f + 0.5f * fsgn(x - b) * (1 - expf(-fabsf(x - b) / c)); }
__device__ __forceinline__ float c_gauss(float x, float b, float c)
{ return 0.5f * (1 + erff((x - b) / (c * 1.41421356237309505f))); }
__device__ __forceinline__ float c_lorentz(float x, float b, float c)
{ return 0.5f + atanf((x - b) / c) / kPi; }
__device__ __forceinline__ float c_sech2(float x, float b, float c)
{ return 0.5f * (1 + tanhf((x - b) / c)); }

__device__ __forceinline__ float lerp_nw(float r, int idx, const float* xs,
                                         const float* ys, int offset)
{   // math.hpp:500-511
    const float x1 = xs[idx - 1], x2 = xs[idx];
    const float y1 = ys[offset + idx - 1], y2 = ys[offset + idx];
    const float t = (r - x1) / (x2 - x1);
    return y1 + t * (y2 - y1);
}

// ---- evaluation point ---------------------------------------------------
struct PPoint {
    int rnidx;
    int nrnodes;
    const float* rnodes;
    float x, y, r;
    float ct, st;      // cos/sin of the in-plane angle
    float theta;       // valid only when need_theta
    float si, ci;      // sin/cos of the inclination
};

__device__ __forceinline__ void rot_lh_z(float& ox, float& oy, float x,
                                         float y, float s, float c)
{   // math.hpp:62-71
    ox = -x * s + y * c;
    oy = -x * c - y * s;
}

__device__ __forceinline__ void rot_lh_x(float& oy, float& oz, float y,
                                         float z, float s, float c)
{   // math.hpp:44-51
    oy = y * c + z * s;
    oz = -y * s + z * c;
}

// Blob mixtures (traits.hpp:199-246): np = 6 or 7 parameters per blob.
__device__ inline float mixture(int fun, int np, const PPoint& p,
                                const float* cv, const float* q)
{
    const int nb = (int)rintf(cv[0]);
    float res = 0;
    for (int i = 0; i < nb; ++i) {
        const float br = q[i];
        const float bt = q[nb + i] * kD2R;
        const float ba = q[2 * nb + i];
        const float bs = q[3 * nb + i];
        const float bb = np == 7 ? q[4 * nb + i] : 0.0f;
        const float bq = q[(np == 7 ? 5 : 4) * nb + i];
        const float bp = q[(np == 7 ? 6 : 5) * nb + i] * kD2R;
        float sbt, cbt, sb, cb;
        sincosf(bt, &sbt, &cbt);
        float xn = p.r * cosf(p.theta) - br * cbt;
        float yn = p.r * sinf(p.theta) - br * sbt;
        sincosf(bt + bp, &sb, &cb);
        float xr, yr;
        rot_lh_z(xr, yr, xn, yn, sb, cb);
        const float rn = sqrtf(xr * xr + (yr * yr) / (bq * bq));
        switch (fun) {
        case 2: res += p_expon(rn, ba, 0, bs); break;
        case 3: res += p_gauss(rn, ba, 0, bs); break;
        case 4: res += p_ggauss(rn, ba, 0, bs, bb); break;
        default: res += p_moffat(rn, ba, 0, bs, bb); break;
        }
    }
    return res;
}

__device__ __forceinline__ float nw_harmonic(const PPoint& p, const float* cv,
                                             const float* q)
{   // traits.hpp:285-295
    const int k = (int)rintf(cv[0]);
    const float a = lerp_nw(p.r, p.rnidx, p.rnodes, q, 0);
    const float ph = k ? lerp_nw(p.r, p.rnidx, p.rnodes, q, p.nrnodes) * kD2R
                       : 0.0f;
    return a * cosf(k * (p.theta - ph));
}

__device__ __forceinline__ float nw_distortion(const PPoint& p, const float* q)
{   // traits.hpp:297-308
    const float a = lerp_nw(p.r, p.rnidx, p.rnodes, q, 0);
    const float ph = lerp_nw(p.r, p.rnidx, p.rnodes, q, p.nrnodes) * kD2R;
    const float s = lerp_nw(p.r, p.rnidx, p.rnodes, q, 2 * p.nrnodes);
    const float t = wrap_angle(p.theta - ph);
    return a * expf(-(t * t * p.r * p.r) / (2 * s * s));
}

// Density / opacity polar traits (traits.hpp:1242-1331).
__device__ __forceinline__ float rp_trait(int uid, const float* c,
                                          const float* q, const PPoint& p)
{
    switch (uid >= 1000 ? uid - 1000 : uid) {
    case 1: return q[0];
    case 2: return p_expon(p.r, q[0], 0, q[1]);
    case 3: return p_gauss(p.r, q[0], 0, q[1]);
    case 4: return p_ggauss(p.r, q[0], 0, q[1], q[2]);
    case 5: return p_lorentz(p.r, q[0], 0, q[1]);
    case 6: return p_moffat(p.r, q[0], 0, q[1], q[2]);
    case 7: return p_sech2(p.r, q[0], 0, q[1]);
    case 51: return mixture(2, 6, p, c, q);
    case 52: return mixture(3, 6, p, c, q);
    case 53: return mixture(4, 7, p, c, q);
    case 54: return mixture(6, 7, p, c, q);
    case 101: return lerp_nw(p.r, p.rnidx, p.rnodes, q, 0);
    case 102: return nw_harmonic(p, c, q);
    case 103: return nw_distortion(p, q);
    default: return __builtin_nanf("");
    }
}

__device__ __forceinline__ bool rp_needs_theta(int uid)
{
    const int u = uid >= 1000 ? uid - 1000 : uid;
    return u >= 51;
}

// Surface-brightness height pdfs (uid 1..7) and opacity height functions
// (1001..1007): traits.hpp:561-896,1437-1514.
__device__ inline float rh_trait(int uid, const float* c, const float* q,
                                 const PPoint& p, float z)
{
    const float trunc = c[0];
    const bool use_nodes = c[1] != 0.0f;
    const int n = p.nrnodes;
    auto nwp = [&](int k) {
        return use_nodes ? lerp_nw(p.r, p.rnidx, p.rnodes, q, k * n) : q[k];
    };
    if (uid < 1000) {
        const float p1 = nwp(0);
        const float p2 = (uid == 4 || uid == 6) ? nwp(1) : 0.0f;
        float v, norm = 1.0f;
        switch (uid) {
        case 1: v = p_uniform(z, 1 / (p1 + p1), -p1, p1); break;
        case 2: v = p_expon(z, 1 / (2 * p1), 0, p1); break;
        case 3: v = p_gauss(z, 1 / (p1 * 2.50662827463100050f), 0, p1); break;
        case 4: v = p_ggauss(z, p2 / (2 * p1 * tgammaf(1 / p2)), 0, p1, p2);
                break;
        case 5: v = p_lorentz(z, 1 / (kPi * p1), 0, p1); break;
        case 6: v = 0.0f; break;          // moffat pdf is a stub (math.hpp:426)
        case 7: v = p_sech2(z, 1 / (2 * p1), 0, p1); break;
        default: return __builtin_nanf("");
        }
        if (trunc == 0) return v;
        const float lo = -trunc * p1, hi = trunc * p1;
        if (!(z >= lo && z <= hi)) return 0.0f;
        switch (uid) {     // truncated pdf normalisation (math.hpp:108-114)
        case 1: norm = c_uniform(hi, -p1, p1) - c_uniform(lo, -p1, p1); break;
        case 2: norm = c_expon(hi, 0, p1) - c_expon(lo, 0, p1); break;
        case 3: norm = c_gauss(hi, 0, p1) - c_gauss(lo, 0, p1); break;
        case 5: norm = c_lorentz(hi, 0, p1) - c_lorentz(lo, 0, p1); break;
        case 7: norm = c_sech2(hi, 0, p1) - c_sech2(lo, 0, p1); break;
        default: norm = 0.0f; break;     // ggauss/moffat cdf stubs return 0
        }
        return v / norm;
    }
    const int u = uid - 1000;
    const float p1 = nwp(0), p2 = nwp(1);
    // 3-parameter forms read params[1] for the shape when not node-wise
    // (traits.hpp:666-668)
    const float p3 = (u == 4 || u == 6)
        ? (use_nodes ? lerp_nw(p.r, p.rnidx, p.rnodes, q, 2 * n) : q[1]) : 0.0f;
    float v;
    switch (u) {
    case 1: v = p_uniform(z, p1, -p2, p2); break;
    case 2: v = p_expon(z, p1, 0, p2); break;
    case 3: v = p_gauss(z, p1, 0, p2); break;
    case 4: v = p_ggauss(z, p1, 0, p2, p3); break;
    case 5: v = p_lorentz(z, p1, 0, p2); break;
    case 6: v = p_moffat(z, p1, 0, p2, p3); break;
    case 7: v = p_sech2(z, p1, 0, p2); break;
    default: return __builtin_nanf("");
    }
    if (trunc == 0) return v;
    return (z >= -trunc * p2 && z <= trunc * p2) ? v : 0.0f;
}

// Velocity polar traits (traits.hpp:898-1078,1592-1673).
__device__ __forceinline__ float vp_trait(int uid, const float* c,
                                          const float* q, const PPoint& p)
{
    const float r = p.r;
    float v;
    switch (uid) {
    case 1: return q[0] * p.ct * p.si;
    case 2: v = q[1] * (2 / kPi) * atanf(r / q[0]); return v * p.ct * p.si;
    case 3: v = q[1] * (1 - expf(-r / q[0])); return v * p.ct * p.si;
    case 4: v = q[1] * powf(r / q[0], q[3]) / (1 + powf(r / q[0], q[2]));
            return v * p.ct * p.si;
    case 5: v = r < q[0] ? q[1] * r / q[0] : q[1]; return v * p.ct * p.si;
    case 6: v = q[1] * tanhf(r / q[0]); return v * p.ct * p.si;
    case 7: v = q[1] * (1 - expf(-r / q[0])) * (1 + q[2] * r / q[0]);
            return v * p.ct * p.si;
    case 8: v = q[1] * powf(1 + r / q[0], q[2]) *
                powf(1 + powf(r / q[0], -q[3]), -1 / q[3]);
            return v * p.ct * p.si;
    case 101: if (r == 0) return 0.0f;
              return lerp_nw(r, p.rnidx, p.rnodes, q, 0) * p.ct * p.si;
    case 102: if (r == 0) return 0.0f; return nw_harmonic(p, c, q) * p.ct * p.si;
    case 103: if (r == 0) return 0.0f;
              return lerp_nw(r, p.rnidx, p.rnodes, q, 0) * p.st * p.si;
    case 104: if (r == 0) return 0.0f; return nw_harmonic(p, c, q) * p.st * p.si;
    case 105: return lerp_nw(r, p.rnidx, p.rnodes, q, 0) * p.ci;
    case 106: return nw_harmonic(p, c, q) * p.ci;
    case 107: return lerp_nw(r, p.rnidx, p.rnodes, q, 0);
    case 108: return nw_harmonic(p, c, q);
    default: return __builtin_nanf("");
    }
}

__device__ __forceinline__ bool vp_needs_theta(int uid)
{
    return uid == 102 || uid == 104 || uid == 106 || uid == 108;
}

// Dispersion polar traits (traits.hpp:1080-1175,1702-1775): the same
// formulas as the density polar traits without the opacity aliases.
__device__ __forceinline__ float dp_trait(int uid, const float* c,
                                          const float* q, const PPoint& p)
{
    if (uid >= 1000) return __builtin_nanf("");
    return rp_trait(uid, c, q, p);
}

__device__ __forceinline__ float zp_trait(int uid, const float* c,
                                          const float* q, const PPoint& p)
{   // traits.hpp:1804-1829
    if (uid == 101) return lerp_nw(p.r, p.rnidx, p.rnodes, q, 0);
    if (uid == 102) return nw_harmonic(p, c, q);
    return __builtin_nanf("");
}

__device__ __forceinline__ float azrange(float theta, float ph, float s)
{   // traits.hpp:1195-1208
    float pmin = ph - s * 0.5f, pmax = ph + s * 0.5f;
    pmin = wrap_angle(pmin);
    pmax = wrap_angle(pmax);
    if (pmin < pmax) return (theta > pmin && theta < pmax) ? 1.0f : 0.0f;
    return (theta < pmin && theta > pmax) ? 0.0f : 1.0f;
}

__device__ __forceinline__ float sp_trait(int uid, const float* c,
                                          const float* q, const PPoint& p)
{   // traits.hpp:1831-1857
    (void)c;
    if (uid == 1) return azrange(p.theta, q[0] * kD2R, q[1] * kD2R);
    if (uid == 101)
        return azrange(p.theta, lerp_nw(p.r, p.rnidx, p.rnodes, q, 0) * kD2R,
                       lerp_nw(p.r, p.rnidx, p.rnodes, q, p.nrnodes) * kD2R);
    return __builtin_nanf("");
}

// ---- trait lists ----------------------------------------------------------
enum TraitKind { TK_RP, TK_VP, TK_DP, TK_ZP, TK_SP };

template<int KIND>
__device__ __forceinline__ float trait_eval(int uid, const float* c,
                                            const float* q, const PPoint& p)
{
    if (KIND == TK_RP) return rp_trait(uid, c, q, p);
    if (KIND == TK_VP) return vp_trait(uid, c, q, p);
    if (KIND == TK_DP) return dp_trait(uid, c, q, p);
    if (KIND == TK_ZP) return zp_trait(uid, c, q, p);
    return sp_trait(uid, c, q, p);
}

// out[i] for each trait of a packed list (traits.hpp:1887-1902).
template<int KIND>
__device__ __forceinline__ void p_traits(const gbkhip_traits& t,
                                         const PPoint& p, float* out)
{
    const float* c = t.cvalues;
    const float* q = t.pvalues;
#pragma unroll
    for (int i = 0; i < kTraitMax; ++i) {
        if (i >= t.n) break;
        out[i] = trait_eval<KIND>(t.uids[i], c, q, p);
        c += t.ccounts[i];
        q += t.pcounts[i];
    }
}

__device__ __forceinline__ void rh_traits(const gbkhip_traits& t,
                                          const PPoint& p, float z, float* out)
{
    const float* c = t.cvalues;
    const float* q = t.pvalues;
#pragma unroll
    for (int i = 0; i < kTraitMax; ++i) {
        if (i >= t.n) break;
        out[i] = rh_trait(t.uids[i], c, q, p, z);
        c += t.ccounts[i];
        q += t.pcounts[i];
    }
}

__device__ __forceinline__ void one_traits(const gbkhip_traits& t, float* out)
{
#pragma unroll
    for (int i = 0; i < kTraitMax; ++i)
        if (i < t.n) out[i] = t.uids[i] == 1 ? 1.0f : __builtin_nanf("");
}

// True when any trait of the disk needs the angle theta itself (the
// others only use cos/sin of it).
__device__ __forceinline__ bool polar_needs_theta(int uid)
{
    const int u = uid >= 1000 ? uid - 1000 : uid;
    return (u >= 51 && u <= 54) || u == 102 || u == 103;
}

__device__ __forceinline__ bool disk_needs_theta(const gbkhip_disk& d)
{
    bool need = d.sp.uids != nullptr;
    for (int i = 0; d.rp.uids && i < d.rp.n && i < kTraitMax; ++i)
        need |= polar_needs_theta(d.rp.uids[i]);
    for (int i = 0; d.dp.uids && i < d.dp.n && i < kTraitMax; ++i)
        need |= polar_needs_theta(d.dp.uids[i]);
    for (int i = 0; d.vp.uids && i < d.vp.n && i < kTraitMax; ++i)
        need |= vp_needs_theta(d.vp.uids[i]);
    for (int i = 0; d.zp.uids && i < d.zp.n && i < kTraitMax; ++i)
        need |= d.zp.uids[i] == 102;
    return need;
}

// ---- geometry ---------------------------------------------------------------
// disk_info (gmodels.hpp:197-215): first i >= 1 with r < nodes[i].
__device__ __forceinline__ bool disk_info(int& index, float r, int n,
                                          const float* nodes)
{
    if (r < nodes[0]) return false;
    if (r >= nodes[n - 1]) return false;
    int lo = 1, hi = n - 1;          // binary search, same answer as the
    while (lo < hi) {                // reference's linear scan
        const int mid = (lo + hi) >> 1;
        if (r < nodes[mid]) hi = mid; else lo = mid + 1;
    }
    index = lo;
    return true;
}

struct NodeGeom {
    float xpos, ypos, sp, cp, si, ci;
};

__device__ __forceinline__ NodeGeom node_geom(const gbkhip_disk& d, int idx)
{
    NodeGeom g;
    g.xpos = d.loose ? d.xpos[idx] : d.xpos[0];
    g.ypos = d.loose ? d.ypos[idx] : d.ypos[0];
    const float pa = (d.tilted ? d.posa[idx] : d.posa[0]) * kD2R;
    const float in = (d.tilted ? d.incl[idx] : d.incl[0]) * kD2R;
    sincosf(pa, &g.sp, &g.cp);
    sincosf(in, &g.si, &g.ci);
    return g;
}

// rnode_radius (gmodels.hpp:167-195).
__device__ __forceinline__ float rnode_radius(const gbkhip_disk& d, int idx,
                                              float x, float y, float z,
                                              bool thick)
{
    const NodeGeom g = node_geom(d, idx);
    x -= g.xpos;
    y -= g.ypos;
    float xr, yr, zr = z;
    rot_lh_z(xr, yr, x, y, g.sp, g.cp);
    if (thick) rot_lh_x(yr, zr, yr, z, g.si, g.ci);
    else yr /= g.ci;
    return sqrtf(xr * xr + yr * yr);
}

// ring_info (gmodels.hpp:217-323): linear scan + "velfi" radius.
__device__ __forceinline__ bool ring_info(int& index, float& radius,
                                          const gbkhip_disk& d, float x,
                                          float y, float z, bool thick)
{
    const int n = d.nrnodes;
    const float* nodes = d.rnodes;
    if (rnode_radius(d, 0, x, y, z, thick) < nodes[0]) return false;
    if (rnode_radius(d, n - 1, x, y, z, thick) >= nodes[n - 1]) return false;
    int i;
    for (i = 1; i < n; ++i)
        if (rnode_radius(d, i, x, y, z, thick) < nodes[i]) break;
    if (i >= n) i = n - 1;
    const float d1 = nodes[i - 1] - rnode_radius(d, i - 1, x, y, z, thick);
    const float d2 = nodes[i] - rnode_radius(d, i, x, y, z, thick);
    radius = (nodes[i] * d1 - nodes[i - 1] * d2) / (d1 - d2);
    index = i;
    return true;
}

// Result of evaluating one spatial voxel of a smooth disk.
struct VoxelLine {
    bool on;           // deposits a line
    float rvalue;      // density (before opacity)
    float orvalue;     // density after opacity
    float v, s;        // line centre and dispersion
    float w;           // weight-trait product
};

// gmodel_smdisk_evaluate_spaxel (gmodels.hpp:645-958) up to, but not
// including, the spectral deposit.  `need_theta` is wave-uniform; `ng` holds
// the per-model geometry of non-loose, non-tilted disks.
__device__ __forceinline__ VoxelLine smdisk_voxel(const gbkhip_disk& d,
                                         const gbkhip_grid& g,
                                         const NodeGeom& ng, int x, int y,
                                         int z, bool need_theta, bool want_w)
{
    VoxelLine L;
    L.on = false;
    const bool thin = d.rh.uids == nullptr;
    float xn = g.spat_zero[0] + x * g.spat_step[0];
    float yn = g.spat_zero[1] + y * g.spat_step[1];
    float zn = g.spat_zero[2] + z * g.spat_step[2];
    int rnidx = 1;
    float rn = 0;
    NodeGeom G = ng;
    float vsysi = d.vsys ? d.vsys[0] : 0.0f;

    if (d.loose || d.tilted) {
        if (!ring_info(rnidx, rn, d, xn, yn, zn, !thin)) return L;
        // vsys follows xpos for loose disks (gmodels.hpp:726, reproduced)
        if (d.loose) {
            vsysi = lerp_nw(rn, rnidx, d.rnodes, d.xpos, 0);
            G.xpos = vsysi;
            G.ypos = lerp_nw(rn, rnidx, d.rnodes, d.ypos, 0);
        }
        if (d.tilted) {
            const float pa = lerp_nw(rn, rnidx, d.rnodes, d.posa, 0) * kD2R;
            const float in = lerp_nw(rn, rnidx, d.rnodes, d.incl, 0) * kD2R;
            sincosf(pa, &G.sp, &G.cp);
            sincosf(in, &G.si, &G.ci);
        }
    }

    xn -= G.xpos;
    yn -= G.ypos;
    float xr, yr, zr = zn;
    rot_lh_z(xr, yr, xn, yn, G.sp, G.cp);
    if (thin) yr /= G.ci;
    else rot_lh_x(yr, zr, yr, zn, G.si, G.ci);

    if (!(d.loose || d.tilted)) {
        rn = sqrtf(xr * xr + yr * yr);
        if (!disk_info(rnidx, rn, d.nrnodes, d.rnodes)) return L;
    }

    PPoint P;
    P.rnidx = rnidx;
    P.nrnodes = d.nrnodes;
    P.rnodes = d.rnodes;
    P.x = xr;
    P.y = yr;
    P.r = rn;
    P.si = G.si;
    P.ci = G.ci;
    const float rxy = sqrtf(xr * xr + yr * yr);
    if (rxy > 0.0f) {
        P.ct = xr / rxy;
        P.st = yr / rxy;
        P.theta = need_theta ? atan2f(yr, xr) : 0.0f;
    } else {
        P.theta = atan2f(yr, xr);
        sincosf(P.theta, &P.st, &P.ct);
    }

    float pt[kTraitMax] = {0, 0, 0, 0}, ht[kTraitMax] = {0, 0, 0, 0};
    if (d.sp.uids) {
        float sv = 0;
        p_traits<TK_SP>(d.sp, P, pt);
#pragma unroll
        for (int i = 0; i < kTraitMax; ++i)
            if (i < d.sp.n) sv += pt[i];
        if (!sv) return L;
    }
    if (d.zp.uids) {
        float zv = 0;
        p_traits<TK_ZP>(d.zp, P, pt);
#pragma unroll
        for (int i = 0; i < kTraitMax; ++i)
            if (i < d.zp.n) zv += pt[i];
        zr += zv;
    }
    float rvalue = 0;
    if (d.rp.uids) p_traits<TK_RP>(d.rp, P, pt);
    if (d.rh.uids) rh_traits(d.rh, P, fabsf(zr), ht);
#pragma unroll
    for (int i = 0; i < kTraitMax; ++i)
        if (i < d.rp.n) rvalue += pt[i] * (thin ? 1.0f : ht[i]);
    if (!rvalue) return L;
    if (thin) rvalue /= G.ci;
    else rvalue *= g.spat_step[2];

    float vvalue = 0;
    if (d.vp.uids) p_traits<TK_VP>(d.vp, P, pt);
    if (d.vh.uids) one_traits(d.vh, ht);
#pragma unroll
    for (int i = 0; i < kTraitMax; ++i)
        if (i < d.vp.n) vvalue += pt[i] * (thin ? 1.0f : ht[i]);
    vvalue += vsysi;

    float dvalue = 0;
    if (d.dp.uids) p_traits<TK_DP>(d.dp, P, pt);
    if (d.dh.uids) one_traits(d.dh, ht);
#pragma unroll
    for (int i = 0; i < kTraitMax; ++i)
        if (i < d.dp.n) dvalue += pt[i] * (thin ? 1.0f : ht[i]);
    dvalue = fabsf(dvalue);

    float wvalue = 1.0f;
    if (d.wp.uids && want_w) {   // wp_trait_axis_range returns 1
#pragma unroll
        for (int i = 0; i < kTraitMax; ++i)
            if (i < d.wp.n)
                wvalue *= d.wp.uids[i] == 1 ? 1.0f : __builtin_nanf("");
    }

    float orvalue = rvalue;
    if (d.opacity) {
        const int nx = g.spat_size[0], ny = g.spat_size[1];
        for (int oz = z + 1; oz < g.spat_size[2]; ++oz)
            orvalue -= orvalue *
                d.opacity[x + (size_t)y * nx + (size_t)oz * nx * ny];
    }
    L.on = true;
    L.rvalue = rvalue;
    L.orvalue = orvalue;
    L.v = vvalue;
    L.s = dvalue;
    L.w = wvalue;
    return L;
}

// Channel window of a line (gmodels.hpp:100-105): [k0, k1], possibly empty.
__device__ __forceinline__ void line_window(const gbkhip_grid& g, float v,
                                            float s, int& k0, int& k1)
{
    const float zmin = v - s * kLineWidth;
    const float zmax = v + s * kLineWidth;
    k0 = (int)fmaxf(rintf((zmin - g.spec_zero) / g.spec_step), 0.0f);
    k1 = (int)fminf(rintf((zmax - g.spec_zero) / g.spec_step),
                    (float)(g.spec_size - 1));
}

// flux * N(zero + k*step; v, s): the spectral line value of channel k
// (gauss_1d_pdf, math.hpp:282-287; no channel-width factor, gmodels.hpp:113).
struct LineEval {
    float amp, v, inv2s2;    // flux / (s sqrt(2 pi)), centre, 1 / (2 s^2)
    float zmv, cc, step;     // spectral zero - v, -log2(e) / (2 s^2), step
    __device__ __forceinline__ float at(int k) const
    {
        // amp exp(-(zero + k step - v)^2 / (2 s^2)) = amp 2^((k step + zmv)^2
        // cc), one v_exp_f32 per channel
        const float dz = fmaf((float)k, step, zmv);
        return amp * __builtin_amdgcn_exp2f(dz * dz * cc);
    }
};

// A line from its amplitude, centre and 1 / (2 s^2) on the spectral grid.
__device__ __forceinline__ LineEval line_from(const gbkhip_grid& g,
                                              float amp, float v,
                                              float inv2s2)
{
    LineEval e;
    e.amp = amp;
    e.v = v;
    e.inv2s2 = inv2s2;
    e.zmv = g.spec_zero - v;
    e.cc = -inv2s2 * 1.44269504088896341f;
    e.step = g.spec_step;
    return e;
}

__device__ __forceinline__ LineEval line_eval(const gbkhip_grid& g,
                                              float flux, float v, float s)
{
    return line_from(g, flux / (s * 2.50662827463100050f), v,
                     1.0f / (2 * s * s));
}

}  // namespace gbkhip
