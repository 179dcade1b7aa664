// Test infrastructure only.  A thin harness that instantiates the REFERENCE
// header-only core (src/gbkfit/driver/native/gbkfit/include/gbkfit/...) from
// the read-only reference tree, so that its own per-voxel / per-cloud /
// per-spaxel templates can produce golden vectors and a CPU baseline.
//
// Nothing of the reference is copied here: the headers are included from
// /root/reference through -I (see oracle/Makefile).  The loops below play the
// role of the reference's OpenMP host driver loops
// (gbkfit_host/include/gbkfit/host/kernels.hpp:55-135,183-306,308-423,484-514),
// which cannot be included because they pull in fftw3.h (absent here).
//
// Built into oracle/_ref/libgbkref.so (git-ignored).  Only tests/, the
// golden-vector generator and bench.py's CPU-baseline leg load it.

#include <cassert>
#include <cmath>
#include <cstdint>
#include <algorithm>
#include <tuple>
#include <type_traits>
#include <omp.h>

#include "gbkfit/host/random.hpp"           // RNG<T> = std::mt19937 + uniform_real
#include "gbkfit/gmodel/gmodels.hpp"
#include "gbkfit/dmodel/dmodels.hpp"
#include "gbkfit/objective/objective.hpp"

namespace {

template<typename T> void plain_add(T* a, T v) { a[0] += v; }
template<typename T> void plain_set(T* a, T v) { a[0] = v; }
template<typename T> void omp_add(T* a, T v) {
    #pragma omp atomic update
    a[0] += v;
}
template<typename T> void omp_set(T* a, T v) {
    #pragma omp atomic write
    a[0] = v;
}

}  // namespace

template<typename T>
struct RefTraits {
    int n;
    const int* uids;
    const T* cvalues;
    const int* ccounts;
    const T* pvalues;
    const int* pcounts;
};

template<typename T>
struct RefDisk {
    int loose, tilted, nrnodes;
    const T* rnodes;
    const T* vsys;
    const T* xpos;
    const T* ypos;
    const T* posa;
    const T* incl;
    RefTraits<T> rp, rh, vp, vh, dp, dh, zp, sp, wp;
    const T* opacity;
};

template<typename T>
struct RefGrid {
    int spat_size[3];
    T spat_step[3];
    T spat_zero[3];
    int spec_size;
    T spec_step;
    T spec_zero;
};

template<typename T>
struct RefOut {
    T* image; T* scube;
    T* wdata; T* wdata_cmp;
    T* rdata; T* rdata_cmp;
    T* ordata; T* ordata_cmp;
    T* vdata_cmp; T* ddata_cmp;
};

template<auto ADD, typename T>
static void smdisk_voxel(int x, int y, int z,
                         const RefDisk<T>& d, const RefGrid<T>& g,
                         const RefOut<T>& o)
{
    gbkfit::gmodel_smdisk_evaluate_spaxel<ADD, T>(
        x, y, z, d.loose != 0, d.tilted != 0, d.nrnodes, d.rnodes,
        d.vsys, d.xpos, d.ypos, d.posa, d.incl,
        d.rp.n, d.rp.uids, d.rp.cvalues, d.rp.ccounts, d.rp.pvalues, d.rp.pcounts,
        d.rh.uids, d.rh.cvalues, d.rh.ccounts, d.rh.pvalues, d.rh.pcounts,
        d.vp.n, d.vp.uids, d.vp.cvalues, d.vp.ccounts, d.vp.pvalues, d.vp.pcounts,
        d.vh.uids, d.vh.cvalues, d.vh.ccounts, d.vh.pvalues, d.vh.pcounts,
        d.dp.n, d.dp.uids, d.dp.cvalues, d.dp.ccounts, d.dp.pvalues, d.dp.pcounts,
        d.dh.uids, d.dh.cvalues, d.dh.ccounts, d.dh.pvalues, d.dh.pcounts,
        d.zp.n, d.zp.uids, d.zp.cvalues, d.zp.ccounts, d.zp.pvalues, d.zp.pcounts,
        d.sp.n, d.sp.uids, d.sp.cvalues, d.sp.ccounts, d.sp.pvalues, d.sp.pcounts,
        d.wp.n, d.wp.uids, d.wp.cvalues, d.wp.ccounts, d.wp.pvalues, d.wp.pcounts,
        d.opacity,
        g.spat_size[0], g.spat_size[1], g.spat_size[2],
        g.spat_step[0], g.spat_step[1], g.spat_step[2],
        g.spat_zero[0], g.spat_zero[1], g.spat_zero[2],
        g.spec_size, g.spec_step, g.spec_zero,
        o.image, o.scube, o.wdata, o.wdata_cmp, o.rdata, o.rdata_cmp,
        o.ordata, o.ordata_cmp, o.vdata_cmp, o.ddata_cmp);
}

template<typename T>
static void smdisk_serial(const RefDisk<T>* d, const RefGrid<T>* g,
                          const RefOut<T>* o)
{
    for (int y = 0; y < g->spat_size[1]; ++y)
    for (int x = 0; x < g->spat_size[0]; ++x)
    for (int z = 0; z < g->spat_size[2]; ++z)
        smdisk_voxel<plain_add<T>, T>(x, y, z, *d, *g, *o);
}

template<typename T>
static void smdisk_omp(const RefDisk<T>* d, const RefGrid<T>* g,
                       const RefOut<T>* o)
{
    const int ny = g->spat_size[1], nx = g->spat_size[0], nz = g->spat_size[2];
    #pragma omp parallel for collapse(3)
    for (int y = 0; y < ny; ++y)
    for (int x = 0; x < nx; ++x)
    for (int z = 0; z < nz; ++z)
        smdisk_voxel<omp_add<T>, T>(x, y, z, *d, *g, *o);
}

template<typename T>
static void mcdisk_serial(T cflux, int nclouds, const int* ncloudscsum,
                          int ncloudscsum_len, const unsigned char* hasordint,
                          const RefDisk<T>* d, const RefGrid<T>* g,
                          const RefOut<T>* o, int seed)
{
    gbkfit::RNG<T> rng(0, 1, seed);
    static_assert(sizeof(bool) == 1, "bool must be one byte");
    const bool* hoi = reinterpret_cast<const bool*>(hasordint);
    for (int ci = 0; ci < nclouds; ++ci) {
        gbkfit::gmodel_mcdisk_evaluate_cloud<plain_set<T>, plain_add<T>, T>(
            rng, ci, cflux, nclouds, ncloudscsum, ncloudscsum_len, hoi,
            d->loose != 0, d->tilted != 0, d->nrnodes, d->rnodes,
            d->vsys, d->xpos, d->ypos, d->posa, d->incl,
            d->rp.n, d->rp.uids, d->rp.cvalues, d->rp.ccounts, d->rp.pvalues, d->rp.pcounts,
            d->rh.uids, d->rh.cvalues, d->rh.ccounts, d->rh.pvalues, d->rh.pcounts,
            d->vp.n, d->vp.uids, d->vp.cvalues, d->vp.ccounts, d->vp.pvalues, d->vp.pcounts,
            d->vh.uids, d->vh.cvalues, d->vh.ccounts, d->vh.pvalues, d->vh.pcounts,
            d->dp.n, d->dp.uids, d->dp.cvalues, d->dp.ccounts, d->dp.pvalues, d->dp.pcounts,
            d->dh.uids, d->dh.cvalues, d->dh.ccounts, d->dh.pvalues, d->dh.pcounts,
            d->zp.n, d->zp.uids, d->zp.cvalues, d->zp.ccounts, d->zp.pvalues, d->zp.pcounts,
            d->sp.n, d->sp.uids, d->sp.cvalues, d->sp.ccounts, d->sp.pvalues, d->sp.pcounts,
            d->wp.n, d->wp.uids, d->wp.cvalues, d->wp.ccounts, d->wp.pvalues, d->wp.pcounts,
            d->opacity,
            g->spat_size[0], g->spat_size[1], g->spat_size[2],
            g->spat_step[0], g->spat_step[1], g->spat_step[2],
            g->spat_zero[0], g->spat_zero[1], g->spat_zero[2],
            g->spec_size, g->spec_step, g->spec_zero,
            o->image, o->scube, o->wdata, o->wdata_cmp, o->rdata, o->rdata_cmp,
            o->ordata, o->ordata_cmp, o->vdata_cmp, o->ddata_cmp);
    }
}

template<typename T>
static void downscale(const int* scale, const int* offset, const int* src_size,
                      const int* dst_size, const T* src, T* dst, int parallel)
{
    #pragma omp parallel for collapse(3) if(parallel)
    for (int z = 0; z < dst_size[2]; ++z)
    for (int y = 0; y < dst_size[1]; ++y)
    for (int x = 0; x < dst_size[0]; ++x)
        gbkfit::dmodel_dcube_downscale<T>(
            x, y, z, scale[0], scale[1], scale[2],
            offset[0], offset[1], offset[2],
            src_size[0], src_size[1], src_size[2],
            dst_size[0], dst_size[1], dst_size[2], src, dst);
}

template<typename T>
static void mask(T cutoff, int apply, const int* size, T* d, T* m, T* w,
                 int parallel)
{
    #pragma omp parallel for collapse(3) if(parallel)
    for (int z = 0; z < size[2]; ++z)
    for (int y = 0; y < size[1]; ++y)
    for (int x = 0; x < size[0]; ++x)
        gbkfit::dmodel_dcube_mask<T>(x, y, z, cutoff, apply != 0,
                                     size[0], size[1], size[2], d, m, w);
}

// Arguments in the core's own order (mmaps_d, mmaps_m, mmaps_w):
// core/dmodel/dmodels.hpp:77-84.
template<typename T>
static void moments(const int* size, const T* step, const T* zero,
                    const T* dcube_d, const T* dcube_w, T cutoff, int norders,
                    const int* orders, T* mmaps_d, T* mmaps_m, T* mmaps_w,
                    int parallel)
{
    #pragma omp parallel for collapse(2) if(parallel)
    for (int y = 0; y < size[1]; ++y)
    for (int x = 0; x < size[0]; ++x)
        gbkfit::dmodel_mmaps_moments<T>(
            x, y, size[0], size[1], size[2], step[0], step[1], step[2],
            zero[0], zero[1], zero[2], dcube_d, dcube_w, cutoff, norders,
            orders, mmaps_d, mmaps_m, mmaps_w);
}

template<typename T>
static void residual(const T* obs_d, const T* obs_e, const T* obs_m,
                     const T* mdl_d, const T* mdl_w, const T* mdl_m,
                     int size, T weight, T* res, int parallel)
{
    #pragma omp parallel for if(parallel)
    for (int i = 0; i < size; ++i)
        gbkfit::objective_residual<T>(i, obs_d, obs_e, obs_m, mdl_d, mdl_w,
                                      mdl_m, size, weight, res);
}

template<typename T>
static T residual_sum(const T* r, int size, int squared, int parallel)
{
    // Same reduction the host driver performs (kernels.hpp:502-514).
    T sum = 0;
    #pragma omp parallel for reduction(+:sum) if(parallel)
    for (int i = 0; i < size; ++i)
        sum += squared ? r[i] * r[i] : std::abs(r[i]);
    return sum;
}

// Spectral weight cube from a spatial one, per spaxel: the reference's
// gmodel_wcube_pixel (gmodels.hpp:52-82) over every (x, y), as its host loop
// does (gbkfit_host/include/gbkfit/host/kernels.hpp, gmodel_wcube_evaluate).
template<typename T>
static void wcube(const int* spat, int spec_z, const T* sw, T* out)
{
    for (int y = 0; y < spat[1]; ++y)
    for (int x = 0; x < spat[0]; ++x)
        gbkfit::gmodel_wcube_pixel<T>(x, y, spat[0], spat[1], spat[2], spec_z,
                                      sw, out);
}

#define REF_EXPORTS(T, S)                                                      \
extern "C" {                                                                   \
void ref_smdisk_##S(const RefDisk<T>* d, const RefGrid<T>* g,                  \
                    const RefOut<T>* o) { smdisk_serial<T>(d, g, o); }         \
void ref_smdisk_omp_##S(const RefDisk<T>* d, const RefGrid<T>* g,              \
                        const RefOut<T>* o) { smdisk_omp<T>(d, g, o); }        \
void ref_mcdisk_##S(T cflux, int nclouds, const int* csum, int csum_len,       \
                    const unsigned char* hoi, const RefDisk<T>* d,             \
                    const RefGrid<T>* g, const RefOut<T>* o, int seed)         \
    { mcdisk_serial<T>(cflux, nclouds, csum, csum_len, hoi, d, g, o, seed); }  \
void ref_downscale_##S(const int* sc, const int* off, const int* ss,           \
                       const int* ds, const T* src, T* dst, int par)           \
    { downscale<T>(sc, off, ss, ds, src, dst, par); }                          \
void ref_mask_##S(T cutoff, int apply, const int* size, T* d, T* m, T* w,      \
                  int par) { mask<T>(cutoff, apply, size, d, m, w, par); }     \
void ref_moments_##S(const int* size, const T* step, const T* zero,            \
                     const T* d, const T* w, T cutoff, int norders,            \
                     const int* orders, T* md, T* mm, T* mw, int par)          \
    { moments<T>(size, step, zero, d, w, cutoff, norders, orders,              \
                 md, mm, mw, par); }                                           \
void ref_residual_##S(const T* od, const T* oe, const T* om, const T* md,      \
                      const T* mw, const T* mm, int size, T weight, T* res,    \
                      int par)                                                 \
    { residual<T>(od, oe, om, md, mw, mm, size, weight, res, par); }           \
T ref_residual_sum_##S(const T* r, int size, int squared, int par)             \
    { return residual_sum<T>(r, size, squared, par); }                         \
void ref_wcube_##S(const int* spat, int spec_z, const T* sw, T* out)           \
    { wcube<T>(spat, spec_z, sw, out); }                                       \
}

REF_EXPORTS(float, f32)
REF_EXPORTS(double, f64)

extern "C" int ref_num_threads(void) { return omp_get_max_threads(); }
