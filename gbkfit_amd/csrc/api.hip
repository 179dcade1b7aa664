// C ABI of libgbkhip (include/gbkhip.h): argument validation, error codes,
// context and cache management.  Kernels live in the other .hip files.
#include "fft_plan.hpp"
#include "launch.hpp"
#include "sepconv.hpp"

#include <cstring>

namespace gbkhip {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

}  // namespace gbkhip

struct gbkhip_ctx {
    gbkhip::Context impl;
};

namespace gbkhip {

Context* ctx_of(gbkhip_ctx* c)
{
    GBK_CHECK(c != nullptr, GBKHIP_ERR_ARG, "null context");
    GBK_HIP(hipSetDevice(c->impl.device));
    return &c->impl;
}

static void check_size3(const int* s, const char* what)
{
    GBK_CHECK(s != nullptr, GBKHIP_ERR_ARG, std::string(what) + " is null");
    GBK_CHECK(s[0] > 0 && s[1] > 0 && s[2] > 0, GBKHIP_ERR_ARG,
              std::string(what) + " must be positive");
}

static void check_ptr(const void* p, const char* what)
{
    GBK_CHECK(p != nullptr, GBKHIP_ERR_ARG, std::string(what) + " is null");
}

static void check_traits(const gbkhip_traits& t, const char* kind)
{
    if (!t.uids) return;
    GBK_CHECK(t.n >= 0 && t.n <= 4, GBKHIP_ERR_ARG,
              std::string("at most 4 traits per kind (") + kind + ")");
    GBK_CHECK(t.cvalues && t.ccounts && t.pvalues && t.pcounts,
              GBKHIP_ERR_ARG, std::string("incomplete trait arrays (") +
              kind + ")");
}

static void check_disk(const gbkhip_disk* d, const gbkhip_grid* g)
{
    check_ptr(d, "disk");
    check_ptr(g, "grid");
    GBK_CHECK(d->nrnodes >= 2, GBKHIP_ERR_ARG, "need at least two rnodes");
    check_ptr(d->rnodes, "rnodes");
    check_ptr(d->xpos, "xpos");
    check_ptr(d->ypos, "ypos");
    check_ptr(d->posa, "posa");
    check_ptr(d->incl, "incl");
    check_traits(d->rp, "rp");
    check_traits(d->rh, "rh");
    check_traits(d->vp, "vp");
    check_traits(d->vh, "vh");
    check_traits(d->dp, "dp");
    check_traits(d->dh, "dh");
    check_traits(d->zp, "zp");
    check_traits(d->sp, "sp");
    check_traits(d->wp, "wp");
    GBK_CHECK(g->spat_size[0] >= 0 && g->spat_size[1] >= 0 &&
              g->spat_size[2] >= 0 && g->spec_size >= 0, GBKHIP_ERR_ARG,
              "negative grid size");
    GBK_CHECK(g->spec_step != 0.0f || g->spec_size <= 1, GBKHIP_ERR_ARG,
              "spectral step must be non-zero");
}

static gbkhip_traits absent_if_empty(gbkhip_traits t)
{
    if (t.n == 0) t.uids = nullptr;
    return t;
}

static gbkhip_disk normalise(const gbkhip_disk& in)
{
    gbkhip_disk d = in;
    d.rp = absent_if_empty(d.rp);
    d.rh = absent_if_empty(d.rh);
    d.vp = absent_if_empty(d.vp);
    d.vh = absent_if_empty(d.vh);
    d.dp = absent_if_empty(d.dp);
    d.dh = absent_if_empty(d.dh);
    d.zp = absent_if_empty(d.zp);
    d.sp = absent_if_empty(d.sp);
    d.wp = absent_if_empty(d.wp);
    return d;
}

}  // namespace gbkhip

using namespace gbkhip;

extern "C" {

const char* gbkhip_last_error(void) { return g_last_error.c_str(); }

int gbkhip_version(void) { return 10000; }

int gbkhip_ctx_create(int device, gbkhip_ctx** out)
{
    return guarded([&] {
        check_ptr(out, "out");
        int n = 0;
        GBK_HIP(hipGetDeviceCount(&n));
        GBK_CHECK(device >= 0 && device < n, GBKHIP_ERR_ARG,
                  "device index out of range");
        GBK_HIP(hipSetDevice(device));
        // the fixed scratch's address on this device (a failure is reported
        // by the entry point that uses it, not here)
        noalloc_scratch_setup(device);
        auto* c = new gbkhip_ctx();
        c->impl.device = device;
        c->impl.stream = nullptr;
        *out = c;
    });
}

int gbkhip_ctx_destroy(gbkhip_ctx* ctx)
{
    return guarded([&] {
        if (!ctx) return;
        (void)hipSetDevice(ctx->impl.device);
        (void)hipStreamSynchronize(ctx->impl.stream);
        delete ctx;
    });
}

int gbkhip_ctx_set_stream(gbkhip_ctx* ctx, void* stream)
{
    return guarded([&] {
        Context* c = ctx_of(ctx);
        c->stream = static_cast<hipStream_t>(stream);
    });
}

int gbkhip_ctx_synchronize(gbkhip_ctx* ctx)
{
    return guarded([&] {
        Context* c = ctx_of(ctx);
        GBK_HIP(hipStreamSynchronize(c->stream));
    });
}

// ---- FFT -------------------------------------------------------------------
int gbkhip_fft_r2c(gbkhip_ctx* ctx, const int size[3], const float* data_r,
                   float* data_c)
{
    return guarded([&] {
        Context* c = ctx_of(ctx);
        check_size3(size, "size");
        check_ptr(data_r, "data_r");
        check_ptr(data_c, "data_c");
        std::lock_guard<std::mutex> lk(c->mu);
        auto p = get_plan(*c, size[0], size[1], size[2]);
        fft_forward(*c, *p, data_r, reinterpret_cast<float2*>(data_c), 1, 0,
                    0);
    });
}

int gbkhip_fft_c2r(gbkhip_ctx* ctx, const int size[3], const float* data_c,
                   float* data_r)
{
    return guarded([&] {
        Context* c = ctx_of(ctx);
        check_size3(size, "size");
        check_ptr(data_r, "data_r");
        check_ptr(data_c, "data_c");
        std::lock_guard<std::mutex> lk(c->mu);
        auto p = get_plan(*c, size[0], size[1], size[2]);
        const size_t nc = (size_t)p->nxh * p->ny * p->nz;
        auto key = std::make_tuple(size[0], size[1], size[2], (uintptr_t)0);
        auto& w = c->work[key];
        if (!w) w = std::make_shared<DevBuf>();
        w->ensure(nc * sizeof(float2));
        fft_inverse(*c, *p, reinterpret_cast<const float2*>(data_c),
                    w->as<float2>(), data_r, 1, 0, 0);
    });
}

int gbkhip_fft_convolve(gbkhip_ctx* ctx, const int size[3], float* data1_r,
                        float* data1_c, const float* data2_c)
{
    return guarded([&] {
        Context* c = ctx_of(ctx);
        check_size3(size, "size");
        check_ptr(data1_r, "data1_r");
        check_ptr(data1_c, "data1_c");
        check_ptr(data2_c, "data2_c");
        std::lock_guard<std::mutex> lk(c->mu);
        auto p = get_plan(*c, size[0], size[1], size[2]);
        fft_convolve_with(*c, *p, data1_r, reinterpret_cast<float2*>(data1_c),
                          reinterpret_cast<const float2*>(data2_c), 1, 0, 0,
                          0);
    });
}

int gbkhip_fft_convolve_cached(gbkhip_ctx* ctx, const int size[3],
                               float* data1_r, const float* data2_r,
                               uint64_t data2_generation)
{
    return guarded([&] {
        Context* c = ctx_of(ctx);
        check_size3(size, "size");
        check_ptr(data1_r, "data1_r");
        check_ptr(data2_r, "data2_r");
        std::lock_guard<std::mutex> lk(c->mu);
        const auto skey = std::make_tuple(size[0], size[1], size[2],
                                          (uintptr_t)data2_r,
                                          data2_generation);
        // a kernel that factors into short 1-D kernels (the PSF x LSF cube
        // DCube builds): three direct circular passes instead of the FFTs
        auto& sc = c->sepconvs[skey];
        if (!sc) {
            sc = std::make_shared<SepConv>();
            sepconv_plan(*c, size, data2_r, *sc);
        }
        if (sc->ok) {
            sepconv_run(*c, size, *sc, data1_r);
            return;
        }
        auto p = get_plan(*c, size[0], size[1], size[2]);
        const size_t nc = (size_t)p->nxh * p->ny * p->nz;
        auto& spec = c->spectra[skey];
        if (!spec) {
            spec = std::make_shared<DevBuf>(nc * sizeof(float2));
            fft_forward(*c, *p, data2_r, spec->as<float2>(), 1, 0, 0);
        }
        const auto wkey = std::make_tuple(size[0], size[1], size[2],
                                          (uintptr_t)data1_r);
        auto& w = c->work[wkey];
        if (!w) w = std::make_shared<DevBuf>();
        w->ensure(nc * sizeof(float2));
        fft_convolve_with(*c, *p, data1_r, w->as<float2>(),
                          spec->as<float2>(), 1, 0, 0, 0);
    });
}

int gbkhip_fft_clear_cache(gbkhip_ctx* ctx)
{
    return guarded([&] {
        Context* c = ctx_of(ctx);
        std::lock_guard<std::mutex> lk(c->mu);
        GBK_HIP(hipStreamSynchronize(c->stream));
        c->spectra.clear();
        c->sepconvs.clear();
        c->work.clear();
        c->plans.clear();
    });
}

// ---- DModel -----------------------------------------------------------------
int gbkhip_dcube_downscale(gbkhip_ctx* ctx, const int scale[3],
                           const int offset[3], const int src_size[3],
                           const int dst_size[3], const float* src, float* dst)
{
    return guarded([&] {
        Context* c = ctx_of(ctx);
        check_size3(scale, "scale");
        check_ptr(offset, "offset");
        check_size3(src_size, "src_size");
        check_size3(dst_size, "dst_size");
        check_ptr(src, "src");
        check_ptr(dst, "dst");
        for (int i = 0; i < 3; ++i)
            GBK_CHECK(offset[i] >= 0 &&
                      offset[i] + dst_size[i] * scale[i] <= src_size[i],
                      GBKHIP_ERR_ARG,
                      "downscale window exceeds the source cube");
        launch_downscale(*c, scale, offset, src_size, dst_size, src, dst, 1,
                         0, 0);
    });
}

int gbkhip_dcube_mask(gbkhip_ctx* ctx, float cutoff, int apply,
                      const int size[3], float* dcube_d, float* dcube_m,
                      float* dcube_w)
{
    (void)dcube_w;  // weights are not touched (dmodels.hpp:55-57)
    return guarded([&] {
        Context* c = ctx_of(ctx);
        check_size3(size, "size");
        check_ptr(dcube_d, "dcube_d");
        launch_mask(*c, cutoff, apply,
                    (long long)size[0] * size[1] * size[2], dcube_d, dcube_m);
    });
}

int gbkhip_mmaps_moments(gbkhip_ctx* ctx, const int size[3],
                         const float step[3], const float zero[3],
                         const float* dcube_d, const float* dcube_w,
                         float cutoff, int norders, const int* orders,
                         float* mmaps_d, float* mmaps_m, float* mmaps_w)
{
    return guarded([&] {
        Context* c = ctx_of(ctx);
        check_size3(size, "size");
        check_ptr(step, "step");
        check_ptr(zero, "zero");
        check_ptr(dcube_d, "dcube_d");
        check_ptr(orders, "orders");
        check_ptr(mmaps_d, "mmaps_d");
        check_ptr(mmaps_m, "mmaps_m");
        GBK_CHECK(norders >= 1 && norders <= 8, GBKHIP_ERR_ARG,
                  "norders must be in [1, 8]");
        GBK_CHECK(!dcube_w || mmaps_w, GBKHIP_ERR_ARG,
                  "mmaps_w is required when dcube_w is given");
        launch_moments(*c, size, step, zero, dcube_d, dcube_w, cutoff,
                       norders, orders, mmaps_d, mmaps_m,
                       dcube_w ? mmaps_w : nullptr, 1, 0, 0);
    });
}

int gbkhip_mmaps_moments_batched(gbkhip_ctx* ctx, const int size[3],
                                 const float step[3], const float zero[3],
                                 const float* dcube_d, float cutoff,
                                 int norders, const int* orders,
                                 float* mmaps_d, float* mmaps_m, int nbatch)
{
    return guarded([&] {
        Context* c = ctx_of(ctx);
        check_size3(size, "size");
        check_ptr(step, "step");
        check_ptr(zero, "zero");
        check_ptr(dcube_d, "dcube_d");
        check_ptr(orders, "orders");
        check_ptr(mmaps_d, "mmaps_d");
        check_ptr(mmaps_m, "mmaps_m");
        GBK_CHECK(norders >= 1 && norders <= 8, GBKHIP_ERR_ARG,
                  "norders must be in [1, 8]");
        GBK_CHECK(nbatch >= 0 && nbatch <= 65535, GBKHIP_ERR_ARG,
                  "nbatch must be within [0, 65535] (one grid row per cube)");
        if (nbatch == 0) return;
        const long long plane = (long long)size[0] * size[1];
        launch_moments(*c, size, step, zero, dcube_d, nullptr, cutoff,
                       norders, orders, mmaps_d, mmaps_m, nullptr, nbatch,
                       plane * size[2], plane);
    });
}

// ---- GModel -----------------------------------------------------------------
int gbkhip_smdisk_evaluate(gbkhip_ctx* ctx, const gbkhip_disk* disk,
                           const gbkhip_grid* grid, const gbkhip_outputs* out)
{
    return guarded([&] {
        Context* c = ctx_of(ctx);
        check_disk(disk, grid);
        check_ptr(out, "outputs");
        launch_smdisk(*c, normalise(*disk), *grid, *out);
    });
}

int gbkhip_mcdisk_evaluate(gbkhip_ctx* ctx, float cflux, int nclouds,
                           const int* ncloudscsum, int ncloudscsum_len,
                           const unsigned char* hasordint,
                           const gbkhip_disk* disk, const gbkhip_grid* grid,
                           const gbkhip_outputs* out, uint64_t seed)
{
    return guarded([&] {
        Context* c = ctx_of(ctx);
        check_disk(disk, grid);
        check_ptr(out, "outputs");
        GBK_CHECK(nclouds >= 0, GBKHIP_ERR_ARG, "negative cloud count");
        if (nclouds == 0) return;
        check_ptr(ncloudscsum, "ncloudscsum");
        check_ptr(hasordint, "hasordint");
        GBK_CHECK(ncloudscsum_len > 0, GBKHIP_ERR_ARG,
                  "empty cloud cumulative sum");
        // cloud i of trait t samples height trait t (gmodels.hpp:406-416)
        GBK_CHECK(disk->rp.n > 0 && disk->rh.n >= disk->rp.n, GBKHIP_ERR_ARG,
                  "mcdisk needs one density height trait per polar trait");
        std::lock_guard<std::mutex> lk(c->mu);
        launch_mcdisk(*c, cflux, nclouds, ncloudscsum, ncloudscsum_len,
                      hasordint, normalise(*disk), *grid, *out, seed);
    });
}

int gbkhip_mcdisk_evaluate_noalloc(gbkhip_ctx* ctx, float cflux,
                                   int nclouds, const int* ncloudscsum,
                                   int ncloudscsum_len,
                                   const unsigned char* hasordint,
                                   const gbkhip_disk* disk,
                                   const gbkhip_grid* grid,
                                   const gbkhip_outputs* out, uint64_t seed)
{
    return guarded([&] {
        Context* c = ctx_of(ctx);
        check_disk(disk, grid);
        check_ptr(out, "outputs");
        GBK_CHECK(nclouds >= 0, GBKHIP_ERR_ARG, "negative cloud count");
        if (nclouds == 0) return;
        check_ptr(ncloudscsum, "ncloudscsum");
        check_ptr(hasordint, "hasordint");
        GBK_CHECK(ncloudscsum_len > 0, GBKHIP_ERR_ARG,
                  "empty cloud cumulative sum");
        GBK_CHECK(disk->rp.n > 0 && disk->rh.n >= disk->rp.n, GBKHIP_ERR_ARG,
                  "mcdisk needs one density height trait per polar trait");
        char* fixed = noalloc_scratch(c->device);
        std::lock_guard<std::mutex> lk(c->mu);
        launch_mcdisk(*c, cflux, nclouds, ncloudscsum, ncloudscsum_len,
                      hasordint, normalise(*disk), *grid, *out, seed, fixed,
                      kNoAllocScratchBytes);
    });
}

int gbkhip_wcube_evaluate(gbkhip_ctx* ctx, const int spat_size[3],
                          int spec_size, const float* spat_data,
                          float* spec_data)
{
    return guarded([&] {
        Context* c = ctx_of(ctx);
        check_size3(spat_size, "spat_size");
        check_ptr(spat_data, "spat_data");
        check_ptr(spec_data, "spec_data");
        GBK_CHECK(spec_size >= 0, GBKHIP_ERR_ARG, "negative spec_size");
        launch_wcube(*c, spat_size, spec_size, spat_data, spec_data);
    });
}

int gbkhip_convolve(gbkhip_ctx* ctx, const int size1[3], const int size2[3],
                    const float* data1, const float* data2, float* result)
{
    return guarded([&] {
        Context* c = ctx_of(ctx);
        check_size3(size1, "size1");
        check_size3(size2, "size2");
        check_ptr(data1, "data1");
        check_ptr(data2, "data2");
        check_ptr(result, "result");
        launch_convolve(*c, size1, size2, data1, data2, result);
    });
}

// ---- Objective -----------------------------------------------------------------
int gbkhip_residual(gbkhip_ctx* ctx, const float* obs_d, const float* obs_e,
                    const float* obs_m, const float* mdl_d, const float* mdl_w,
                    const float* mdl_m, long long size, float weight,
                    float* res)
{
    return guarded([&] {
        Context* c = ctx_of(ctx);
        GBK_CHECK(size >= 0, GBKHIP_ERR_ARG, "negative size");
        if (size == 0) return;
        check_ptr(obs_d, "obs_d");
        check_ptr(mdl_d, "mdl_d");
        check_ptr(res, "res");
        launch_residual(*c, obs_d, obs_e, obs_m, mdl_d, mdl_w, mdl_m, size,
                        weight, res);
    });
}

int gbkhip_residual_sum(gbkhip_ctx* ctx, const float* residual, long long size,
                        int squared, float* sum_out)
{
    return guarded([&] {
        Context* c = ctx_of(ctx);
        GBK_CHECK(size >= 0, GBKHIP_ERR_ARG, "negative size");
        GBK_CHECK(size == 0 || residual, GBKHIP_ERR_ARG, "residual is null");
        check_ptr(sum_out, "sum_out");
        std::lock_guard<std::mutex> lk(c->mu);
        launch_sum(*c, residual, size, squared, sum_out, nullptr);
    });
}

int gbkhip_count_pixels(gbkhip_ctx* ctx, const float* data1,
                        const float* data2, long long size, float epsilon,
                        int* counts)
{
    return guarded([&] {
        Context* c = ctx_of(ctx);
        GBK_CHECK(size >= 0, GBKHIP_ERR_ARG, "negative size");
        check_ptr(counts, "counts");
        if (size == 0) return;
        check_ptr(data1, "data1");
        check_ptr(data2, "data2");
        launch_count(*c, data1, data2, size, epsilon, counts);
    });
}

}  // extern "C"
