/*
 * gbkhip.h -- C ABI of libgbkhip, the MI355X (gfx950) driver for gbkfit's
 * model-evaluation hot path.
 *
 * Every function returns GBKHIP_OK (0) or an error code; the message of the
 * last error of the calling thread is available from gbkhip_last_error().
 * Array arguments are DEVICE pointers unless stated otherwise; absent optional
 * arrays are NULL (the reference passes 0, drivers/host/_backend.py:18-22).
 * Sizes are (x, y, z) -- the reversed C-order shape, as gbkfit's native
 * marshalling passes them (src/gbkfit/driver/drivers/_detail/native.py:70-131)
 * -- and cubes are C-order (z, y, x) with x fastest
 * (core/utilities/indexutils.hpp:96-101).
 *
 * Work is enqueued on the context's stream (gbkhip_ctx_set_stream); calls
 * are asynchronous with respect to the host unless stated otherwise.
 *
 * Reference paths are relative to src/gbkfit/driver/native/ (core/ =
 * gbkfit/include/gbkfit/, host/ = gbkfit_host/).
 */
#ifndef GBKHIP_H
#define GBKHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    GBKHIP_OK = 0,
    GBKHIP_ERR_ARG = 1,          /* invalid argument (shape, null, range) */
    GBKHIP_ERR_HIP = 2,          /* HIP runtime failure */
    GBKHIP_ERR_ALLOC = 3,        /* allocation failure */
    GBKHIP_ERR_UNSUPPORTED = 4,  /* valid but unsupported request */
    GBKHIP_ERR_INTERNAL = 5
};

typedef struct gbkhip_ctx gbkhip_ctx;

/* ---- library / context ------------------------------------------------- */
const char* gbkhip_last_error(void);
int gbkhip_version(void);
/* One context per device; owns FFT plans, cached kernel spectra and scratch
 * (replaces the FFT plan/data caches of host/include/gbkfit/host/fft.hpp:
 * 20-69, which are per FFT-backend object in the reference). */
int gbkhip_ctx_create(int device, gbkhip_ctx** out);
int gbkhip_ctx_destroy(gbkhip_ctx* ctx);
/* `stream` is a hipStream_t (NULL = the null stream). */
int gbkhip_ctx_set_stream(gbkhip_ctx* ctx, void* stream);
int gbkhip_ctx_synchronize(gbkhip_ctx* ctx);

/* ---- FFT backend: FFTf32 (host/src/bindings.cpp:35-40,
 *      host/include/gbkfit/host/fft.hpp:71-125) ---------------------------- */
/* Unnormalised forward r2c: (nz, ny, nx) real -> (nz, ny, nx/2+1) complex
 * (interleaved float pairs). */
int gbkhip_fft_r2c(gbkhip_ctx* ctx, const int size[3], const float* data_r,
                   float* data_c);
/* Unnormalised inverse c2r; data_c is not modified. */
int gbkhip_fft_c2r(gbkhip_ctx* ctx, const int size[3], const float* data_c,
                   float* data_r);
/* data1_r <- irfft(rfft(data1_r) * data2_c) / N, using data1_c as the
 * spectrum buffer (fft.hpp:86-96,199-207). */
int gbkhip_fft_convolve(gbkhip_ctx* ctx, const int size[3], float* data1_r,
                        float* data1_c, const float* data2_c);
/* As above with the spectra of data2_r and the work spectrum of data1_r
 * cached in the context (fft.hpp:98-125).  The reference keys the cache by
 * (size, pointer) only; here `data2_generation` is part of the key, so a
 * caller that rewrites data2_r in place bumps it to invalidate. */
int gbkhip_fft_convolve_cached(gbkhip_ctx* ctx, const int size[3],
                               float* data1_r, const float* data2_r,
                               uint64_t data2_generation);
/* Drop every cached spectrum and plan. */
int gbkhip_fft_clear_cache(gbkhip_ctx* ctx);

/* ---- DModel backend: DModelf32 (host/src/dmodels.cpp:7-59) -------------- */
/* Block mean over scale[0]*scale[1]*scale[2] source voxels starting at
 * offset (core/dmodel/dmodels.hpp:7-46). */
int gbkhip_dcube_downscale(gbkhip_ctx* ctx, const int scale[3],
                           const int offset[3], const int src_size[3],
                           const int dst_size[3], const float* src,
                           float* dst);
/* m = |d| > cutoff; d = NaN where masked if apply (dmodels.hpp:48-74).
 * Covers the whole x*y*z cube (the CUDA launcher's x*y bug is not kept). */
int gbkhip_dcube_mask(gbkhip_ctx* ctx, float cutoff, int apply,
                      const int size[3], float* dcube_d, float* dcube_m,
                      float* dcube_w);
/* Moment maps (dmodels.hpp:76-228).  `orders` is a device array of norders
 * sorted orders in [0, 7]; mmaps_d (norders, ny, nx), mmaps_m (ny, nx),
 * mmaps_w (norders, ny, nx; only plane 0 written, NULL allowed when dcube_w
 * is NULL).  Pointer order (d, m, w) is the semantic one: the reference's
 * Python passes (d, w, m) into a (d, m, w) binding that swaps them back
 * (_detail/native.py:122-131, host/include/gbkfit/host/kernels.hpp:124-131). */
int gbkhip_mmaps_moments(gbkhip_ctx* ctx, const int size[3],
                         const float step[3], const float zero[3],
                         const float* dcube_d, const float* dcube_w,
                         float cutoff, int norders, const int* orders,
                         float* mmaps_d, float* mmaps_m, float* mmaps_w);

/* Moment maps of nbatch cubes in one launch: no counterpart in the
 * reference (which computes one model's maps per call); for batched callers
 * such as a serving loop.  dcube_d (nbatch, nz, ny, nx) contiguous; mmaps_d
 * (nbatch, norders, ny, nx); mmaps_m (nbatch, ny, nx); no weight cube.  Each
 * cube's maps equal gbkhip_mmaps_moments' of that cube. */
int gbkhip_mmaps_moments_batched(gbkhip_ctx* ctx, const int size[3],
                                 const float step[3], const float zero[3],
                                 const float* dcube_d, float cutoff,
                                 int norders, const int* orders,
                                 float* mmaps_d, float* mmaps_m, int nbatch);

/* ---- GModel backend: GModelf32 (host/src/gmodels.cpp, 83/88 positional
 *      arguments, host/include/gbkfit/host/gmodels.hpp:25-132) ------------ */
typedef struct {
    int n;                   /* traits of this kind (0 when absent) */
    const int* uids;         /* device; NULL when absent */
    const float* cvalues;    /* device, packed constants */
    const int* ccounts;      /* device, constants per trait */
    const float* pvalues;    /* device, packed parameters */
    const int* pcounts;      /* device, parameters per trait */
} gbkhip_traits;

typedef struct {
    int loose, tilted;
    int nrnodes;             /* number of sub-ring nodes */
    const float* rnodes;     /* device (nrnodes) */
    const float* vsys;       /* device, 1 or nrnodes values, may be NULL */
    const float* xpos;
    const float* ypos;
    const float* posa;       /* degrees */
    const float* incl;       /* degrees */
    gbkhip_traits rp, rh, vp, vh, dp, dh, zp, sp, wp;
    const float* opacity;    /* device spatial cube or NULL */
} gbkhip_disk;

typedef struct {
    int spat_size[3];
    float spat_step[3];
    float spat_zero[3];
    int spec_size;
    float spec_step;
    float spec_zero;
} gbkhip_grid;

typedef struct {
    float* image;            /* (ny, nx) */
    float* scube;            /* (spec_size, ny, nx) */
    float* wdata;            /* spatial cubes (nz, ny, nx) */
    float* wdata_cmp;
    float* rdata;
    float* rdata_cmp;
    float* ordata;
    float* ordata_cmp;
    float* vdata_cmp;
    float* ddata_cmp;
} gbkhip_outputs;

/* Smooth-disk evaluation over every spatial voxel; outputs are accumulated
 * (+=) like the reference (core/gmodel/gmodels.hpp:645-958).  `disk`, `grid`
 * and `out` are host structs holding device pointers. */
int gbkhip_smdisk_evaluate(gbkhip_ctx* ctx, const gbkhip_disk* disk,
                           const gbkhip_grid* grid, const gbkhip_outputs* out);
/* Monte Carlo cloud disk (gmodels.hpp:325-643).  Clouds draw from a
 * Philox4x32-10 stream keyed by (seed, cloud id), so results do not depend
 * on launch shape; the spectral cube is accumulated in int64 fixed point and
 * converted once, so reruns are bit-identical.  ncloudscsum and hasordint
 * are device arrays (ncloudscsum_len entries / rp.n entries); ncloudscsum
 * holds cumulative counts, so it is non-decreasing (a cloud's ring is found
 * by binary search). */
int gbkhip_mcdisk_evaluate(gbkhip_ctx* ctx, float cflux, int nclouds,
                           const int* ncloudscsum, int ncloudscsum_len,
                           const unsigned char* hasordint,
                           const gbkhip_disk* disk, const gbkhip_grid* grid,
                           const gbkhip_outputs* out, uint64_t seed);
/* gbkhip_mcdisk_evaluate that allocates nothing.  It has no counterpart in
 * the reference; it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results, bit for bit; the binned deposit
 * buffers live in a fixed 64 MiB device array that each GPU holds until the
 * process ends (about 1.4 M clouds on a 256x256x128 cube).  Requests that
 * need more scratch, or outputs other than the spectral cube, or more than
 * 8192 channels, are rejected (GBKHIP_ERR_ARG / GBKHIP_ERR_UNSUPPORTED)
 * before anything is enqueued.  Calls that overlap on one GPU (from several
 * contexts, streams or threads) share the array and race. */
int gbkhip_mcdisk_evaluate_noalloc(gbkhip_ctx* ctx, float cflux, int nclouds,
                                   const int* ncloudscsum,
                                   int ncloudscsum_len,
                                   const unsigned char* hasordint,
                                   const gbkhip_disk* disk,
                                   const gbkhip_grid* grid,
                                   const gbkhip_outputs* out, uint64_t seed);
/* Spectral weight cube from a spatial weight cube (gmodels.hpp:52-82). */
int gbkhip_wcube_evaluate(gbkhip_ctx* ctx, const int spat_size[3],
                          int spec_size, const float* spat_data,
                          float* spec_data);
/* Direct zero-padded convolution (gmodels.hpp:9-50). */
int gbkhip_convolve(gbkhip_ctx* ctx, const int size1[3], const int size2[3],
                    const float* data1, const float* data2, float* result);

/* ---- Objective backend: Objectivef32 (host/src/objective.cpp) ----------- */
/* res = (mdl - obs) [/ obs_e] [* mdl_w] [* obs_m] [* mdl_m] * weight
 * (core/objective/objective.hpp:7-33). */
int gbkhip_residual(gbkhip_ctx* ctx, const float* obs_d, const float* obs_e,
                    const float* obs_m, const float* mdl_d,
                    const float* mdl_w, const float* mdl_m, long long size,
                    float weight, float* res);
/* sum r^2 (squared) or |r| into sum_out[0] (device float); accumulated in
 * FP64 with a fixed reduction order, then rounded (host kernels.hpp:502-514
 * uses an FP32 OpenMP reduction). */
int gbkhip_residual_sum(gbkhip_ctx* ctx, const float* residual,
                        long long size, int squared, float* sum_out);
/* counts[0..2] += (data1 only, data2 only, both) beyond epsilon
 * (host kernels.hpp:425-482). */
int gbkhip_count_pixels(gbkhip_ctx* ctx, const float* data1,
                        const float* data2, long long size, float epsilon,
                        int* counts);

/* ---- Batched evaluation: no counterpart in the reference, which runs one
 *      parameter set per Objective.log_likelihood call
 *      (src/gbkfit/fitting/fitters/emcee/sampler.py:34-57).  Many sets of one
 *      smdisk galaxy model share geometry, PSF x LSF spectrum and data; each
 *      set is a flat float vector laid out as described below. ------------ */
typedef struct {
    int n;                   /* traits of this kind (0 when absent) */
    const int* uids;         /* host */
    const int* ccounts;      /* host */
    const int* pcounts;      /* host */
    int off_cvalues;         /* offset of this kind's constants in a set */
    int off_pvalues;         /* offset of this kind's parameters in a set */
} gbkhip_batch_traits;

typedef struct gbkhip_batch_desc_s {
    int loose, tilted;
    int nrnodes;
    const float* rnodes;     /* host, sub-ring nodes */
    int vsys_count;          /* 0 when the disk has no vsys */
    int off_vsys, off_xpos, off_ypos, off_posa, off_incl;
    gbkhip_batch_traits traits[9];   /* rp rh vp vh dp dh zp sp wp */
    int pstride;             /* floats per parameter set */
    gbkhip_grid grid;        /* gmodel grid; must cover the padded cube */
    int size_lo[3];          /* low-res data cube (x, y, z) */
    int size_hi[3];          /* padded hi-res cube */
    int scale[3];
    int edge_hi[3];          /* downscale offset (DCube edge) */
    int has_conv;            /* PSF and/or LSF present */
    const float* pcube;      /* device, padded PSF x LSF cube, centre at 0 */
    const float* obs;        /* device, observed low-res cube; NULL when the
                                objective is moment maps (set_mmaps) */
    const float* err;        /* device error cube or NULL */
    const float* mask;       /* device mask cube or NULL */
    float err_scalar;        /* error used when err is NULL (1: none) */
    float weight;
    int max_batch;           /* sets evaluated per chunk (buffer sizing) */
    int path;                /* 0: automatic, 1: FFT of the padded cube,
                                2: spectral-first fused FFT pipeline,
                                3: direct separable convolution; 2 and 3 need
                                a thin disk and a PSF x LSF cube that
                                factors within the DCube edge; automatic
                                takes the cheaper of 2 and 3, else 1 */
    /* Monte Carlo cloud disk (gmodels.hpp:325-643) instead of smdisk: runs
     * the padded-cube pipeline; evaluate with gbkhip_batch_evaluate_clouds */
    int mcdisk;
    float cflux;             /* flux of one cloud */
    int csum_len;            /* length of each set's cumulative cloud counts */
    const unsigned char* hoi;   /* device, per density trait: analytic
                                   integral (one sub-ring) or not */
    unsigned long long seed; /* Philox key of the cloud stream */
    /* further components of the gmodel (smdisk; all thin or all thick like
     * the first), summed into the model cube after the first one
     * (gmodels/_detail.py:405-422): each a desc of which only the disk
     * fields (loose .. traits) are read, their offsets into the same
     * parameter vector.  Read at creation only.  With further components
     * the direct pipeline deposits a hi-res cube and the fused FFT
     * pipeline is not used. */
    int nextra;
    const struct gbkhip_batch_desc_s* extra;
} gbkhip_batch_desc;

typedef struct gbkhip_batch gbkhip_batch;

/* Builds the plan: copies the layout, computes the kernel spectrum and sizes
 * the chunk buffers (max_batch padded cubes and spectra).  max_batch must
 * not exceed the device's grid y/z limit (one grid row per set); passes
 * whose grid holds several rows per set are launched over slices of sets. */
int gbkhip_batch_create(gbkhip_ctx* ctx, const gbkhip_batch_desc* desc,
                        gbkhip_batch** out);
int gbkhip_batch_destroy(gbkhip_batch* batch);
/* chi2[s] (device, FP64) = sum of squared residuals of set s; cubes (device,
 * nsets x low-res cube) receives the model cubes when not NULL.  With
 * use_graph the launch sequence for these pointers is captured once in a
 * HIP graph and replayed on later calls. */
int gbkhip_batch_evaluate(gbkhip_batch* batch, const float* params,
                          int nsets, double* chi2, float* cubes,
                          int use_graph);
/* evaluate() of an mcdisk batch.  csum (device, nsets x csum_len int32):
 * each set's cumulative (non-decreasing) cloud counts per sub-ring, the
 * integral of its density traits over cflux (_mcdisk.py:16-189; the host
 * computes it, as the reference does); qscale (device, nsets FP64): each
 * set's fixed-point scale (gbkhip_mcdisk_fixed_scale of its total cloud
 * count).  A set's result
 * depends only on its own parameters and counts. */
int gbkhip_batch_evaluate_clouds(gbkhip_batch* batch, const float* params,
                                 const int* csum, const double* qscale,
                                 int nsets, double* chi2, float* cubes,
                                 int use_graph);
/* Cloud capacity per set of an mcdisk batch (its binned deposits keep one
 * record per cloud): at least the largest total cloud count of the sets
 * evaluated next.  Growing it frees the work buffers (re-allocated by the
 * next evaluation) and the cached graphs.  A set whose total cloud count
 * exceeds the capacity deposits only the first `capacity` clouds, and the
 * evaluation raises the batch's overflow flag (gbkhip_batch_check). */
int gbkhip_batch_reserve_clouds(gbkhip_batch* batch, long long clouds);
/* Waits for the batch's stream, then returns in *overflow whether an
 * evaluation since the last check had a set with more clouds than the
 * reserved capacity (its results are then not valid), and clears the flag. */
int gbkhip_batch_check(gbkhip_batch* batch, int* overflow);
/* Fixed-point scale 2^S of the int64 cube of an mcdisk evaluation of nclouds
 * clouds of flux cflux on grid: the value gbkhip_mcdisk_evaluate uses. */
int gbkhip_mcdisk_fixed_scale(float cflux, int nclouds,
                              const gbkhip_grid* grid, double* out);
/* Tile settings of the direct separable pipeline (the autotuner's knobs;
 * other pipelines accept only 1, 1): lanes = chunks in flight on parallel
 * streams (1 or 2), ysplit = blocks the Y outputs of a y-pass column strip
 * are split over (1, 2 or 4).  Takes effect on the next evaluate(). */
int gbkhip_batch_set_tiles(gbkhip_batch* batch, int lanes, int ysplit);

/* ---- Ragged groups: galaxies of different cube sizes, models and data
 *      (each a batch of its own) evaluated in one launch -- one HIP graph
 *      per set of call pointers, members spread over parallel streams.  No
 *      counterpart in the reference, which evaluates one model per call. */
typedef struct gbkhip_group gbkhip_group;
int gbkhip_group_create(gbkhip_ctx* ctx, gbkhip_batch* const* members, int n,
                        gbkhip_group** out);
int gbkhip_group_destroy(gbkhip_group* group);
/* Member i evaluates nsets[i] sets from params[i] into chi2[i] (and its model
 * cubes into cubes[i] when cubes and cubes[i] are not NULL); csum/qscale
 * hold the cloud budgets of mcdisk members (NULL entries for the others,
 * NULL arrays when there are none).  Each member's results equal those of
 * its own gbkhip_batch_evaluate(_clouds). */
int gbkhip_group_evaluate(gbkhip_group* group, const float* const* params,
                          const int* const* csum,
                          const double* const* qscale, const int* nsets,
                          double* const* chi2, float* const* cubes,
                          int use_graph);

/* Direct pipeline, last chunk evaluated on the first lane: counts[0] of
 * counts[1] x-pass tiles (one hi-res row x one channel chunk of one set)
 * that some line reaches, counts[2] of counts[3] y-pass output groups (one
 * channel x 16 (or 4) rows of one set) with a nonzero input, and counts[4]
 * of counts[5] z-pass blocks (16 channels x 4 rows of one set) with a
 * nonzero input; the others are skipped (their outputs are zero; a skipped
 * z-pass block takes its precomputed data-only partial sum).  Waits for
 * the batch's stream.  Zeros for other pipelines. */
int gbkhip_batch_tile_stats(gbkhip_batch* batch, long long* counts);
/* Direct pipeline from line parameters (thin disks), last chunk evaluated
 * on the first lane: counts[0] of counts[1] hi-res spaxels lie in 64-spaxel
 * row segments that some line reaches, the only line parameters
 * batch_lines_thin writes and the x pass reads.  Waits for the batch's
 * stream.  Zeros for other pipelines. */
int gbkhip_batch_line_stats(gbkhip_batch* batch, long long* counts);
/* Stage timing: with timing on, evaluate() launches directly (no graph) and
 * records events around the smdisk, convolution and downscale+chi^2 stages
 * of every chunk; stage_times() waits for them and returns the summed stage
 * durations (ms[3], milliseconds) of the last evaluate() and its chunk
 * count. */
/* Stream of later evaluate() calls (NULL: the context's stream). */
int gbkhip_batch_set_stream(gbkhip_batch* batch, void* stream);
/* Work arena (no counterpart in the reference): the batch's work buffers
 * (everything sized by max_batch) become views into `arena`, a device
 * buffer the caller owns, of at least gbkhip_batch_arena_bytes bytes, laid
 * out the same way every time.  Batches that never run at the same time
 * (one serving slot's stream) can share one arena, so the work memory of
 * many resident galaxies is that of the slots, not of every galaxy.  A
 * graph captured with an arena stays cached for it: rebinding a batch to an
 * arena it used before costs nothing.  NULL returns to buffers of the
 * batch's own.  The caller keeps an arena allocated while any batch may
 * run with it, and does not run two batches with one arena at once. */
int gbkhip_batch_arena_bytes(gbkhip_batch* batch, unsigned long long* bytes);
int gbkhip_batch_bind_work(gbkhip_batch* batch, void* arena,
                           unsigned long long bytes);
/* Moment-map objective instead of the cube residual: the reference's
 * DModelMMaps + Objective (src/gbkfit/model/dmodels/mmaps.py:136-201,
 * objective.py:279-308; moments core/dmodel/dmodels.hpp:76-228).  chi2[s]
 * becomes sum over the orders and spaxels of r^2, r = (moment - data) /
 * error, zero where a data mask entry is 0 or the model's moments are
 * invalid (|sum I dv| <= cutoff), times the order's weight.  orders (host):
 * norders sorted distinct values 0-7; obs (device, norders x ny x nx): the
 * data maps; err (device, same shape) or err_scalars (host, one per
 * order); mask (device, same shape) or NULL; weights (host, one per order)
 * or NULL (1).  zero_z / step_z: the low-res spectral axis.  norders = 0
 * returns to the cube objective.  Frees the work buffers and graphs. */
int gbkhip_batch_set_mmaps(gbkhip_batch* batch, int norders,
                           const int* orders, float cutoff, float zero_z,
                           float step_z, const float* obs, const float* err,
                           const float* err_scalars, const float* mask,
                           const float* weights);
/* Frees the work buffers (everything sized by max_batch) and the cached
 * graphs that use them; the next evaluate() allocates them again.  Graphs
 * captured with a work arena (gbkhip_batch_bind_work) stay cached for it.
 * Resident state (layout, kernel spectra, plans) stays. */
int gbkhip_batch_release(gbkhip_batch* batch);
/* Bytes of resident state and of work buffers (needed whether allocated or
 * not), and whether the work buffers are allocated now. */
int gbkhip_batch_memory(gbkhip_batch* batch, unsigned long long* resident,
                        unsigned long long* work, int* allocated);
/* info[0] = pipeline (0 padded-cube FFT, 1 fused FFT, 2 direct separable),
 * info[1..3] = its transform sizes (x, y, z) (for the direct pipeline: the
 * extent of the inputs that reach the outputs), info[4] = chi^2 partial sums
 * per set. */
int gbkhip_batch_info(gbkhip_batch* batch, int* info);
/* Direct pipeline: taps[0..2] = the filter taps an output uses along x, y
 * and z (the block-mean-folded kernel without its negligible ends: each
 * end's dropped |taps| sum to at most 2^-31 of the total), taps[3..5] = the
 * inputs skipped before an output's first tap.  No reference counterpart (the reference convolves with the
 * whole kernel by FFT, host/include/gbkfit/host/fft.hpp:199-207). */
int gbkhip_batch_taps(gbkhip_batch* batch, int* taps);
int gbkhip_batch_set_timing(gbkhip_batch* batch, int enable);
int gbkhip_batch_stage_times(gbkhip_batch* batch, double* ms, int* chunks);
/* Per-kernel times of the last timed evaluate(), summed over its chunks:
 * *n = 3 on the padded-cube path (smdisk, FFT convolution, downscale +
 * chi^2), 7 on the fused FFT path (lines, deposit + z FFT, y FFT, x FFT
 * convolution, inverse y FFT, inverse z FFT + downscale + chi^2, chi^2
 * final) and 5 on the direct path (lines, x pass, y pass, z pass + chi^2,
 * chi^2 final).  cap = length of ms. */
int gbkhip_batch_kernel_times(gbkhip_batch* batch, double* ms, int cap,
                              int* n);

#ifdef __cplusplus
}
#endif

#endif /* GBKHIP_H */
