// Internal launchers shared between the kernel files and the C ABI layer.
#pragma once

#include "common.hpp"

namespace gbkhip {

struct FFTPlan;

// smdisk.hip
void launch_smdisk(Context& c, const gbkhip_disk& d, const gbkhip_grid& g,
                   const gbkhip_outputs& o);
// mcdisk.hip.  With fixed_scratch (fixed_bytes long) the binned deposit
// buffers live there and nothing is allocated; calls that need more, or the
// non-binned deposit, are rejected before anything is enqueued.
void launch_mcdisk(Context& c, float cflux, int nclouds, const int* csum,
                   int csum_len, const unsigned char* hoi,
                   const gbkhip_disk& d, const gbkhip_grid& g,
                   const gbkhip_outputs& o, uint64_t seed,
                   char* fixed_scratch = nullptr, size_t fixed_bytes = 0);
// size of the fixed scratch of gbkhip_mcdisk_evaluate_noalloc (64 MiB)
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
// its device address on `device` (the current device), looked up once
void noalloc_scratch_setup(int device);
char* noalloc_scratch(int device);
double mcdisk_fixed_scale(float cflux, int nclouds, const gbkhip_grid& g);
// dmodel.hip
void launch_downscale(Context& c, const int* scale, const int* offset,
                      const int* src_size, const int* dst_size,
                      const float* src, float* dst, int batch, long long bsrc,
                      long long bdst);
void launch_mask(Context& c, float cutoff, int apply, long long n, float* d,
                 float* m);
void launch_moments(Context& c, const int* size, const float* step,
                    const float* zero, const float* d, const float* w,
                    float cutoff, int norders, const int* orders, float* md,
                    float* mm, float* mw, int batch, long long bcube,
                    long long bmap);
void launch_wcube(Context& c, const int* spat_size, int spec_size,
                  const float* spat, float* spec);
void launch_convolve(Context& c, const int* s1, const int* s2,
                     const float* a, const float* k, float* out);
// objective.hip
void launch_residual(Context& c, const float* od, const float* oe,
                     const float* om, const float* md, const float* mw,
                     const float* mm, long long n, float weight, float* res);
void launch_sum(Context& c, const float* r, long long n, int squared,
                float* out_f, double* out_d);
void launch_count(Context& c, const float* a, const float* b, long long n,
                  float eps, int* counts);
// fft.hip
std::shared_ptr<FFTPlan> get_plan(Context& c, int nx, int ny, int nz);
void fft_forward(Context& c, const FFTPlan& p, const float* in, float2* out,
                 int batch, long long br, long long bc);
void fft_inverse(Context& c, const FFTPlan& p, const float2* in, float2* work,
                 float* out, int batch, long long bc, long long br);
void fft_prepare_attributes();
void fft_convolve_with(Context& c, const FFTPlan& p, float* data_r,
                       float2* spec, const float2* K, int batch, long long br,
                       long long bc, long long kbc);
// register-tiled engine: shapes, twiddles and the strided line pass
struct FastShape;
FastShape fast_shape(int n);
FastShape fast_shape_lines(int n, int lshift);
FastShape fast_shape_rows(int n);
unsigned div_magic(unsigned d);
void upload_twiddles(DevBuf& buf, int n, int count);
void upload_stage_twiddles(DevBuf& buf, int n);
void launch_fx_lines(Context& c, const FastShape& f, const float2* tw,
                     const float2* tws, const float2* src, float2* dst,
                     const float2* K,
                     int nlines, int inner, long long base0,
                     long long ostride, long long stride, int st_lo,
                     int st_hi, int mode, float scale, int batch,
                     long long bstride, long long kbstride);

}  // namespace gbkhip
