// Shared host/device helpers of libgbkhip: error plumbing, the per-context
// state and small device utilities.  HIP for gfx950 only.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/gbkhip.h"

namespace gbkhip {

// ---------------------------------------------------------------------------
// Errors: every C entry point returns a status code and records a message
// retrievable with gbkhip_last_error() (thread local).
// ---------------------------------------------------------------------------
void set_error(const std::string& msg);

struct Error {
    int code;
    std::string msg;
};

#define GBK_HIP(expr)                                                        \
    do {                                                                     \
        hipError_t e_ = (expr);                                              \
        if (e_ != hipSuccess)                                                \
            throw ::gbkhip::Error{GBKHIP_ERR_HIP,                            \
                std::string(#expr ": ") + hipGetErrorString(e_)};            \
    } while (0)

#define GBK_CHECK(cond, code, msg)                                           \
    do {                                                                     \
        if (!(cond)) throw ::gbkhip::Error{(code), (msg)};                   \
    } while (0)

template<typename F>
int guarded(F&& f)
{
    try {
        f();
        return GBKHIP_OK;
    } catch (const Error& e) {
        set_error(e.msg);
        return e.code;
    } catch (const std::bad_alloc&) {
        set_error("host allocation failed");
        return GBKHIP_ERR_ALLOC;
    } catch (const std::exception& e) {
        set_error(e.what());
        return GBKHIP_ERR_INTERNAL;
    }
}

// ---------------------------------------------------------------------------
// Device buffer owned by the library (plans, cached spectra, scratch).
// ---------------------------------------------------------------------------
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    bool owned = true;      // false: a view into memory someone else owns
    DevBuf() = default;
    explicit DevBuf(size_t n) { alloc(n); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void alloc(size_t n)
    {
        release();
        if (n) GBK_HIP(hipMalloc(&ptr, n));
        bytes = n;
    }
    void view(void* p, size_t n)
    {
        release();
        ptr = n ? p : nullptr;
        bytes = n;
        owned = false;
    }
    void ensure(size_t n)
    {
        if (n > bytes) alloc(n);
    }
    void release()
    {
        if (ptr && owned) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
        owned = true;
    }
    template<typename T> T* as() const { return static_cast<T*>(ptr); }
};

struct FFTPlan;   // fft.hip
struct SepConv;   // sepconv.hpp

struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    // FFT plans keyed by (nx, ny, nz, batch-independent)
    std::map<std::tuple<int, int, int>, std::shared_ptr<FFTPlan>> plans;
    // cached kernel spectra keyed by (nx, ny, nz, pointer, generation)
    std::map<std::tuple<int, int, int, uintptr_t, uint64_t>,
             std::shared_ptr<DevBuf>> spectra;
    // separable-kernel analyses, keyed like the spectra
    std::map<std::tuple<int, int, int, uintptr_t, uint64_t>,
             std::shared_ptr<SepConv>> sepconvs;
    // per-pointer scratch spectra for fft_convolve_cached
    std::map<std::tuple<int, int, int, uintptr_t>, std::shared_ptr<DevBuf>>
        work;
    DevBuf reduce_scratch;   // partial sums of reductions
    DevBuf int_scratch;      // int64 fixed-point accumulation (mcdisk)
    DevBuf mc_scratch;       // binned mcdisk: records, ranks, tile offsets
    std::mutex mu;
};

Context* ctx_of(gbkhip_ctx* c);

inline int ceil_div(long long a, long long b) { return (int)((a + b - 1) / b); }

}  // namespace gbkhip
