// Host-side declarations shared by the C++ translation units of the library
// (gridder object, operators, request server).  Not part of the public ABI.
#pragma once

#include "../../include/mgg.h"
#include "mgg_internal.h"

#include <hip/hip_runtime.h>

#include <cstddef>
#include <string>
#include <vector>

namespace mgg {

// ---- errors: every entry point returns a status; the message is per thread
extern thread_local std::string g_last_error;
int fail(int code, const char *fmt, ...);
[[noreturn]] void raise(int code, const char *fmt, ...);

struct Failure {
    int code;
};

#define HIP_TRY(expr)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess) {                                                              \
            fail(e_ == hipErrorOutOfMemory ? MGG_ERR_OUT_OF_MEMORY : MGG_ERR_HIP,            \
                 "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
            throw Failure{e_ == hipErrorOutOfMemory ? MGG_ERR_OUT_OF_MEMORY : MGG_ERR_HIP};  \
        }                                                                                    \
    } while (0)

#define FFT_TRY(expr)                                                                     \
    do {                                                                                  \
        rocfft_status s_ = (expr);                                                        \
        if (s_ != rocfft_status_success) {                                                \
            fail(MGG_ERR_FFT, "%s failed: rocfft status %d (%s:%d)", #expr, (int)s_,      \
                 __FILE__, __LINE__);                                                     \
            throw Failure{MGG_ERR_FFT};                                                   \
        }                                                                                 \
    } while (0)

template <typename F>
int guarded(F &&f)
{
    try {
        f();
        return MGG_OK;
    } catch (const Failure &e) {
        return e.code;
    } catch (const std::bad_alloc &) {
        return fail(MGG_ERR_OUT_OF_MEMORY, "host allocation failed");
    } catch (const std::exception &e) {
        return fail(MGG_ERR_HIP, "unexpected error: %s", e.what());
    }
}

// Restores the caller's current device when a call returns.
struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev)
    {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (dev >= 0 && dev != prev) HIP_TRY(hipSetDevice(dev));
    }
    ~DeviceGuard()
    {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

// ---------------------------------------------------------------------------
// Owned device buffer (grows on demand, freed in the destructor).
// ---------------------------------------------------------------------------
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release()
    {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    void *reserve(size_t n)
    {
        if (n <= bytes && ptr) return ptr;
        release();
        if (n == 0) return nullptr;
        hipError_t e = hipMalloc(&ptr, n);
        if (e != hipSuccess) {
            ptr = nullptr;
            raise(MGG_ERR_OUT_OF_MEMORY, "hipMalloc of %zu bytes failed: %s", n,
                  hipGetErrorString(e));
        }
        bytes = n;
        return ptr;
    }
    template <typename T> T *as() const { return static_cast<T *>(ptr); }
};

// Largest batch handed to rocFFT in one plan; bigger requests are split so the
// plan cache stays small.
constexpr size_t kFftBatch = 256;

// ---- rocFFT (plans bound to one stream; batches split to kFftBatch)
int current_device();
// kind: 0 = 2-D R2C, 1 = 2-D C2R, 2 = 3-D R2C, 3 = 3-D C2R
void fft_exec(int kind, int m, size_t batch, void *in, void *out, hipStream_t s);
void fft2d_batched(int kind, int m, size_t n, void *in, size_t inStride, void *out,
                   size_t outStride, hipStream_t s);
// Drops the cached plans bound to a stream that is about to be destroyed.
void fft_drop_stream(hipStream_t s);

// ---- host math
double kernel_alpha(int hw);
void kb_table_host(std::vector<double> &w, int size, int hw);
void precomp_host(std::vector<double> &out, int m, int hw);
void axes_from_normal(const double n[3], double phi_r, float out[9]);

// ---- device chains
void check_launch(hipError_t e, const char *what);
const int *brick_order(const ProjGeom &g, int &len);
hipError_t back_project_launch(float *vol, float *dens, const ProjGeom &g, const float *axes, int n,
                               const float *casImgs, hipStream_t s);
void chain_cas_to_images(float *casImgs, float2 *herm, float *images, int n, int m, int nn,
                         hipStream_t s);
// scale / zeroNyq: see launch_herm2d_to_cas (1 / 0 for ImgsToCASImgs itself).
void chain_images_to_cas(const float *images, float *casImgs, float2 *herm, int n, int m, int nn,
                         hipStream_t s, float scale = 1.f, int zeroNyq = 0);
void chain_volume_to_cas(const float *vol, int nn, float *casVol, int m, int pad,
                         const float *dPrecomp, float *real3, float2 *herm3, hipStream_t s);
// transpose: no 1/m^3 (with zeroNyq = 0 this is the exact transpose of
// chain_volume_to_cas).
void chain_cas_to_volume(const float *casVol, int m, int pad, float *vol, int nn,
                         const float *dPrecomp, int zeroNyq, float *real3, float2 *herm3,
                         hipStream_t s, bool transpose = false);
void validate_geom(int v, int m, int kerSize, float hw);

}  // namespace mgg
