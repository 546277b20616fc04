// Host helpers shared by the gridder object, the operators and the request
// server: per-thread errors, device buffers, the rocFFT plan cache, host math
// (Kaiser-Bessel table, precompensation, direction frames) and the device
// chains (image / volume FFT pipelines, back-projection dispatch).
#include "mgg_common.h"

#include <rocfft/rocfft.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

namespace mgg {


// ---------------------------------------------------------------------------
// Errors: every entry point returns a status; the message is kept per thread.
// ---------------------------------------------------------------------------
thread_local std::string g_last_error;

int fail(int code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}




void raise(int code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    throw Failure{code};
}




// ---------------------------------------------------------------------------
// rocFFT plans.  A plan and its execution info are bound to one stream, and
// all plans of a stream share one work buffer (work on a stream runs in
// order), grown when a new plan needs more.  Plans without a buffer of their
// own are small, so the cache holds every batch size fft2d_batched uses.
// ---------------------------------------------------------------------------
std::once_flag g_fft_once;

struct FftPlan {
    rocfft_plan plan = nullptr;
    rocfft_execution_info info = nullptr;
    ~FftPlan()
    {
        if (info) rocfft_execution_info_destroy(info);
        if (plan) rocfft_plan_destroy(plan);
    }
};

// kind: 0 = 2-D R2C, 1 = 2-D C2R, 2 = 3-D R2C, 3 = 3-D C2R
using PlanKey = std::tuple<int, int, int, size_t, hipStream_t>;
using StreamKey = std::pair<int, hipStream_t>;

struct StreamWork {
    void *buf = nullptr;
    size_t bytes = 0;
};

struct PlanCache {
    std::mutex mu;
    std::map<PlanKey, std::unique_ptr<FftPlan>> plans;
    std::vector<PlanKey> order;  // insertion order, for eviction
    std::map<StreamKey, StreamWork> work;
    static constexpr size_t kMaxPlans = 1024;

    FftPlan *get(int device, int kind, int m, size_t batch, hipStream_t stream)
    {
        std::call_once(g_fft_once, [] { rocfft_setup(); });
        std::lock_guard<std::mutex> lk(mu);
        PlanKey key{device, kind, m, batch, stream};
        auto it = plans.find(key);
        if (it != plans.end()) return it->second.get();
        if (plans.size() >= kMaxPlans) {
            // The oldest plan may still be queued on its stream: wait for it.
            PlanKey old = order.front();
            order.erase(order.begin());
            (void)hipStreamSynchronize(std::get<4>(old));
            plans.erase(old);
        }
        auto p = std::make_unique<FftPlan>();
        const size_t len2[2] = {(size_t)m, (size_t)m};
        const size_t len3[3] = {(size_t)m, (size_t)m, (size_t)m};
        const bool threeD = kind >= 2;
        const rocfft_transform_type tt = (kind == 0 || kind == 2) ? rocfft_transform_type_real_forward
                                                                  : rocfft_transform_type_real_inverse;
        FFT_TRY(rocfft_plan_create(&p->plan, rocfft_placement_notinplace, tt,
                                   rocfft_precision_single, threeD ? 3 : 2, threeD ? len3 : len2,
                                   batch, nullptr));
        size_t wbytes = 0;
        FFT_TRY(rocfft_plan_get_work_buffer_size(p->plan, &wbytes));
        FFT_TRY(rocfft_execution_info_create(&p->info));
        FFT_TRY(rocfft_execution_info_set_stream(p->info, stream));
        StreamWork &w = work[StreamKey{device, stream}];
        if (wbytes > w.bytes) {
            // The stream's queued transforms may still use the old buffer.
            if (w.buf) {
                HIP_TRY(hipStreamSynchronize(stream));
                (void)hipFree(w.buf);
                w.buf = nullptr;
                w.bytes = 0;
            }
            hipError_t e = hipMalloc(&w.buf, wbytes);
            if (e != hipSuccess)
                raise(MGG_ERR_OUT_OF_MEMORY, "rocFFT work buffer (%zu bytes): %s", wbytes,
                      hipGetErrorString(e));
            w.bytes = wbytes;
            for (auto &kv : plans)
                if (std::get<0>(kv.first) == device && std::get<4>(kv.first) == stream)
                    FFT_TRY(rocfft_execution_info_set_work_buffer(kv.second->info, w.buf, w.bytes));
        }
        if (w.bytes) FFT_TRY(rocfft_execution_info_set_work_buffer(p->info, w.buf, w.bytes));
        FftPlan *raw = p.get();
        plans.emplace(key, std::move(p));
        order.push_back(key);
        return raw;
    }

    // Called for a stream that is about to be destroyed (its work is done).
    void drop_stream(hipStream_t s)
    {
        std::lock_guard<std::mutex> lk(mu);
        for (auto it = order.begin(); it != order.end();) {
            if (std::get<4>(*it) == s) {
                plans.erase(*it);
                it = order.erase(it);
            } else {
                ++it;
            }
        }
        for (auto it = work.begin(); it != work.end();) {
            if (it->first.second == s) {
                if (it->second.buf) (void)hipFree(it->second.buf);
                it = work.erase(it);
            } else {
                ++it;
            }
        }
    }
};

PlanCache &plan_cache()
{
    static PlanCache *c = new PlanCache();  // leaked on purpose: outlives static teardown
    return *c;
}

int current_device()
{
    int d = 0;
    HIP_TRY(hipGetDevice(&d));
    return d;
}

void fft_exec(int kind, int m, size_t batch, void *in, void *out, hipStream_t s)
{
    FftPlan *p = plan_cache().get(current_device(), kind, m, batch, s);
    void *ib[1] = {in};
    void *ob[1] = {out};
    FFT_TRY(rocfft_execute(p->plan, ib, ob, p->info));
}


// Batches of kFftBatch (a power of two), then the remainder as a sum of powers
// of two: at most log2(kFftBatch) + 1 plan sizes exist per transform shape and
// stream, where a plan per remainder would mean a new plan (and work buffer)
// for most call sizes.
void fft2d_batched(int kind, int m, size_t n, void *in, size_t inStride, void *out,
                   size_t outStride, hipStream_t s)
{
    for (size_t a = 0; a < n;) {
        size_t b = std::min(kFftBatch, n - a);
        b = size_t(1) << (63 - __builtin_clzll((unsigned long long)b));  // largest power of two <= b
        fft_exec(kind, m, b, (char *)in + a * inStride, (char *)out + a * outStride, s);
        a += b;
    }
}

// ---------------------------------------------------------------------------
// Host math (the reference computes these in Matlab).
// ---------------------------------------------------------------------------
double bessel_i0(double x)
{
    // Power series sum_k ((x/2)^2)^k / (k!)^2; converges fast for the
    // arguments used here (x <= 17).
    const double q = 0.25 * x * x;
    double term = 1.0, sum = 1.0;
    for (int k = 1; k < 500; k++) {
        term *= q / ((double)k * (double)k);
        sum += term;
        if (term < sum * 1e-18) break;
    }
    return sum;
}

double kernel_alpha(int hw)
{
    static const double alphas[5] = {5.5, 10.0, 10.2, 13.0, 17.0};  // utils/kernelAlpha.m:6
    if (hw < 1 || hw > 5) return 0.0;
    return alphas[hw - 1];
}

double kaiser(int hw, double alpha, double x)
{
    // utils/gridLibKaiser.m:1-15
    if (std::fabs(x) > hw) return 0.0;
    const double r = std::sqrt(1.0 - (x / hw) * (x / hw));
    return bessel_i0(alpha * r) / bessel_i0(alpha);
}

void kb_table_host(std::vector<double> &w, int size, int hw)
{
    // Matlab/getKernelFiltTable.m:1-22.  Sample offsets are (k - (size-1)/2)*delta
    // so the table is exactly symmetric.
    const double alpha = kernel_alpha(hw);
    const double delta = 2.0 * hw / (size - 1);
    w.assign(size, 0.0);
    double integral = 0.0;
    for (int k = 0; k < size; k++) {
        w[k] = kaiser(hw, alpha, (k - (size - 1) / 2.0) * delta);
        integral += w[k] * delta;
    }
    for (int k = 0; k < size; k++) w[k] /= integral;
    const double w0 = w[0];
    for (int k = 0; k < size; k++) w[k] -= w0;
    w[size - 1] = 0.0;
    // Divide by the cube root of the sum of the separable 3-D kernel at the
    // integer offsets -hw..hw, i.e. by the sum of the 1-D integer taps.
    const int center = (size + 1) / 2;  // 1-based
    const double scale = (center - 1) / (double)hw;
    double s = 0.0;
    for (int x = -hw; x <= hw; x++) s += w[(int)llround(x * scale + center) - 1];
    const double denom = s * s * s;
    const double c = std::cbrt(denom);
    for (int k = 0; k < size; k++) w[k] /= c;
}

void precomp_host(std::vector<double> &out, int m, int hw)
{
    // Matlab/getPreComp.m:1-28.  The 4x-oversampled window is symmetric, so
    // real(fftshift(fft(fftshift(line)))) at frequency k is the cosine sum
    // sum_t w_t cos(2 pi k t / (4m)).
    const int ovr = 4;
    const double alpha = kernel_alpha(hw);
    const int half = hw * ovr;
    std::vector<double> w(2 * half + 1);
    for (int t = -half; t <= half; t++) w[t + half] = kaiser(hw, alpha, (double)t / ovr);
    const double w0 = w[0];
    for (double &x : w) x -= w0;
    w[2 * half] = 0.0;
    const double L = (double)m * ovr;
    auto ft = [&](int k) {
        double s = 0.0;
        for (int t = -half; t <= half; t++) s += w[t + half] * std::cos(2.0 * M_PI * k * t / L);
        return s;
    };
    const double fc0 = ft(0);
    out.resize(m);
    for (int i = 0; i < m; i++) out[i] = fc0 / ft(i - m / 2);
}

void axes_from_normal(const double n[3], double phi_r, float out[9])
{
    // utils/get_axes_from_n_rot.m:7-37: T = (T1*T2)*T3, columns stacked.
    double a[3] = {-n[1], n[0], 0.0};  // cross([0 0 1], n)
    double T1[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}, T2[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    if (std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]) >= 1e-4) {
        const double phi = std::atan2(n[1], n[0]);
        const double th = std::atan2(std::sqrt(n[0] * n[0] + n[1] * n[1]), n[2]);
        double t1[3][3] = {{std::cos(phi), -std::sin(phi), 0}, {std::sin(phi), std::cos(phi), 0}, {0, 0, 1}};
        double t2[3][3] = {{std::cos(th), 0, std::sin(th)}, {0, 1, 0}, {-std::sin(th), 0, std::cos(th)}};
        std::memcpy(T1, t1, sizeof T1);
        std::memcpy(T2, t2, sizeof T2);
    }
    double T3[3][3] = {{std::cos(phi_r), -std::sin(phi_r), 0}, {std::sin(phi_r), std::cos(phi_r), 0}, {0, 0, 1}};
    double A[3][3], T[3][3];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            A[r][c] = 0;
            for (int k = 0; k < 3; k++) A[r][c] += T1[r][k] * T2[k][c];
        }
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            T[r][c] = 0;
            for (int k = 0; k < 3; k++) T[r][c] += A[r][k] * T3[k][c];
        }
    for (int c = 0; c < 3; c++)
        for (int r = 0; r < 3; r++) out[3 * c + r] = (float)T[r][c];
}

// ---------------------------------------------------------------------------
// Device-pointer chains shared by the operators and the gridder object.
// ---------------------------------------------------------------------------
void check_launch(hipError_t e, const char *what)
{
    if (e != hipSuccess) raise(MGG_ERR_HIP, "%s launch failed: %s", what, hipGetErrorString(e));
}

// Centre-out brick order of the fast back projection, one device copy per
// (device, V), built on first use and kept for the process.
std::mutex g_order_mu;
std::map<std::pair<int, int>, std::unique_ptr<DevBuf>> g_order;

const int *brick_order(const ProjGeom &g, int &len)
{
    len = 0;
    if (!bp_uses_fast_path(g)) return nullptr;
    len = bp_brick_count(g.V, true);
    std::lock_guard<std::mutex> lk(g_order_mu);
    auto &b = g_order[std::make_pair(current_device(), g.V)];
    if (!b) {
        std::vector<int> host(len);
        bp_brick_order(g.V, host.data());
        auto nb = std::make_unique<DevBuf>();
        nb->reserve(sizeof(int) * len);
        HIP_TRY(hipMemcpy(nb->ptr, host.data(), sizeof(int) * len, hipMemcpyHostToDevice));
        b = std::move(nb);
    }
    return b->as<int>();
}

hipError_t back_project_launch(float *vol, float *dens, const ProjGeom &g, const float *axes, int n,
                               const float *casImgs, hipStream_t s)
{
    int len = 0;
    const int *order = brick_order(g, len);
    return launch_back_project(vol, dens, g, axes, n, casImgs, s, order, len);
}

// CAS images (n x m^2) -> images (n x nn^2); casImgs is reused as the C2R output.
void chain_cas_to_images(float *casImgs, float2 *herm, float *images, int n, int m, int nn,
                         hipStream_t s)
{
    if (n <= 0) return;
    if (fused_chain_supported(m, nn)) {
        check_launch(launch_fused_cas_to_images(casImgs, herm, images, m, nn, n, 1.0f / ((float)m * (float)m), s),
                     "fused_cas_to_images");
        return;
    }
    check_launch(launch_cas_to_herm2d(casImgs, herm, m, n, s), "cas_to_herm2d");
    fft2d_batched(1, m, n, herm, sizeof(float2) * (size_t)m * (m / 2 + 1), casImgs,
                  sizeof(float) * (size_t)m * m, s);
    check_launch(launch_real2d_to_imgs(casImgs, images, m, nn, n, 1.0f / ((float)m * (float)m), s),
                 "real2d_to_imgs");
}

// images (n x nn^2) -> CAS images (n x m^2); casImgs doubles as the R2C input.
void chain_images_to_cas(const float *images, float *casImgs, float2 *herm, int n, int m, int nn,
                         hipStream_t s, float scale, int zeroNyq)
{
    if (n <= 0) return;
    if (fused_chain_supported(m, nn)) {
        check_launch(launch_fused_images_to_cas(images, herm, casImgs, m, nn, n, scale, zeroNyq, s),
                     "fused_images_to_cas");
        return;
    }
    check_launch(launch_imgs_to_real2d(images, casImgs, m, nn, n, s), "imgs_to_real2d");
    fft2d_batched(0, m, n, casImgs, sizeof(float) * (size_t)m * m, herm,
                  sizeof(float2) * (size_t)m * (m / 2 + 1), s);
    check_launch(launch_herm2d_to_cas(herm, casImgs, m, n, s, scale, zeroNyq), "herm2d_to_cas");
}

void chain_volume_to_cas(const float *vol, int nn, float *casVol, int m, int pad,
                         const float *dPrecomp, float *real3, float2 *herm3, hipStream_t s)
{
    check_launch(launch_vol_to_real3d(vol, nn, real3, m, dPrecomp, s), "vol_to_real3d");
    fft_exec(2, m, 1, real3, herm3, s);
    check_launch(launch_herm3d_to_casvol(herm3, m, casVol, m + 2 * pad, pad, s), "herm3d_to_casvol");
}

void chain_cas_to_volume(const float *casVol, int m, int pad, float *vol, int nn,
                         const float *dPrecomp, int zeroNyq, float *real3, float2 *herm3,
                         hipStream_t s, bool transpose)
{
    check_launch(launch_casvol_to_herm3d(casVol, m + 2 * pad, pad, herm3, m, zeroNyq, s),
                 "casvol_to_herm3d");
    fft_exec(3, m, 1, herm3, real3, s);
    const double scale = transpose ? 1.0 : 1.0 / ((double)m * (double)m * (double)m);
    check_launch(launch_real3d_to_vol(real3, m, vol, nn, dPrecomp, (float)scale, s), "real3d_to_vol");
}

void validate_geom(int v, int m, int kerSize, float hw)
{
    if (v <= 0 || m <= 0) raise(MGG_ERR_INVALID_ARGUMENT, "sizes must be positive (v=%d m=%d)", v, m);
    if (m % 2) raise(MGG_ERR_INVALID_ARGUMENT, "CAS image size m=%d must be even", m);
    if (kerSize < 2 || kerSize > kMaxKerSize)
        raise(MGG_ERR_INVALID_ARGUMENT, "kernel table size %d outside [2, %d]", kerSize, kMaxKerSize);
    if (!(hw > 0.f)) raise(MGG_ERR_INVALID_ARGUMENT, "kernel half width must be positive");
}

void fft_drop_stream(hipStream_t s) { plan_cache().drop_stream(s); }

}  // namespace mgg
