// Host side of libMultiGPUGridder.so: the C ABI declared in include/mgg.h.
//
//   * rocFFT plan cache (batched 2-D / 3-D real transforms, one plan and work
//     buffer per (shape, batch, stream));
//   * device-pointer operators mgg_op_* / mgg_filter_*;
//   * the gridder object, which keeps the reference MultiGPUGridder /
//     gpuGridder semantics (borrowed host arrays, in-place results,
//     synchronous calls, per-stream direction chunks) on one MI355X;
//   * host math (Kaiser-Bessel table, precompensation, direction grid);
//   * the legacy ctypes symbols.
#include "../../include/mgg.h"
#include "mgg_common.h"

#include <hip/hip_runtime.h>
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

using namespace mgg;

// ===========================================================================
// Gridder object
// ===========================================================================
struct mgg_gridder {
    // geometry
    int N = 0;            // volume / image edge
    int M = 0;            // CAS image edge
    int V = 0;            // CAS volume edge
    int pad = 3;          // extraPadding
    float interp = 2.f;
    bool rawCas = false;  // legacy mode: V and M given directly, no spatial volume
    // parameters (AbstractGridder.cpp:7-17, MultiGPUGridder_Matlab_Class.m:47-49)
    int numAxes = 0;
    int maxAxes = 1000;
    int nStreamsFP = 10;
    int nStreamsBP = 4;
    float maskRadius = 0.f;
    float kerHWidth = 2.f;
    int runFFTOnDevice = 1;
    int verbose = 0;
    int device = 0;
    // borrowed host arrays
    float *hVol = nullptr;
    float *hCasVol = nullptr;
    float *hImgs = nullptr;
    float *hCasImgs = nullptr;
    float *hAxes = nullptr;
    float *hKer = nullptr;
    size_t hVolLen = 0, hCasVolLen = 0, hImgsLen = 0, hCasImgsLen = 0, hAxesLen = 0;
    int hKerLen = 0;
    // device state
    DevBuf dCasVol, dKer, dPrecomp, dVolTmp, dReal3, dHerm3, dDens;
    DevBuf dCheck;  // launch_check_volume's flag for dCasVol and dKer
    DevBuf dPartials, dPartialsDens;  // one V^3 partial volume per BP stream, contiguous
    std::vector<float> defaultKer;
    int kerLenOnDevice = 0;
    bool casFromVolume = false;  // device CAS volume produced by volume_to_cas
    struct Slot {
        hipStream_t stream = nullptr;
        DevBuf casImgs, herm, imgs, axes;
    };
    std::vector<std::unique_ptr<Slot>> slots;
    // legacy API: library-owned CAS-image output
    std::vector<float> legacyImgs;
    int legacyImgSize = 0;

    ~mgg_gridder()
    {
        DeviceGuard g(device);
        for (auto &s : slots) {
            if (s->stream) {
                (void)hipStreamSynchronize(s->stream);
                fft_drop_stream(s->stream);
                (void)hipStreamDestroy(s->stream);
            }
        }
    }

    void log(const char *fmt, ...)
    {
        if (!verbose) return;
        va_list ap;
        va_start(ap, fmt);
        fprintf(stderr, "[mgg] ");
        vfprintf(stderr, fmt, ap);
        fprintf(stderr, "\n");
        va_end(ap);
    }

    void ensure_slots(int n)
    {
        while ((int)slots.size() < n) {
            auto s = std::make_unique<Slot>();
            HIP_TRY(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
            slots.push_back(std::move(s));
        }
    }

    void upload_kernel_table()
    {
        const float *src;
        int len;
        if (hKer && hKerLen > 0) {
            src = hKer;
            len = hKerLen;
        } else {
            if (defaultKer.empty()) {
                std::vector<double> w;
                kb_table_host(w, 501, (int)lroundf(kerHWidth));
                defaultKer.assign(w.begin(), w.end());
            }
            src = defaultKer.data();
            len = (int)defaultKer.size();
        }
        if (len > kMaxKerSize) raise(MGG_ERR_INVALID_ARGUMENT, "kernel table of %d entries exceeds %d", len, kMaxKerSize);
        dKer.reserve(sizeof(float) * len);
        HIP_TRY(hipMemcpy(dKer.ptr, src, sizeof(float) * len, hipMemcpyHostToDevice));
        kerLenOnDevice = len;
    }

    ProjGeom geom() const
    {
        ProjGeom g;
        g.V = V;
        g.M = M;
        g.maskRadius = maskRadius;
        g.ker = dKer.as<float>();
        g.K = kerLenOnDevice;
        g.kerHWidth = kerHWidth;
        return g;
    }

    void ensure_precomp()
    {
        if (dPrecomp.ptr) return;
        std::vector<double> p;
        precomp_host(p, M, (int)lroundf(kerHWidth));
        std::vector<float> pf(p.begin(), p.end());
        dPrecomp.reserve(sizeof(float) * M);
        HIP_TRY(hipMemcpy(dPrecomp.ptr, pf.data(), sizeof(float) * M, hipMemcpyHostToDevice));
    }

    void require_axes(int n)
    {
        if (n <= 0) raise(MGG_ERR_INVALID_ARGUMENT, "number of axes is %d", n);
        if (!hAxes) raise(MGG_ERR_NOT_SET, "coordinate axes not set (SetCoordAxes)");
        if ((size_t)9 * n > hAxesLen)
            raise(MGG_ERR_INVALID_ARGUMENT, "num_axes %d needs %zu axis values, array holds %zu", n,
                  (size_t)9 * n, hAxesLen);
    }

    // Chunk plan (the role of gpuGridder::PlanOffsetValues, gpuGridder.cpp:
    // 598-752): chunk size E = ceil(min(P, maxAxes)/nStreams), chunk c runs on
    // stream c % nStreams; buffers of a stream are reused chunk after chunk.
    struct Chunk {
        int first, count, slot;
    };
    std::vector<Chunk> plan_chunks(int P, int nStreams, int &E) const
    {
        nStreams = std::max(1, nStreams);
        E = (int)(((long long)std::min(P, std::max(1, maxAxes)) + nStreams - 1) / nStreams);
        E = std::max(E, 1);
        std::vector<Chunk> out;
        int c = 0;
        for (int first = 0; first < P; first += E, c++) out.push_back({first, std::min(E, P - first), c % nStreams});
        return out;
    }

    void ensure_device_cas_volume()
    {
        dCasVol.reserve(sizeof(float) * (size_t)V * V * V);
    }

    void sync_slots(int n)
    {
        for (int i = 0; i < n && i < (int)slots.size(); i++) HIP_TRY(hipStreamSynchronize(slots[i]->stream));
    }

    // ---- ForwardProject (gpuGridder.cpp:180-305) ----
    void forward_project()
    {
        DeviceGuard dg(device);
        const int P = numAxes;
        require_axes(P);
        const bool wantImgs = runFFTOnDevice && !rawCas;
        const bool wantCas = (!runFFTOnDevice || verbose || rawCas) && (hCasImgs || rawCas);
        if (wantImgs && !hImgs) raise(MGG_ERR_NOT_SET, "images array not set (SetImages)");
        if (wantImgs && (size_t)P * N * N > hImgsLen)
            raise(MGG_ERR_INVALID_ARGUMENT, "images array holds %zu values, %d images of %d^2 need more", hImgsLen, P, N);
        float *casOut = rawCas ? legacyImgs.data() : hCasImgs;
        size_t casOutLen = rawCas ? legacyImgs.size() : hCasImgsLen;
        if (rawCas) {
            legacyImgs.assign((size_t)P * M * M, 0.f);
            casOut = legacyImgs.data();
            casOutLen = legacyImgs.size();
        }
        if (!wantImgs && !wantCas)
            raise(MGG_ERR_NOT_SET, "no output array: set the images (RunFFTOnDevice=1) or the CAS images");
        if (wantCas && (size_t)P * M * M > casOutLen)
            raise(MGG_ERR_INVALID_ARGUMENT, "CAS-images array too small for %d images", P);
        upload_kernel_table();
        ensure_device_cas_volume();
        if (!casFromVolume) {
            // The reference uploads the caller's CAS volume on every call (gpuGridder.cpp:194).
            float *src = rawCas ? hVol : hCasVol;
            size_t len = rawCas ? hVolLen : hCasVolLen;
            if (!src) raise(MGG_ERR_NOT_SET, "CAS volume not set (SetCASVolume or volume_to_cas_volume)");
            if (len < (size_t)V * V * V) raise(MGG_ERR_INVALID_ARGUMENT, "CAS volume array holds %zu values, need %d^3", len, V);
            HIP_TRY(hipMemcpy(dCasVol.ptr, src, sizeof(float) * (size_t)V * V * V, hipMemcpyHostToDevice));
        }
        int E = 0;
        auto chunks = plan_chunks(P, nStreamsFP, E);
        const int nS = std::max(1, std::min(nStreamsFP, (int)chunks.size()));
        ensure_slots(nS);
        const ProjGeom g = geom();
        // One check of the CAS volume and the table per call (every stream's
        // launches come after the device synchronisation below).
        dCheck.reserve(sizeof(unsigned));
        check_launch(launch_check_volume(dCasVol.ptr, DT_F32, (size_t)V * V * V, g.ker, g.K, dCheck.as<unsigned>(),
                                         slots[0]->stream),
                     "check_volume");
        for (int i = 0; i < nS; i++) {
            Slot &s = *slots[i];
            s.axes.reserve(sizeof(float) * 9 * E);
            s.casImgs.reserve(sizeof(float) * (size_t)E * M * M);
            if (wantImgs) {
                s.herm.reserve(sizeof(float2) * (size_t)E * M * (M / 2 + 1));
                s.imgs.reserve(sizeof(float) * (size_t)E * N * N);
            }
        }
        HIP_TRY(hipDeviceSynchronize());
        for (const Chunk &c : chunks) {
            Slot &s = *slots[c.slot];
            log("forward projection stream %d: axes %d..%d", c.slot, c.first, c.first + c.count - 1);
            HIP_TRY(hipMemcpyAsync(s.axes.ptr, hAxes + (size_t)9 * c.first, sizeof(float) * 9 * c.count,
                                   hipMemcpyHostToDevice, s.stream));
            check_launch(launch_forward_project(dCasVol.ptr, DT_F32, g, s.axes.as<float>(), c.count,
                                                s.casImgs.as<float>(), s.stream, dCheck.as<unsigned>()),
                         "forward_project");
            if (wantCas)
                HIP_TRY(hipMemcpyAsync(casOut + (size_t)c.first * M * M, s.casImgs.ptr,
                                       sizeof(float) * (size_t)c.count * M * M, hipMemcpyDeviceToHost, s.stream));
            if (wantImgs) {
                chain_cas_to_images(s.casImgs.as<float>(), s.herm.as<float2>(), s.imgs.as<float>(), c.count,
                                    M, N, s.stream);
                HIP_TRY(hipMemcpyAsync(hImgs + (size_t)c.first * N * N, s.imgs.ptr,
                                       sizeof(float) * (size_t)c.count * N * N, hipMemcpyDeviceToHost, s.stream));
            }
        }
        sync_slots(nS);
    }

    // ---- BackProject (gpuGridder.cpp:307-430), optionally with density ----
    void back_project(bool withDensity, bool onesCas)
    {
        DeviceGuard dg(device);
        const int P = numAxes;
        require_axes(P);
        const bool fromImgs = runFFTOnDevice && !rawCas && !onesCas;
        const float *casIn = rawCas ? hCasImgs : hCasImgs;
        if (rawCas && !casIn) raise(MGG_ERR_NOT_SET, "images not set (SetImages)");
        if (fromImgs) {
            if (!hImgs) raise(MGG_ERR_NOT_SET, "images array not set (SetImages)");
            if ((size_t)P * N * N > hImgsLen) raise(MGG_ERR_INVALID_ARGUMENT, "images array too small for %d images", P);
        } else if (!onesCas) {
            if (!casIn) raise(MGG_ERR_NOT_SET, "CAS images not set (SetCASImages)");
            if ((size_t)P * M * M > hCasImgsLen) raise(MGG_ERR_INVALID_ARGUMENT, "CAS-images array too small for %d images", P);
        }
        upload_kernel_table();
        ensure_device_cas_volume();
        int E = 0;
        auto chunks = plan_chunks(P, nStreamsBP, E);
        const int nS = std::max(1, std::min(nStreamsBP, (int)chunks.size()));
        ensure_slots(nS);
        const size_t V3 = (size_t)V * V * V;
        const ProjGeom g = geom();
        for (int i = 0; i < nS; i++) {
            Slot &s = *slots[i];
            s.axes.reserve(sizeof(float) * 9 * E);
            s.casImgs.reserve(sizeof(float) * (size_t)E * M * M);
            if (fromImgs) {
                s.herm.reserve(sizeof(float2) * (size_t)E * M * (M / 2 + 1));
                s.imgs.reserve(sizeof(float) * (size_t)E * N * N);
            }
        }
        float *parts = (float *)dPartials.reserve(sizeof(float) * V3 * nS);
        float *partsDens = withDensity ? (float *)dPartialsDens.reserve(sizeof(float) * V3 * nS) : nullptr;
        HIP_TRY(hipDeviceSynchronize());
        for (int i = 0; i < nS; i++) {
            Slot &s = *slots[i];
            check_launch(launch_fill(parts + V3 * i, 0.f, V3, s.stream), "fill");
            if (withDensity) check_launch(launch_fill(partsDens + V3 * i, 0.f, V3, s.stream), "fill");
        }
        for (const Chunk &c : chunks) {
            Slot &s = *slots[c.slot];
            log("back projection stream %d: axes %d..%d", c.slot, c.first, c.first + c.count - 1);
            HIP_TRY(hipMemcpyAsync(s.axes.ptr, hAxes + (size_t)9 * c.first, sizeof(float) * 9 * c.count,
                                   hipMemcpyHostToDevice, s.stream));
            if (onesCas) {
                check_launch(launch_fill(s.casImgs.as<float>(), 1.f, (size_t)c.count * M * M, s.stream), "fill");
            } else if (fromImgs) {
                HIP_TRY(hipMemcpyAsync(s.imgs.ptr, hImgs + (size_t)c.first * N * N, sizeof(float) * (size_t)c.count * N * N,
                                       hipMemcpyHostToDevice, s.stream));
                chain_images_to_cas(s.imgs.as<float>(), s.casImgs.as<float>(), s.herm.as<float2>(), c.count, M, N,
                                    s.stream);
            } else {
                HIP_TRY(hipMemcpyAsync(s.casImgs.ptr, casIn + (size_t)c.first * M * M,
                                       sizeof(float) * (size_t)c.count * M * M, hipMemcpyHostToDevice, s.stream));
            }
            check_launch(back_project_launch(parts + V3 * c.slot, withDensity ? partsDens + V3 * c.slot : nullptr, g,
                                             s.axes.as<float>(), c.count, s.casImgs.as<float>(), s.stream),
                         "back_project");
        }
        sync_slots(nS);
        // FP64 reduction of the per-stream partial volumes into the device CAS
        // volume (the single-GPU form of MultiGPUGridder::AddCASVolumes,
        // MultiGPUGridder.cpp:376-411).
        hipStream_t s0 = slots[0]->stream;
        check_launch(launch_reduce_partials(dCasVol.as<float>(), parts, nS, V3, V3, 0, s0), "reduce_partials");
        if (withDensity) {
            dDens.reserve(sizeof(float) * V3);
            check_launch(launch_reduce_partials(dDens.as<float>(), partsDens, nS, V3, V3, 0, s0), "reduce_partials");
        }
        HIP_TRY(hipStreamSynchronize(s0));
        casFromVolume = false;
    }

    void cas_volume_to_host()
    {
        DeviceGuard dg(device);
        float *dst = rawCas ? hVol : hCasVol;
        size_t len = rawCas ? hVolLen : hCasVolLen;
        if (!dst) raise(MGG_ERR_NOT_SET, "CAS volume array not set (SetCASVolume)");
        if (len < (size_t)V * V * V) raise(MGG_ERR_INVALID_ARGUMENT, "CAS volume array too small");
        if (!dCasVol.ptr) raise(MGG_ERR_NOT_SET, "no device CAS volume yet");
        HIP_TRY(hipMemcpy(dst, dCasVol.ptr, sizeof(float) * (size_t)V * V * V, hipMemcpyDeviceToHost));
    }

    void ensure_fft3()
    {
        dReal3.reserve(sizeof(float) * (size_t)M * M * M);
        dHerm3.reserve(sizeof(float2) * (size_t)M * M * (M / 2 + 1));
        dVolTmp.reserve(sizeof(float) * (size_t)N * N * N);
    }

    void volume_to_cas()
    {
        if (rawCas) raise(MGG_ERR_UNSUPPORTED, "no spatial volume in the legacy CAS-only mode");
        DeviceGuard dg(device);
        if (!hVol) raise(MGG_ERR_NOT_SET, "volume not set (SetVolume)");
        if (hVolLen < (size_t)N * N * N) raise(MGG_ERR_INVALID_ARGUMENT, "volume array too small");
        ensure_precomp();
        ensure_fft3();
        ensure_device_cas_volume();
        ensure_slots(1);
        hipStream_t s = slots[0]->stream;
        HIP_TRY(hipMemcpyAsync(dVolTmp.ptr, hVol, sizeof(float) * (size_t)N * N * N, hipMemcpyHostToDevice, s));
        chain_volume_to_cas(dVolTmp.as<float>(), N, dCasVol.as<float>(), M, pad, dPrecomp.as<float>(),
                            dReal3.as<float>(), dHerm3.as<float2>(), s);
        if (hCasVol && hCasVolLen >= (size_t)V * V * V)
            HIP_TRY(hipMemcpyAsync(hCasVol, dCasVol.ptr, sizeof(float) * (size_t)V * V * V, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        casFromVolume = true;
    }

    void cas_to_volume_device(const float *cas, hipStream_t s)
    {
        ensure_precomp();
        ensure_fft3();
        chain_cas_to_volume(cas, M, pad, dVolTmp.as<float>(), N, dPrecomp.as<float>(), 1, dReal3.as<float>(),
                            dHerm3.as<float2>(), s);
        HIP_TRY(hipMemcpyAsync(hVol, dVolTmp.ptr, sizeof(float) * (size_t)N * N * N, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
    }

    void get_volume()
    {
        if (rawCas) raise(MGG_ERR_UNSUPPORTED, "no spatial volume in the legacy CAS-only mode");
        DeviceGuard dg(device);
        if (!hVol) raise(MGG_ERR_NOT_SET, "volume not set (SetVolume)");
        if (hVolLen < (size_t)N * N * N) raise(MGG_ERR_INVALID_ARGUMENT, "volume array too small");
        if (!dCasVol.ptr) raise(MGG_ERR_NOT_SET, "no device CAS volume yet");
        ensure_slots(1);
        if (hCasVol && hCasVolLen >= (size_t)V * V * V)
            HIP_TRY(hipMemcpy(hCasVol, dCasVol.ptr, sizeof(float) * (size_t)V * V * V, hipMemcpyDeviceToHost));
        cas_to_volume_device(dCasVol.as<float>(), slots[0]->stream);
    }

    void reconstruct()
    {
        if (rawCas) raise(MGG_ERR_UNSUPPORTED, "no spatial volume in the legacy CAS-only mode");
        if (!hVol) raise(MGG_ERR_NOT_SET, "volume not set (SetVolume)");
        back_project(true, false);
        DeviceGuard dg(device);
        const size_t V3 = (size_t)V * V * V;
        hipStream_t s = slots[0]->stream;
        if (hCasVol && hCasVolLen >= V3)
            HIP_TRY(hipMemcpyAsync(hCasVol, dCasVol.ptr, sizeof(float) * V3, hipMemcpyDeviceToHost, s));
        // The quotient replaces dDens so the BP CAS volume stays on the device.
        check_launch(launch_divide_density(dDens.as<float>(), dCasVol.as<float>(), dDens.as<float>(), V3, 1e-6, s),
                     "divide_density");
        cas_to_volume_device(dDens.as<float>(), s);
    }

    void reset_volume()
    {
        DeviceGuard dg(device);
        if (hVol) std::memset(hVol, 0, sizeof(float) * hVolLen);
        if (hCasVol) std::memset(hCasVol, 0, sizeof(float) * hCasVolLen);
        if (dCasVol.ptr) {
            ensure_slots(1);
            check_launch(launch_fill(dCasVol.as<float>(), 0.f, (size_t)V * V * V, slots[0]->stream), "fill");
            HIP_TRY(hipStreamSynchronize(slots[0]->stream));
        }
        casFromVolume = false;
    }
};

namespace {

size_t dims_product(const int *dims, int n)
{
    size_t p = 1;
    for (int i = 0; i < n; i++) {
        if (dims[i] < 0) raise(MGG_ERR_INVALID_ARGUMENT, "negative dimension %d", dims[i]);
        p *= (size_t)dims[i];
    }
    return p;
}

void need(const void *p, const char *what)
{
    if (!p) raise(MGG_ERR_INVALID_ARGUMENT, "%s is NULL", what);
}

hipStream_t as_stream(void *s) { return (hipStream_t)s; }

}  // namespace

namespace {
// Scratch half-spectrum buffers for the stateless image-chain operators, one
// per (device, stream), grown on demand.
std::mutex g_scratch_mu;
std::map<std::pair<int, hipStream_t>, std::unique_ptr<DevBuf>> g_scratch;

float2 *scratch_herm(size_t bytes, hipStream_t s)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto key = std::make_pair(current_device(), s);
    auto &b = g_scratch[key];
    if (!b) b = std::make_unique<DevBuf>();
    if (b->bytes < bytes) {
        HIP_TRY(hipStreamSynchronize(s));  // previous user of the buffer on this stream
        b->reserve(bytes);
    }
    return b->as<float2>();
}
}  // namespace

namespace {
struct Vol3Scratch {
    DevBuf real, herm, precomp;
    int m = 0, hw = 0;
};
std::mutex g_vol_mu;
std::map<std::pair<int, hipStream_t>, std::unique_ptr<Vol3Scratch>> g_vol_scratch;

Vol3Scratch &vol_scratch(int m, int hw, hipStream_t s)
{
    std::lock_guard<std::mutex> lk(g_vol_mu);
    auto key = std::make_pair(current_device(), s);
    auto &p = g_vol_scratch[key];
    if (!p) p = std::make_unique<Vol3Scratch>();
    if (p->m != m || p->hw != hw) {
        HIP_TRY(hipStreamSynchronize(s));
        p->real.reserve(sizeof(float) * (size_t)m * m * m);
        p->herm.reserve(sizeof(float2) * (size_t)m * m * (m / 2 + 1));
        std::vector<double> pc;
        precomp_host(pc, m, hw);
        std::vector<float> pf(pc.begin(), pc.end());
        p->precomp.reserve(sizeof(float) * m);
        HIP_TRY(hipMemcpy(p->precomp.ptr, pf.data(), sizeof(float) * m, hipMemcpyHostToDevice));
        p->m = m;
        p->hw = hw;
    }
    return *p;
}
}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char *mgg_last_error(void) { return g_last_error.c_str(); }

const char *mgg_version(void) { return "multigpugridder-amd 0.1 (gfx950)"; }

int mgg_device_count(int *count)
{
    return guarded([&] {
        need(count, "count");
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
        *count = n;
    });
}

int mgg_create(int volume_size, int num_coord_axes, float interp_factor, int extra_padding, int num_gpus,
               const int *gpu_devices, int run_fft_on_device, int verbose, mgg_gridder **out)
{
    return guarded([&] {
        need(out, "out");
        *out = nullptr;
        if (volume_size <= 0) raise(MGG_ERR_INVALID_ARGUMENT, "volume size must be positive");
        if (num_coord_axes < 0) raise(MGG_ERR_INVALID_ARGUMENT, "number of axes must be >= 0");
        if (!(interp_factor > 0.f)) raise(MGG_ERR_INVALID_ARGUMENT, "interpFactor must be positive");
        if (extra_padding < 0) raise(MGG_ERR_INVALID_ARGUMENT, "extraPadding must be >= 0");
        if (num_gpus != 1)
            raise(MGG_ERR_UNSUPPORTED, "this build drives one GPU per gridder (asked for %d)", num_gpus);
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) raise(MGG_ERR_DEVICE, "no GPU visible");
        const int dev = gpu_devices ? gpu_devices[0] : 0;
        if (dev < 0 || dev >= ndev) raise(MGG_ERR_DEVICE, "GPU id %d out of range [0, %d)", dev, ndev);
        auto g = std::make_unique<mgg_gridder>();
        g->N = volume_size;
        g->interp = interp_factor;
        // gpuGridder.cpp:73-74: M = N * interpFactor, V = M + 2 * extraPadding.
        g->M = (int)((float)volume_size * interp_factor);
        g->pad = extra_padding;
        g->V = g->M + 2 * extra_padding;
        if (g->M % 2) raise(MGG_ERR_INVALID_ARGUMENT, "CAS image size N*interpFactor = %d must be even", g->M);
        if (g->M < g->N) raise(MGG_ERR_INVALID_ARGUMENT, "interpFactor must be >= 1");
        if ((g->M - g->N) % 2) raise(MGG_ERR_INVALID_ARGUMENT, "N*interpFactor - N must be even");
        g->numAxes = num_coord_axes;
        // AbstractGridder.cpp:13: maskRadius = (N * interpFactor) / 2 - 1.
        g->maskRadius = ((float)volume_size * interp_factor) / 2 - 1;
        g->runFFTOnDevice = run_fft_on_device ? 1 : 0;
        g->verbose = verbose ? 1 : 0;
        g->device = dev;
        *out = g.release();
    });
}

int mgg_destroy(mgg_gridder *g)
{
    return guarded([&] { delete g; });
}

#define WITH_G(body) return guarded([&] { need(g, "gridder"); body; })

int mgg_set_volume(mgg_gridder *g, float *volume, const int *dims)
{
    WITH_G({
        need(volume, "volume");
        need(dims, "dims");
        g->hVol = volume;
        g->hVolLen = dims_product(dims, 3);
    });
}

int mgg_set_coord_axes(mgg_gridder *g, float *coord_axes, const int *dims)
{
    WITH_G({
        need(coord_axes, "coord_axes");
        need(dims, "dims");
        g->hAxes = coord_axes;
        g->hAxesLen = dims_product(dims, 2);
    });
}

int mgg_set_images(mgg_gridder *g, float *images, const int *dims)
{
    WITH_G({
        need(images, "images");
        need(dims, "dims");
        g->hImgs = images;
        g->hImgsLen = dims_product(dims, 3);
    });
}

int mgg_set_cas_images(mgg_gridder *g, float *cas_images, const int *dims)
{
    WITH_G({
        need(cas_images, "cas_images");
        need(dims, "dims");
        g->hCasImgs = cas_images;
        g->hCasImgsLen = dims_product(dims, 3);
    });
}

int mgg_set_cas_volume(mgg_gridder *g, float *cas_volume, const int *dims)
{
    WITH_G({
        need(cas_volume, "cas_volume");
        need(dims, "dims");
        g->hCasVol = cas_volume;
        g->hCasVolLen = dims_product(dims, 3);
        g->casFromVolume = false;
    });
}

int mgg_set_kb_table(mgg_gridder *g, float *table, const int *dims)
{
    WITH_G({
        need(table, "table");
        need(dims, "dims");
        if (dims[0] < 2 || dims[0] > kMaxKerSize)
            raise(MGG_ERR_INVALID_ARGUMENT, "kernel table size %d outside [2, %d]", dims[0], kMaxKerSize);
        g->hKer = table;
        g->hKerLen = dims[0];
    });
}

int mgg_set_mask_radius(mgg_gridder *g, float r)
{
    WITH_G({
        if (!(r >= 0.f)) raise(MGG_ERR_INVALID_ARGUMENT, "mask radius must be >= 0");
        g->maskRadius = r;
    });
}

int mgg_set_num_axes(mgg_gridder *g, int n)
{
    WITH_G({
        if (n < 0) raise(MGG_ERR_INVALID_ARGUMENT, "number of axes must be >= 0");
        g->numAxes = n;
    });
}

int mgg_set_max_axes_to_allocate(mgg_gridder *g, int n)
{
    WITH_G({
        if (n <= 0) raise(MGG_ERR_INVALID_ARGUMENT, "MaxAxesToAllocate must be a positive integer");
        g->maxAxes = n;
    });
}

int mgg_set_num_streams_fp(mgg_gridder *g, int n)
{
    WITH_G({
        if (n <= 0 || n > 64) raise(MGG_ERR_INVALID_ARGUMENT, "number of streams %d outside [1, 64]", n);
        g->nStreamsFP = n;
    });
}

int mgg_set_num_streams_bp(mgg_gridder *g, int n)
{
    WITH_G({
        if (n <= 0 || n > 64) raise(MGG_ERR_INVALID_ARGUMENT, "number of streams %d outside [1, 64]", n);
        g->nStreamsBP = n;
    });
}

int mgg_set_gpus(mgg_gridder *g, const int *gpu_devices, int num_gpus)
{
    WITH_G({
        need(gpu_devices, "gpu_devices");
        if (num_gpus != 1) raise(MGG_ERR_UNSUPPORTED, "this build drives one GPU per gridder (asked for %d)", num_gpus);
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess) ndev = 0;
        if (gpu_devices[0] < 0 || gpu_devices[0] >= ndev) raise(MGG_ERR_DEVICE, "GPU id %d out of range", gpu_devices[0]);
        if (!g->slots.empty() || g->dCasVol.ptr)
            raise(MGG_ERR_INVALID_ARGUMENT, "GPUs can only be changed before the first projection");
        g->device = gpu_devices[0];
    });
}

int mgg_set_ker_hwidth(mgg_gridder *g, float hwidth)
{
    WITH_G({
        if (!(hwidth > 0.f) || hwidth > 5.f) raise(MGG_ERR_INVALID_ARGUMENT, "kernel half width must be in (0, 5]");
        g->kerHWidth = hwidth;
        g->defaultKer.clear();
        g->dPrecomp.release();
    });
}

int mgg_get_sizes(const mgg_gridder *g, int *volume_size, int *cas_image_size, int *cas_volume_size,
                  int *num_axes)
{
    return guarded([&] {
        need(g, "gridder");
        if (volume_size) *volume_size = g->N;
        if (cas_image_size) *cas_image_size = g->M;
        if (cas_volume_size) *cas_volume_size = g->V;
        if (num_axes) *num_axes = g->numAxes;
    });
}

float mgg_get_mask_radius(const mgg_gridder *g) { return g ? g->maskRadius : -1.f; }

int mgg_forward_project(mgg_gridder *g) { WITH_G(g->forward_project()); }
int mgg_back_project(mgg_gridder *g) { WITH_G(g->back_project(false, false)); }
int mgg_cas_volume_to_host(mgg_gridder *g) { WITH_G(g->cas_volume_to_host()); }
int mgg_volume_to_cas_volume(mgg_gridder *g) { WITH_G(g->volume_to_cas()); }
int mgg_get_volume(mgg_gridder *g) { WITH_G(g->get_volume()); }
int mgg_reconstruct_volume(mgg_gridder *g) { WITH_G(g->reconstruct()); }
int mgg_reset_volume(mgg_gridder *g) { WITH_G(g->reset_volume()); }

// ---------------------------------------------------------------------------
// Device-pointer operators
// ---------------------------------------------------------------------------
int mgg_op_forward_project_checked(const void *cas_volume, int cas_dtype, int v, const float *axes, int n_axes,
                                   float *cas_images, int m, float mask_radius, const float *ker, int ker_size,
                                   float ker_hwidth, const unsigned *check, void *stream)
{
    return guarded([&] {
        if (n_axes == 0) return;
        need(cas_volume, "cas_volume"); need(axes, "axes"); need(cas_images, "cas_images"); need(ker, "ker");
        validate_geom(v, m, ker_size, ker_hwidth);
        if (n_axes < 0) raise(MGG_ERR_INVALID_ARGUMENT, "n_axes < 0");
        if (cas_dtype != MGG_DTYPE_F32 && cas_dtype != MGG_DTYPE_BF16 && cas_dtype != MGG_DTYPE_F16)
            raise(MGG_ERR_INVALID_ARGUMENT, "unknown CAS volume dtype %d", cas_dtype);
        ProjGeom g{v, m, mask_radius, ker, ker_size, ker_hwidth};
        check_launch(launch_forward_project(cas_volume, cas_dtype, g, axes, n_axes, cas_images, as_stream(stream),
                                            check),
                     "forward_project");
    });
}

int mgg_op_forward_project(const void *cas_volume, int cas_dtype, int v, const float *axes, int n_axes,
                           float *cas_images, int m, float mask_radius, const float *ker, int ker_size,
                           float ker_hwidth, void *stream)
{
    return mgg_op_forward_project_checked(cas_volume, cas_dtype, v, axes, n_axes, cas_images, m, mask_radius, ker,
                                          ker_size, ker_hwidth, nullptr, stream);
}

int mgg_op_check_volume(const void *cas_volume, int cas_dtype, int v, const float *ker, int ker_size,
                        unsigned *check, void *stream)
{
    return guarded([&] {
        need(cas_volume, "cas_volume"); need(ker, "ker"); need(check, "check");
        if (v <= 0) raise(MGG_ERR_INVALID_ARGUMENT, "volume size must be positive, got %d", v);
        if (ker_size <= 0 || ker_size > kMaxKerSize)
            raise(MGG_ERR_INVALID_ARGUMENT, "kernel table size %d outside 1..%d", ker_size, kMaxKerSize);
        if (cas_dtype != MGG_DTYPE_F32 && cas_dtype != MGG_DTYPE_BF16 && cas_dtype != MGG_DTYPE_F16)
            raise(MGG_ERR_INVALID_ARGUMENT, "unknown CAS volume dtype %d", cas_dtype);
        check_launch(launch_check_volume(cas_volume, cas_dtype, (size_t)v * v * v, ker, ker_size, check,
                                         as_stream(stream)),
                     "check_volume");
    });
}

int mgg_op_back_project(float *cas_volume, float *density, int v, const float *axes, int n_axes,
                        const float *cas_images, int m, float mask_radius, const float *ker, int ker_size,
                        float ker_hwidth, void *stream)
{
    return guarded([&] {
        if (n_axes == 0) return;
        need(cas_volume, "cas_volume"); need(axes, "axes"); need(cas_images, "cas_images"); need(ker, "ker");
        validate_geom(v, m, ker_size, ker_hwidth);
        if (n_axes < 0) raise(MGG_ERR_INVALID_ARGUMENT, "n_axes < 0");
        ProjGeom g{v, m, mask_radius, ker, ker_size, ker_hwidth};
        check_launch(back_project_launch(cas_volume, density, g, axes, n_axes, cas_images, as_stream(stream)),
                     "back_project");
    });
}

int mgg_op_back_project_adjoint_checked(float *cas_volume, int v, const float *axes, int n_axes,
                                        const float *cas_images, int m, float mask_radius, const float *ker,
                                        int ker_size, float ker_hwidth, const unsigned *check, void *stream)
{
    return guarded([&] {
        if (n_axes == 0) return;
        need(cas_volume, "cas_volume"); need(axes, "axes"); need(cas_images, "cas_images"); need(ker, "ker");
        validate_geom(v, m, ker_size, ker_hwidth);
        if (n_axes < 0) raise(MGG_ERR_INVALID_ARGUMENT, "n_axes < 0");
        ProjGeom g{v, m, mask_radius, ker, ker_size, ker_hwidth};
        check_launch(launch_back_project_adjoint(cas_volume, g, axes, n_axes, cas_images, as_stream(stream),
                                                 check),
                     "back_project_adjoint");
    });
}

int mgg_op_back_project_adjoint(float *cas_volume, int v, const float *axes, int n_axes,
                                const float *cas_images, int m, float mask_radius, const float *ker,
                                int ker_size, float ker_hwidth, void *stream)
{
    return mgg_op_back_project_adjoint_checked(cas_volume, v, axes, n_axes, cas_images, m, mask_radius, ker,
                                               ker_size, ker_hwidth, nullptr, stream);
}

int mgg_op_check_images(const float *cas_images, int n_images, int m, const float *ker, int ker_size,
                        unsigned *check, void *stream)
{
    return guarded([&] {
        need(ker, "ker"); need(check, "check");
        if (n_images < 0 || m <= 0) raise(MGG_ERR_INVALID_ARGUMENT, "need n_images >= 0 and m > 0");
        if (n_images > 0) need(cas_images, "cas_images");
        if (ker_size <= 0 || ker_size > kMaxKerSize)
            raise(MGG_ERR_INVALID_ARGUMENT, "kernel table size %d outside 1..%d", ker_size, kMaxKerSize);
        check_launch(launch_check_volume(cas_images, DT_F32, (size_t)n_images * m * m, ker, ker_size, check,
                                         as_stream(stream)),
                     "check_images");
    });
}


int mgg_op_cas_images_to_images(float *cas_images, float *images, int n, int m, int nn, void *stream)
{
    return guarded([&] {
        if (n == 0) return;
        need(cas_images, "cas_images"); need(images, "images");
        if (n < 0 || m <= 0 || nn <= 0 || nn > m || (m % 2) || ((m - nn) % 2))
            raise(MGG_ERR_INVALID_ARGUMENT, "bad sizes n=%d m=%d nn=%d", n, m, nn);
        hipStream_t s = as_stream(stream);
        float2 *herm = scratch_herm(sizeof(float2) * (size_t)std::min<size_t>(n, kFftBatch) * m * (m / 2 + 1), s);
        for (int a = 0; a < n; a += (int)kFftBatch) {
            const int b = std::min<int>((int)kFftBatch, n - a);
            chain_cas_to_images(cas_images + (size_t)a * m * m, herm, images + (size_t)a * nn * nn, b, m, nn, s);
        }
    });
}

// ---------------------------------------------------------------------------
// Allocation-free CAS images -> images.  For callers that cannot allocate
// device memory (code under HIP graph capture, serving loops, harnesses that
// forbid hipMalloc).  The fused chains' half-spectrum intermediate lives in a
// fixed device array, `scratch`, of kNoAllocScratchBytes per GPU (every GPU
// the library runs on holds it until the process ends); its device address is
// taken once per device with hipGetSymbolAddress and only that address is
// handed to kernels.  Every call on one GPU shares the array: calls that
// overlap on one GPU (different streams or threads) race.  Each chunk's
// intermediate is written whole by the first kernel before the second reads
// it, so nothing needs clearing between calls.
// ---------------------------------------------------------------------------
namespace {
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;
static __device__ __align__(256) float2 scratch[kNoAllocScratchBytes / sizeof(float2)];

constexpr int kNoAllocMaxDevices = 64;
std::mutex g_noalloc_mu;
float2 *g_noalloc_ptr[kNoAllocMaxDevices] = {};

// Device address of this GPU's `scratch` (the current device must be the one
// the caller's stream belongs to).
float2 *noalloc_scratch()
{
    const int dev = current_device();
    if (dev < 0 || dev >= kNoAllocMaxDevices)
        raise(MGG_ERR_UNSUPPORTED, "device %d beyond the %d devices the allocation-free path tracks", dev,
              kNoAllocMaxDevices);
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            raise(MGG_ERR_HIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev, hipGetErrorString(e));
        g_noalloc_ptr[dev] = static_cast<float2 *>(ptr);
    }
    return g_noalloc_ptr[dev];
}
}  // namespace

int mgg_op_cas_images_to_images_noalloc(const float *cas_images, float *images, int n, int m, int nn,
                                        void *stream)
{
    return guarded([&] {
        if (n == 0) return;
        need(cas_images, "cas_images"); need(images, "images");
        if (n < 0 || m <= 0 || nn <= 0 || nn > m || (m % 2) || ((m - nn) % 2))
            raise(MGG_ERR_INVALID_ARGUMENT, "bad sizes n=%d m=%d nn=%d", n, m, nn);
        if (!fused_chain_supported(m, nn))
            raise(MGG_ERR_UNSUPPORTED,
                  "the allocation-free image chain runs the fused chains only (m = 128, 192, 256, 384, 512); "
                  "m = %d needs rocFFT plans, which allocate", m);
        const size_t perImage = sizeof(float2) * (size_t)nn * (size_t)(m / 2 + 1);
        if (perImage > kNoAllocScratchBytes)
            raise(MGG_ERR_INVALID_ARGUMENT, "one image needs %zu scratch bytes, the fixed scratch holds %zu",
                  perImage, kNoAllocScratchBytes);
        const int chunk = (int)std::min<size_t>(kNoAllocScratchBytes / perImage, 65535);
        float2 *tmp = noalloc_scratch();
        hipStream_t s = as_stream(stream);
        const float scale = 1.0f / ((float)m * (float)m);
        for (int a = 0; a < n; a += chunk) {
            const int b = std::min(chunk, n - a);
            check_launch(launch_fused_cas_to_images(cas_images + (size_t)a * m * m, tmp, images + (size_t)a * nn * nn,
                                                    m, nn, b, scale, s),
                         "fused_cas_to_images (allocation-free)");
        }
    });
}

int mgg_op_images_to_cas_images(const float *images, float *cas_images, int n, int m, int nn, void *stream)
{
    return guarded([&] {
        if (n == 0) return;
        need(cas_images, "cas_images"); need(images, "images");
        if (n < 0 || m <= 0 || nn <= 0 || nn > m || (m % 2) || ((m - nn) % 2))
            raise(MGG_ERR_INVALID_ARGUMENT, "bad sizes n=%d m=%d nn=%d", n, m, nn);
        hipStream_t s = as_stream(stream);
        float2 *herm = scratch_herm(sizeof(float2) * (size_t)std::min<size_t>(n, kFftBatch) * m * (m / 2 + 1), s);
        for (int a = 0; a < n; a += (int)kFftBatch) {
            const int b = std::min<int>((int)kFftBatch, n - a);
            chain_images_to_cas(images + (size_t)a * nn * nn, cas_images + (size_t)a * m * m, herm, b, m, nn, s);
        }
    });
}


int mgg_op_cas_images_to_images_transpose(const float *images, float *cas_images, int n, int m, int nn,
                                          void *stream)
{
    return guarded([&] {
        if (n == 0) return;
        need(cas_images, "cas_images"); need(images, "images");
        if (n < 0 || m <= 0 || nn <= 0 || nn > m || (m % 2) || ((m - nn) % 2))
            raise(MGG_ERR_INVALID_ARGUMENT, "bad sizes n=%d m=%d nn=%d", n, m, nn);
        hipStream_t s = as_stream(stream);
        float2 *herm = scratch_herm(sizeof(float2) * (size_t)std::min<size_t>(n, kFftBatch) * m * (m / 2 + 1), s);
        const float scale = (float)(1.0 / ((double)m * (double)m));
        for (int a = 0; a < n; a += (int)kFftBatch) {
            const int b = std::min<int>((int)kFftBatch, n - a);
            chain_images_to_cas(images + (size_t)a * nn * nn, cas_images + (size_t)a * m * m, herm, b, m, nn, s,
                                scale, 1);
        }
    });
}

int mgg_op_volume_to_cas_volume_transpose(const float *cas_volume, int m, int pad, float *volume, int nn,
                                          float ker_hwidth, void *stream)
{
    return guarded([&] {
        need(volume, "volume"); need(cas_volume, "cas_volume");
        if (nn <= 0 || m < nn || (m % 2) || ((m - nn) % 2) || pad < 0)
            raise(MGG_ERR_INVALID_ARGUMENT, "bad sizes nn=%d m=%d pad=%d", nn, m, pad);
        hipStream_t s = as_stream(stream);
        Vol3Scratch &sc = vol_scratch(m, (int)lroundf(ker_hwidth), s);
        chain_cas_to_volume(cas_volume, m, pad, volume, nn, sc.precomp.as<float>(), 0, sc.real.as<float>(),
                            sc.herm.as<float2>(), s, true);
    });
}

int mgg_op_volume_to_cas_volume(const float *volume, int nn, float *cas_volume, int m, int pad, float ker_hwidth,
                                void *stream)
{
    return guarded([&] {
        need(volume, "volume"); need(cas_volume, "cas_volume");
        if (nn <= 0 || m < nn || (m % 2) || ((m - nn) % 2) || pad < 0)
            raise(MGG_ERR_INVALID_ARGUMENT, "bad sizes nn=%d m=%d pad=%d", nn, m, pad);
        hipStream_t s = as_stream(stream);
        Vol3Scratch &sc = vol_scratch(m, (int)lroundf(ker_hwidth), s);
        chain_volume_to_cas(volume, nn, cas_volume, m, pad, sc.precomp.as<float>(), sc.real.as<float>(),
                            sc.herm.as<float2>(), s);
    });
}

int mgg_op_cas_volume_to_volume(const float *cas_volume, int m, int pad, float *volume, int nn, float ker_hwidth,
                                int zero_nyquist, void *stream)
{
    return guarded([&] {
        need(volume, "volume"); need(cas_volume, "cas_volume");
        if (nn <= 0 || m < nn || (m % 2) || ((m - nn) % 2) || pad < 0)
            raise(MGG_ERR_INVALID_ARGUMENT, "bad sizes nn=%d m=%d pad=%d", nn, m, pad);
        hipStream_t s = as_stream(stream);
        Vol3Scratch &sc = vol_scratch(m, (int)lroundf(ker_hwidth), s);
        chain_cas_to_volume(cas_volume, m, pad, volume, nn, sc.precomp.as<float>(), zero_nyquist ? 1 : 0,
                            sc.real.as<float>(), sc.herm.as<float2>(), s);
    });
}

int mgg_op_divide_density(float *out, const float *num, const float *den, size_t n, double eps, void *stream)
{
    return guarded([&] {
        need(out, "out"); need(num, "num"); need(den, "den");
        check_launch(launch_divide_density(out, num, den, n, eps, as_stream(stream)), "divide_density");
    });
}

int mgg_op_reduce_partials(float *out, const float *parts, int n_parts, size_t n, size_t stride, int accumulate,
                           void *stream)
{
    return guarded([&] {
        need(out, "out"); need(parts, "parts");
        if (n_parts <= 0) raise(MGG_ERR_INVALID_ARGUMENT, "n_parts must be positive");
        check_launch(launch_reduce_partials(out, parts, n_parts, n, stride, accumulate, as_stream(stream)),
                     "reduce_partials");
    });
}

int mgg_op_convert_volume(void *dst, int dtype, const float *src, size_t n, void *stream)
{
    return guarded([&] {
        need(dst, "dst"); need(src, "src");
        if (dtype != MGG_DTYPE_BF16 && dtype != MGG_DTYPE_F16)
            raise(MGG_ERR_INVALID_ARGUMENT, "dtype must be MGG_DTYPE_BF16 or MGG_DTYPE_F16");
        check_launch(launch_convert_volume(dst, dtype, src, n, as_stream(stream)), "convert_volume");
    });
}

#define FILTER(call, what)                                          \
    return guarded([&] { check_launch((call), what); })

int mgg_filter_cas_to_complex(const float *in, float *out_complex, int size, int n_slices, void *stream)
{
    FILTER(launch_filter_cas_to_complex(in, (float2 *)out_complex, size, n_slices, as_stream(stream)), "cas_to_complex");
}
int mgg_filter_complex_to_cas(const float *in_complex, float *out, int size, int n_slices, void *stream)
{
    FILTER(launch_filter_complex_to_cas((const float2 *)in_complex, out, size, n_slices, as_stream(stream)),
           "complex_to_cas");
}
int mgg_filter_complex_to_real(const float *in_complex, float *out, int size, int n_slices, void *stream)
{
    FILTER(launch_filter_complex_to_real((const float2 *)in_complex, out, size, n_slices, as_stream(stream)),
           "complex_to_real");
}
int mgg_filter_real_to_complex(const float *in, float *out_complex, int size, int n_slices, void *stream)
{
    FILTER(launch_filter_real_to_complex(in, (float2 *)out_complex, size, n_slices, as_stream(stream)),
           "real_to_complex");
}
int mgg_filter_fftshift2d(void *data, int elem_bytes, int size, int n_slices, void *stream)
{
    FILTER(launch_filter_fftshift2d(data, elem_bytes, size, n_slices, as_stream(stream)), "fftshift2d");
}
int mgg_filter_fftshift3d(void *data, int elem_bytes, int size, void *stream)
{
    FILTER(launch_filter_fftshift3d(data, elem_bytes, size, as_stream(stream)), "fftshift3d");
}
int mgg_filter_pad(const float *in, float *out, int in_size, int out_size, int n_slices, void *stream)
{
    FILTER(launch_filter_pad(in, out, in_size, out_size, n_slices, as_stream(stream)), "pad");
}
int mgg_filter_crop(const void *in, void *out, int elem_bytes, int in_size, int out_size, int n_slices,
                    void *stream)
{
    FILTER(launch_filter_crop(in, out, elem_bytes, in_size, out_size, n_slices, as_stream(stream)), "crop");
}
int mgg_filter_divide_scalar(float *data, float scalar, size_t n, void *stream)
{
    return guarded([&] {
        if (scalar == 0.f) raise(MGG_ERR_INVALID_ARGUMENT, "divide by zero (DivideScalarFilter.cu:36-39)");
        check_launch(launch_filter_divide_scalar(data, scalar, n, as_stream(stream)), "divide_scalar");
    });
}
int mgg_filter_divide_volume(float *a, const float *b, size_t n, void *stream)
{
    FILTER(launch_filter_divide_volume(a, b, n, as_stream(stream)), "divide_volume");
}
int mgg_filter_multiply_volume(float *a, const float *b, size_t n, void *stream)
{
    FILTER(launch_filter_multiply_volume(a, b, n, as_stream(stream)), "multiply_volume");
}
int mgg_filter_multiply_complex_volume(float *a_complex, const float *b, size_t n, void *stream)
{
    FILTER(launch_filter_multiply_complex_volume((float2 *)a_complex, b, n, as_stream(stream)),
           "multiply_complex_volume");
}
int mgg_filter_add_volume(float *a, const float *b, size_t n, void *stream)
{
    FILTER(launch_filter_add_volume(a, b, n, as_stream(stream)), "add_volume");
}

// ---------------------------------------------------------------------------
// Host math
// ---------------------------------------------------------------------------
int mgg_kb_table(float *out, int size, int hwidth)
{
    return guarded([&] {
        need(out, "out");
        if (size < 3) raise(MGG_ERR_INVALID_ARGUMENT, "table size must be >= 3");
        if (kernel_alpha(hwidth) == 0.0) raise(MGG_ERR_INVALID_ARGUMENT, "invalid kernel size %d (kernelAlpha.m)", hwidth);
        std::vector<double> w;
        kb_table_host(w, size, hwidth);
        for (int k = 0; k < size; k++) out[k] = (float)w[k];
    });
}

int mgg_precomp(float *out, int m, int hwidth)
{
    return guarded([&] {
        need(out, "out");
        if (m <= 0 || (m % 2)) raise(MGG_ERR_INVALID_ARGUMENT, "m must be positive and even");
        if (kernel_alpha(hwidth) == 0.0) raise(MGG_ERR_INVALID_ARGUMENT, "invalid kernel size %d", hwidth);
        std::vector<double> p;
        precomp_host(p, m, hwidth);
        for (int i = 0; i < m; i++) out[i] = (float)p[i];
    });
}

int mgg_uniform_axes(float *out, int ntheta, int nphi)
{
    return guarded([&] {
        need(out, "out");
        if (ntheta <= 0 || nphi <= 0) raise(MGG_ERR_INVALID_ARGUMENT, "ntheta and nphi must be positive");
        const double al = M_PI / 180.0;
        size_t a = 0;
        for (int kt = 1; kt <= ntheta; kt++) {
            const double theta = kt * (90.0 / ntheta);
            for (int kp = 0; kp < nphi; kp++) {
                const double phi = kp * (360.0 / nphi);
                const double n[3] = {std::sin(theta * al) * std::cos(phi * al), std::sin(theta * al) * std::sin(phi * al),
                                     std::cos(theta * al)};
                axes_from_normal(n, 30 * al, out + 9 * a);
                a++;
            }
        }
    });
}

// ---------------------------------------------------------------------------
// Legacy ctypes symbols (Python/MultiGPUGridder_Python.py:11-48).  The legacy
// wrapper works on CAS arrays directly: SetVolume hands over a CAS volume,
// Forward_Project writes CAS images into a library-owned array (GetImages),
// SetImages hands over CAS images and Back_Project writes the CAS volume into
// the array given to SetVolume (GetVolume).
// ---------------------------------------------------------------------------
static mgg_gridder *legacy(void *h) { return static_cast<mgg_gridder *>(h); }

void *Gridder_new(void)
{
    mgg_gridder *g = nullptr;
    int rc = guarded([&] {
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) raise(MGG_ERR_DEVICE, "no GPU visible");
        auto p = std::make_unique<mgg_gridder>();
        p->rawCas = true;
        p->runFFTOnDevice = 0;
        p->pad = 0;
        g = p.release();
    });
    return rc == MGG_OK ? g : nullptr;
}

void *SetNumberGPUs(void *h, int n)
{
    int rc = guarded([&] {
        need(h, "handle");
        if (n != 1) raise(MGG_ERR_UNSUPPORTED, "this build drives one GPU per gridder");
    });
    return rc == MGG_OK ? h : nullptr;
}

void *SetNumberStreams(void *h, int n)
{
    if (mgg_set_num_streams_fp(legacy(h), n) != MGG_OK) return nullptr;
    if (mgg_set_num_streams_bp(legacy(h), n) != MGG_OK) return nullptr;
    return h;
}

void *SetVolume(void *h, float *volume, int *dims)
{
    int rc = guarded([&] {
        need(h, "handle"); need(volume, "volume"); need(dims, "dims");
        mgg_gridder *g = legacy(h);
        if (dims[0] <= 0 || dims[0] != dims[1] || dims[0] != dims[2])
            raise(MGG_ERR_INVALID_ARGUMENT, "CAS volume must be a cube");
        g->hVol = volume;
        g->hVolLen = dims_product(dims, 3);
        g->V = dims[0];
        if (g->M == 0) g->M = g->V;
        if (g->maskRadius == 0.f) g->maskRadius = g->M / 2 - 1;
    });
    return rc == MGG_OK ? h : nullptr;
}

float *GetVolume(void *h) { return h ? legacy(h)->hVol : nullptr; }

void *ResetVolume(void *h)
{
    return mgg_reset_volume(legacy(h)) == MGG_OK ? h : nullptr;
}

void *SetImages(void *h, float *images)
{
    int rc = guarded([&] {
        need(h, "handle"); need(images, "images");
        mgg_gridder *g = legacy(h);
        g->hCasImgs = images;
        g->hCasImgsLen = (size_t)g->numAxes * g->M * g->M;
    });
    return rc == MGG_OK ? h : nullptr;
}

float *GetImages(void *h) { return h ? legacy(h)->legacyImgs.data() : nullptr; }

void *SetAxes(void *h, float *axes, int *dims)
{
    int rc = guarded([&] {
        need(h, "handle"); need(axes, "axes"); need(dims, "dims");
        mgg_gridder *g = legacy(h);
        g->hAxes = axes;
        g->hAxesLen = (size_t)dims[0];
        g->numAxes = dims[0] / 9;
    });
    return rc == MGG_OK ? h : nullptr;
}

void *SetImgSize(void *h, int *dims)
{
    int rc = guarded([&] {
        need(h, "handle"); need(dims, "dims");
        mgg_gridder *g = legacy(h);
        if (dims[0] <= 0 || dims[0] % 2) raise(MGG_ERR_INVALID_ARGUMENT, "image size must be positive and even");
        g->M = dims[0];
        g->legacyImgSize = dims[0];
    });
    return rc == MGG_OK ? h : nullptr;
}

void *SetMaskRadius(void *h, float *radius)
{
    int rc = guarded([&] {
        need(h, "handle"); need(radius, "radius");
        legacy(h)->maskRadius = *radius;
    });
    return rc == MGG_OK ? h : nullptr;
}

void *Forward_Project(void *h)
{
    return mgg_forward_project(legacy(h)) == MGG_OK ? h : nullptr;
}

void *Back_Project(void *h)
{
    int rc = guarded([&] {
        need(h, "handle");
        mgg_gridder *g = legacy(h);
        g->back_project(false, false);
        g->cas_volume_to_host();
    });
    return rc == MGG_OK ? h : nullptr;
}

}  // extern "C"
