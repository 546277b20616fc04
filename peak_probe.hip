// Measured peaks of the three resources the projection kernels use, on the
// GPU it runs on: the FP32 VALU issue rate (non-packed FMA, wave64
// instructions per CU per ns), the rate of wave64 LDS dword reads at random
// addresses (the back projection's kernel-table lookups), and device-memory
// copy bandwidth.  Prints one JSON line.  Used by bench.py to put the
// projectors' measured instruction rates beside the issue peak they are bound
// by (the HBM roofline alone says little about a gather kernel).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CHECK(x)                                                                        \
    do {                                                                                \
        hipError_t e_ = (x);                                                            \
        if (e_ != hipSuccess) {                                                         \
            std::fprintf(stderr, "peak_probe: %s: %s\n", #x, hipGetErrorString(e_));    \
            std::exit(1);                                                               \
        }                                                                               \
    } while (0)

namespace {

constexpr int kFmaChains = 8;

// kFmaChains independent FMA chains per lane: issue-bound, not latency-bound.
__global__ __launch_bounds__(256) void fma_chains(float *out, float a, float b, int iters)
{
    float x[kFmaChains];
#pragma unroll
    for (int c = 0; c < kFmaChains; c++) x[c] = threadIdx.x * 1e-3f + c;
    for (int i = 0; i < iters; i++) {
#pragma unroll
        for (int c = 0; c < kFmaChains; c++) x[c] = fmaf(x[c], a, b);
    }
    float s = 0.f;
#pragma unroll
    for (int c = 0; c < kFmaChains; c++) s += x[c];
    out[blockIdx.x * 256 + threadIdx.x] = s;
}

// Same with two-operand adds (VOP2): the issue rate of simple VALU ops, which
// a three-source FMA can fall short of.
__global__ __launch_bounds__(256) void add_chains(float *out, float b, int iters)
{
    float x[kFmaChains];
#pragma unroll
    for (int c = 0; c < kFmaChains; c++) x[c] = threadIdx.x * 1e-3f + c;
    for (int i = 0; i < iters; i++) {
#pragma unroll
        for (int c = 0; c < kFmaChains; c++) x[c] = x[c] + b;
    }
    float s = 0.f;
#pragma unroll
    for (int c = 0; c < kFmaChains; c++) s += x[c];
    out[blockIdx.x * 256 + threadIdx.x] = s;
}

// Eight independent LDS dword reads per iteration at per-lane pseudo-random
// addresses inside a 1024-entry table (two VALU per read for the address).
constexpr int kLdsReads = 8;
__global__ __launch_bounds__(256) void lds_random(float *out, const unsigned *seed, int iters)
{
    __shared__ float t[1024];
    for (int i = threadIdx.x; i < 1024; i += 256) t[i] = (float)i;
    __syncthreads();
    unsigned base[kLdsReads];
#pragma unroll
    for (int k = 0; k < kLdsReads; k++) base[k] = seed[(blockIdx.x * 256 + threadIdx.x + 977 * k) & 4095];
    float s = 0.f;
    for (int i = 0; i < iters; i++) {
#pragma unroll
        for (int k = 0; k < kLdsReads; k++) {
            base[k] += 0x9E3779B9u;
            s += t[(base[k] >> 7) & 1023];
        }
    }
    out[blockIdx.x * 256 + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void copy4(const float4 *__restrict__ src, float4 *__restrict__ dst, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) dst[i] = src[i];
}

template <typename F>
float time_ms(F &&launch, int reps)
{
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    launch();  // warm-up
    CHECK(hipEventRecord(e0));
    for (int r = 0; r < reps; r++) launch();
    CHECK(hipEventRecord(e1));
    CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    CHECK(hipEventDestroy(e0));
    CHECK(hipEventDestroy(e1));
    return ms / reps;
}

}  // namespace

int main()
{
    hipDeviceProp_t prop;
    CHECK(hipGetDeviceProperties(&prop, 0));
    const int ncu = prop.multiProcessorCount;
    const int blocks = ncu * 16, iters = 4096;
    float *out = nullptr;
    unsigned *seed = nullptr;
    CHECK(hipMalloc(&out, sizeof(float) * (size_t)blocks * 256));
    CHECK(hipMalloc(&seed, sizeof(unsigned) * 4096));
    unsigned h[4096], s = 12345u;
    for (int i = 0; i < 4096; i++) {
        s = s * 1664525u + 1013904223u;
        h[i] = s;
    }
    CHECK(hipMemcpy(seed, h, sizeof(h), hipMemcpyHostToDevice));

    const float fma_ms = time_ms([&] { hipLaunchKernelGGL(fma_chains, dim3(blocks), dim3(256), 0, 0, out, 0.999f, 1e-3f, iters); }, 3);
    CHECK(hipGetLastError());
    const double fma_inst = (double)blocks * 4 * iters * kFmaChains;  // wave64 FMA instructions
    const float add_ms = time_ms([&] { hipLaunchKernelGGL(add_chains, dim3(blocks), dim3(256), 0, 0, out, 1e-3f, iters); }, 3);
    CHECK(hipGetLastError());
    const float lds_ms = time_ms([&] { hipLaunchKernelGGL(lds_random, dim3(blocks), dim3(256), 0, 0, out, seed, iters); }, 3);
    CHECK(hipGetLastError());
    const double lds_inst = (double)blocks * 4 * iters * kLdsReads;  // wave64 ds_read_b32

    const size_t bytes = (size_t)1 << 30;
    float4 *a = nullptr, *b = nullptr;
    CHECK(hipMalloc(&a, bytes));
    CHECK(hipMalloc(&b, bytes));
    CHECK(hipMemset(a, 0, bytes));
    const size_t n4 = bytes / sizeof(float4);
    const float copy_ms = time_ms([&] { hipLaunchKernelGGL(copy4, dim3(ncu * 32), dim3(256), 0, 0, a, b, n4); }, 5);
    CHECK(hipGetLastError());

    std::printf("{\"cus\": %d, \"valu_fma_wave_inst_per_cu_ns\": %.4f, \"valu_add_wave_inst_per_cu_ns\": %.4f, "
                "\"valu_fp32_fma_tflops\": %.2f, \"lds_random_read_wave_inst_per_cu_ns\": %.4f, \"copy_gbs\": %.1f}\n",
                ncu, fma_inst / ncu / (fma_ms * 1e6), fma_inst / ncu / (add_ms * 1e6),
                fma_inst * 64 * 2 / (fma_ms * 1e-3) / 1e12, lds_inst / ncu / (lds_ms * 1e6),
                2.0 * (double)bytes / (copy_ms * 1e-3) / 1e9);
    CHECK(hipFree(a));
    CHECK(hipFree(b));
    CHECK(hipFree(out));
    CHECK(hipFree(seed));
    return 0;
}
