// HIP kernels for the Fourier-domain projection gridder, written for gfx950
// (CDNA4, wave64).
//
// Numerics contract.  This file is compiled with -ffp-contract=off, so every
// fused multiply-add is written explicitly (fmaf).  The FP32 expressions that
// feed discrete decisions (rotated coordinates, roundf of them, kernel-table
// indices, the mask and slab tests) follow the reference kernels as nvcc
// compiles them by default (--fmad=true; the reference's build passes no
// flags): a multiply feeding an add is contracted, without reassociation,
// each add of two products fusing its first operand:
//   a*x + b*y + c           -> fl(fma(a, x, fl(b*y)) + c)       (ref_sum2c)
//   x0*a0 + x1*a1 + x2*a2   -> fma(x2, a2, fma(x0, a0, fl(x1*a1)))  (ref_sum3)
//   r*kerScale + kerCenter  -> fma(r, kerScale, kerCenter)
// The CPU oracle restates the same reading, so rounding decisions match it
// bit for bit; accumulation order is free (FP32 here, FP64 in the oracle).
//
// Layout: column-major, x fastest.  vol[x + y*V + z*V*V];
// CAS image a: img[a*M*M + i + j*M]; a direction is 9 floats [u; v; n].
#include "mgg_internal.h"

#include <hip/hip_fp16.h>

#include <algorithm>
#include <type_traits>

namespace mgg {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef unsigned int u4 __attribute__((ext_vector_type(4)));
typedef unsigned int u3 __attribute__((ext_vector_type(3)));

constexpr int kThreads = 256;

// ---------------------------------------------------------------------------
// Decision arithmetic as contracted by nvcc (see the numerics contract above).
// ---------------------------------------------------------------------------
// a*x + b*y + c: a pixel's rotated coordinate (gpuForwardProject.cu:53-55) or
// a back-projection tap position (gpuBackProject.cu:77,84,91).
__device__ __forceinline__ float ref_sum2c(float a, float x, float b, float y, float c)
{
    return fmaf(a, x, b * y) + c;
}
// x0*a0 + x1*a1 + x2*a2: a voxel's image-frame coordinate before its constant
// (gpuBackProject.cu:62-64).
__device__ __forceinline__ float ref_sum3(float x0, float a0, float x1, float a1, float x2, float a2)
{
    return fmaf(x2, a2, fmaf(x0, a0, x1 * a1));
}

// Kernel-table index (gpuForwardProject.cu:63-67, gpuBackProject.cu:79-82):
// clamp to [-convW, convW], scale and offset (one FMA), round half away from
// zero, clamp to table.
__device__ __forceinline__ int kb_index(float r, float fW, float kerScale, float kerCenter,
                                        int K)
{
    r = fminf(fmaxf(r, -fW), fW);
    const float s = fmaf(r, kerScale, kerCenter);
    int idx = (int)roundf(s);
    idx = min(max(idx, 0), K - 1);
    return idx;
}

// XCD-aware work mapping: hardware hands consecutive workgroup ids to the 8
// XCDs round-robin; remapping gives each XCD a contiguous run of work items so
// neighbouring tiles of one direction share that XCD's L2.
__device__ __forceinline__ unsigned xcd_work(unsigned perXcd)
{
    unsigned b = blockIdx.x;
    return (b & 7u) * perXcd + (b >> 3);
}

// ---------------------------------------------------------------------------
// Volume loads for the three storage types.  row5 fetches the 5 contiguous
// x-taps of one (y, z) row with unaligned vector loads.
// ---------------------------------------------------------------------------
template <int DT> struct VolLoad;

template <> struct VolLoad<DT_F32> {
    typedef float T;
    static constexpr int kRowSlack = 0;  // elements read past the 5th tap
    static __device__ __forceinline__ float one(const float *p) { return *p; }
    static __device__ __forceinline__ void row5(const float *p, float (&r)[5])
    {
        f4 a;
        __builtin_memcpy(&a, p, 16);
        r[0] = a.x; r[1] = a.y; r[2] = a.z; r[3] = a.w;
        r[4] = p[4];
    }
};

template <> struct VolLoad<DT_BF16> {
    typedef uint16_t T;
    static constexpr int kRowSlack = 1;
    static __device__ __forceinline__ float one(const uint16_t *p)
    {
        return __uint_as_float((uint32_t)(*p) << 16);
    }
    static __device__ __forceinline__ void row5(const uint16_t *p, float (&r)[5])
    {
        u3 a;
        __builtin_memcpy(&a, p, 12);
        r[0] = __uint_as_float(a.x << 16); r[1] = __uint_as_float(a.x & 0xffff0000u);
        r[2] = __uint_as_float(a.y << 16); r[3] = __uint_as_float(a.y & 0xffff0000u);
        r[4] = __uint_as_float(a.z << 16);
    }
};

template <> struct VolLoad<DT_F16> {
    typedef uint16_t T;
    static constexpr int kRowSlack = 1;
    static __device__ __forceinline__ float h2f(uint32_t bits)
    {
        __half_raw raw;
        raw.x = (unsigned short)(bits & 0xffffu);
        return __half2float(__half(raw));
    }
    static __device__ __forceinline__ float one(const uint16_t *p) { return h2f(*p); }
    static __device__ __forceinline__ void row5(const uint16_t *p, float (&r)[5])
    {
        u3 a;
        __builtin_memcpy(&a, p, 12);
        r[0] = h2f(a.x); r[1] = h2f(a.x >> 16);
        r[2] = h2f(a.y); r[3] = h2f(a.y >> 16);
        r[4] = h2f(a.z);
    }
};

// ---------------------------------------------------------------------------
// Forward projection (gpuForwardProject.cu:4-101).
// Work item = (direction, 16x16 pixel tile); one pixel per thread; wave = 16x4
// pixels.  The KB table sits in LDS.  For kerHWidth 2 the 5x5x5 gather is
// unrolled with one unaligned vector load per x-row; other widths use the
// generic loop.
// ---------------------------------------------------------------------------
struct FPArgs {
    const void *vol;
    const float *axes;
    float *out;
    const float *ker;
    int V, M, K, nAxes;
    float maskR, kerScale, kerCenter, fW;
    int convW, tilesPerRow, tilesPerImg;
    unsigned perXcd, totalWork;
    int maskD2;  // fp_lds_kernel: largest integer d2 with sqrtf((float)d2) <= maskR (-1: none)
    const unsigned *check;  // fp_lds_kernel with the 501-entry table: launch_check_volume's flag, or null
};

constexpr int KFP_FAST = 501;  // the table the mixed-volume launch requires (kerScale 125, kerCenter 250)

template <int DT, bool W2>
__global__ __launch_bounds__(kThreads) void fp_kernel(FPArgs p)
{
    using L = VolLoad<DT>;
    using T = typename L::T;
    __shared__ float sker[kMaxKerSize];

    const unsigned work = xcd_work(p.perXcd);
    if (work >= p.totalWork) return;  // uniform per block
    for (int t = threadIdx.x; t < p.K; t += kThreads) sker[t] = p.ker[t];
    __syncthreads();

    const int a = (int)(work / (unsigned)p.tilesPerImg);
    const int tile = (int)(work - (unsigned)a * (unsigned)p.tilesPerImg);
    const int ty = tile / p.tilesPerRow, tx = tile - ty * p.tilesPerRow;
    const int i = tx * 16 + (threadIdx.x & 15), j = ty * 16 + (threadIdx.x >> 4);
    const int M = p.M, V = p.V;
    if (i >= M || j >= M) return;
    const int ci = M / 2, cv = V / 2;

    float r = sqrtf((float)(i - ci) * (i - ci) + (j - ci) * (j - ci));
    float acc = 0.f;
    if (r <= p.maskR) {
        const float *ax = p.axes + (size_t)9 * a;
        const float u0 = ax[0], u1 = ax[1], u2 = ax[2];
        const float v0 = ax[3], v1 = ax[4], v2 = ax[5];
        const float fi = (float)(i - ci), fj = (float)(j - ci);
        const float x = ref_sum2c(u0, fi, v0, fj, (float)cv);
        const float y = ref_sum2c(u1, fi, v1, fj, (float)cv);
        const float z = ref_sum2c(u2, fi, v2, fj, (float)cv);
        const int X = (int)roundf(x), Y = (int)roundf(y), Z = (int)roundf(z);
        const T *vol = (const T *)p.vol;
        const long long V3 = (long long)V * V * V;

        if (W2) {
            float wi[5], wj[5], wk[5];
#pragma unroll
            for (int t = 0; t < 5; t++) {
                wi[t] = sker[kb_index((float)(X - 2 + t) - x, p.fW, p.kerScale, p.kerCenter, p.K)];
                wj[t] = sker[kb_index((float)(Y - 2 + t) - y, p.fW, p.kerScale, p.kerCenter, p.K)];
                wk[t] = sker[kb_index((float)(Z - 2 + t) - z, p.fW, p.kerScale, p.kerCenter, p.K)];
            }
            float wij[5][5];
#pragma unroll
            for (int aa = 0; aa < 5; aa++)
#pragma unroll
                for (int bb = 0; bb < 5; bb++) wij[aa][bb] = wi[aa] * wj[bb];

            const bool fast = X >= 2 && X + 2 + L::kRowSlack < V && Y >= 2 && Y + 2 < V &&
                              Z >= 2 && Z + 2 < V;
            if (fast) {
                const T *base = vol + (((size_t)(Z - 2) * V + (size_t)(Y - 2)) * V + (X - 2));
#pragma unroll
                for (int cc = 0; cc < 5; cc++) {
#pragma unroll
                    for (int bb = 0; bb < 5; bb++) {
                        float rv[5];
                        L::row5(base + ((size_t)cc * V + bb) * V, rv);
#pragma unroll
                        for (int aa = 0; aa < 5; aa++) acc = fmaf(wij[aa][bb] * wk[cc], rv[aa], acc);
                    }
                }
            } else {
                // Neighbourhood leaves the volume (only with a user mask radius
                // beyond the default or non-unit axes): the reference's
                // linear-index test (gpuForwardProject.cu:90), plus a lower
                // bound so nothing is read before the volume.
#pragma unroll
                for (int aa = 0; aa < 5; aa++)
#pragma unroll
                    for (int bb = 0; bb < 5; bb++)
#pragma unroll
                        for (int cc = 0; cc < 5; cc++) {
                            long long lin = (long long)(Z - 2 + cc) * V * V +
                                            (long long)(Y - 2 + bb) * V + (X - 2 + aa);
                            if (lin < V3 && lin >= 0)
                                acc = fmaf(wij[aa][bb] * wk[cc], L::one(vol + lin), acc);
                        }
            }
        } else {
            const int W = p.convW;
            for (int i1 = X - W; i1 <= X + W; i1++) {
                float wi = sker[kb_index((float)i1 - x, p.fW, p.kerScale, p.kerCenter, p.K)];
                for (int j1 = Y - W; j1 <= Y + W; j1++) {
                    float wj = sker[kb_index((float)j1 - y, p.fW, p.kerScale, p.kerCenter, p.K)];
                    for (int k1 = Z - W; k1 <= Z + W; k1++) {
                        float wk = sker[kb_index((float)k1 - z, p.fW, p.kerScale, p.kerCenter, p.K)];
                        float w = wi * wj * wk;
                        long long lin = (long long)k1 * V * V + (long long)j1 * V + i1;
                        if (lin < V3 && lin >= 0) acc = fmaf(w, L::one(vol + lin), acc);
                    }
                }
            }
        }
    }
    p.out[(size_t)a * M * M + (size_t)j * M + i] = acc;
}

// ---------------------------------------------------------------------------
// Forward projection, LDS-staged (kerHWidth 2).  A workgroup is 4 waves; each
// wave owns one 8x8 pixel tile of one direction.  The wave stages into LDS the
// axis-aligned box of volume voxels its pixels' 5x5x5 stencils can touch
// (bounded from the tile's four corner positions, about 30 voxels per pixel),
// reading each box row as unaligned 16-byte segments, and gathers the 125 taps
// from LDS instead of issuing 25 scattered row loads per pixel to the texture
// path.  A tile whose box exceeds the per-wave LDS budget is staged and
// gathered as two 8x4 halves, and a half whose box still does not fit gathers
// from global memory.  The gather is arranged as
// acc += wk[c] * sum_ab (wi[a]*wj[b]) * vol, i.e. the reference's product
// (wi*wj)*wk regrouped: rounding decisions are unchanged, only FP32
// accumulation rounding differs.
// ---------------------------------------------------------------------------
constexpr int FPT = 8;          // tile edge (pixels)
// Per-wave staged box budget (floats).  Four boxes take 40,960 bytes of LDS,
// a quarter of the CU's 160 KB, so four workgroups share a CU (the kernel's
// 121 VGPRs allow four waves per SIMD, at most 128 do); 73% of the 8x8 tiles' boxes fit whole
// on the grid and random workloads at N=128 (median 2,400 floats), the rest
// are staged as two 8x4 halves.  Occupancy matters more than the halves'
// extra work: three workgroups per CU with a 3,392-float budget (nearly every
// box whole) ran 31% slower, and 2,304 floats (45% whole) 8% slower.
constexpr int FPBOX = 2560;     // per-wave staged box budget (floats)
constexpr int FPKER = 640;      // generic-table path: table floats kept in LDS

__device__ __forceinline__ void wave_lds_sync_fp()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// floor(x + 0.5) in one instruction; equals roundf(x) for x >= 0.5 (below,
// an FP32 x + 0.5 can round up to the next integer).
__device__ __forceinline__ int cvt_rpi(float x)
{
    int i;
    asm("v_cvt_rpi_i32_f32 %0, %1" : "=v"(i) : "v"(x));
    return i;
}

// 501-entry table lookup (see bp_fast_kernel for the exactness argument),
// read from the table in global memory (2 KB, cache-resident): the forward
// kernel does only 15 lookups per pixel, and keeping the table out of LDS
// leaves room for a fourth workgroup per CU.  Clamping the index to [0, 500]
// picks the entry the reference's clamp of r would (|r| <= 2.5 here), and
// keeps any index, even from a non-finite coordinate, inside the table.
__device__ __forceinline__ float kb_fast_fp(const float *__restrict__ ker, float r)
{
    const float s = fmaf(r, 125.0f, 250.0f);
    const int i = cvt_rpi(s);
    return *(const float *)((const char *)ker + (unsigned)min(max(i, 0), 500) * 4u);
}

template <typename L>
__device__ __forceinline__ void load4(const typename L::T *p, float (&v)[4]);
template <>
__device__ __forceinline__ void load4<VolLoad<DT_F32>>(const float *p, float (&v)[4])
{
    f4 a;
    __builtin_memcpy(&a, p, 16);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
}
template <>
__device__ __forceinline__ void load4<VolLoad<DT_BF16>>(const uint16_t *p, float (&v)[4])
{
    typedef unsigned int u2 __attribute__((ext_vector_type(2)));
    u2 a;
    __builtin_memcpy(&a, p, 8);
    v[0] = __uint_as_float(a.x << 16); v[1] = __uint_as_float(a.x & 0xffff0000u);
    v[2] = __uint_as_float(a.y << 16); v[3] = __uint_as_float(a.y & 0xffff0000u);
}
template <>
__device__ __forceinline__ void load4<VolLoad<DT_F16>>(const uint16_t *p, float (&v)[4])
{
    typedef unsigned int u2 __attribute__((ext_vector_type(2)));
    u2 a;
    __builtin_memcpy(&a, p, 8);
    v[0] = VolLoad<DT_F16>::h2f(a.x); v[1] = VolLoad<DT_F16>::h2f(a.x >> 16);
    v[2] = VolLoad<DT_F16>::h2f(a.y); v[3] = VolLoad<DT_F16>::h2f(a.y >> 16);
}

// Tile body.  W4 = false: every pixel gathers its full 5x5x5 stencil from the
// box of all its stencils.  W4 = true (501-entry table whose end entries are 0,
// volume known to be finite): along each axis the reference's taps X-2..X+2
// (X = roundf(x)) have r = tap - x clamped to [-2, 2], and the tap with r <= -2
// reads table entry 0 and the one with r >= 2 entry 500 (fma(-2, 125, 250) = 0
// and fma(2, 125, 250) = 500 exactly).  If x >= X, then X - 2 - x <= -2; if
// x < X, then X + 2 - x > 2: one of the two end taps always has weight 0, and
// the others are the four taps from Xs = X - 2 + (x >= X) = floor(x) - 1.  The
// zero taps add fma(0, v, s) = s for finite v, so gathering the 4x4x4 window
// gives the same sums bit for bit, and the box shrinks to the voxels the
// windows touch: [floor(min) - 1, floor(max) + 2] per axis instead of
// [floor(min) - 2, ceil(max) + 2], about 40% fewer voxels to stage.
template <int DT, bool FAST, bool W4, bool BIG>
__device__ __forceinline__ void fp_lds_tile_body(const FPArgs &p, const float *__restrict__ sker,
                                                 float *__restrict__ box, int a, int i0, int j0, int ln,
                                                 bool active, bool pix, float *out)
{
    using L = VolLoad<DT>;
    using T = typename L::T;
    constexpr int NW = W4 ? 4 : 5;  // taps per axis
    const int i = i0 + (ln & 7), j = j0 + (ln >> 3);
    const int M = p.M, V = p.V, ci = M / 2, cv = V / 2;
    const float *ax = p.axes + (size_t)9 * a;
    const float u0 = ax[0], u1 = ax[1], u2 = ax[2];
    const float v0 = ax[3], v1 = ax[4], v2 = ax[5];
    const float fi = (float)(i - ci), fj = (float)(j - ci);
    const float x = ref_sum2c(u0, fi, v0, fj, (float)cv);
    const float y = ref_sum2c(u1, fi, v1, fj, (float)cv);
    const float z = ref_sum2c(u2, fi, v2, fj, (float)cv);
    const int X = (int)roundf(x), Y = (int)roundf(y), Z = (int)roundf(z);
    // first tap of each axis' window
    const int Xs = X - 2 + ((W4 && x >= (float)X) ? 1 : 0);
    const int Ys = Y - 2 + ((W4 && y >= (float)Y) ? 1 : 0);
    const int Zs = Z - 2 + ((W4 && z >= (float)Z) ? 1 : 0);

    const long long V3 = (long long)V * V * V;
    const T *vol = (const T *)p.vol;

    float wi[NW], wj[NW], wk[NW];
#pragma unroll
    for (int t = 0; t < NW; t++) {
        if (FAST) {
            wi[t] = kb_fast_fp(p.ker, (float)(Xs + t) - x);
            wj[t] = kb_fast_fp(p.ker, (float)(Ys + t) - y);
            wk[t] = kb_fast_fp(p.ker, (float)(Zs + t) - z);
        } else {
            wi[t] = sker[kb_index((float)(Xs + t) - x, p.fW, p.kerScale, p.kerCenter, p.K)];
            wj[t] = sker[kb_index((float)(Ys + t) - y, p.fW, p.kerScale, p.kerCenter, p.K)];
            wk[t] = sker[kb_index((float)(Zs + t) - z, p.fW, p.kerScale, p.kerCenter, p.K)];
        }
    }

    // Stages the voxel box of pixel rows j0 + rlo .. j0 + rhi of the tile and
    // gathers the pixels of the lanes with mine set; false (nothing done) when
    // the box does not fit the wave's LDS budget.  Uniform across the wave.
    float acc = 0.f;
    auto stageAndGather = [&](int rlo, int rhi, bool mine, auto splitT) -> bool {
        // Box from the corners: positions are affine in (i, j), so every
        // pixel's coordinate lies in [min corner, max corner] (the 1e-3 margin
        // covers FP32 evaluation error), widened by the window.
        const float ai = (float)(i0 - ci), bi = (float)(i0 + FPT - 1 - ci);
        const float aj = (float)(j0 + rlo - ci), bj = (float)(j0 + rhi - ci);
        auto lo = [&](float uu, float vv) {
            return (int)floorf(fminf(uu * ai, uu * bi) + fminf(vv * aj, vv * bj) + (float)cv - 1e-3f) - (W4 ? 1 : 2);
        };
        auto hi = [&](float uu, float vv) {
            const float e = fmaxf(uu * ai, uu * bi) + fmaxf(vv * aj, vv * bj) + (float)cv + 1e-3f;
            return (int)(W4 ? floorf(e) : ceilf(e)) + 2;
        };
        const int bx0 = __builtin_amdgcn_readfirstlane(lo(u0, v0));
        const int by0 = __builtin_amdgcn_readfirstlane(lo(u1, v1));
        const int bz0 = __builtin_amdgcn_readfirstlane(lo(u2, v2));
        const int Dx = __builtin_amdgcn_readfirstlane(hi(u0, v0)) - bx0 + 1;
        const int Dy = __builtin_amdgcn_readfirstlane(hi(u1, v1)) - by0 + 1;
        const int Dz = __builtin_amdgcn_readfirstlane(hi(u2, v2)) - bz0 + 1;
        // Volume addressing by 32-bit byte offsets: from the volume's base
        // (below 4 GB), or (BIG) from the box's first element, whose
        // linear-index span is below 64 V^2 elements.
        if (!(Dx >= NW && Dy >= NW && Dz >= NW && Dx <= 64 && Dy <= 64 && Dz <= 64 &&
              (BIG ? 64LL * V * V : V3) * (long long)sizeof(T) < (1LL << 32)))
            return false;
        // 16-byte segments per box row: enough that the 8-wide read of a window
        // (below) stays inside the row's staged segments.
        const int segs = (Dx + 8 - NW) >> 2;
        // LDS pitches (floats), multiples of 4 (16-byte rows).  When the box
        // still fits, the row and plane pitches are padded to an odd number
        // of 16-byte units so that the rows and planes a tile's pixels read
        // spread over the banks (a pitch of 16 puts every row at the same
        // bank offset); otherwise the rows are packed.
        int RP = (segs | 1) * 4;
        int PP = RP * Dy + ((Dy & 1) ? 0 : RP);
        if (PP * Dz > FPBOX) {
            RP = segs * 4;
            PP = RP * Dy;
        }
        if (PP * Dz > FPBOX) return false;

        // Stage.  A plane of the box holds P = Dy * segs row segments.  With
        // P <= 64 one pass covers G = 64 / P planes (lane -> plane offset pl,
        // row ey, segment sg), else G = 1 and a plane takes ceil(P / 64) lane
        // passes; each lane then steps through its planes with one add for
        // the volume offset and one for the LDS pointer.  The lane -> task
        // divisions use the reciprocal: (q + 0.5) / P is at least 0.5 / P
        // away from an integer, far more than its error for q, P <= 1024.
        const int VV = V * V, V3i = (int)V3;
        const int P = Dy * segs;
        const float rP = __builtin_amdgcn_rcpf((float)P), rSegs = __builtin_amdgcn_rcpf((float)segs);
        const int G = P <= 64 ? __builtin_amdgcn_readfirstlane((int)(64.5f * rP)) : 1;
        const int nLanePass = (P + 63) >> 6;
        // The reference bounds only the linear index (gpuForwardProject.cu:90):
        // when the box's whole index range is inside the volume no element
        // needs a check.  lin grows with ez, ey and the segment, so the range
        // is spanned by the first and last segment of the box.
        const long long linLo = (long long)bz0 * VV + (long long)by0 * V + bx0;
        const long long linHi = (long long)(bz0 + Dz - 1) * VV + (long long)(by0 + Dy - 1) * V + bx0 + 4 * segs - 1;
        const bool boxIn = linLo >= 0 && linHi < V3;
        // Six planes per batch: all loads of a batch are in flight before the
        // first LDS store waits on them (2, 4 and 6 planes: 19.9, 18.5 and
        // 18.3 ms per 10,000 directions at N=128 with the 5-wide windows; 8
        // needs more than 128 VGPRs, three waves per SIMD instead of four,
        // 22.0 ms).  The checked and unchecked variants are separate loops
        // (boxIn is wave-uniform).  The unchecked one addresses the volume by
        // a 32-bit byte offset from the volume's base, or (BIG: volumes of
        // 4 GB and more) from the box's first element; the checked one, which
        // runs only on boxes that cross the volume's index range, by the
        // int linear index, or (BIG) a 64-bit one.
        const T *volBox = vol + ((BIG && boxIn) ? linLo : 0);
        using Lin = std::conditional_t<BIG, long long, int>;
        constexpr int SB = 6;
        const int lStep = G * VV, dStep = G * PP;
        for (int c = 0; c < nLanePass; c++) {
            const int q = ln + 64 * c;
            const int pl = (int)(((float)q + 0.5f) * rP);
            const int pq = q - pl * P;
            const int ey = (int)(((float)pq + 0.5f) * rSegs);
            const int sg = pq - ey * segs;
            if (pl >= G) continue;
            const Lin lin0 = (Lin)(bz0 + pl) * VV + (Lin)(by0 + ey) * V + (bx0 + 4 * sg);
            float *d0 = box + pl * PP + ey * RP + 4 * sg;
            auto stage = [&](auto inBox) {
                // offset from volBox (BIG and boxIn: inside the box's span,
                // below 2^30) or the linear index
                using Idx = std::conditional_t<decltype(inBox)::value, int, Lin>;
                Idx l = (Idx)((BIG && decltype(inBox)::value) ? lin0 - (Lin)linLo : lin0);
                float *dp = d0;
                for (int ez = pl; ez < Dz; ez += SB * G) {
                    float v4[SB][4];
#pragma unroll
                    for (int k = 0; k < SB; k++) {
                        if (ez + k * G < Dz) {
                            const Idx lk = l + (Idx)k * lStep;
                            if (decltype(inBox)::value) {
                                load4<L>((const T *)((const char *)volBox + (unsigned)lk * (unsigned)sizeof(T)), v4[k]);
                            } else {
#pragma unroll
                                for (int e = 0; e < 4; e++)
                                    v4[k][e] = (lk + e >= 0 && lk + e < (BIG ? (Lin)V3 : (Lin)V3i)) ? L::one(vol + lk + e) : 0.f;
                            }
                        }
                    }
#pragma unroll
                    for (int k = 0; k < SB; k++) {
                        if (ez + k * G < Dz) {
                            f4 w4 = {v4[k][0], v4[k][1], v4[k][2], v4[k][3]};
                            *(f4 *)(dp + k * dStep) = w4;
                        }
                    }
                    l += SB * lStep;
                    dp += SB * dStep;
                }
            };
            if (boxIn) {
                stage(std::true_type{});
            } else {
                stage(std::false_type{});
            }
        }
        // 16-byte gather: each NW-tap row is read as the two aligned 16-byte
        // units that contain it, [xo - o, xo - o + 8) with o = xo & 3, and the
        // x weights are shifted by o with zeros in the unused slots.  Since
        // xo <= Dx - NW, xo - o + 8 <= 4 * floor((Dx - NW) / 4) + 8, which
        // is at most 4 * segs, so the read stays inside the row's staged
        // segments.  The unused values are staged volume values times 0,
        // which leaves every sum bit-identical to the NW-tap read when they
        // are finite.  With the 5-wide window a non-finite one makes the
        // pixel's sum NaN (0 * NaN and 0 * Inf are NaN), so a pixel whose sum
        // is not finite is gathered again with the 5-tap read, and NaN and Inf
        // reach exactly the pixels whose stencil holds them, as in the
        // reference; the 4-wide window runs only on finite volumes.
        wave_lds_sync_fp();
        // x weights shifted by o (zeros in the unused slots) times the y weights.
        auto shifted = [&](const float (&wx)[NW], const float (&wy)[NW], int o, float (&w8)[8][NW]) {
            float ws[8];
#pragma unroll
            for (int t = 0; t < 8; t++) {
                float w = 0.f;
#pragma unroll
                for (int sft = 0; sft < 4; sft++)
                    if (t - sft >= 0 && t - sft < NW) w = (o == sft) ? wx[t - sft] : w;
                ws[t] = w;
            }
#pragma unroll
            for (int t = 0; t < 8; t++)
#pragma unroll
                for (int bb = 0; bb < NW; bb++) w8[t][bb] = ws[t] * wy[bb];
        };
        // One z plane of a window: NW rows of the 16-byte read.
        auto planeSum = [&](const float *bq, const float (&w8)[8][NW], int cc) {
            float sc = 0.f;
#pragma unroll
            for (int bb = 0; bb < NW; bb++) {
                const float *rp = bq + cc * PP + bb * RP;
                const f4 q0 = *(const f4 *)rp, q1 = *(const f4 *)(rp + 4);
                sc = fmaf(w8[0][bb], q0.x, sc);
                sc = fmaf(w8[1][bb], q0.y, sc);
                sc = fmaf(w8[2][bb], q0.z, sc);
                sc = fmaf(w8[3][bb], q0.w, sc);
                sc = fmaf(w8[4][bb], q1.x, sc);
                sc = fmaf(w8[5][bb], q1.y, sc);
                sc = fmaf(w8[6][bb], q1.z, sc);
                sc = fmaf(w8[7][bb], q1.w, sc);
            }
            return sc;
        };
        float res = 0.f;
        bool gathered = false;
        if constexpr (!decltype(splitT)::value) {
            if (active && mine) {
                const int xo = Xs - bx0, o = xo & 3;
                const float *bq = box + (Zs - bz0) * PP + (Ys - by0) * RP + (xo - o);
                float w8[8][NW];
                shifted(wi, wj, o, w8);
#pragma unroll
                for (int cc = 0; cc < NW; cc++) res = fmaf(wk[cc], planeSum(bq, w8, cc), res);
                gathered = true;
            }
        } else {
            // Half tile: the lanes of the other half (lane ^ 32; their own
            // pixels are not in this box) gather the last two planes of their
            // partner's window while the owner gathers the first NW - 2, so
            // the wave reads 3 * NW rows per pixel instead of NW * NW; the
            // owner adds the parts.
            constexpr int H = NW - 2;  // planes the owner gathers
            const int pX = __shfl_xor(Xs, 32), pY = __shfl_xor(Ys, 32), pZ = __shfl_xor(Zs, 32);
            const int pAct = __shfl_xor(active ? 1 : 0, 32);
            float gwi[NW], gwj[NW];
#pragma unroll
            for (int t = 0; t < NW; t++) {
                const float a5 = __shfl_xor(wi[t], 32), b5 = __shfl_xor(wj[t], 32);
                gwi[t] = mine ? wi[t] : a5;
                gwj[t] = mine ? wj[t] : b5;
            }
            const float pkA = __shfl_xor(wk[H], 32), pkB = __shfl_xor(wk[H + 1], 32);
            const int gX = mine ? Xs : pX, gY = mine ? Ys : pY, gZ = mine ? Zs : pZ + H;
            const float k0 = mine ? wk[0] : pkA, k1 = mine ? wk[1] : pkB;
            float part = 0.f;
            if (mine ? active : pAct != 0) {
                const int xo = gX - bx0, o = xo & 3;
                const float *bq = box + (gZ - bz0) * PP + (gY - by0) * RP + (xo - o);
                float w8[8][NW];
                shifted(gwi, gwj, o, w8);
                part = fmaf(k0, planeSum(bq, w8, 0), part);
                part = fmaf(k1, planeSum(bq, w8, 1), part);
                if (!W4 && mine) part = fmaf(wk[2], planeSum(bq, w8, 2), part);
            }
            const float other = __shfl_xor(part, 32);
            res = part + other;
            gathered = active && mine;
        }
        if (gathered) {
            if (!W4 && !__builtin_isfinite(res)) {
                const float *bp = box + (Z - 2 - bz0) * PP + (Y - 2 - by0) * RP + (X - 2 - bx0);
                res = 0.f;
#pragma unroll
                for (int cc = 0; cc < NW; cc++) {
                    float sc = 0.f;
#pragma unroll
                    for (int bb = 0; bb < NW; bb++) {
                        const float *rp = bp + cc * PP + bb * RP;
#pragma unroll
                        for (int aa = 0; aa < NW; aa++) sc = fmaf(wi[aa] * wj[bb], rp[aa], sc);
                    }
                    res = fmaf(wk[cc], sc, res);
                }
            }
            acc = res;
        }
        wave_lds_sync_fp();  // the box is read before it is staged again
        return true;
    };

    // The whole 8x8 tile; if its box is too large, its two 8x4 halves one
    // after the other; a half that still does not fit gathers from global
    // memory.
    bool done = stageAndGather(0, FPT - 1, true, std::false_type{});
    if (!done) {
        const bool top = (ln >> 3) < FPT / 2;
        const bool d0 = stageAndGather(0, FPT / 2 - 1, top, std::true_type{});
        const bool d1 = stageAndGather(FPT / 2, FPT - 1, !top, std::true_type{});
        done = top ? d0 : d1;
    }
    if (!done && active) {
#pragma unroll
        for (int cc = 0; cc < NW; cc++) {
            float sc = 0.f;
#pragma unroll
            for (int bb = 0; bb < NW; bb++)
#pragma unroll
                for (int aa = 0; aa < NW; aa++) {
                    const long long lin = (long long)(Zs + cc) * V * V + (long long)(Ys + bb) * V + (Xs + aa);
                    if (lin < V3 && lin >= 0) sc = fmaf(wi[aa] * wj[bb], L::one(vol + lin), sc);
                }
            acc = fmaf(wk[cc], sc, acc);
        }
    }
    if (pix) *out = active ? acc : 0.f;
}

// One 8x8 tile of one direction; W4 (only with the 501-entry table) selects
// the 4-wide windows.
template <int DT, bool FAST, bool W4, bool BIG = false>
__device__ __forceinline__ void fp_lds_tile(const FPArgs &p, const float *__restrict__ sker, float *__restrict__ box,
                                            unsigned work, int ln)
{
    const int a = (int)(work / (unsigned)p.tilesPerImg);
    const int tile = (int)(work - (unsigned)a * (unsigned)p.tilesPerImg);
    const int ty = tile / p.tilesPerRow, tx = tile - ty * p.tilesPerRow;
    const int i0 = tx * FPT, j0 = ty * FPT;
    const int i = i0 + (ln & 7), j = j0 + (ln >> 3);
    const int M = p.M, ci = M / 2;
    const bool pix = i < M && j < M;

    // sqrtf(d2) <= maskR for the exact integer d2, as one integer compare
    // (sqrtf is monotone; the host finds the largest passing d2).
    const bool active = pix && (i - ci) * (i - ci) + (j - ci) * (j - ci) <= p.maskD2;
    float *out = p.out + (size_t)a * M * M + (size_t)j * M + i;
    if (__ballot(active) == 0ull) {
        if (pix) *out = 0.f;
        return;
    }
    fp_lds_tile_body<DT, FAST, FAST && W4, BIG>(p, sker, box, a, i0, j0, ln, active, pix, out);
}

// The 4-wide windows need a finite volume and a finite 501-entry table whose
// end entries are 0 (the reference's table sets them so): launch_check_volume
// writes *check = 1 when they hold.  Wave-uniform.
__device__ __forceinline__ bool fp_win4(const unsigned *check)
{
    return check != nullptr && __builtin_amdgcn_readfirstlane(*check) == 1u;
}

template <int DT, bool FAST, bool BIG>
__global__ __launch_bounds__(kThreads) void fp_lds_kernel(FPArgs p)
{
    __shared__ float sker[FPKER];
    __shared__ __attribute__((aligned(16))) float sbox[4][FPBOX];

    if (!FAST) {  // the fast lookup reads the table from global memory
        for (int t = threadIdx.x; t < p.K; t += kThreads) sker[t] = p.ker[t];
        __syncthreads();
    }
    // Each wave takes work items wv + 4*blockIdx.x, + 4*gridDim.x, ...  (wave
    // uniform, so the work index and its divisions stay scalar)
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), ln = threadIdx.x & 63;
    // The window width is chosen once, outside the tile loop, so that each
    // loop is register-allocated on its own.
    if (FAST && fp_win4(p.check)) {
        for (unsigned work = blockIdx.x * 4u + (unsigned)wv; work < p.totalWork; work += gridDim.x * 4u)
            fp_lds_tile<DT, FAST, true, BIG>(p, sker, sbox[wv], work, ln);
    } else {
        for (unsigned work = blockIdx.x * 4u + (unsigned)wv; work < p.totalWork; work += gridDim.x * 4u)
            fp_lds_tile<DT, FAST, false, BIG>(p, sker, sbox[wv], work, ln);
    }
}

// ---------------------------------------------------------------------------
// Exact transpose of the forward gather: each in-mask pixel scatters its
// value through the same (voxel, weight) pairs with FP32 atomics.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void adj_kernel(FPArgs p, const float *__restrict__ img,
                                                       float *__restrict__ volAcc)
{
    __shared__ float sker[kMaxKerSize];
    const unsigned work = xcd_work(p.perXcd);
    if (work >= p.totalWork) return;
    for (int t = threadIdx.x; t < p.K; t += kThreads) sker[t] = p.ker[t];
    __syncthreads();

    const int a = (int)(work / (unsigned)p.tilesPerImg);
    const int tile = (int)(work - (unsigned)a * (unsigned)p.tilesPerImg);
    const int ty = tile / p.tilesPerRow, tx = tile - ty * p.tilesPerRow;
    const int i = tx * 16 + (threadIdx.x & 15), j = ty * 16 + (threadIdx.x >> 4);
    const int M = p.M, V = p.V;
    if (i >= M || j >= M) return;
    const int ci = M / 2, cv = V / 2;
    float r = sqrtf((float)(i - ci) * (i - ci) + (j - ci) * (j - ci));
    if (!(r <= p.maskR)) return;
    const float val = img[(size_t)a * M * M + (size_t)j * M + i];
    if (val == 0.f) return;
    const float *ax = p.axes + (size_t)9 * a;
    const float u0 = ax[0], u1 = ax[1], u2 = ax[2];
    const float v0 = ax[3], v1 = ax[4], v2 = ax[5];
    const float fi = (float)(i - ci), fj = (float)(j - ci);
    const float x = ref_sum2c(u0, fi, v0, fj, (float)cv);
    const float y = ref_sum2c(u1, fi, v1, fj, (float)cv);
    const float z = ref_sum2c(u2, fi, v2, fj, (float)cv);
    const int X = (int)roundf(x), Y = (int)roundf(y), Z = (int)roundf(z);
    const long long V3 = (long long)V * V * V;
    const int W = p.convW;
    for (int i1 = X - W; i1 <= X + W; i1++) {
        float wi = sker[kb_index((float)i1 - x, p.fW, p.kerScale, p.kerCenter, p.K)];
        for (int j1 = Y - W; j1 <= Y + W; j1++) {
            float wj = sker[kb_index((float)j1 - y, p.fW, p.kerScale, p.kerCenter, p.K)];
            for (int k1 = Z - W; k1 <= Z + W; k1++) {
                float wk = sker[kb_index((float)k1 - z, p.fW, p.kerScale, p.kerCenter, p.K)];
                float w = wi * wj * wk;
                long long lin = (long long)k1 * V * V + (long long)j1 * V + i1;
                if (lin < V3 && lin >= 0 && w != 0.f) unsafeAtomicAdd(volAcc + lin, w * val);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Forward projection over several volumes in one launch (the request
// server's batches mix volumes of different N).  Each segment is a run of
// directions on one volume with its own geometry and output; work items are
// the segments' 8x8 tiles, numbered segment after segment, and every wave
// finds its segment by a binary search over the segments' first work items
// (wave-uniform, so the search and the descriptor reads are scalar).  The
// tile code is fp_lds_tile's, so each image is bit-identical to a launch of
// its own volume alone.  A batch of requests on several volumes is one launch;
// its workgroups walk the work list three tiles per wave.
// ---------------------------------------------------------------------------
template <int DT, bool BIG>
__global__ __launch_bounds__(kThreads) void fp_lds_multi_kernel(FPMultiArgs m)
{
    __shared__ __attribute__((aligned(16))) float sbox[4][FPBOX];
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), ln = threadIdx.x & 63;
    for (unsigned work = blockIdx.x * 4u + (unsigned)wv; work < m.totalWork; work += gridDim.x * 4u) {
        int lo = 0, hi = m.nSegs - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (m.seg[mid].workStart <= work) lo = mid; else hi = mid - 1;
        }
        const FPSegment &sg = m.seg[lo];
        FPArgs a;
        a.vol = sg.vol; a.axes = sg.axes; a.out = sg.out; a.ker = sg.ker;
        a.V = sg.V; a.M = sg.M; a.K = KFP_FAST; a.nAxes = sg.nAxes;
        a.maskR = 0.f; a.kerScale = 125.0f; a.kerCenter = 250.0f; a.fW = 2.0f;
        a.convW = 2; a.tilesPerRow = sg.tilesPerRow; a.tilesPerImg = sg.tilesPerImg;
        a.perXcd = 0; a.totalWork = 0; a.maskD2 = sg.maskD2; a.check = sg.check;
        if (fp_win4(sg.check)) fp_lds_tile<DT, true, true, BIG>(a, nullptr, sbox[wv], work - sg.workStart, ln);
        else fp_lds_tile<DT, true, false, BIG>(a, nullptr, sbox[wv], work - sg.workStart, ln);
    }
}

// ---------------------------------------------------------------------------
// Exact transpose, LDS-box scatter (kerHWidth 2, tables of at most FPKER
// entries).  A wave owns one 8x8 pixel tile of one direction, as in
// fp_lds_tile: it bounds the voxel box its pixels' 5x5x5 stencils touch (the
// same corner bound), zeroes that box in LDS, adds every tap's w * value into
// it (ordered plain read-modify-write in rounds that keep colliding lanes
// apart; LDS float atomics for frames that are not orthonormal), and adds each touched
// box voxel to the volume with one global atomic: about 11 global atomics per
// pixel at N=128 instead of one per tap (125).  Box voxels map to the volume through the reference's
// linear index (gpuForwardProject.cu:90), so the transpose keeps the forward
// gather's row wrap.  A tile whose box exceeds the LDS budget runs as two 8x4
// halves; a half that still does not fit scatters with global atomics.  With
// launch_check_volume's flag over the CAS images and the table (all finite,
// table ends 0) it scatters only the 4x4x4 windows of the taps that can have
// nonzero weight, from the windows' smaller box, as fp_lds_tile_body gathers
// them: the skipped taps add w * value = 0.
// ---------------------------------------------------------------------------
template <bool FAST, bool W4>
__device__ __forceinline__ void adj_lds_tile(const FPArgs &p, const float *__restrict__ sker,
                                             float *__restrict__ box, const float *__restrict__ img,
                                             float *__restrict__ volAcc, unsigned work, int ln)
{
    const int a = (int)(work / (unsigned)p.tilesPerImg);
    const int tile = (int)(work - (unsigned)a * (unsigned)p.tilesPerImg);
    const int ty = tile / p.tilesPerRow, tx = tile - ty * p.tilesPerRow;
    const int i0 = tx * FPT, j0 = ty * FPT;
    const int i = i0 + (ln & 7), j = j0 + (ln >> 3);
    const int M = p.M, V = p.V, ci = M / 2, cv = V / 2;
    const bool pix = i < M && j < M;
    const bool inMask = pix && (i - ci) * (i - ci) + (j - ci) * (j - ci) <= p.maskD2;
    const float val = inMask ? img[(size_t)a * M * M + (size_t)j * M + i] : 0.f;
    const bool act = inMask && val != 0.f;
    if (__ballot(act) == 0ull) return;

    const float *ax = p.axes + (size_t)9 * a;
    const float u0 = ax[0], u1 = ax[1], u2 = ax[2];
    const float v0 = ax[3], v1 = ax[4], v2 = ax[5];
    const float fi = (float)(i - ci), fj = (float)(j - ci);
    const float x = ref_sum2c(u0, fi, v0, fj, (float)cv);
    const float y = ref_sum2c(u1, fi, v1, fj, (float)cv);
    const float z = ref_sum2c(u2, fi, v2, fj, (float)cv);
    const int X = (int)roundf(x), Y = (int)roundf(y), Z = (int)roundf(z);
    // Window of each axis: all five taps, or (W4, checked images and table)
    // the four that can have nonzero weight, from floor(x) - 1 (as in
    // fp_lds_tile_body).
    constexpr int NW = W4 ? 4 : 5;
    const int Xs = X - 2 + ((W4 && x >= (float)X) ? 1 : 0);
    const int Ys = Y - 2 + ((W4 && y >= (float)Y) ? 1 : 0);
    const int Zs = Z - 2 + ((W4 && z >= (float)Z) ? 1 : 0);
    const long long V3 = (long long)V * V * V;
    const long long VV = (long long)V * V;
    // [u v] orthonormal to 1e-4 (wave-uniform: one direction per tile)
    const bool orthoFrame = fabsf(u0 * u0 + u1 * u1 + u2 * u2 - 1.f) <= 1e-4f &&
                            fabsf(v0 * v0 + v1 * v1 + v2 * v2 - 1.f) <= 1e-4f &&
                            fabsf(u0 * v0 + u1 * v1 + u2 * v2) <= 1e-4f;

    float wi[NW], wj[NW], wk[NW];
#pragma unroll
    for (int t = 0; t < NW; t++) {
        if (FAST) {
            wi[t] = kb_fast_fp(p.ker, (float)(Xs + t) - x);
            wj[t] = kb_fast_fp(p.ker, (float)(Ys + t) - y);
            wk[t] = kb_fast_fp(p.ker, (float)(Zs + t) - z);
        } else {
            wi[t] = sker[kb_index((float)(Xs + t) - x, p.fW, p.kerScale, p.kerCenter, p.K)];
            wj[t] = sker[kb_index((float)(Ys + t) - y, p.fW, p.kerScale, p.kerCenter, p.K)];
            wk[t] = sker[kb_index((float)(Zs + t) - z, p.fW, p.kerScale, p.kerCenter, p.K)];
        }
    }

    // The reference's bounds test on the linear index, one global atomic per
    // tap (gpuForwardProject.cu:90); every tap of the window, as the box path.
    auto scatterGlobal = [&]() {
#pragma unroll
        for (int cc = 0; cc < NW; cc++)
#pragma unroll
            for (int bb = 0; bb < NW; bb++)
#pragma unroll
                for (int aa = 0; aa < NW; aa++) {
                    const long long lin = (long long)(Zs + cc) * VV + (long long)(Ys + bb) * V + (Xs + aa);
                    const float w = wi[aa] * wj[bb] * wk[cc];
                    if (lin < V3 && lin >= 0) unsafeAtomicAdd(volAcc + lin, w * val);
                }
    };

    // Scatters the pixels of the tile rows j0 + rlo .. j0 + rhi (lanes with
    // mine set) through an LDS box; false (nothing done) when the box does not
    // fit the wave's budget.  Uniform across the wave.
    auto scatter = [&](int rlo, int rhi, bool mine) -> bool {
        const float ai = (float)(i0 - ci), bi = (float)(i0 + FPT - 1 - ci);
        const float aj = (float)(j0 + rlo - ci), bj = (float)(j0 + rhi - ci);
        auto lo = [&](float uu, float vv) {
            return (int)floorf(fminf(uu * ai, uu * bi) + fminf(vv * aj, vv * bj) + (float)cv - 1e-3f) - (W4 ? 1 : 2);
        };
        auto hi = [&](float uu, float vv) {
            const float e = fmaxf(uu * ai, uu * bi) + fmaxf(vv * aj, vv * bj) + (float)cv + 1e-3f;
            return (int)(W4 ? floorf(e) : ceilf(e)) + 2;
        };
        const int bx0 = __builtin_amdgcn_readfirstlane(lo(u0, v0));
        const int by0 = __builtin_amdgcn_readfirstlane(lo(u1, v1));
        const int bz0 = __builtin_amdgcn_readfirstlane(lo(u2, v2));
        const int Dx = __builtin_amdgcn_readfirstlane(hi(u0, v0)) - bx0 + 1;
        const int Dy = __builtin_amdgcn_readfirstlane(hi(u1, v1)) - by0 + 1;
        const int Dz = __builtin_amdgcn_readfirstlane(hi(u2, v2)) - bz0 + 1;
        if (!(Dx >= NW && Dy >= NW && Dz >= NW && Dx <= 64 && Dy <= 64 && Dz <= 64)) return false;
        const int RP = Dx, PP = Dx * Dy, total = PP * Dz;
        if (total > FPBOX) return false;
        for (int e = ln; e < total; e += 64) box[e] = 0.f;
        wave_lds_sync_fp();
        // Every finite in-row pixel's stencil lies inside the box (the corner
        // bound); a lane whose does not (non-finite directions) scatters
        // through the checked global path instead of outside its box.
        const bool inBox = (unsigned)(Xs - bx0) <= (unsigned)(Dx - NW) &&
                           (unsigned)(Ys - by0) <= (unsigned)(Dy - NW) &&
                           (unsigned)(Zs - bz0) <= (unsigned)(Dz - NW);
        if (act && mine && !inBox) scatterGlobal();
        const int key = (Zs - bz0) * PP + (Ys - by0) * RP + (Xs - bx0);
        // Tap value (wi * val) * (wj * wk): the forward weight times the pixel
        // value, grouped so that only NW + NW * NW products are live across
        // the rounds instead of one per tap (the per-tap products, hoisted
        // out of the rounds loop, held the kernel at 225 VGPRs).
        float wiv[NW];
#pragma unroll
        for (int aa = 0; aa < NW; aa++) wiv[aa] = wi[aa] * val;
        auto addStencil = [&](bool atomic) {
            float *bq = box + key;
#pragma unroll 1
            for (int cc = 0; cc < NW; cc++)
#pragma unroll
                for (int bb = 0; bb < NW; bb++) {
                    float *rp = bq + cc * PP + bb * RP;
                    const float wjk = wj[bb] * wk[cc];
#pragma unroll
                    for (int aa = 0; aa < NW; aa++) {
                        const float t = wiv[aa] * wjk;
                        if (atomic) {
                            atomicAdd(rp + aa, t);
                        } else {
                            rp[aa] = rp[aa] + t;
                            // Keep this tap's write before the next tap's read:
                            // another lane's stencil may cover this word at a
                            // later tap, and its read must see this write.
                            __asm__ volatile("" ::: "memory");
                        }
                    }
                }
        };
        bool pend = act && mine && inBox;
        if (orthoFrame) {
            // Plain read-modify-write, one tap after the other (LDS float
            // atomics ran 3.6x slower: 40 vs 11 ms at N=96 with 2,048
            // directions).  One wave's LDS instructions execute in order and
            // each tap's read follows the previous tap's write (compiler
            // barrier), so a lane's read sees every earlier write of every
            // lane; only lanes of the same instruction can lose an update,
            // and those target the same word at every tap exactly when their
            // windows start at the same voxel (key).  The start is X - 2 with
            // X = roundf(x), or floor(x) - 1 with the 4-wide windows, per axis;
            // for an orthonormal frame two pixels whose starts coincide have
            // positions less than 1 apart in every coordinate, so they are at
            // most sqrt(3) apart: tile neighbours.  A lane waits while a lower
            // pending neighbour (left, up, up-left, up-right) has its key;
            // every round serves at least the lowest pending lane of each key.
            const int ix = ln & 7, jy = ln >> 3;
            while (__ballot(pend) != 0ull) {
                const int k = pend ? key : -1 - ln;  // distinct for lanes not pending
                const int kl = __shfl(k, ln - 1), ku = __shfl(k, ln - 8);
                const int kul = __shfl(k, ln - 9), kur = __shfl(k, ln - 7);
                const bool wait = pend && ((ix >= 1 && kl == k) || (jy >= 1 && ku == k) ||
                                           (ix >= 1 && jy >= 1 && kul == k) || (ix <= 6 && jy >= 1 && kur == k));
                if (pend && !wait) addStencil(false);
                pend = wait;
                wave_lds_sync_fp();
            }
        } else if (pend) {
            addStencil(true);
        }
        wave_lds_sync_fp();
        // Flush the touched voxels plane by plane; a lane takes elements
        // q = ln, ln + 64, ... of a plane (row ey = q / RP by the reciprocal,
        // exact for q, RP <= 4096).
        const float rRP = __builtin_amdgcn_rcpf((float)RP);
        for (int ez = 0; ez < Dz; ez++) {
            const float *bp = box + ez * PP;
            const long long zbase = (long long)(bz0 + ez) * VV + (long long)by0 * V + bx0;
            for (int q = ln; q < PP; q += 64) {
                const float s = bp[q];
                if (s != 0.f) {
                    const int ey = (int)(((float)q + 0.5f) * rRP);
                    const int ex = q - ey * RP;
                    const long long lin = zbase + (long long)ey * V + ex;
                    if (lin >= 0 && lin < V3) unsafeAtomicAdd(volAcc + lin, s);
                }
            }
        }
        wave_lds_sync_fp();  // the box is read before it is zeroed again
        return true;
    };

    bool done = scatter(0, FPT - 1, true);
    if (!done) {
        const bool top = (ln >> 3) < FPT / 2;
        const bool d0 = scatter(0, FPT / 2 - 1, top);
        const bool d1 = scatter(FPT / 2, FPT - 1, !top);
        done = top ? d0 : d1;
    }
    if (!done && act) scatterGlobal();
}

template <bool FAST>
__global__ __launch_bounds__(kThreads) void adj_lds_kernel(FPArgs p, const float *__restrict__ img,
                                                           float *__restrict__ volAcc)
{
    __shared__ float sker[FPKER];
    __shared__ float sbox[4][FPBOX];
    if (!FAST) {
        for (int t = threadIdx.x; t < p.K; t += kThreads) sker[t] = p.ker[t];
        __syncthreads();
    }
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), ln = threadIdx.x & 63;
    if (FAST && fp_win4(p.check)) {
        for (unsigned work = blockIdx.x * 4u + (unsigned)wv; work < p.totalWork; work += gridDim.x * 4u)
            adj_lds_tile<FAST, true>(p, sker, sbox[wv], img, volAcc, work, ln);
    } else {
        for (unsigned work = blockIdx.x * 4u + (unsigned)wv; work < p.totalWork; work += gridDim.x * 4u)
            adj_lds_tile<FAST, false>(p, sker, sbox[wv], img, volAcc, work, ln);
    }
}

// ---------------------------------------------------------------------------
// Back projection, reference semantics (gpuBackProject.cu:4-113), as an
// owner-computes gather: a workgroup owns an 8x8x4 voxel brick, culls the
// directions whose slab |n.(v-c)| <= kerHWidth cannot touch the brick, keeps
// the survivors in LDS in direction order, and each thread accumulates its own
// voxel.  No atomics: launches on one volume must be stream-ordered.
// ---------------------------------------------------------------------------
struct BPArgs {
    float *vol;
    float *dens;
    const float *axes;
    const float *img;
    const float *ker;
    int V, M, K, nAxes;
    float maskR, kerScale, kerCenter, fW, hw;
    int convW, bricksX, bricksXY;
    int maskD2;  // largest integer d2 with sqrtf((float)d2) <= maskR (-1: none)
    int posRound;  // bp_fast_kernel: in-mask voxels of orthonormal frames project to fi, fj >= 0.5
    int maskLo;    // smallest coordinate of an in-mask voxel (V/2 - floor(sqrt(maskD2)))
    const int *order;  // bp_fast_kernel: brick id per work item, heaviest first
    unsigned perXcd, totalWork;
};

constexpr int BPX = 8, BPY = 8, BPZ = 4;

template <bool DENS>
__global__ __launch_bounds__(kThreads) void bp_kernel(BPArgs p)
{
    __shared__ float sker[kMaxKerSize];
    __shared__ float sax[kThreads * 9];
    __shared__ int sidx[kThreads];
    __shared__ int swave[kThreads / 64];

    const unsigned work = xcd_work(p.perXcd);
    if (work >= p.totalWork) return;
    const int bz = (int)(work / (unsigned)p.bricksXY);
    const int rem = (int)(work - (unsigned)bz * (unsigned)p.bricksXY);
    const int by = rem / p.bricksX, bx = rem - by * p.bricksX;
    const int ox = bx * BPX, oy = by * BPY, oz = bz * BPZ;
    const int V = p.V, M = p.M, cv = V / 2, ci = M / 2;

    // Whole brick outside the mask sphere: nothing to do (uniform exit).
    {
        float dx = fmaxf(fmaxf((float)(ox - cv), (float)(cv - (ox + BPX - 1))), 0.f);
        float dy = fmaxf(fmaxf((float)(oy - cv), (float)(cv - (oy + BPY - 1))), 0.f);
        float dz = fmaxf(fmaxf((float)(oz - cv), (float)(cv - (oz + BPZ - 1))), 0.f);
        if (dx * dx + dy * dy + dz * dz > (p.maskR + 1.f) * (p.maskR + 1.f)) return;
    }
    const int t = threadIdx.x;
    for (int k = t; k < p.K; k += kThreads) sker[k] = p.ker[k];

    const int vi = ox + (t & 7), vj = oy + ((t >> 3) & 7), vk = oz + (t >> 6);
    const bool inVol = vi < V && vj < V && vk < V;
    const float rr = sqrtf((float)(vi - cv) * (vi - cv) + (vj - cv) * (vj - cv) + (vk - cv) * (vk - cv));
    const bool active = inVol && (rr <= p.maskR);
    const float dvi = (float)vi - cv, dvj = (float)vj - cv, dvk = (float)vk - cv;
    const float fvi = (float)vi, fvj = (float)vj, fvk = (float)vk, cvf = (float)cv;
    const float bcx = (float)ox + 0.5f * (BPX - 1) - (float)cv;
    const float bcy = (float)oy + 0.5f * (BPY - 1) - (float)cv;
    const float bcz = (float)oz + 0.5f * (BPZ - 1) - (float)cv;
    const int M2 = M * M;

    double accAll = 0.0, densAll = 0.0;
    const int wv = t >> 6, ln = t & 63;

    for (int base = 0; base < p.nAxes; base += kThreads) {
        const int n = min(kThreads, p.nAxes - base);
        __syncthreads();  // previous list consumed (and table loaded)
        bool pass = false;
        float a9[9];
        if (t < n) {
            const float *ax = p.axes + (size_t)9 * (base + t);
#pragma unroll
            for (int k = 0; k < 9; k++) a9[k] = ax[k];
            const float d = bcx * a9[6] + bcy * a9[7] + bcz * a9[8];
            const float sup = 0.5f * (BPX - 1) * fabsf(a9[6]) + 0.5f * (BPY - 1) * fabsf(a9[7]) +
                              0.5f * (BPZ - 1) * fabsf(a9[8]);
            pass = fabsf(d) <= p.hw + sup + 1e-3f;
        }
        const unsigned long long bal = __ballot(pass);
        if (ln == 0) swave[wv] = __popcll(bal);
        __syncthreads();
        int off = 0, total = 0;
#pragma unroll
        for (int w = 0; w < kThreads / 64; w++) {
            off += (w < wv) ? swave[w] : 0;
            total += swave[w];
        }
        if (pass) {
            const int pos = off + __popcll(bal & ((1ull << ln) - 1ull));
            sidx[pos] = base + t;
#pragma unroll
            for (int k = 0; k < 9; k++) sax[pos * 9 + k] = a9[k];
        }
        __syncthreads();

        for (int q = 0; q < total; q++) {
            const float *ax = sax + q * 9;
            const float n0 = ax[6], n1 = ax[7], n2 = ax[8];
            const float fk = ref_sum3(dvi, n0, dvj, n1, dvk, n2);
            if (!(active && fabsf(fk) <= p.hw)) continue;
            const float u0 = ax[0], u1 = ax[1], u2 = ax[2];
            const float v0 = ax[3], v1 = ax[4], v2 = ax[5];
            const float fi = ref_sum3(dvi, u0, dvj, u1, dvk, u2) + (float)ci;
            const float fj = ref_sum3(dvi, v0, dvj, v1, dvk, v2) + (float)ci;
            const int I = (int)roundf(fi), J = (int)roundf(fj);
            const float *im = p.img + (size_t)sidx[q] * M2;
            float cum = 0.f, dsum = 0.f;
            if (p.convW == 2) {
                // Tap position fl(fma(di, u, fl(dj*v)) + c): the products
                // dj*v are rounded once per row, di stays exact.
                float Di[5], Bx[5], By[5], Bz[5];
#pragma unroll
                for (int s = 0; s < 5; s++) {
                    const float dj = (float)(J - 2 + s - ci);
                    Di[s] = (float)(I - 2 + s - ci);
                    Bx[s] = dj * v0; By[s] = dj * v1; Bz[s] = dj * v2;
                }
#pragma unroll
                for (int jj = 0; jj < 5; jj++) {
#pragma unroll
                    for (int ii = 0; ii < 5; ii++) {
                        const float px = fmaf(Di[ii], u0, Bx[jj]) + cvf;
                        float w = sker[kb_index(px - fvi, p.fW, p.kerScale, p.kerCenter, p.K)];
                        const float py = fmaf(Di[ii], u1, By[jj]) + cvf;
                        w = w * sker[kb_index(py - fvj, p.fW, p.kerScale, p.kerCenter, p.K)];
                        const float pz = fmaf(Di[ii], u2, Bz[jj]) + cvf;
                        w = w * sker[kb_index(pz - fvk, p.fW, p.kerScale, p.kerCenter, p.K)];
                        const int lin = (J - 2 + jj) * M + (I - 2 + ii);
                        if (lin >= 0 && lin < M2) {
                            cum = cum + im[lin] * w;
                            if (DENS) dsum = dsum + w;
                        }
                    }
                }
            } else {
                const int W = p.convW;
                for (int j1 = J - W; j1 <= J + W; j1++) {
                    for (int i1 = I - W; i1 <= I + W; i1++) {
                        const float ti = (float)(i1 - ci), tj = (float)(j1 - ci);
                        const float px = ref_sum2c(ti, u0, tj, v0, cvf);
                        float w = sker[kb_index(px - fvi, p.fW, p.kerScale, p.kerCenter, p.K)];
                        const float py = ref_sum2c(ti, u1, tj, v1, cvf);
                        w = w * sker[kb_index(py - fvj, p.fW, p.kerScale, p.kerCenter, p.K)];
                        const float pz = ref_sum2c(ti, u2, tj, v2, cvf);
                        w = w * sker[kb_index(pz - fvk, p.fW, p.kerScale, p.kerCenter, p.K)];
                        const int lin = j1 * M + i1;
                        if (lin >= 0 && lin < M2) {
                            cum = cum + im[lin] * w;
                            if (DENS) dsum = dsum + w;
                        }
                    }
                }
            }
            accAll += (double)cum;
            if (DENS) densAll += (double)dsum;
        }
    }
    if (active) {
        const size_t vidx = ((size_t)vk * V + vj) * V + vi;
        p.vol[vidx] = (float)((double)p.vol[vidx] + accAll);
        if (DENS) p.dens[vidx] = (float)((double)p.dens[vidx] + densAll);
    }
}

// ---------------------------------------------------------------------------
// Back projection fast path (kerHWidth 2, 501-entry table: kerScale 125,
// kerCenter 250), same reference semantics as bp_kernel.
//
// Work layout.  A workgroup owns an 8x8x8 brick and the list of directions
// whose slab can touch it (culled and compacted in direction order, as in
// bp_kernel).  Wave w takes list entries w, w+4, w+8, ...  For each direction
// the wave first finds the brick voxels inside the slab |fk| <= 2 (lane l
// tests the 8 voxels of column (l & 7, l >> 3)), compacts them into a
// per-wave queue in LDS, and then runs the 25-tap gather on 64 queued voxels
// at a time, so every lane carries an active voxel (a fixed voxel-per-lane
// layout leaves most lanes idle: a slab activates ~1/3 of a brick).
// The queue spans directions (entries carry a 7-bit direction slot), so the
// tail of one direction fills up with the next and groups stay full.  Two
// entries of one group can name the same voxel (different directions), so
// contributions go to a per-wave LDS accumulator with LDS atomic adds; the
// four wave accumulators are combined in a fixed order at the end and added
// to the volume in FP64.
//
// Table index.  The reference, as contracted, takes
// k = roundf(fma(clamp(r, -2, 2), 125, 250)) with r = fl(p - v).  For every
// FP32 r in [-4.5, 4.5], floor(fma(r, 125, 250) + 0.5) saturated to [0, 500]
// equals k (checked exhaustively, with floor(s + 0.5) exact and with s + 0.5
// first rounded to FP32, so the result does not depend on how the instruction
// rounds inside), and v_cvt_rpi_i32_f32 computes floor(x + 0.5) in one
// instruction.  The LDS table is padded on both sides with copies of its end
// entries, so offsets beyond +-2 land on the entry the reference's clamp
// would pick and r needs no clamp.
//
// Mask.  sqrtf(d2) <= maskRadius with d2 a small integer is monotone in d2,
// so the host passes the largest passing d2 and the test is an integer
// compare.
// ---------------------------------------------------------------------------
constexpr int BB = 8;                  // brick edge
constexpr int BV = BB * BB * BB;       // voxels per brick
constexpr int KPAD = 320;              // padding entries on each side of the table
constexpr int KFAST = 501;
constexpr int KTOT = KFAST + 2 * KPAD; // padded table length
constexpr int QCAP = 1024;             // per-wave queue ring (>= BV + 64, power of 2)
constexpr int BPCH = 128;              // directions culled per pass (queue entries hold 9-bit voxel + 7-bit slot)

__device__ __forceinline__ float kb_fast(const float *__restrict__ sk, float r)
{
    int i = cvt_rpi(fmaf(r, 125.0f, 250.0f));  // floor(s + 0.5)
    i = min(max(i, -KPAD), KFAST - 1 + KPAD);
    return sk[KPAD + i];
}

// Same lookup without the clamp, for directions whose frame is orthonormal
// (checked per chunk): then every tap has |r| <= 4.13 and the index stays in
// [-267, 767], inside the padding.
__device__ __forceinline__ float kb_fast_nc(const float *__restrict__ sk, float r)
{
    return sk[KPAD + cvt_rpi(fmaf(r, 125.0f, 250.0f))];
}

// Fused form of the same unclamped lookup for a tap at position p and an
// integer voxel coordinate v >= 3: where the reference's clamp is inactive
// (|p - v| <= 2) p >= 1, so p and v are multiples of ulp(p) and p - v is
// exact; then fma(r, 125, 250) = RN(125 p + 250 - 125 v) = fma(p, 125, c) with
// c = 250 - 125 v an exact integer.  Where |p - v| > 2 both sides saturate at
// table entry 0 or 500.  Checked for every float p in [v - 4.5, v + 4.5] and
// v in 3..4 and 9..2,100 (0 mismatches); v = 0, 1, 2 have mismatches.
constexpr int BP_FORM_A_MIN = 3;

__device__ __forceinline__ float kb_form(const float *__restrict__ sk, float pos, float c)
{
    return sk[KPAD + cvt_rpi(fmaf(pos, 125.0f, c))];
}

__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

template <bool DENS>
__global__ __launch_bounds__(kThreads) void bp_fast_kernel(BPArgs p)
{
    __shared__ float sker[KTOT];
    __shared__ float sax[BPCH * 9];
    __shared__ int sidx[BPCH];
    __shared__ int swave[kThreads / 64];
    __shared__ int snon[kThreads / 64];
    __shared__ float sacc[4][BV];
    __shared__ float sden[DENS ? 4 : 1][DENS ? BV : 1];
    __shared__ unsigned short squeue[4][QCAP];

    // Bricks near the centre meet every direction and bricks near the mask edge
    // few, so the order table puts the heavy (central) bricks first and
    // consecutive work items go round-robin to the XCDs.
    if (blockIdx.x >= p.totalWork) return;
    const unsigned work = (unsigned)p.order[blockIdx.x];
    const int bz = (int)(work / (unsigned)p.bricksXY);
    const int rem = (int)(work - (unsigned)bz * (unsigned)p.bricksXY);
    const int by = rem / p.bricksX, bx = rem - by * p.bricksX;
    const int ox = bx * BB, oy = by * BB, oz = bz * BB;
    const int V = p.V, M = p.M, cv = V / 2, ci = M / 2;
    {
        float dx = fmaxf(fmaxf((float)(ox - cv), (float)(cv - (ox + BB - 1))), 0.f);
        float dy = fmaxf(fmaxf((float)(oy - cv), (float)(cv - (oy + BB - 1))), 0.f);
        float dz = fmaxf(fmaxf((float)(oz - cv), (float)(cv - (oz + BB - 1))), 0.f);
        if (dx * dx + dy * dy + dz * dz > (p.maskR + 1.f) * (p.maskR + 1.f)) return;
    }
    const int t = threadIdx.x;
    for (int k = t; k < KTOT; k += kThreads) sker[k] = p.ker[min(max(k - KPAD, 0), KFAST - 1)];
    for (int k = t; k < 4 * BV; k += kThreads) {
        (&sacc[0][0])[k] = 0.f;
        if (DENS) (&sden[0][0])[k] = 0.f;
    }

    // Orthonormal chunks use the fused lookup (kb_form) when every active
    // voxel of the brick has coordinates >= BP_FORM_A_MIN: the brick's corner
    // does, or the mask keeps every in-mask voxel there.
    const bool brickA = p.posRound &&
                        (min(ox, min(oy, oz)) >= BP_FORM_A_MIN || p.maskLo >= BP_FORM_A_MIN);

    const float cvf = (float)cv, cif = (float)ci;
    const float bc = 0.5f * (BB - 1);
    const float bcx = (float)ox + bc - cvf, bcy = (float)oy + bc - cvf, bcz = (float)oz + bc - cvf;
    const int M2 = M * M;
    const int wv = t >> 6, ln = t & 63;
    // This lane's voxel column for the slab test: (ox + cx, oy + cy, oz + 0..7).
    const int cx = ln & 7, cy = ln >> 3;
    const int vi0 = ox + cx, vj0 = oy + cy;
    unsigned vmask = 0;  // inside the volume and the mask sphere (direction-independent)
#pragma unroll
    for (int k = 0; k < BB; k++) {
        const int vk = oz + k;
        const int d2 = (vi0 - cv) * (vi0 - cv) + (vj0 - cv) * (vj0 - cv) + (vk - cv) * (vk - cv);
        if (vi0 < V && vj0 < V && vk < V && d2 <= p.maskD2) vmask |= 1u << k;
    }
    const float dvi0 = (float)vi0 - cv, dvj0 = (float)vj0 - cv;

    for (int base = 0; base < p.nAxes; base += BPCH) {
        const int n = min(BPCH, p.nAxes - base);
        __syncthreads();
        bool pass = false, orth = true;
        float a9[9];
        if (t < n) {
            const float *ax = p.axes + (size_t)9 * (base + t);
#pragma unroll
            for (int k = 0; k < 9; k++) a9[k] = ax[k];
            const float d = bcx * a9[6] + bcy * a9[7] + bcz * a9[8];
            const float sup = bc * (fabsf(a9[6]) + fabsf(a9[7]) + fabsf(a9[8]));
            pass = fabsf(d) <= p.hw + sup + 1e-3f;
            // [u v n] orthonormal to 1e-4 (rotation matrices are, to ~1e-7)
            const float uu = a9[0] * a9[0] + a9[1] * a9[1] + a9[2] * a9[2];
            const float vv = a9[3] * a9[3] + a9[4] * a9[4] + a9[5] * a9[5];
            const float nn = a9[6] * a9[6] + a9[7] * a9[7] + a9[8] * a9[8];
            const float uv = a9[0] * a9[3] + a9[1] * a9[4] + a9[2] * a9[5];
            const float un = a9[0] * a9[6] + a9[1] * a9[7] + a9[2] * a9[8];
            const float vn = a9[3] * a9[6] + a9[4] * a9[7] + a9[5] * a9[8];
            orth = fabsf(uu - 1.f) <= 1e-4f && fabsf(vv - 1.f) <= 1e-4f && fabsf(nn - 1.f) <= 1e-4f &&
                   fabsf(uv) <= 1e-4f && fabsf(un) <= 1e-4f && fabsf(vn) <= 1e-4f;
        }
        const unsigned long long bal = __ballot(pass);
        const unsigned long long nonOrth = __ballot(!orth);
        if (ln == 0) {
            swave[wv] = __popcll(bal);
            snon[wv] = nonOrth != 0ull;
        }
        __syncthreads();
        const bool chunkOrtho = !(snon[0] | snon[1] | snon[2] | snon[3]);
        int off = 0, total = 0;
#pragma unroll
        for (int w = 0; w < kThreads / 64; w++) {
            off += (w < wv) ? swave[w] : 0;
            total += swave[w];
        }
        if (pass) {
            const int pos = off + __popcll(bal & ((1ull << ln) - 1ull));
            sidx[pos] = base + t;
#pragma unroll
            for (int k = 0; k < 9; k++) sax[pos * 9 + k] = a9[k];
        }
        __syncthreads();

        // Directions of this chunk that wave wv handles: q = wv, wv+4, ...
        // Their active voxels stream through the wave's FIFO queue; every
        // full group of 64 entries is gathered with all lanes busy.
        int head = 0, tail = 0;  // queue indices (uniform), ring of size QCAP
        for (int q = wv; q < total || tail > head; q += 4) {
            if (q < total) {
                const float *ax = sax + q * 9;
                const float n0 = ax[6], n1 = ax[7], n2 = ax[8];
                // Slab test of the lane's column (the reference's contracted
                // fk = fma(dvk, n2, fma(dvi, n0, fl(dvj * n1)))), appended to
                // the queue in (z, lane) order.
                const float fk0 = fmaf(dvi0, n0, dvj0 * n1);
#pragma unroll
                for (int k = 0; k < BB; k++) {
                    const float dvk = (float)(oz + k) - cv;
                    const float fk = fmaf(dvk, n2, fk0);
                    const bool on = ((vmask >> k) & 1u) && fabsf(fk) <= p.hw;
                    const unsigned long long b = __ballot(on);
                    if (on)
                        squeue[wv][(tail + __popcll(b & ((1ull << ln) - 1ull))) & (QCAP - 1)] =
                            (unsigned short)((cx + BB * (cy + BB * k)) | (q << 9));
                    tail += __popcll(b);
                }
            }
            const bool last = q + 4 >= total;
            wave_lds_sync();
            // Gather full groups (and, after the last direction, the rest).
            auto gather = [&](auto formT) {
            while (tail - head >= 64 || (last && tail > head)) {
                const int e = head + ln;
                head = min(head + 64, tail);
                if (e < tail) {
                    const unsigned ent = squeue[wv][e & (QCAP - 1)];
                    const int vox = ent & 511, qq = ent >> 9;
                    const float *ax = sax + qq * 9;
                    const float u0 = ax[0], u1 = ax[1], u2 = ax[2];
                    const float v0 = ax[3], v1 = ax[4], v2 = ax[5];
                    const float *im = p.img + (size_t)sidx[qq] * M2;
                    const int lx = vox & 7, ly = (vox >> 3) & 7, lz = vox >> 6;
                    const int vi = ox + lx, vj = oy + ly, vk = oz + lz;
                    const float dvi = (float)vi - cv, dvj = (float)vj - cv, dvk = (float)vk - cv;
                    const float fi = ref_sum3(dvi, u0, dvj, u1, dvk, u2) + cif;
                    const float fj = ref_sum3(dvi, v0, dvj, v1, dvk, v2) + cif;
                    constexpr int KF = decltype(formT)::value;
                    // The fused variant runs only when posRound holds (fi, fj >= 0.5),
                    // where floor(x + 0.5) is roundf(x) and one instruction.
                    const int I = KF > 0 ? cvt_rpi(fi) : (int)roundf(fi);
                    const int J = KF > 0 ? cvt_rpi(fj) : (int)roundf(fj);
                    const float fvi = (float)vi, fvj = (float)vj, fvk = (float)vk;
                    // Form constants 250 - 125 v (exact; used by the fused variant).
                    const float cX = fmaf(-125.0f, fvi, 250.0f);
                    const float cY = fmaf(-125.0f, fvj, 250.0f);
                    const float cZ = fmaf(-125.0f, fvk, 250.0f);
                    float Di[5];
#pragma unroll
                    for (int s5 = 0; s5 < 5; s5++) Di[s5] = (float)(I - 2 + s5 - ci);
                    float cum = 0.f, dsum = 0.f;
                    const bool inside = I >= 2 && I + 2 < M && J >= 2 && J + 2 < M;
#pragma unroll 1
                    for (int jj = 0; jj < 5; jj++) {
                        const float dj = (float)(J - 2 + jj - ci);
                        const float Bx = dj * v0, By = dj * v1, Bz = dj * v2;
                        float val[5], vin[5];  // image values; 1 where the tap is in bounds, else 0
                        const int row = (J - 2 + jj) * M + (I - 2);
                        if (inside) {
                            f4 a;
                            __builtin_memcpy(&a, im + row, 16);
                            val[0] = a.x; val[1] = a.y; val[2] = a.z; val[3] = a.w;
                            val[4] = im[row + 4];
#pragma unroll
                            for (int ii = 0; ii < 5; ii++) vin[ii] = 1.f;
                        } else {
                            // Reference bound on the linear index only
                            // (gpuBackProject.cu:99): wraps across rows.
#pragma unroll
                            for (int ii = 0; ii < 5; ii++) {
                                const int lin = row + ii;
                                const bool ok = lin >= 0 && lin < M2;
                                val[ii] = ok ? im[lin] : 0.f;
                                vin[ii] = ok ? 1.f : 0.f;
                            }
                        }
#pragma unroll
                        for (int ii = 0; ii < 5; ii++) {
                            // fl(fma(di, u, fl(dj*v)) + c), the reference's contracted
                            // tap position (gpuBackProject.cu:77,84,91).
                            const float px = fmaf(Di[ii], u0, Bx) + cvf;
                            const float py = fmaf(Di[ii], u1, By) + cvf;
                            const float pz = fmaf(Di[ii], u2, Bz) + cvf;
                            float w, wy, wz;
                            if constexpr (KF < 0) {
                                w = kb_fast(sker, px - fvi);
                                wy = kb_fast(sker, py - fvj);
                                wz = kb_fast(sker, pz - fvk);
                            } else if constexpr (KF == 0) {
                                w = kb_fast_nc(sker, px - fvi);
                                wy = kb_fast_nc(sker, py - fvj);
                                wz = kb_fast_nc(sker, pz - fvk);
                            } else {
                                w = kb_form(sker, px, cX);
                                wy = kb_form(sker, py, cY);
                                wz = kb_form(sker, pz, cZ);
                            }
                            w = w * wy;
                            w = w * wz;
                            cum = fmaf(val[ii], w, cum);
                            // the density: the back projection of all-ones images
                            // (fma(1, w, d) = d + w, fma(0, w, d) = d)
                            if (DENS) dsum = fmaf(vin[ii], w, dsum);
                        }
                    }
                    // Two queued entries of one group can name the same voxel
                    // (different directions): LDS atomic add.
                    atomicAdd(&sacc[wv][vox], cum);
                    if (DENS) atomicAdd(&sden[wv][vox], dsum);
                }
            }
            };
            if (!chunkOrtho) {
                gather(std::integral_constant<int, -1>{});
            } else if (brickA) {
                gather(std::integral_constant<int, 1>{});
            } else {
                gather(std::integral_constant<int, 0>{});
            }
            wave_lds_sync();  // queue slots consumed before they are refilled
            if (last) break;
        }
    }
    __syncthreads();
    for (int k = t; k < BV; k += kThreads) {
        const int lx = k & 7, ly = (k >> 3) & 7, lz = k >> 6;
        const int vi = ox + lx, vj = oy + ly, vk = oz + lz;
        if (vi >= V || vj >= V || vk >= V) continue;
        const int d2 = (vi - cv) * (vi - cv) + (vj - cv) * (vj - cv) + (vk - cv) * (vk - cv);
        if (d2 > p.maskD2) continue;
        const size_t vidx = ((size_t)vk * V + vj) * V + vi;
        double s = (double)p.vol[vidx];
#pragma unroll
        for (int w = 0; w < 4; w++) s += (double)sacc[w][k];
        p.vol[vidx] = (float)s;
        if (DENS) {
            double d = (double)p.dens[vidx];
#pragma unroll
            for (int w = 0; w < 4; w++) d += (double)sden[w][k];
            p.dens[vidx] = (float)d;
        }
    }
}

// ---------------------------------------------------------------------------
// Image chains.  With even M, fftshift o DFT o fftshift equals the plain DFT
// with a (-1)^(x+y) checkerboard on input and output (the 2-D global sign
// (-1)^M is +1), so the shifts of gpuGridder.cpp:463-488 and :562-587 fold
// into these loads and stores.  The CAS-derived spectrum is exactly Hermitian,
// so the inverse uses C2R and the forward R2C (half the FFT work).
// ---------------------------------------------------------------------------
__device__ __forceinline__ float chk(int s) { return (s & 1) ? -1.f : 1.f; }

// The 2-D chain kernels run one workgroup per (row, image): blockIdx.x is the
// row, blockIdx.y strides over the images, the threads over the row, so every
// index is 32-bit arithmetic on the row and column (no per-element division)
// and each row is read and written contiguously.

// CASToComplexFilter.cu:53-75 (2-D stack mode) + fftshift, half spectrum.
__global__ void k_cas_to_herm2d(const float *__restrict__ cas, float2 *__restrict__ herm, int M, int n)
{
    const int H = M / 2 + 1;
    const int cy = blockIdx.x;
    for (int a = blockIdx.y; a < n; a += gridDim.y) {
        const float *c = cas + (size_t)a * M * M;
        float2 *h = herm + ((size_t)a * M + cy) * H;
        for (int cx = threadIdx.x; cx < H; cx += blockDim.x) {
            float2 z = make_float2(0.f, 0.f);
            if (cx != 0 && cy != 0) {
                const float c1 = c[cy * M + cx];
                const float c2 = c[(M - cy) * M + (M - cx)];
                const float sg = chk(cx + cy);
                z.x = sg * (0.5f * (c1 + c2));
                z.y = sg * (0.5f * (c1 - c2));
            }
            h[cx] = z;
        }
    }
}

// fftshift + ComplexToReal + CropVolume + DivideScalar (gpuGridder.cpp:485-517).
__global__ void k_real2d_to_imgs(const float *__restrict__ real, float *__restrict__ imgs, int M,
                                 int N, float scale, int n)
{
    const int off = (M - N) / 2;
    const int py = blockIdx.x, my = py + off;
    for (int a = blockIdx.y; a < n; a += gridDim.y) {
        const float *r = real + ((size_t)a * M + my) * M + off;
        float *o = imgs + ((size_t)a * N + py) * N;
        for (int px = threadIdx.x; px < N; px += blockDim.x)
            o[px] = chk(px + off + my) * r[px] * scale;
    }
}

// PadVolume + RealToComplex + fftshift (gpuGridder.cpp:543-566).
__global__ void k_imgs_to_real2d(const float *__restrict__ imgs, float *__restrict__ real, int M,
                                 int N, int n)
{
    const int off = (M - N) / 2;
    const int my = blockIdx.x, py = my - off;
    const bool rowIn = py >= 0 && py < N;
    for (int a = blockIdx.y; a < n; a += gridDim.y) {
        const float *im = imgs + ((size_t)a * N + (rowIn ? py : 0)) * N;
        float *o = real + ((size_t)a * M + my) * M;
        for (int mx = threadIdx.x; mx < M; mx += blockDim.x) {
            const int px = mx - off;
            float v = 0.f;
            if (rowIn && px >= 0 && px < N) v = chk(mx + my) * im[px];
            o[mx] = v;
        }
    }
}

// fftshift + ComplexToCAS (gpuGridder.cpp:584-595) from the R2C half spectrum.
__global__ void k_herm2d_to_cas(const float2 *__restrict__ herm, float *__restrict__ cas, int M, int n,
                                float scale, int zeroNyq)
{
    const int H = M / 2 + 1;
    const int cy = blockIdx.x, my = (M - cy) % M;
    for (int a = blockIdx.y; a < n; a += gridDim.y) {
        const float2 *h = herm + (size_t)a * M * H;
        const float2 *hr = h + (size_t)cy * H, *hm = h + (size_t)my * H;
        float *o = cas + ((size_t)a * M + cy) * M;
        for (int cx = threadIdx.x; cx < M; cx += blockDim.x) {
            float re, im;
            if (cx <= M / 2) {
                const float2 f = hr[cx];
                re = f.x; im = f.y;
            } else {
                const float2 f = hm[M - cx];
                re = f.x; im = -f.y;
            }
            const float sg = chk(cx + cy);
            // zeroNyq: the index-0 row / column FromCAS discards (transpose of
            // the inverse image chain); scale is exact (1) on the forward chain.
            o[cx] = (zeroNyq && (cx == 0 || cy == 0)) ? 0.f : (sg * re + sg * im) * scale;
        }
    }
}

// ---------------------------------------------------------------------------
// Fused image chains for M = 128, 192, 256, 384 and 512.  The general chains above make
// five or six HBM passes per image (CAS -> Hermitian half spectrum, two rocFFT
// passes, crop; or the reverse).  These make two each, with the M-point FFT
// done in LDS as four-step 16 x R (M = 16 R; 16 threads per transform, the
// R-point and 16-point DFTs in registers, one twiddled exchange through LDS;
// the 32-point DFT is two 16-point DFTs and a radix-2 step, the 8-point DFT
// two 4-point DFTs and a radix-2 step, the 12- and 24-point DFTs 4- and
// 8-point DFTs and a radix-3 step), and are ordered
// so that CAS images are read and written as whole rows:
//   CAS images -> images: (1) k_inv_rows: per spectrum row ky <= M/2 the full
//   complex row is built from CAS rows ky and M - ky as they are read,
//   inverse-transformed along x, and the N cropped columns are kept
//   ((M/2 + 1) x N complex per image; rows ky > M/2 follow by Hermitian
//   symmetry); (2) k_inv_cols: 32 columns per workgroup, two Hermitian
//   columns per complex transform along y (its real and imaginary outputs are
//   the two real columns), checkerboard and 1/M^2, written through LDS.
//   images -> CAS images: (1) k_fwd_cols: two zero-padded image columns per
//   complex transform along y, split into two half spectra; (2) k_fwd_rows:
//   per spectrum row ky <= M/2 the transform along x gives the full row, and
//   CAS rows ky and M - ky are written whole.
// Same arithmetic as the general chains up to FP32 rounding in the FFT
// (twiddles from sincospif, accurate to ~1 ulp).
// ---------------------------------------------------------------------------
struct F16Tw { float c, s; };
__device__ __constant__ const F16Tw kTw16[16] = {
    {1.0f, 0.0f},
    {0.92387953251128674f, 0.38268343236508977f},
    {0.70710678118654752f, 0.70710678118654752f},
    {0.38268343236508977f, 0.92387953251128674f},
    {0.0f, 1.0f},
    {-0.38268343236508977f, 0.92387953251128674f},
    {-0.70710678118654752f, 0.70710678118654752f},
    {-0.92387953251128674f, 0.38268343236508977f},
    {-1.0f, 0.0f},
    {-0.92387953251128674f, -0.38268343236508977f},
    {-0.70710678118654752f, -0.70710678118654752f},
    {-0.38268343236508977f, -0.92387953251128674f},
    {0.0f, -1.0f},
    {0.38268343236508977f, -0.92387953251128674f},
    {0.70710678118654752f, -0.70710678118654752f},
    {0.92387953251128674f, -0.38268343236508977f}};

__device__ __forceinline__ float2 cmulw(float2 a, float c, float s)
{
    return make_float2(fmaf(a.x, c, -a.y * s), fmaf(a.x, s, a.y * c));
}

// 4-point DFT in place, X[k] = sum_n x[n] e^{S 2 pi i n k / 4}.
template <int S>
__device__ __forceinline__ void fft4(float2 &a, float2 &b, float2 &c, float2 &d)
{
    const float2 t0 = make_float2(a.x + c.x, a.y + c.y), t1 = make_float2(a.x - c.x, a.y - c.y);
    const float2 t2 = make_float2(b.x + d.x, b.y + d.y), t3 = make_float2(b.x - d.x, b.y - d.y);
    const float2 r = S > 0 ? make_float2(-t3.y, t3.x) : make_float2(t3.y, -t3.x);
    a = make_float2(t0.x + t2.x, t0.y + t2.y);
    c = make_float2(t0.x - t2.x, t0.y - t2.y);
    b = make_float2(t1.x + r.x, t1.y + r.y);
    d = make_float2(t1.x - r.x, t1.y - r.y);
}

// 16-point DFT in registers (radix 4 x 4).  Input v[n] in natural order;
// output X[k2 + 4 k1] is left in v[4 k2 + k1].
template <int S>
__device__ __forceinline__ void fft16(float2 (&v)[16])
{
#pragma unroll
    for (int n1 = 0; n1 < 4; n1++) fft4<S>(v[n1], v[n1 + 4], v[n1 + 8], v[n1 + 12]);
#pragma unroll
    for (int n1 = 1; n1 < 4; n1++)
#pragma unroll
        for (int k2 = 1; k2 < 4; k2++) {
            const F16Tw w = kTw16[n1 * k2];
            v[n1 + 4 * k2] = cmulw(v[n1 + 4 * k2], w.c, S > 0 ? w.s : -w.s);
        }
#pragma unroll
    for (int k2 = 0; k2 < 4; k2++) fft4<S>(v[4 * k2], v[4 * k2 + 1], v[4 * k2 + 2], v[4 * k2 + 3]);
}

__device__ __forceinline__ constexpr int fft16_slot(int k) { return 4 * (k & 3) + (k >> 2); }

// cos, sin of 2 pi j / 32, j = 0..15
__device__ __constant__ const F16Tw kTw32[16] = {
    {1.0f, 0.0f},
    {0.98078528040323043f, 0.19509032201612825f},
    {0.92387953251128674f, 0.38268343236508977f},
    {0.83146961230254524f, 0.55557023301960218f},
    {0.70710678118654752f, 0.70710678118654752f},
    {0.55557023301960218f, 0.83146961230254524f},
    {0.38268343236508977f, 0.92387953251128674f},
    {0.19509032201612825f, 0.98078528040323043f},
    {0.0f, 1.0f},
    {-0.19509032201612825f, 0.98078528040323043f},
    {-0.38268343236508977f, 0.92387953251128674f},
    {-0.55557023301960218f, 0.83146961230254524f},
    {-0.70710678118654752f, 0.70710678118654752f},
    {-0.83146961230254524f, 0.55557023301960218f},
    {-0.92387953251128674f, 0.38268343236508977f},
    {-0.98078528040323043f, 0.19509032201612825f}};

// 32-point DFT in registers: two 16-point DFTs (even and odd inputs) and a
// radix-2 combination.  X[k] is left in v[fft16_slot(k)], X[k + 16] in
// v[16 + fft16_slot(k)].
template <int S>
__device__ __forceinline__ void fft32(float2 (&v)[32])
{
    float2 e[16], o[16];
#pragma unroll
    for (int i = 0; i < 16; i++) {
        e[i] = v[2 * i];
        o[i] = v[2 * i + 1];
    }
    fft16<S>(e);
    fft16<S>(o);
#pragma unroll
    for (int k = 0; k < 16; k++) {
        const F16Tw w = kTw32[k];
        const float2 E = e[fft16_slot(k)], O = cmulw(o[fft16_slot(k)], w.c, S > 0 ? w.s : -w.s);
        v[fft16_slot(k)] = make_float2(E.x + O.x, E.y + O.y);
        v[16 + fft16_slot(k)] = make_float2(E.x - O.x, E.y - O.y);
    }
}

// 8-point DFT in registers: two 4-point DFTs (even and odd inputs) and a
// radix-2 combination; input and output in natural order.
template <int S>
__device__ __forceinline__ void fft8(float2 (&v)[8])
{
    float2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    float2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    fft4<S>(e0, e1, e2, e3);
    fft4<S>(o0, o1, o2, o3);
    constexpr float h = 0.70710678118654752f;  // cos(pi / 4)
    o1 = cmulw(o1, h, S > 0 ? h : -h);
    o2 = S > 0 ? make_float2(-o2.y, o2.x) : make_float2(o2.y, -o2.x);
    o3 = cmulw(o3, -h, S > 0 ? h : -h);
    v[0] = make_float2(e0.x + o0.x, e0.y + o0.y);
    v[4] = make_float2(e0.x - o0.x, e0.y - o0.y);
    v[1] = make_float2(e1.x + o1.x, e1.y + o1.y);
    v[5] = make_float2(e1.x - o1.x, e1.y - o1.y);
    v[2] = make_float2(e2.x + o2.x, e2.y + o2.y);
    v[6] = make_float2(e2.x - o2.x, e2.y - o2.y);
    v[3] = make_float2(e3.x + o3.x, e3.y + o3.y);
    v[7] = make_float2(e3.x - o3.x, e3.y - o3.y);
}

// cos, sin of 2 pi j / 24, j = 0..14
__device__ __constant__ const F16Tw kTw24[15] = {
    {1.00000000000000000f, 0.00000000000000000f},
    {0.96592582628906831f, 0.25881904510252074f},
    {0.86602540378443871f, 0.49999999999999994f},
    {0.70710678118654757f, 0.70710678118654746f},
    {0.50000000000000011f, 0.86602540378443860f},
    {0.25881904510252074f, 0.96592582628906831f},
    {0.00000000000000006f, 1.00000000000000000f},
    {-0.25881904510252063f, 0.96592582628906831f},
    {-0.49999999999999978f, 0.86602540378443871f},
    {-0.70710678118654746f, 0.70710678118654757f},
    {-0.86602540378443871f, 0.49999999999999994f},
    {-0.96592582628906820f, 0.25881904510252102f},
    {-1.00000000000000000f, 0.00000000000000012f},
    {-0.96592582628906831f, -0.25881904510252079f},
    {-0.86602540378443882f, -0.49999999999999972f}};

// 3-point DFT in place (natural order).
template <int S>
__device__ __forceinline__ void fft3(float2 &a, float2 &b, float2 &c)
{
    constexpr float s3 = 0.86602540378443865f;  // sin(2 pi / 3)
    const float2 t1 = make_float2(b.x + c.x, b.y + c.y);
    const float2 t2 = make_float2(a.x - 0.5f * t1.x, a.y - 0.5f * t1.y);
    const float2 d = make_float2(b.x - c.x, b.y - c.y);
    const float2 r = S > 0 ? make_float2(-s3 * d.y, s3 * d.x) : make_float2(s3 * d.y, -s3 * d.x);
    a = make_float2(a.x + t1.x, a.y + t1.y);
    b = make_float2(t2.x + r.x, t2.y + r.y);
    c = make_float2(t2.x - r.x, t2.y - r.y);
}

// 12- and 24-point DFTs in registers, natural order in and out: n = 3 p + q,
// k = k1 + (R / 3) k2; a (R / 3)-point DFT over p for each q, twiddles
// w_R^(q k1), then a 3-point DFT over q for each k1.
template <int S>
__device__ __forceinline__ void fft12(float2 (&v)[12])
{
    float2 b[3][4];
#pragma unroll
    for (int q = 0; q < 3; q++) {
        float2 a0 = v[q], a1 = v[3 + q], a2 = v[6 + q], a3 = v[9 + q];
        fft4<S>(a0, a1, a2, a3);
        b[q][0] = a0; b[q][1] = a1; b[q][2] = a2; b[q][3] = a3;
#pragma unroll
        for (int k1 = 1; k1 < 4; k1++) {
            if (q == 0) continue;
            const F16Tw w = kTw24[2 * q * k1];
            b[q][k1] = cmulw(b[q][k1], w.c, S > 0 ? w.s : -w.s);
        }
    }
#pragma unroll
    for (int k1 = 0; k1 < 4; k1++) {
        float2 x0 = b[0][k1], x1 = b[1][k1], x2 = b[2][k1];
        fft3<S>(x0, x1, x2);
        v[k1] = x0; v[k1 + 4] = x1; v[k1 + 8] = x2;
    }
}

template <int S>
__device__ __forceinline__ void fft24(float2 (&v)[24])
{
    float2 b[3][8];
#pragma unroll
    for (int q = 0; q < 3; q++) {
#pragma unroll
        for (int p8 = 0; p8 < 8; p8++) b[q][p8] = v[3 * p8 + q];
        fft8<S>(b[q]);
#pragma unroll
        for (int k1 = 1; k1 < 8; k1++) {
            if (q == 0) continue;
            const F16Tw w = kTw24[q * k1];
            b[q][k1] = cmulw(b[q][k1], w.c, S > 0 ? w.s : -w.s);
        }
    }
#pragma unroll
    for (int k1 = 0; k1 < 8; k1++) {
        float2 x0 = b[0][k1], x1 = b[1][k1], x2 = b[2][k1];
        fft3<S>(x0, x1, x2);
        v[k1] = x0; v[k1 + 8] = x1; v[k1 + 16] = x2;
    }
}

template <int R, int S>
__device__ __forceinline__ void fft_regs(float2 (&v)[R])
{
    if constexpr (R == 8) fft8<S>(v);
    else if constexpr (R == 12) fft12<S>(v);
    else if constexpr (R == 16) fft16<S>(v);
    else if constexpr (R == 24) fft24<S>(v);
    else fft32<S>(v);
}

// Position of X[k] in v[] after fft_regs<R>.
template <int R>
__device__ __forceinline__ constexpr int fftR_slot(int k)
{
    return (R == 8 || R == 12 || R == 24) ? k
         : (R == 16 ? fft16_slot(k) : (k < 16 ? fft16_slot(k) : 16 + fft16_slot(k - 16)));
}

// After fft_m<R>, thread t holds X[out_idx<R>(t, m)] in v[out_slot<R>(m)],
// m < R (see there).
template <int R>
__device__ __forceinline__ constexpr int out_slot(int m)
{
    return (R == 8 || R == 12 || R == 24) ? m : (R == 16 ? fft16_slot(m) : 16 * (m & 1) + fft16_slot(m >> 1));
}

template <int R>
__device__ __forceinline__ int out_idx(int t, int m)
{
    if constexpr (R == 12 || R == 24) {
        constexpr int D = R == 12 ? 4 : 2, L = 16 / D;
        const int P = t + 16 * (m / L), mm = m % L;
        return P % R + R * (D * mm + P / R);
    } else {
        return t + 16 * m;
    }
}

// Index of the mirrored row or column, (M - k) mod M, for 0 <= k < M.
template <int M>
__device__ __forceinline__ int mirror_idx(int k)
{
    if constexpr ((M & (M - 1)) == 0) return (M - k) & (M - 1);
    else return k == 0 ? 0 : M - k;
}

constexpr int FFP = 17;  // exchange pitch (float2)
template <int R> struct FftGeom {
    static constexpr int M = 16 * R, H = M / 2 + 1;
    // LDS float2 per transform (>= M); the odd extra float2 puts the
    // transforms of a workgroup on different banks, where 16 * 17 * 2 floats
    // (0 mod 32) put them all on one (the column fills write one element of
    // each of 16 transforms per row; 3-5% faster chains at M = 256).
    static constexpr int FFB = R * FFP + 1;
};

// M-point DFT (M = 16 R) by the 16 threads t = 0..15 of one transform,
// four-step 16 x R: R-point DFTs in registers, twiddles, exchange through
// LDS, 16-point DFTs (outputs X[k2 + R k1], k2 < R, k1 < 16).  On entry v[n2]
// holds x[t + 16 n2] (n2 < R); on exit v[out_slot<R>(m)] holds
// X[out_idx<R>(t, m)] (m < R).  With R = 16, 32 thread t does the 16-point
// DFTs of k2 = t + 16 j (out_idx = t + 16 m).  Otherwise each 16-point DFT is
// shared by D threads through one radix-D decimation in frequency: the share
// h < D takes the outputs k1 = D mm + h, an L = 16 / D point DFT over n of
// w16^(n h) sum_e z[n + L e] w_D^(e h).  R = 8: D = 2, thread t takes
// k2 = t % 8, h = t / 8 (out_idx = t + 16 m again); R = 12 (D = 4) and
// R = 24 (D = 2): thread t takes the three (k2, h) = (P % R, P / R),
// P = t + 16 j, j < 3.
// ex is the transform's FFB-float2 LDS exchange area, tw the M twiddles
// e^{2 pi i q / M}.  Contains workgroup barriers: every thread of the block
// must call it.
template <int R, int S>
__device__ __forceinline__ void fft_m(float2 (&v)[R], int t, float2 *ex, const float2 *tw)
{
    fft_regs<R, S>(v);
#pragma unroll
    for (int k2 = 0; k2 < R; k2++) {
        const float2 w = tw[t * k2];
        ex[k2 * FFP + t] = cmulw(v[fftR_slot<R>(k2)], w.x, S > 0 ? w.y : -w.y);
    }
    __syncthreads();
    if constexpr (R == 8) {
        const int k2 = t & 7, hf = t >> 3;
        const float sg = hf ? -1.f : 1.f;
#pragma unroll
        for (int n = 0; n < 8; n++) {
            const float2 z0 = ex[k2 * FFP + n], z1 = ex[k2 * FFP + n + 8];
            const float2 d = make_float2(fmaf(sg, z1.x, z0.x), fmaf(sg, z1.y, z0.y));
            const float2 w = tw[((16 * R) / 16) * n * hf];  // w16^(n h)
            v[n] = cmulw(d, w.x, S > 0 ? w.y : -w.y);
        }
        fft8<S>(v);
    } else if constexpr (R == 12 || R == 24) {
        constexpr int D = R == 12 ? 4 : 2, L = 16 / D;
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const int P = t + 16 * j, k2 = P % R, h = P / R;
            float2 b[L];
#pragma unroll
            for (int n = 0; n < L; n++) {
                float2 sm;
                if constexpr (D == 2) {
                    const float2 z0 = ex[k2 * FFP + n], z1 = ex[k2 * FFP + n + 8];
                    const float sg = h ? -1.f : 1.f;
                    sm = make_float2(fmaf(sg, z1.x, z0.x), fmaf(sg, z1.y, z0.y));
                } else {
                    sm = ex[k2 * FFP + n];
#pragma unroll
                    for (int e = 1; e < 4; e++) {
                        // z * (S i)^(e h): quarter turns, counter-clockwise for S > 0
                        const float2 z = ex[k2 * FFP + n + 4 * e];
                        const int ph = (S > 0 ? e * h : 4 * e - e * h) & 3;
                        const float2 r = ph == 0 ? z : ph == 1 ? make_float2(-z.y, z.x)
                                       : ph == 2 ? make_float2(-z.x, -z.y) : make_float2(z.y, -z.x);
                        sm = make_float2(sm.x + r.x, sm.y + r.y);
                    }
                }
                const float2 w = tw[R * n * h];  // w16^(n h)
                b[n] = cmulw(sm, w.x, S > 0 ? w.y : -w.y);
            }
            if constexpr (L == 4) fft4<S>(b[0], b[1], b[2], b[3]);
            else fft8<S>(b);
#pragma unroll
            for (int mm = 0; mm < L; mm++) v[j * L + mm] = b[mm];
        }
    } else {
#pragma unroll
        for (int j = 0; j < R / 16; j++) {
            float2 u[16];
#pragma unroll
            for (int n1 = 0; n1 < 16; n1++) u[n1] = ex[(t + 16 * j) * FFP + n1];
            fft16<S>(u);
#pragma unroll
            for (int q = 0; q < 16; q++) v[16 * j + q] = u[q];
        }
    }
    __syncthreads();  // the exchange area may be reused
}

template <int R>
__device__ __forceinline__ void load_tw(float2 *tw)
{
    constexpr int M = 16 * R;
    for (int q = threadIdx.x; q < M; q += blockDim.x) {
        float sn, cs;
        sincospif((float)q / (float)(M / 2), &sn, &cs);
        tw[q] = make_float2(cs, sn);
    }
}

// CAS images -> R[a][ky][mx'] (ky in [0, M/2 + 1), mx' in [0, N)): per spectrum
// row ky the full complex row is built from CAS rows ky and M - ky as they are
// read (whole rows), transformed along x, and the N kept columns are stored.
// Rows ky > M/2 are not needed: R[M - ky] = conj R[ky].  Grid (row groups of
// 16, images).
template <int R>
__global__ __launch_bounds__(256) void k_inv_rows(const float *__restrict__ cas, float2 *__restrict__ tmp,
                                                  int n, int N)
{
    using G = FftGeom<R>;
    constexpr int M = G::M, H = G::H;
    __shared__ float2 sx[16 * G::FFB];
    __shared__ float2 tw[M];
    load_tw<R>(tw);
    const int tid = threadIdx.x, f = tid >> 4, t = tid & 15;
    const int off = (M - N) / 2;
    const int ky = 16 * blockIdx.x + f;
    const bool rowOk = ky <= M / 2;
    float2 *ex = sx + f * G::FFB;
    for (int a = blockIdx.y; a < n; a += gridDim.y) {
        const float *rowA = cas + (size_t)a * M * M + (size_t)(rowOk ? ky : 0) * M;
        const float *rowB = cas + (size_t)a * M * M + (size_t)mirror_idx<M>(ky) * M;
        // Unconditional loads (every address is inside the image), so that all
        // 2R are in flight before the first use; the CAS index 0 row and column
        // and rows beyond M/2 select zero afterwards.
        float c1[R], c2[R];
#pragma unroll
        for (int n2 = 0; n2 < R; n2++) {
            const int kx = t + 16 * n2;
            c1[n2] = rowA[kx];
            c2[n2] = rowB[mirror_idx<M>(kx)];
        }
        float2 v[R];
#pragma unroll
        for (int n2 = 0; n2 < R; n2++) {
            const int kx = t + 16 * n2;
            const float sg = (rowOk && kx != 0 && ky != 0) ? chk(kx + ky) : 0.f;
            v[n2] = make_float2(sg * (0.5f * (c1[n2] + c2[n2])), sg * (0.5f * (c1[n2] - c2[n2])));
        }
        __syncthreads();
        fft_m<R, 1>(v, t, ex, tw);
        if (rowOk) {
            float2 *o = tmp + ((size_t)a * H + ky) * N;
#pragma unroll
            for (int m = 0; m < R; m++) {
                const int px = out_idx<R>(t, m) - off;
                if (px >= 0 && px < N) o[px] = v[out_slot<R>(m)];
            }
        }
    }
}

// R -> images: 32 columns per workgroup, two Hermitian columns packed into one
// complex transform along y (its real and imaginary outputs are the two real
// columns); output through LDS as whole row segments.  Grid (ceil(N / 32),
// images).
template <int R>
__global__ __launch_bounds__(256) void k_inv_cols(const float2 *__restrict__ tmp, float *__restrict__ imgs,
                                                  int n, int N, float scale)
{
    using G = FftGeom<R>;
    constexpr int M = G::M, H = G::H;
    __shared__ float2 sx[16 * G::FFB];
    __shared__ float2 tw[M];
    load_tw<R>(tw);
    const int tid = threadIdx.x, f = tid >> 4, t = tid & 15;
    const int off = (M - N) / 2;
    const int c0 = 32 * blockIdx.x;
    float2 *ex = sx + f * G::FFB;
    float *so = reinterpret_cast<float *>(sx);  // output staging, pitch 33 floats (N * 33 <= 32 * FFB)
    for (int a = blockIdx.y; a < n; a += gridDim.y) {
        __syncthreads();
        // Z_f[ky] = R[ky][c0 + 2f] + i R[ky][c0 + 2f + 1], and by symmetry
        // Z_f[M - ky] = conj R[ky][2f] + i conj R[ky][2f + 1].  Eight loads per
        // thread in flight before the first exchange and LDS store (lanes 2j
        // and 2j + 1 hold neighbouring columns of one row: both in range or
        // both not, since H * 32 is even).
        constexpr int LB = 8;
        for (int base = tid; base < H * 32; base += 256 * LB) {
            float2 vals[LB];
#pragma unroll
            for (int q = 0; q < LB; q++) {
                // unconditional load from a clamped (in-range) address
                const int idx = base + 256 * q, ky = idx >> 5, px = c0 + (idx & 31);
                const float2 raw = tmp[((size_t)a * H + min(ky, H - 1)) * N + min(px, N - 1)];
                vals[q] = (idx < H * 32 && px < N) ? raw : make_float2(0.f, 0.f);
            }
#pragma unroll
            for (int q = 0; q < LB; q++) {
                const int idx = base + 256 * q, ky = idx >> 5, col = idx & 31;
                const float2 val = vals[q];
                const float ox = __shfl_xor(val.x, 1), oy = __shfl_xor(val.y, 1);
                if (idx < H * 32 && !(col & 1)) {
                    float2 *zf = sx + (col >> 1) * G::FFB;
                    zf[ky] = make_float2(val.x - oy, val.y + ox);
                    if (ky != 0 && ky != M / 2) zf[M - ky] = make_float2(val.x + oy, ox - val.y);
                }
            }
        }
        __syncthreads();
        float2 v[R];
#pragma unroll
        for (int n2 = 0; n2 < R; n2++) v[n2] = ex[t + 16 * n2];
        __syncthreads();
        fft_m<R, 1>(v, t, ex, tw);
#pragma unroll
        for (int m = 0; m < R; m++) {
                const int my = out_idx<R>(t, m), py = my - off;
                if (py >= 0 && py < N) {
                    const float2 z = v[out_slot<R>(m)];
                    const int mx = c0 + 2 * f + off;
                    so[py * 33 + 2 * f] = chk(mx + my) * z.x * scale;
                    so[py * 33 + 2 * f + 1] = chk(mx + 1 + my) * z.y * scale;
                }
            }
        __syncthreads();
        for (int idx = tid; idx < N * 32; idx += 256) {
            const int py = idx >> 5, col = idx & 31, px = c0 + col;
            if (px < N) imgs[((size_t)a * N + py) * N + px] = so[py * 33 + col];
        }
    }
}

// Images -> C[a][ky][px] (ky in [0, M/2 + 1)): 32 image columns per workgroup,
// two zero-padded real columns per complex transform along y, split into the
// two half spectra; input through LDS from whole row segments.  Grid
// (ceil(N / 32), images).
template <int R>
__global__ __launch_bounds__(256) void k_fwd_cols(const float *__restrict__ imgs, float2 *__restrict__ tmp,
                                                  int n, int N)
{
    using G = FftGeom<R>;
    constexpr int M = G::M, H = G::H;
    __shared__ float2 sx[16 * G::FFB];
    __shared__ float2 tw[M];
    load_tw<R>(tw);
    const int tid = threadIdx.x, f = tid >> 4, t = tid & 15;
    const int off = (M - N) / 2;
    const int c0 = 32 * blockIdx.x;
    float2 *ex = sx + f * G::FFB;
    float *sf = reinterpret_cast<float *>(sx);
    for (int a = blockIdx.y; a < n; a += gridDim.y) {
        __syncthreads();
        // Column c0 + col goes to transform col / 2, real (even) or imaginary
        // part.  The N image rows are loaded eight per thread at a time, all
        // eight in flight before the first LDS store; the M - N padding rows
        // are zeros.
        constexpr int LB = 8;
        for (int base = tid; base < N * 32; base += 256 * LB) {
            float xv[LB];
#pragma unroll
            for (int q = 0; q < LB; q++) {
                // unconditional load from a clamped (in-image) address, so that
                // no branch separates the eight loads
                const int idx = base + 256 * q, py = idx >> 5, col = idx & 31, px = c0 + col;
                const float raw = imgs[((size_t)a * N + min(py, N - 1)) * N + min(px, N - 1)];
                xv[q] = (idx < N * 32 && px < N) ? chk(px + off + py + off) * raw : 0.f;
            }
#pragma unroll
            for (int q = 0; q < LB; q++) {
                const int idx = base + 256 * q, col = idx & 31;
                if (idx < N * 32) sf[((col >> 1) * G::FFB + (idx >> 5) + off) * 2 + (col & 1)] = xv[q];
            }
        }
        for (int idx = tid; idx < (M - N) * 32; idx += 256) {
            const int r = idx >> 5, col = idx & 31, my = r < off ? r : r + N;
            sf[((col >> 1) * G::FFB + my) * 2 + (col & 1)] = 0.f;
        }
        __syncthreads();
        float2 v[R];
#pragma unroll
        for (int n2 = 0; n2 < R; n2++) v[n2] = ex[t + 16 * n2];
        __syncthreads();
        fft_m<R, -1>(v, t, ex, tw);
#pragma unroll
        for (int m = 0; m < R; m++) ex[out_idx<R>(t, m)] = v[out_slot<R>(m)];
        __syncthreads();
        // C_even[ky] = (Z[ky] + conj Z[M-ky]) / 2, C_odd[ky] = (Z[ky] - conj Z[M-ky]) / (2i)
        for (int idx = tid; idx < H * 32; idx += 256) {
            const int ky = idx >> 5, col = idx & 31, px = c0 + col;
            if (px >= N) continue;
            const float2 *zf = sx + (col >> 1) * G::FFB;
            const float2 z = zf[ky], zm = zf[mirror_idx<M>(ky)];
            tmp[((size_t)a * H + ky) * N + px] = (col & 1)
                ? make_float2(0.5f * (z.y + zm.y), -0.5f * (z.x - zm.x))
                : make_float2(0.5f * (z.x + zm.x), 0.5f * (z.y - zm.y));
        }
    }
}

// C -> CAS images: per spectrum row ky in [0, M/2 + 1) the transform along x
// gives the full row F[ky][.]; CAS row ky and (Hermitian) row M - ky are
// written whole.  Grid (row groups of 16, images).
template <int R>
__global__ __launch_bounds__(256) void k_fwd_rows(const float2 *__restrict__ tmp, float *__restrict__ cas,
                                                  int n, int N, float scale, int zeroNyq)
{
    using G = FftGeom<R>;
    constexpr int M = G::M, H = G::H;
    __shared__ float2 sx[16 * G::FFB];
    __shared__ float2 tw[M];
    load_tw<R>(tw);
    const int tid = threadIdx.x, f = tid >> 4, t = tid & 15;
    const int off = (M - N) / 2;
    const int ky = 16 * blockIdx.x + f;
    const bool rowOk = ky <= M / 2;
    float2 *ex = sx + f * G::FFB;
    for (int a = blockIdx.y; a < n; a += gridDim.y) {
        const float2 *crow = tmp + ((size_t)a * H + (rowOk ? ky : 0)) * N;
        float2 v[R];
#pragma unroll
        for (int n2 = 0; n2 < R; n2++) {
            const int px = t + 16 * n2 - off;
            v[n2] = (rowOk && px >= 0 && px < N) ? crow[px] : make_float2(0.f, 0.f);
        }
        __syncthreads();
        fft_m<R, -1>(v, t, ex, tw);
        if (rowOk) {
            float *o = cas + (size_t)a * M * M;
            const bool mirror = ky != 0 && ky != M / 2;
#pragma unroll
            for (int m = 0; m < R; m++) {
                    const int kx = out_idx<R>(t, m);
                    const float2 F = v[out_slot<R>(m)];
                    const float sd = chk(kx + ky);
                    o[ky * M + kx] = (zeroNyq && (kx == 0 || ky == 0)) ? 0.f : (sd * F.x + sd * F.y) * scale;
                    if (mirror) {
                        const int x = mirror_idx<M>(kx);  // CAS[M - ky][x] from conj F[ky][(M - x) % M]
                        const float sm = chk(x + ky);
                        o[(M - ky) * M + x] = (zeroNyq && x == 0) ? 0.f : (sm * F.x - sm * F.y) * scale;
                    }
                }
        }
    }
}

// ---------------------------------------------------------------------------
// Volume chains (utils/CASFromVol_Gridder.m:34-60, utils/volFromCAS_Gridder.m:
// 37-56).  In 3-D the shift identity carries a global sign (-1)^(3M/2).
// ---------------------------------------------------------------------------
__global__ void k_vol_to_real3d(const float *__restrict__ vol, int N, float *__restrict__ real,
                                int M, const float *__restrict__ pc, size_t total)
{
    const int off = (M - N) / 2;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (size_t)gridDim.x * blockDim.x) {
        const int mx = (int)(e % M);
        const size_t rest = e / M;
        const int my = (int)(rest % M);
        const int mz = (int)(rest / M);
        const int qx = mx - off, qy = my - off, qz = mz - off;
        float v = 0.f;
        if (qx >= 0 && qx < N && qy >= 0 && qy < N && qz >= 0 && qz < N) {
            const double w = (double)pc[mx] * (double)pc[my] * (double)pc[mz];
            v = chk(mx + my + mz) *
                (float)(w * (double)vol[((size_t)qz * N + qy) * N + qx]);
        }
        real[e] = v;
    }
}

__global__ void k_herm3d_to_casvol(const float2 *__restrict__ herm, int M,
                                   float *__restrict__ casVol, int V, int pad, float gsign,
                                   size_t total)
{
    const int H = M / 2 + 1;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (size_t)gridDim.x * blockDim.x) {
        const int vx = (int)(e % V);
        const size_t rest = e / V;
        const int vy = (int)(rest % V);
        const int vz = (int)(rest / V);
        const int cx = vx - pad, cy = vy - pad, cz = vz - pad;
        float out = 0.f;
        if (cx >= 0 && cx < M && cy >= 0 && cy < M && cz >= 0 && cz < M) {
            float re, im;
            if (cx <= M / 2) {
                const float2 f = herm[((size_t)cz * M + cy) * H + cx];
                re = f.x; im = f.y;
            } else {
                const float2 f = herm[((size_t)((M - cz) % M) * M + (M - cy) % M) * H + (M - cx)];
                re = f.x; im = -f.y;
            }
            const float s = gsign * chk(cx + cy + cz);
            out = s * re + s * im;
        }
        casVol[e] = out;
    }
}

__global__ void k_casvol_to_herm3d(const float *__restrict__ casVol, int V, int pad,
                                   float2 *__restrict__ herm, int M, int zeroNyq, size_t total)
{
    const int H = M / 2 + 1;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (size_t)gridDim.x * blockDim.x) {
        const int cx = (int)(e % H);
        const size_t rest = e / H;
        const int cy = (int)(rest % M);
        const int cz = (int)(rest / M);
        float2 z = make_float2(0.f, 0.f);
        if (!(zeroNyq && (cx == 0 || cy == 0 || cz == 0))) {
            const float c1 = casVol[((size_t)(cz + pad) * V + (cy + pad)) * V + (cx + pad)];
            const int mx = (M - cx) % M, my = (M - cy) % M, mz = (M - cz) % M;
            const float c2 = casVol[((size_t)(mz + pad) * V + (my + pad)) * V + (mx + pad)];
            const float s = chk(cx + cy + cz);
            z.x = s * (0.5f * (c1 + c2));
            z.y = s * (0.5f * (c1 - c2));
        }
        herm[e] = z;
    }
}

__global__ void k_real3d_to_vol(const float *__restrict__ real, int M, float *__restrict__ vol,
                                int N, const float *__restrict__ pc, double scale, size_t total)
{
    const int off = (M - N) / 2;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (size_t)gridDim.x * blockDim.x) {
        const int qx = (int)(e % N);
        const size_t rest = e / N;
        const int qy = (int)(rest % N);
        const int qz = (int)(rest / N);
        const int mx = qx + off, my = qy + off, mz = qz + off;
        const double w = (double)pc[mx] * (double)pc[my] * (double)pc[mz] * scale;
        vol[e] = chk(mx + my + mz) * (float)(w * (double)real[((size_t)mz * M + my) * M + mx]);
    }
}

// ---------------------------------------------------------------------------
// Reductions, conversions, fills.
// ---------------------------------------------------------------------------
__global__ void k_reduce_partials(float *__restrict__ out, const float *__restrict__ parts,
                                  int nParts, size_t n, size_t stride, int accumulate)
{
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x) {
        double s = accumulate ? (double)out[e] : 0.0;
        for (int q = 0; q < nParts; q++) s += (double)parts[(size_t)q * stride + e];
        out[e] = (float)s;
    }
}

__global__ void k_divide_density(float *__restrict__ out, const float *__restrict__ num,
                                 const float *__restrict__ den, size_t n, double eps)
{
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x)
        out[e] = (float)((double)num[e] / ((double)den[e] + eps));
}

__global__ void k_to_bf16(uint16_t *__restrict__ dst, const float *__restrict__ src, size_t n)
{
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x) {
        uint32_t u = __float_as_uint(src[e]);
        if ((u & 0x7f800000u) == 0x7f800000u) {  // inf / nan: truncate, keep nan quiet
            dst[e] = (uint16_t)((u >> 16) | ((u & 0x007fffffu) ? 0x40u : 0u));
        } else {
            u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
            dst[e] = (uint16_t)(u >> 16);
        }
    }
}

__global__ void k_to_f16(uint16_t *__restrict__ dst, const float *__restrict__ src, size_t n)
{
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x) {
        __half h = __float2half_rn(src[e]);
        dst[e] = __half_raw(h).x;
    }
}

__global__ void k_fill(float *__restrict__ p, float v, size_t n)
{
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x)
        p[e] = v;
}

// ---------------------------------------------------------------------------
// Check behind the forward projector's 4-wide windows: *check = 1 when the
// table has the 501 entries, all finite, with both end entries 0, and every
// value of the volume is finite; else 0.  k_check_table (one workgroup) writes
// the table's verdict, then k_check_volume clears it on any non-finite voxel
// (all finding workgroups store the same 0).
// ---------------------------------------------------------------------------
__global__ void k_check_table(const float *__restrict__ ker, int K, unsigned *__restrict__ check)
{
    bool ok = true;
    for (int t = threadIdx.x; t < K; t += blockDim.x) ok = ok && __builtin_isfinite(ker[t]);
    ok = __syncthreads_and(ok);
    if (threadIdx.x == 0) *check = (ok && K == KFP_FAST && ker[0] == 0.f && ker[K - 1] == 0.f) ? 1u : 0u;
}

// Non-finite = exponent field all ones (FP32 0x7f800000, BF16 0x7f80, FP16 0x7c00).
template <int DT>
__global__ void k_check_volume(const void *__restrict__ vol, size_t n, unsigned *__restrict__ check)
{
    const size_t t0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x, st = (size_t)gridDim.x * blockDim.x;
    bool bad = false;
    if (DT == DT_F32) {
        const unsigned *w = (const unsigned *)vol;
        size_t e0 = 0;
        if (((uintptr_t)w & 15u) == 0) {  // 16-byte loads over the aligned body
            const size_t n4 = n / 4;
            const u4 *w4 = (const u4 *)w;
            for (size_t e = t0; e < n4; e += st) {
                const u4 q = w4[e];
                bad |= ((q.x & 0x7f800000u) == 0x7f800000u) | ((q.y & 0x7f800000u) == 0x7f800000u) |
                       ((q.z & 0x7f800000u) == 0x7f800000u) | ((q.w & 0x7f800000u) == 0x7f800000u);
            }
            e0 = n4 * 4;
        }
        for (size_t e = e0 + t0; e < n; e += st) bad |= (w[e] & 0x7f800000u) == 0x7f800000u;
    } else {
        const uint16_t *h = (const uint16_t *)vol;
        const unsigned m = DT == DT_BF16 ? 0x7f80u : 0x7c00u;
        for (size_t e = t0; e < n; e += st) bad |= ((unsigned)h[e] & m) == m;
    }
    if (__syncthreads_or(bad) && threadIdx.x == 0) *check = 0u;
}

// ---------------------------------------------------------------------------
// Standalone filters (Filters/*.cu).  Same element semantics as the reference
// kernels; grid-stride 1-D launches instead of one thread per (x, y) column.
// ---------------------------------------------------------------------------
__global__ void k_f_cas_to_complex(const float *__restrict__ in, float2 *__restrict__ out, int S,
                                   int nSlices, int volumeMode, size_t total)
{
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(e % S);
        const size_t rest = e / S;
        const int j = (int)(rest % S);
        const int k = (int)(rest / S);
        float2 z = make_float2(0.f, 0.f);
        const bool skip = volumeMode ? (i == 0 || j == 0 || k == 0) : (i == 0 || j == 0);
        if (!skip) {
            const int k2 = volumeMode ? S - k : k;
            const size_t n2 = (size_t)(S - i) + (size_t)(S - j) * S + (size_t)k2 * S * S;
            z.x = 0.5f * (in[e] + in[n2]);
            z.y = 0.5f * (in[e] - in[n2]);
        }
        out[e] = z;
    }
}

__global__ void k_f_complex_to_cas(const float2 *__restrict__ in, float *__restrict__ out, size_t n)
{
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x)
        out[e] = in[e].x + in[e].y;
}

__global__ void k_f_complex_to_real(const float2 *__restrict__ in, float *__restrict__ out, size_t n)
{
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x)
        out[e] = in[e].x;
}

__global__ void k_f_real_to_complex(const float *__restrict__ in, float2 *__restrict__ out, size_t n)
{
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x)
        out[e] = make_float2(in[e], 0.f);
}

// In-place quadrant swap per slice (even S): element (x, y) with y < S/2
// swaps with ((x + S/2) mod S, y + S/2).
template <typename T>
__global__ void k_f_fftshift2d(T *__restrict__ d, int S, size_t total)
{
    const int h = S / 2;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (size_t)gridDim.x * blockDim.x) {
        const int x = (int)(e % S);
        const size_t rest = e / S;
        const int y = (int)(rest % h);
        const size_t k = rest / h;
        const size_t a = k * S * S + (size_t)y * S + x;
        const size_t b = k * S * S + (size_t)(y + h) * S + ((x + h) % S);
        T t = d[a];
        d[a] = d[b];
        d[b] = t;
    }
}

// In-place octant swap (even S): element (x, y, z) with z < S/2 swaps with
// ((x + S/2) mod S, (y + S/2) mod S, z + S/2).
template <typename T>
__global__ void k_f_fftshift3d(T *__restrict__ d, int S, size_t total)
{
    const int h = S / 2;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (size_t)gridDim.x * blockDim.x) {
        const int x = (int)(e % S);
        const size_t rest = e / S;
        const int y = (int)(rest % S);
        const int z = (int)(rest / S);
        const size_t a = ((size_t)z * S + y) * S + x;
        const size_t b = ((size_t)(z + h) * S + (y + h) % S) * S + (x + h) % S;
        T t = d[a];
        d[a] = d[b];
        d[b] = t;
    }
}

__global__ void k_f_pad(const float *__restrict__ in, float *__restrict__ out, int inS, int outS,
                        size_t total)
{
    const int off = (outS - inS) / 2;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (size_t)gridDim.x * blockDim.x) {
        const int x = (int)(e % outS);
        const size_t rest = e / outS;
        const int y = (int)(rest % outS);
        const size_t k = rest / outS;
        const int ix = x - off, iy = y - off;
        float v = 0.f;
        if (ix >= 0 && ix < inS && iy >= 0 && iy < inS) v = in[k * inS * inS + (size_t)iy * inS + ix];
        out[e] = v;
    }
}

template <typename T>
__global__ void k_f_crop(const T *__restrict__ in, T *__restrict__ out, int inS, int outS,
                         size_t total)
{
    const int off = (inS - outS) / 2;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (size_t)gridDim.x * blockDim.x) {
        const int x = (int)(e % outS);
        const size_t rest = e / outS;
        const int y = (int)(rest % outS);
        const size_t k = rest / outS;
        out[e] = in[k * inS * inS + (size_t)(y + off) * inS + (x + off)];
    }
}

__global__ void k_f_divide_scalar(float *__restrict__ d, float s, size_t n)
{
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x)
        d[e] = d[e] / s;
}

__global__ void k_f_divide_volume(float *__restrict__ a, const float *__restrict__ b, size_t n)
{
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x) {
        const float x = a[e], y = b[e];
        a[e] = (x != 0.f && y != 0.f) ? x / y : 0.f;
    }
}

__global__ void k_f_multiply_volume(float *__restrict__ a, const float *__restrict__ b, size_t n)
{
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x)
        a[e] = a[e] * b[e];
}

__global__ void k_f_multiply_complex_volume(float2 *__restrict__ a, const float *__restrict__ b,
                                            size_t n)
{
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x) {
        float2 z = a[e];
        z.x = z.x * b[e];
        z.y = z.y * b[e];
        a[e] = z;
    }
}

__global__ void k_f_add_volume(float *__restrict__ a, const float *__restrict__ b, size_t n)
{
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x)
        a[e] = a[e] + b[e];
}

// 1-D grid-stride launch geometry: enough blocks to fill 256 CUs many times
// over, capped so huge arrays loop inside the kernel.
inline unsigned grid_for(size_t n)
{
    size_t g = (n + kThreads - 1) / kThreads;
    if (g > 65536) g = 65536;
    if (g == 0) g = 1;
    return (unsigned)g;
}

// Largest integer d2 (a squared voxel distance, at most 3*(v/2)^2) with
// sqrtf((float)d2) <= r, i.e. the integer form of the reference's mask test
// sqrtf(...) <= maskRadius (gpuBackProject.cu:48-50).  -1 if none passes.
inline int mask_d2(float r, int v)
{
    const long long hi = 3LL * (v / 2 + 1) * (v / 2 + 1);
    if (!(sqrtf(0.f) <= r)) return -1;
    long long lo = 0, top = hi;  // sqrtf is monotone: binary search the last passing d2
    if (sqrtf((float)top) <= r) return (int)top;
    while (top - lo > 1) {
        const long long mid = (lo + top) / 2;
        if (sqrtf((float)mid) <= r) lo = mid; else top = mid;
    }
    return (int)lo;
}

// The mask radius as the reference's kernels see it: RunKernel takes it as an
// int (gpuForwardProject.h:57, gpuBackProject.h:56) from the float member
// (gpuGridder.cpp:262,422), so it is truncated toward zero.  NaN and values
// beyond the int range (no meaningful conversion) pass through.
inline float ref_mask_radius(float r)
{
    if (!(r > -2147483648.0f && r < 2147483648.0f)) return r;
    return (float)(int)r;
}

inline void proj_consts(const ProjGeom &g, float &kerScale, float &kerCenter, float &fW,
                        int &convW)
{
    // gpuForwardProject.cu:20-23: kerCenter = (K-1)/2, kerScale = kerCenter/kerHWidth.
    kerCenter = ((float)g.K - 1) / 2;
    kerScale = kerCenter / g.kerHWidth;
    convW = (int)roundf(g.kerHWidth);
    fW = (float)convW;
}

}  // namespace

// ===========================================================================
// Launchers
// ===========================================================================
hipError_t launch_forward_project(const void *vol, int dtype, const ProjGeom &g,
                                  const float *axes, int nAxes, float *casImgs, hipStream_t s,
                                  const unsigned *check)
{
    if (nAxes <= 0) return hipSuccess;
    FPArgs a;
    a.vol = vol; a.axes = axes; a.out = casImgs; a.ker = g.ker; a.check = nullptr;
    a.V = g.V; a.M = g.M; a.K = g.K; a.nAxes = nAxes;
    a.maskR = ref_mask_radius(g.maskRadius);
    proj_consts(g, a.kerScale, a.kerCenter, a.fW, a.convW);
    const bool w2 = a.convW == 2 && g.K <= FPKER;
    if (w2) {
        // LDS-staged kernel: 8x8 tiles, one per wave, 4 per workgroup.
        const bool fast = a.kerScale == 125.0f && a.kerCenter == 250.0f && g.K == 501;
        a.check = fast ? check : nullptr;
        a.tilesPerRow = (g.M + FPT - 1) / FPT;
        a.tilesPerImg = a.tilesPerRow * a.tilesPerRow;
        a.maskD2 = mask_d2(a.maskR, g.V);
        const unsigned long long work = (unsigned long long)a.tilesPerImg * nAxes;
        if (work > 0xffffffffull / 2) return hipErrorInvalidValue;
        a.totalWork = (unsigned)work;
        a.perXcd = 0;
        // Three tiles per wave, a grid-stride apart (one workgroup per twelve
        // tiles): 10.8 ms per 10,000 directions at N=128 against 12.1 ms with
        // one tile per wave, 11.4 / 11.9 / 11.3 / 11.6 ms with 2 / 4 / 6 / 8
        // tiles per wave, 13.5 ms with a persistent 1,024-workgroup grid.
        const dim3 grid((a.totalWork + 11) / 12), block(kThreads);
        // volumes of 4 GB and more: box addressing from the box's first element
        const int esz = dtype == DT_F32 ? 4 : 2;
        const bool big = (long long)g.V * g.V * g.V * esz >= (1LL << 32);
#define MGG_FP_LAUNCH(DTV)                                                                   \
    if (big) {                                                                               \
        if (fast) hipLaunchKernelGGL((fp_lds_kernel<DTV, true, true>), grid, block, 0, s, a);  \
        else hipLaunchKernelGGL((fp_lds_kernel<DTV, false, true>), grid, block, 0, s, a);      \
    } else if (fast) hipLaunchKernelGGL((fp_lds_kernel<DTV, true, false>), grid, block, 0, s, a); \
    else hipLaunchKernelGGL((fp_lds_kernel<DTV, false, false>), grid, block, 0, s, a);
        switch (dtype) {
        case DT_F32: MGG_FP_LAUNCH(DT_F32) break;
        case DT_BF16: MGG_FP_LAUNCH(DT_BF16) break;
        case DT_F16: MGG_FP_LAUNCH(DT_F16) break;
        default: return hipErrorInvalidValue;
        }
#undef MGG_FP_LAUNCH
        return hipGetLastError();
    }
    a.tilesPerRow = (g.M + 15) / 16;
    a.tilesPerImg = a.tilesPerRow * a.tilesPerRow;
    const unsigned long long work = (unsigned long long)a.tilesPerImg * nAxes;
    if (work > 0xffffffffull / 2) return hipErrorInvalidValue;
    a.totalWork = (unsigned)work;
    a.perXcd = (a.totalWork + 7) / 8;
    const dim3 grid(a.perXcd * 8), block(kThreads);
    const bool rows = a.convW == 2;  // 5-tap row loads (tables too large for the staged kernel)
    switch (dtype) {
    case DT_F32:
        if (rows) hipLaunchKernelGGL((fp_kernel<DT_F32, true>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((fp_kernel<DT_F32, false>), grid, block, 0, s, a);
        break;
    case DT_BF16:
        if (rows) hipLaunchKernelGGL((fp_kernel<DT_BF16, true>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((fp_kernel<DT_BF16, false>), grid, block, 0, s, a);
        break;
    case DT_F16:
        if (rows) hipLaunchKernelGGL((fp_kernel<DT_F16, true>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((fp_kernel<DT_F16, false>), grid, block, 0, s, a);
        break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

bool fp_segment_supported(const ProjGeom &g)
{
    float kerScale, kerCenter, fW;
    int convW;
    proj_consts(g, kerScale, kerCenter, fW, convW);
    return convW == 2 && kerScale == 125.0f && kerCenter == 250.0f && g.K == KFP_FAST && g.K <= FPKER;
}

unsigned fp_segment(FPSegment &sg, const void *vol, const ProjGeom &g, const float *axes, int nAxes, float *casImgs,
                    unsigned workStart, const unsigned *check)
{
    sg.vol = vol; sg.axes = axes; sg.out = casImgs; sg.ker = g.ker; sg.check = check;
    sg.V = g.V; sg.M = g.M; sg.nAxes = nAxes;
    sg.maskD2 = mask_d2(ref_mask_radius(g.maskRadius), g.V);
    sg.tilesPerRow = (g.M + FPT - 1) / FPT;
    sg.tilesPerImg = sg.tilesPerRow * sg.tilesPerRow;
    sg.workStart = workStart;
    return (unsigned)sg.tilesPerImg * (unsigned)nAxes;
}

hipError_t launch_forward_project_multi(const FPMultiArgs &m, hipStream_t s)
{
    if (m.nSegs <= 0 || m.totalWork == 0) return hipSuccess;
    if (m.nSegs > kFpMaxSegments) return hipErrorInvalidValue;
    // The single-volume launch's grid: three tiles per wave, a grid-stride
    // apart.  A persistent grid (no more workgroups than are resident, each
    // walking the list) ran 25% slower on the same tiles.
    const unsigned grid = (m.totalWork + 11) / 12;
    bool big = false;  // a volume of 4 GB or more
    for (int i = 0; i < m.nSegs; i++) big |= (long long)m.seg[i].V * m.seg[i].V * m.seg[i].V * 4 >= (1LL << 32);
    if (big) hipLaunchKernelGGL((fp_lds_multi_kernel<DT_F32, true>), dim3(grid), dim3(kThreads), 0, s, m);
    else hipLaunchKernelGGL((fp_lds_multi_kernel<DT_F32, false>), dim3(grid), dim3(kThreads), 0, s, m);
    return hipGetLastError();
}

bool bp_uses_fast_path(const ProjGeom &g)
{
    float kerScale, kerCenter, fW;
    int convW;
    proj_consts(g, kerScale, kerCenter, fW, convW);
    return convW == 2 && kerScale == 125.0f && kerCenter == 250.0f && g.K == KFAST && g.V <= 8192;
}

int bp_brick_count(int V, bool fast)
{
    const int b = fast ? (V + BB - 1) / BB : 0;
    return b * b * b;
}

void bp_brick_order(int V, int *out)
{
    // Bricks of edge BB sorted by the distance of their centre from the volume
    // centre (ties by index): central bricks, which every direction's slab
    // crosses, are dispatched first.
    const int nb = (V + BB - 1) / BB;
    const double c = V / 2;
    struct Item { double d; int id; };
    Item *items = new Item[(size_t)nb * nb * nb];
    int n = 0;
    for (int z = 0; z < nb; z++)
        for (int y = 0; y < nb; y++)
            for (int x = 0; x < nb; x++) {
                const double dx = x * BB + 0.5 * (BB - 1) - c, dy = y * BB + 0.5 * (BB - 1) - c,
                             dz = z * BB + 0.5 * (BB - 1) - c;
                items[n] = {dx * dx + dy * dy + dz * dz, (z * nb + y) * nb + x};
                n++;
            }
    std::sort(items, items + n, [](const Item &a, const Item &b) { return a.d < b.d || (a.d == b.d && a.id < b.id); });
    for (int i = 0; i < n; i++) out[i] = items[i].id;
    delete[] items;
}

hipError_t launch_back_project_adjoint(float *volAcc, const ProjGeom &g, const float *axes,
                                       int nAxes, const float *casImgs, hipStream_t s, const unsigned *check)
{
    if (nAxes <= 0) return hipSuccess;
    FPArgs a;
    a.vol = nullptr; a.axes = axes; a.out = nullptr; a.ker = g.ker; a.check = nullptr;
    a.V = g.V; a.M = g.M; a.K = g.K; a.nAxes = nAxes;
    a.maskR = ref_mask_radius(g.maskRadius);
    proj_consts(g, a.kerScale, a.kerCenter, a.fW, a.convW);
    if (a.convW == 2 && g.K <= FPKER) {
        // LDS-box scatter: 8x8 tiles, one per wave, 4 per workgroup.
        const bool fast = a.kerScale == 125.0f && a.kerCenter == 250.0f && g.K == 501;
        a.check = fast ? check : nullptr;
        a.tilesPerRow = (g.M + FPT - 1) / FPT;
        a.tilesPerImg = a.tilesPerRow * a.tilesPerRow;
        a.maskD2 = mask_d2(a.maskR, g.V);
        const unsigned long long work = (unsigned long long)a.tilesPerImg * nAxes;
        if (work > 0xffffffffull / 2) return hipErrorInvalidValue;
        a.totalWork = (unsigned)work;
        a.perXcd = 0;
        const dim3 grid((a.totalWork + 3) / 4), block(kThreads);
        if (fast) hipLaunchKernelGGL(adj_lds_kernel<true>, grid, block, 0, s, a, casImgs, volAcc);
        else hipLaunchKernelGGL(adj_lds_kernel<false>, grid, block, 0, s, a, casImgs, volAcc);
        return hipGetLastError();
    }
    a.tilesPerRow = (g.M + 15) / 16;
    a.tilesPerImg = a.tilesPerRow * a.tilesPerRow;
    const unsigned long long work = (unsigned long long)a.tilesPerImg * nAxes;
    if (work > 0xffffffffull / 2) return hipErrorInvalidValue;
    a.totalWork = (unsigned)work;
    a.perXcd = (a.totalWork + 7) / 8;
    hipLaunchKernelGGL(adj_kernel, dim3(a.perXcd * 8), dim3(kThreads), 0, s, a, casImgs, volAcc);
    return hipGetLastError();
}

hipError_t launch_back_project(float *volAcc, float *densAcc, const ProjGeom &g,
                               const float *axes, int nAxes, const float *casImgs, hipStream_t s,
                               const int *brickOrder, int orderLen)
{
    if (nAxes <= 0) return hipSuccess;
    BPArgs a;
    a.vol = volAcc; a.dens = densAcc; a.axes = axes; a.img = casImgs; a.ker = g.ker;
    a.V = g.V; a.M = g.M; a.K = g.K; a.nAxes = nAxes;
    a.maskR = ref_mask_radius(g.maskRadius); a.hw = g.kerHWidth;
    proj_consts(g, a.kerScale, a.kerCenter, a.fW, a.convW);
    const bool fast = bp_uses_fast_path(g);
    a.maskD2 = mask_d2(a.maskR, g.V);
    {
        int r = 0;  // floor(sqrt(maskD2))
        while ((long long)(r + 1) * (r + 1) <= (long long)a.maskD2) r++;
        a.maskLo = a.maskD2 < 0 ? g.V : g.V / 2 - r;
    }
    // In-mask voxel: |v - c| <= maskR; orthonormal frame (to 1e-4): |u|, |v| <=
    // 1.00005, and the FP32 evaluation of fi adds < 1e-3.  Then fi, fj >=
    // M/2 - 1.00005 maskR - 1e-3, which must be >= 0.5 for floor(fi + 0.5) to be
    // roundf(fi) (for 0.25 <= x < 0.5 an FP32 x + 0.5 can round up to 1).
    a.posRound = ((double)(g.M / 2) - 1.00005 * (double)a.maskR - 1e-3 >= 0.5) ? 1 : 0;
    const int ex = fast ? BB : BPX, ey = fast ? BB : BPY, ez = fast ? BB : BPZ;
    a.bricksX = (g.V + ex - 1) / ex;
    const int bricksY = (g.V + ey - 1) / ey;
    const int bricksZ = (g.V + ez - 1) / ez;
    a.bricksXY = a.bricksX * bricksY;
    a.totalWork = (unsigned)a.bricksXY * (unsigned)bricksZ;
    a.perXcd = (a.totalWork + 7) / 8;
    a.order = brickOrder;
    const dim3 grid(a.perXcd * 8), block(kThreads);
    if (fast) {
        if (!brickOrder || orderLen != (int)a.totalWork) return hipErrorInvalidValue;
        const dim3 g1(a.totalWork);
        if (densAcc) hipLaunchKernelGGL(bp_fast_kernel<true>, g1, block, 0, s, a);
        else hipLaunchKernelGGL(bp_fast_kernel<false>, g1, block, 0, s, a);
    } else {
        if (densAcc) hipLaunchKernelGGL(bp_kernel<true>, grid, block, 0, s, a);
        else hipLaunchKernelGGL(bp_kernel<false>, grid, block, 0, s, a);
    }
    return hipGetLastError();
}

// Threads per row workgroup of the 2-D chain kernels: the row length rounded
// up to a wave, at most 256 (longer rows loop).
static int row_threads(int len) { return std::min(256, std::max(64, (len + 63) / 64 * 64)); }
static dim3 rows_grid(int rows, int n) { return dim3((unsigned)rows, (unsigned)std::min(n, 65535)); }

hipError_t launch_cas_to_herm2d(const float *cas, float2 *herm, int M, int n, hipStream_t s)
{
    if (n <= 0 || M <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_cas_to_herm2d, rows_grid(M, n), dim3(row_threads(M / 2 + 1)), 0, s, cas, herm, M, n);
    return hipGetLastError();
}

hipError_t launch_real2d_to_imgs(const float *real, float *imgs, int M, int N, int n, float scale,
                                 hipStream_t s)
{
    if (n <= 0 || N <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_real2d_to_imgs, rows_grid(N, n), dim3(row_threads(N)), 0, s, real, imgs, M, N, scale, n);
    return hipGetLastError();
}

hipError_t launch_imgs_to_real2d(const float *imgs, float *real, int M, int N, int n, hipStream_t s)
{
    if (n <= 0 || M <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_imgs_to_real2d, rows_grid(M, n), dim3(row_threads(M)), 0, s, imgs, real, M, N, n);
    return hipGetLastError();
}

hipError_t launch_herm2d_to_cas(const float2 *herm, float *cas, int M, int n, hipStream_t s, float scale,
                                int zeroNyq)
{
    if (n <= 0 || M <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_herm2d_to_cas, rows_grid(M, n), dim3(row_threads(M)), 0, s, herm, cas, M, n, scale,
                       zeroNyq);
    return hipGetLastError();
}

bool fused_chain_supported(int M, int N)
{
    return (M == 128 || M == 192 || M == 256 || M == 384 || M == 512) && N >= 1 && N <= M;
}

template <int R>
static void fused_inverse(const float *cas, float2 *tmp, float *imgs, int N, int n, float scale, hipStream_t s)
{
    constexpr int H = 8 * R + 1;
    const unsigned ny = (unsigned)std::min(n, 65535);
    hipLaunchKernelGGL(k_inv_rows<R>, dim3((H + 15) / 16, ny), dim3(256), 0, s, cas, tmp, n, N);
    hipLaunchKernelGGL(k_inv_cols<R>, dim3((unsigned)((N + 31) / 32), ny), dim3(256), 0, s, tmp, imgs, n, N,
                       scale);
}

template <int R>
static void fused_forward(const float *imgs, float2 *tmp, float *cas, int N, int n, float scale, int zeroNyq,
                          hipStream_t s)
{
    constexpr int H = 8 * R + 1;
    const unsigned ny = (unsigned)std::min(n, 65535);
    hipLaunchKernelGGL(k_fwd_cols<R>, dim3((unsigned)((N + 31) / 32), ny), dim3(256), 0, s, imgs, tmp, n, N);
    hipLaunchKernelGGL(k_fwd_rows<R>, dim3((H + 15) / 16, ny), dim3(256), 0, s, tmp, cas, n, N, scale, zeroNyq);
}

hipError_t launch_fused_cas_to_images(const float *cas, float2 *tmp, float *imgs, int M, int N, int n,
                                      float scale, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    if (!fused_chain_supported(M, N)) return hipErrorInvalidValue;
    switch (M) {
    case 128: fused_inverse<8>(cas, tmp, imgs, N, n, scale, s); break;
    case 192: fused_inverse<12>(cas, tmp, imgs, N, n, scale, s); break;
    case 256: fused_inverse<16>(cas, tmp, imgs, N, n, scale, s); break;
    case 384: fused_inverse<24>(cas, tmp, imgs, N, n, scale, s); break;
    default: fused_inverse<32>(cas, tmp, imgs, N, n, scale, s); break;
    }
    return hipGetLastError();
}

hipError_t launch_fused_images_to_cas(const float *imgs, float2 *tmp, float *cas, int M, int N, int n,
                                      float scale, int zeroNyq, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    if (!fused_chain_supported(M, N)) return hipErrorInvalidValue;
    switch (M) {
    case 128: fused_forward<8>(imgs, tmp, cas, N, n, scale, zeroNyq, s); break;
    case 192: fused_forward<12>(imgs, tmp, cas, N, n, scale, zeroNyq, s); break;
    case 256: fused_forward<16>(imgs, tmp, cas, N, n, scale, zeroNyq, s); break;
    case 384: fused_forward<24>(imgs, tmp, cas, N, n, scale, zeroNyq, s); break;
    default: fused_forward<32>(imgs, tmp, cas, N, n, scale, zeroNyq, s); break;
    }
    return hipGetLastError();
}

hipError_t launch_vol_to_real3d(const float *vol, int N, float *real, int M, const float *precomp,
                                hipStream_t s)
{
    const size_t total = (size_t)M * M * M;
    hipLaunchKernelGGL(k_vol_to_real3d, dim3(grid_for(total)), dim3(kThreads), 0, s, vol, N, real, M,
                       precomp, total);
    return hipGetLastError();
}

hipError_t launch_herm3d_to_casvol(const float2 *herm, int M, float *casVol, int V, int pad,
                                   hipStream_t s)
{
    const size_t total = (size_t)V * V * V;
    const float gsign = ((3 * (M / 2)) & 1) ? -1.f : 1.f;
    hipLaunchKernelGGL(k_herm3d_to_casvol, dim3(grid_for(total)), dim3(kThreads), 0, s, herm, M, casVol,
                       V, pad, gsign, total);
    return hipGetLastError();
}

hipError_t launch_casvol_to_herm3d(const float *casVol, int V, int pad, float2 *herm, int M,
                                   int zeroNyquist, hipStream_t s)
{
    const size_t total = (size_t)M * M * (M / 2 + 1);
    hipLaunchKernelGGL(k_casvol_to_herm3d, dim3(grid_for(total)), dim3(kThreads), 0, s, casVol, V, pad,
                       herm, M, zeroNyquist, total);
    return hipGetLastError();
}

hipError_t launch_real3d_to_vol(const float *real, int M, float *vol, int N, const float *precomp,
                                float scale, hipStream_t s)
{
    const size_t total = (size_t)N * N * N;
    const double gsign = ((3 * (M / 2)) & 1) ? -1.0 : 1.0;
    hipLaunchKernelGGL(k_real3d_to_vol, dim3(grid_for(total)), dim3(kThreads), 0, s, real, M, vol, N,
                       precomp, gsign * (double)scale, total);
    return hipGetLastError();
}

hipError_t launch_reduce_partials(float *out, const float *parts, int nParts, size_t n,
                                  size_t stride, int accumulate, hipStream_t s)
{
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_reduce_partials, dim3(grid_for(n)), dim3(kThreads), 0, s, out, parts, nParts, n,
                       stride, accumulate);
    return hipGetLastError();
}

hipError_t launch_divide_density(float *out, const float *num, const float *den, size_t n,
                                 double eps, hipStream_t s)
{
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_divide_density, dim3(grid_for(n)), dim3(kThreads), 0, s, out, num, den, n, eps);
    return hipGetLastError();
}

hipError_t launch_convert_volume(void *dst, int dtype, const float *src, size_t n, hipStream_t s)
{
    if (!n) return hipSuccess;
    if (dtype == DT_BF16)
        hipLaunchKernelGGL(k_to_bf16, dim3(grid_for(n)), dim3(kThreads), 0, s, (uint16_t *)dst, src, n);
    else if (dtype == DT_F16)
        hipLaunchKernelGGL(k_to_f16, dim3(grid_for(n)), dim3(kThreads), 0, s, (uint16_t *)dst, src, n);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_check_volume(const void *vol, int dtype, size_t n, const float *ker, int K, unsigned *check,
                               hipStream_t s)
{
    hipLaunchKernelGGL(k_check_table, dim3(1), dim3(kThreads), 0, s, ker, K, check);
    if (n == 0) return hipGetLastError();
    const dim3 grid(std::min(grid_for(n), 4096u)), block(kThreads);
    switch (dtype) {
    case DT_F32: hipLaunchKernelGGL(k_check_volume<DT_F32>, grid, block, 0, s, vol, n, check); break;
    case DT_BF16: hipLaunchKernelGGL(k_check_volume<DT_BF16>, grid, block, 0, s, vol, n, check); break;
    case DT_F16: hipLaunchKernelGGL(k_check_volume<DT_F16>, grid, block, 0, s, vol, n, check); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_fill(float *p, float value, size_t n, hipStream_t s)
{
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_fill, dim3(grid_for(n)), dim3(kThreads), 0, s, p, value, n);
    return hipGetLastError();
}

hipError_t launch_filter_cas_to_complex(const float *in, float2 *out, int size, int nSlices,
                                        hipStream_t s)
{
    const int volumeMode = nSlices <= 0;
    const int slices = volumeMode ? size : nSlices;
    const size_t total = (size_t)size * size * slices;
    if (!total) return hipSuccess;
    hipLaunchKernelGGL(k_f_cas_to_complex, dim3(grid_for(total)), dim3(kThreads), 0, s, in, out, size,
                       slices, volumeMode, total);
    return hipGetLastError();
}

hipError_t launch_filter_complex_to_cas(const float2 *in, float *out, int size, int nSlices,
                                        hipStream_t s)
{
    const size_t n = (size_t)size * size * nSlices;
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_f_complex_to_cas, dim3(grid_for(n)), dim3(kThreads), 0, s, in, out, n);
    return hipGetLastError();
}

hipError_t launch_filter_complex_to_real(const float2 *in, float *out, int size, int nSlices,
                                         hipStream_t s)
{
    const size_t n = (size_t)size * size * nSlices;
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_f_complex_to_real, dim3(grid_for(n)), dim3(kThreads), 0, s, in, out, n);
    return hipGetLastError();
}

hipError_t launch_filter_real_to_complex(const float *in, float2 *out, int size, int nSlices,
                                         hipStream_t s)
{
    const size_t n = (size_t)size * size * nSlices;
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_f_real_to_complex, dim3(grid_for(n)), dim3(kThreads), 0, s, in, out, n);
    return hipGetLastError();
}

hipError_t launch_filter_fftshift2d(void *data, int elemBytes, int size, int nSlices, hipStream_t s)
{
    if (size % 2) return hipErrorInvalidValue;
    const size_t total = (size_t)size * (size / 2) * nSlices;
    if (!total) return hipSuccess;
    if (elemBytes == 4)
        hipLaunchKernelGGL(k_f_fftshift2d<float>, dim3(grid_for(total)), dim3(kThreads), 0, s,
                           (float *)data, size, total);
    else if (elemBytes == 8)
        hipLaunchKernelGGL(k_f_fftshift2d<float2>, dim3(grid_for(total)), dim3(kThreads), 0, s,
                           (float2 *)data, size, total);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_filter_fftshift3d(void *data, int elemBytes, int size, hipStream_t s)
{
    if (size % 2) return hipErrorInvalidValue;
    const size_t total = (size_t)size * size * (size / 2);
    if (!total) return hipSuccess;
    if (elemBytes == 4)
        hipLaunchKernelGGL(k_f_fftshift3d<float>, dim3(grid_for(total)), dim3(kThreads), 0, s,
                           (float *)data, size, total);
    else if (elemBytes == 8)
        hipLaunchKernelGGL(k_f_fftshift3d<float2>, dim3(grid_for(total)), dim3(kThreads), 0, s,
                           (float2 *)data, size, total);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_filter_pad(const float *in, float *out, int inSize, int outSize, int nSlices,
                             hipStream_t s)
{
    if (outSize < inSize) return hipErrorInvalidValue;
    const size_t total = (size_t)outSize * outSize * nSlices;
    if (!total) return hipSuccess;
    hipLaunchKernelGGL(k_f_pad, dim3(grid_for(total)), dim3(kThreads), 0, s, in, out, inSize, outSize,
                       total);
    return hipGetLastError();
}

hipError_t launch_filter_crop(const void *in, void *out, int elemBytes, int inSize, int outSize,
                              int nSlices, hipStream_t s)
{
    if (outSize > inSize) return hipErrorInvalidValue;
    const size_t total = (size_t)outSize * outSize * nSlices;
    if (!total) return hipSuccess;
    if (elemBytes == 4)
        hipLaunchKernelGGL(k_f_crop<float>, dim3(grid_for(total)), dim3(kThreads), 0, s,
                           (const float *)in, (float *)out, inSize, outSize, total);
    else if (elemBytes == 8)
        hipLaunchKernelGGL(k_f_crop<float2>, dim3(grid_for(total)), dim3(kThreads), 0, s,
                           (const float2 *)in, (float2 *)out, inSize, outSize, total);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_filter_divide_scalar(float *data, float scalar, size_t n, hipStream_t s)
{
    if (scalar == 0.f) return hipErrorInvalidValue;
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_f_divide_scalar, dim3(grid_for(n)), dim3(kThreads), 0, s, data, scalar, n);
    return hipGetLastError();
}

hipError_t launch_filter_divide_volume(float *a, const float *b, size_t n, hipStream_t s)
{
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_f_divide_volume, dim3(grid_for(n)), dim3(kThreads), 0, s, a, b, n);
    return hipGetLastError();
}

hipError_t launch_filter_multiply_volume(float *a, const float *b, size_t n, hipStream_t s)
{
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_f_multiply_volume, dim3(grid_for(n)), dim3(kThreads), 0, s, a, b, n);
    return hipGetLastError();
}

hipError_t launch_filter_multiply_complex_volume(float2 *a, const float *b, size_t n, hipStream_t s)
{
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_f_multiply_complex_volume, dim3(grid_for(n)), dim3(kThreads), 0, s, a, b, n);
    return hipGetLastError();
}

hipError_t launch_filter_add_volume(float *a, const float *b, size_t n, hipStream_t s)
{
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_f_add_volume, dim3(grid_for(n)), dim3(kThreads), 0, s, a, b, n);
    return hipGetLastError();
}

}  // namespace mgg
