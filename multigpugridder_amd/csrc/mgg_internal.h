// Internal declarations shared by the HIP kernels and the C++ host object.
// Not part of the public C ABI (include/mgg.h is).
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace mgg {

// Storage type of a CAS volume that the forward projector gathers from.
// Gather arithmetic and accumulation are FP32 for every storage type.
enum VolDtype : int { DT_F32 = 0, DT_BF16 = 1, DT_F16 = 2 };

// Largest kernel table the projectors accept (the table lives in LDS).
constexpr int kMaxKerSize = 2048;

// Geometry shared by every projector launch.
struct ProjGeom {
    int V;              // CAS volume edge (M + 2*extraPadding)
    int M;              // CAS image edge (N * interpFactor)
    float maskRadius;   // pixel / voxel mask radius
    const float *ker;   // device pointer to the Kaiser-Bessel table
    int K;              // table length
    float kerHWidth;    // table half width (2)
};

// ---- projectors --------------------------------------------------------
// Forward projection, reference semantics: casImgs[a] (M*M) is fully written
// (0 outside the mask disc).
// With the 501-entry table and check pointing at a flag launch_check_volume
// set to 1 (same stream, after the volume's last change), the gather reads
// only each axis' four taps that can have nonzero weight (same sums, smaller
// voxel boxes); check == nullptr gathers the full 5x5x5 stencils.
hipError_t launch_forward_project(const void *vol, int dtype, const ProjGeom &g,
                                  const float *axes, int nAxes, float *casImgs,
                                  hipStream_t s, const unsigned *check = nullptr);
// *check = 1 when the table has 501 finite entries with both ends 0 and the
// volume's n values (storage dtype) are all finite, else 0 (two launches on s).
hipError_t launch_check_volume(const void *vol, int dtype, size_t n, const float *ker, int K, unsigned *check,
                               hipStream_t s);
// Back projection, reference semantics (voxel-owner gather, no atomics).
// volAcc (V^3) and, if non-null, densAcc (V^3, the plane density = back
// projection of all-ones CAS images) are accumulated into.
hipError_t launch_back_project(float *volAcc, float *densAcc, const ProjGeom &g,
                               const float *axes, int nAxes, const float *casImgs,
                               hipStream_t s, const int *brickOrder, int orderLen);
// Forward projection over several volumes in one launch (FP32 storage, the
// 501-entry table): each segment is a run of directions on one volume,
// written to its own CAS images; work items are the segments' 8x8 pixel
// tiles, numbered segment after segment from workStart.
constexpr int kFpMaxSegments = 16;
struct FPSegment {
    const void *vol;
    const float *axes;
    float *out;
    const float *ker;
    const unsigned *check;  // launch_check_volume's flag for vol and ker, or nullptr
    int V, M, nAxes, maskD2, tilesPerRow, tilesPerImg;
    unsigned workStart;
};
struct FPMultiArgs {
    FPSegment seg[kFpMaxSegments];
    int nSegs;
    unsigned totalWork;
};
// True when a volume of this geometry can be a segment (kernel half width 2,
// the 501-entry table).
bool fp_segment_supported(const ProjGeom &g);
// Fills sg for nAxes directions of the volume vol (FP32) and returns its work
// items (tiles); the caller sets the next segment's workStart past them.
unsigned fp_segment(FPSegment &sg, const void *vol, const ProjGeom &g, const float *axes, int nAxes,
                    float *casImgs, unsigned workStart, const unsigned *check = nullptr);
hipError_t launch_forward_project_multi(const FPMultiArgs &m, hipStream_t s);
// The fast back projection walks bricks in a centre-out order; these give the
// table length (0 when the generic kernel is used) and fill it (host memory).
int bp_brick_count(int V, bool fast);
void bp_brick_order(int V, int *out);
// True when launch_back_project takes the fast path for this geometry.
bool bp_uses_fast_path(const ProjGeom &g);
// Exact transpose of launch_forward_project (atomic scatter into volAcc).
// With check pointing at launch_check_volume's flag for the nAxes * M * M CAS
// image values and the table, the scatter skips the taps of weight 0 (same
// sums, smaller boxes).
hipError_t launch_back_project_adjoint(float *volAcc, const ProjGeom &g, const float *axes,
                                       int nAxes, const float *casImgs, hipStream_t s,
                                       const unsigned *check = nullptr);

// ---- image chains (2-D, batched) ----------------------------------------
// CAS images (n x M x M) -> Hermitian half spectra (n x M x (M/2+1)) ready for
// an unnormalised C2R transform; fftshift folded in as a checkerboard sign.
hipError_t launch_cas_to_herm2d(const float *cas, float2 *herm, int M, int n, hipStream_t s);
// C2R output (n x M x M) -> images (n x N x N): sign, crop, scale.
hipError_t launch_real2d_to_imgs(const float *real, float *imgs, int M, int N, int n,
                                 float scale, hipStream_t s);
// images (n x N x N) -> zero-padded, sign-modulated real input (n x M x M).
hipError_t launch_imgs_to_real2d(const float *imgs, float *real, int M, int N, int n,
                                 hipStream_t s);
// R2C output (n x M x (M/2+1)) -> CAS images (n x M x M), times scale; with
// zeroNyq the index-0 row and column of every CAS image are written as 0.
hipError_t launch_herm2d_to_cas(const float2 *herm, float *cas, int M, int n, hipStream_t s,
                                float scale = 1.f, int zeroNyq = 0);
// Fused two-pass image chains (M = 256): tmp holds n * N * (M/2 + 1) float2.
bool fused_chain_supported(int M, int N);
hipError_t launch_fused_cas_to_images(const float *cas, float2 *tmp, float *imgs, int M, int N, int n,
                                      float scale, hipStream_t s);
hipError_t launch_fused_images_to_cas(const float *imgs, float2 *tmp, float *cas, int M, int N, int n,
                                      float scale, int zeroNyq, hipStream_t s);

// ---- volume chains (3-D) -----------------------------------------------
// volume (N^3) -> padded, precompensated, sign-modulated real input (M^3).
hipError_t launch_vol_to_real3d(const float *vol, int N, float *real, int M,
                                const float *precomp, hipStream_t s);
// R2C output (M x M x (M/2+1)) -> CAS volume (V^3, embedded at offset pad).
hipError_t launch_herm3d_to_casvol(const float2 *herm, int M, float *casVol, int V, int pad,
                                   hipStream_t s);
// CAS volume (V^3) -> Hermitian half spectrum for an unnormalised C2R.
// zeroNyquist = 1 reproduces FromCAS (index-0 planes set to 0).
hipError_t launch_casvol_to_herm3d(const float *casVol, int V, int pad, float2 *herm, int M,
                                   int zeroNyquist, hipStream_t s);
// C2R output (M^3) -> volume (N^3): sign, scale, precompensation, crop.
hipError_t launch_real3d_to_vol(const float *real, int M, float *vol, int N,
                                const float *precomp, float scale, hipStream_t s);

// ---- reductions, conversions, fills --------------------------------------
// out[i] = sum_p parts[p*stride + i] accumulated in FP64 (partial volumes).
hipError_t launch_reduce_partials(float *out, const float *parts, int nParts, size_t n,
                                  size_t stride, int accumulate, hipStream_t s);
// out[i] = num[i] / (den[i] + eps) in FP64 (plane-density division).
hipError_t launch_divide_density(float *out, const float *num, const float *den, size_t n,
                                 double eps, hipStream_t s);
// FP32 -> BF16 / FP16 storage copy of a CAS volume (round to nearest even).
hipError_t launch_convert_volume(void *dst, int dtype, const float *src, size_t n, hipStream_t s);
hipError_t launch_fill(float *p, float value, size_t n, hipStream_t s);

// ---- standalone filters (Filters/*.cu semantics) -------------------------
hipError_t launch_filter_cas_to_complex(const float *in, float2 *out, int size, int nSlices,
                                        hipStream_t s);
hipError_t launch_filter_complex_to_cas(const float2 *in, float *out, int size, int nSlices,
                                        hipStream_t s);
hipError_t launch_filter_complex_to_real(const float2 *in, float *out, int size, int nSlices,
                                         hipStream_t s);
hipError_t launch_filter_real_to_complex(const float *in, float2 *out, int size, int nSlices,
                                         hipStream_t s);
hipError_t launch_filter_fftshift2d(void *data, int elemBytes, int size, int nSlices,
                                    hipStream_t s);
hipError_t launch_filter_fftshift3d(void *data, int elemBytes, int size, hipStream_t s);
hipError_t launch_filter_pad(const float *in, float *out, int inSize, int outSize, int nSlices,
                             hipStream_t s);
hipError_t launch_filter_crop(const void *in, void *out, int elemBytes, int inSize, int outSize,
                              int nSlices, hipStream_t s);
hipError_t launch_filter_divide_scalar(float *data, float scalar, size_t n, hipStream_t s);
hipError_t launch_filter_divide_volume(float *a, const float *b, size_t n, hipStream_t s);
hipError_t launch_filter_multiply_volume(float *a, const float *b, size_t n, hipStream_t s);
hipError_t launch_filter_multiply_complex_volume(float2 *a, const float *b, size_t n,
                                                 hipStream_t s);
hipError_t launch_filter_add_volume(float *a, const float *b, size_t n, hipStream_t s);

}  // namespace mgg
