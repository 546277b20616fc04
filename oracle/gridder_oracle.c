/*
 * CPU oracle for the Fourier-domain projection gridder.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this file: it is the checker used by tests/, by the smoke entry point
 * and by the cpu_baseline leg of bench.py.
 *
 * It restates the reference's projection kernels on the CPU:
 *   - orc_forward_project      : src/Projection/gpuForwardProject.cu:4-101
 *   - orc_back_project         : src/Projection/gpuBackProject.cu:4-113
 *   - orc_back_project_adjoint : exact transpose of orc_forward_project
 *                                (no reference counterpart; pins the autograd
 *                                backward)
 *   - cpu_forward_project_f32 / cpu_back_project_f32 : the same two kernels
 *     with the reference's FP32 accumulation and loop structure, OpenMP
 *     parallel, used as the timed CPU path.
 *
 * Numerics contract.  Every quantity that feeds a discrete decision (the
 * rotated coordinate, roundf of it, the kernel-table index, the mask test and
 * the back-projection slab gate) is evaluated in FP32 with the operation order
 * written in the reference source and with NO fused multiply-add contraction
 * (this file must be compiled with -ffp-contract=off and without -ffast-math;
 * oracle/Makefile does so).  The HIP kernels follow the same pinned order, so
 * both sides take identical decisions.  Accumulation here is FP64.
 *
 * Array layout (all arrays): column-major with x fastest,
 *   vol[x + y*V + z*V*V], img[a*M*M + i + j*M]   (gpuForwardProject.cu:90-93).
 * A direction is 9 floats [u; v; n]: u = image x-axis, v = image y-axis,
 * n = projection normal, all in volume coordinates (gpuForwardProject.cu:50-55).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Kernel-table lookup index for a signed offset r (gpuForwardProject.cu:63-67,
 * gpuBackProject.cu:79-82): clamp r to [-convW, convW], scale, round half
 * away from zero, clamp to the table. */
static inline int kb_index(float r, float convW, float kerScale, float kerCenter, int K)
{
    r = fminf(fmaxf(r, -convW), convW);
    float s = r * kerScale;
    s = s + kerCenter;
    int idx = (int)roundf(s);
    if (idx < 0) idx = 0;
    if (idx > K - 1) idx = K - 1;
    return idx;
}

/* Forward projection, reference semantics (gpuForwardProject.cu:4-101).
 * out: nAxes*M*M doubles; pixels outside the mask disc are written 0 (the
 * reference leaves them at the memset value 0, gpuGridder.cpp:232-236). */
void orc_forward_project(const float *vol, int V, const float *axes, int nAxes,
                         const float *ker, int K, float kerHWidth, int M,
                         float maskRadius, double *out)
{
    const int cv = V / 2, ci = M / 2;
    const int convW = (int)roundf(kerHWidth);
    const float fW = (float)convW;
    const float kerCenter = ((float)K - 1) / 2;
    const float kerScale = kerCenter / kerHWidth;
    const long long V3 = (long long)V * V * V;

    /* directions are independent: one thread per direction, same sums */
#pragma omp parallel for schedule(dynamic, 1)
    for (int a = 0; a < nAxes; a++) {
        const float *nx = axes + 9 * a;
        const float *ny = nx + 3;
        double *img = out + (long long)a * M * M;
        for (int j = 0; j < M; j++) {
            for (int i = 0; i < M; i++) {
                float r = sqrtf((float)(i - ci) * (i - ci) + (j - ci) * (j - ci));
                double acc = 0.0;
                if (r <= maskRadius) {
                    float fi = (float)(i - ci), fj = (float)(j - ci);
                    float x = nx[0] * fi + ny[0] * fj + (float)cv;
                    float y = nx[1] * fi + ny[1] * fj + (float)cv;
                    float z = nx[2] * fi + ny[2] * fj + (float)cv;
                    int X = (int)roundf(x), Y = (int)roundf(y), Z = (int)roundf(z);
                    for (int i1 = X - convW; i1 <= X + convW; i1++) {
                        float wi = ker[kb_index((float)i1 - x, fW, kerScale, kerCenter, K)];
                        for (int j1 = Y - convW; j1 <= Y + convW; j1++) {
                            float wj = ker[kb_index((float)j1 - y, fW, kerScale, kerCenter, K)];
                            for (int k1 = Z - convW; k1 <= Z + convW; k1++) {
                                float wk = ker[kb_index((float)k1 - z, fW, kerScale, kerCenter, K)];
                                float w = wi * wj * wk;
                                long long lin = (long long)k1 * V * V + (long long)j1 * V + i1;
                                if (lin < V3 && lin >= 0)
                                    acc += (double)w * (double)vol[lin];
                            }
                        }
                    }
                }
                img[(long long)j * M + i] = acc;
            }
        }
    }
}

/* Exact transpose of orc_forward_project: every (pixel, voxel, weight) triple
 * the forward gather uses is scattered back.  volOut (V^3 doubles) is
 * accumulated into, not cleared. */
void orc_back_project_adjoint(double *volOut, int V, const float *axes, int nAxes,
                              const float *img, int M, float maskRadius,
                              const float *ker, int K, float kerHWidth)
{
    const int cv = V / 2, ci = M / 2;
    const int convW = (int)roundf(kerHWidth);
    const float fW = (float)convW;
    const float kerCenter = ((float)K - 1) / 2;
    const float kerScale = kerCenter / kerHWidth;
    const long long V3 = (long long)V * V * V;

    for (int a = 0; a < nAxes; a++) {
        const float *nx = axes + 9 * a;
        const float *ny = nx + 3;
        const float *im = img + (long long)a * M * M;
        for (int j = 0; j < M; j++) {
            for (int i = 0; i < M; i++) {
                float r = sqrtf((float)(i - ci) * (i - ci) + (j - ci) * (j - ci));
                if (!(r <= maskRadius)) continue;
                double p = (double)im[(long long)j * M + i];
                if (p == 0.0) continue;
                float fi = (float)(i - ci), fj = (float)(j - ci);
                float x = nx[0] * fi + ny[0] * fj + (float)cv;
                float y = nx[1] * fi + ny[1] * fj + (float)cv;
                float z = nx[2] * fi + ny[2] * fj + (float)cv;
                int X = (int)roundf(x), Y = (int)roundf(y), Z = (int)roundf(z);
                for (int i1 = X - convW; i1 <= X + convW; i1++) {
                    float wi = ker[kb_index((float)i1 - x, fW, kerScale, kerCenter, K)];
                    for (int j1 = Y - convW; j1 <= Y + convW; j1++) {
                        float wj = ker[kb_index((float)j1 - y, fW, kerScale, kerCenter, K)];
                        for (int k1 = Z - convW; k1 <= Z + convW; k1++) {
                            float wk = ker[kb_index((float)k1 - z, fW, kerScale, kerCenter, K)];
                            float w = wi * wj * wk;
                            long long lin = (long long)k1 * V * V + (long long)j1 * V + i1;
                            if (lin < V3 && lin >= 0)
                                volOut[lin] += (double)w * p;
                        }
                    }
                }
            }
        }
    }
}

/* Back projection, reference semantics (gpuBackProject.cu:4-113): each voxel
 * inside the sphere gathers, for every direction whose plane passes within
 * kerHWidth of it, the 5x5 CAS-image window around its rounded projection.
 * The image bound is the reference's linear-index test (gpuBackProject.cu:99),
 * so i1 = -1 reads the last pixel of the previous row.  volOut is accumulated
 * into (the caller zeroes it, as gpuGridder.cpp:330-333 does). */
void orc_back_project(double *volOut, int V, const float *axes, int nAxes,
                      const float *img, int M, float maskRadius,
                      const float *ker, int K, float kerHWidth)
{
    const int cv = V / 2, ci = M / 2;
    const int convW = (int)roundf(kerHWidth);
    const float fW = (float)convW;
    const float kerCenter = ((float)K - 1) / 2;
    const float kerScale = kerCenter / kerHWidth;
    const long long M2 = (long long)M * M;

    /* each voxel is owned by one thread and sums its directions in order */
#pragma omp parallel for schedule(dynamic, 1)
    for (int vk = 0; vk < V; vk++)
    for (int vj = 0; vj < V; vj++)
    for (int vi = 0; vi < V; vi++) {
        float r = sqrtf((float)(vi - cv) * (vi - cv) + (vj - cv) * (vj - cv) + (vk - cv) * (vk - cv));
        if (!(r <= maskRadius)) continue;
        double accAll = 0.0;
        for (int a = 0; a < nAxes; a++) {
            const float *nx = axes + 9 * a;
            const float *ny = nx + 3;
            const float *nz = ny + 3;
            float fi = ((float)vi - cv) * nx[0] + ((float)vj - cv) * nx[1] + ((float)vk - cv) * nx[2] + ci;
            float fj = ((float)vi - cv) * ny[0] + ((float)vj - cv) * ny[1] + ((float)vk - cv) * ny[2] + ci;
            float fk = ((float)vi - cv) * nz[0] + ((float)vj - cv) * nz[1] + ((float)vk - cv) * nz[2];
            if (!(fabsf(fk) <= kerHWidth)) continue;
            int I = (int)roundf(fi), J = (int)roundf(fj);
            const float *im = img + (long long)a * M2;
            double cum = 0.0;
            for (int j1 = J - convW; j1 <= J + convW; j1++) {
                for (int i1 = I - convW; i1 <= I + convW; i1++) {
                    float px = (i1 - ci) * nx[0] + (j1 - ci) * ny[0] + cv;
                    float w = ker[kb_index(px - vi, fW, kerScale, kerCenter, K)];
                    float py = (i1 - ci) * nx[1] + (j1 - ci) * ny[1] + cv;
                    w = w * ker[kb_index(py - vj, fW, kerScale, kerCenter, K)];
                    float pz = (i1 - ci) * nx[2] + (j1 - ci) * ny[2] + cv;
                    w = w * ker[kb_index(pz - vk, fW, kerScale, kerCenter, K)];
                    long long lin = (long long)j1 * M + i1;
                    if (lin < M2 && lin >= 0)
                        cum += (double)im[lin] * (double)w;
                }
            }
            accAll += cum;
        }
        volOut[(long long)vk * V * V + (long long)vj * V + vi] += accAll;
    }
}

/* ---------------------------------------------------------------------------
 * The timed CPU path: the two reference kernels with FP32 accumulation in
 * the reference's loop structure (forward: pixel-parallel over directions,
 * accumulating into the output pixel; back: voxel-owner over the sphere),
 * OpenMP-parallel over output rows / voxel slabs.
 * ------------------------------------------------------------------------- */
void cpu_forward_project_f32(const float *vol, int V, const float *axes, int nAxes,
                             const float *ker, int K, float kerHWidth, int M,
                             float maskRadius, float *out, int nthreads)
{
    const int cv = V / 2, ci = M / 2;
    const int convW = (int)roundf(kerHWidth);
    const float fW = (float)convW;
    const float kerCenter = ((float)K - 1) / 2;
    const float kerScale = kerCenter / kerHWidth;
    const long long V3 = (long long)V * V * V;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
#endif
    for (int a = 0; a < nAxes; a++) {
        for (int j = 0; j < M; j++) {
            const float *nx = axes + 9 * a;
            const float *ny = nx + 3;
            float *img = out + (long long)a * M * M;
            for (int i = 0; i < M; i++) {
                float r = sqrtf((float)(i - ci) * (i - ci) + (j - ci) * (j - ci));
                float acc = 0.0f;
                if (r <= maskRadius) {
                    float fi = (float)(i - ci), fj = (float)(j - ci);
                    float x = nx[0] * fi + ny[0] * fj + (float)cv;
                    float y = nx[1] * fi + ny[1] * fj + (float)cv;
                    float z = nx[2] * fi + ny[2] * fj + (float)cv;
                    int X = (int)roundf(x), Y = (int)roundf(y), Z = (int)roundf(z);
                    for (int i1 = X - convW; i1 <= X + convW; i1++) {
                        float wi = ker[kb_index((float)i1 - x, fW, kerScale, kerCenter, K)];
                        for (int j1 = Y - convW; j1 <= Y + convW; j1++) {
                            float wj = ker[kb_index((float)j1 - y, fW, kerScale, kerCenter, K)];
                            for (int k1 = Z - convW; k1 <= Z + convW; k1++) {
                                float wk = ker[kb_index((float)k1 - z, fW, kerScale, kerCenter, K)];
                                float w = wi * wj * wk;
                                long long lin = (long long)k1 * V * V + (long long)j1 * V + i1;
                                if (lin < V3 && lin >= 0)
                                    acc = acc + w * vol[lin];
                            }
                        }
                    }
                }
                img[(long long)j * M + i] = acc;
            }
        }
    }
}

void cpu_back_project_f32(float *volOut, int V, const float *axes, int nAxes,
                          const float *img, int M, float maskRadius,
                          const float *ker, int K, float kerHWidth, int nthreads)
{
    const int cv = V / 2, ci = M / 2;
    const int convW = (int)roundf(kerHWidth);
    const float fW = (float)convW;
    const float kerCenter = ((float)K - 1) / 2;
    const float kerScale = kerCenter / kerHWidth;
    const long long M2 = (long long)M * M;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for collapse(2) schedule(dynamic, 1)
#endif
    for (int vk = 0; vk < V; vk++)
    for (int vj = 0; vj < V; vj++) {
        for (int vi = 0; vi < V; vi++) {
            float r = sqrtf((float)(vi - cv) * (vi - cv) + (vj - cv) * (vj - cv) + (vk - cv) * (vk - cv));
            if (!(r <= maskRadius)) continue;
            float accAll = 0.0f;
            for (int a = 0; a < nAxes; a++) {
                const float *nx = axes + 9 * a;
                const float *ny = nx + 3;
                const float *nz = ny + 3;
                float fi = ((float)vi - cv) * nx[0] + ((float)vj - cv) * nx[1] + ((float)vk - cv) * nx[2] + ci;
                float fj = ((float)vi - cv) * ny[0] + ((float)vj - cv) * ny[1] + ((float)vk - cv) * ny[2] + ci;
                float fk = ((float)vi - cv) * nz[0] + ((float)vj - cv) * nz[1] + ((float)vk - cv) * nz[2];
                if (!(fabsf(fk) <= kerHWidth)) continue;
                int I = (int)roundf(fi), J = (int)roundf(fj);
                const float *im = img + (long long)a * M2;
                float cum = 0.0f;
                for (int j1 = J - convW; j1 <= J + convW; j1++) {
                    for (int i1 = I - convW; i1 <= I + convW; i1++) {
                        float px = (i1 - ci) * nx[0] + (j1 - ci) * ny[0] + cv;
                        float w = ker[kb_index(px - vi, fW, kerScale, kerCenter, K)];
                        float py = (i1 - ci) * nx[1] + (j1 - ci) * ny[1] + cv;
                        w = w * ker[kb_index(py - vj, fW, kerScale, kerCenter, K)];
                        float pz = (i1 - ci) * nx[2] + (j1 - ci) * ny[2] + cv;
                        w = w * ker[kb_index(pz - vk, fW, kerScale, kerCenter, K)];
                        long long lin = (long long)j1 * M + i1;
                        if (lin < M2 && lin >= 0)
                            cum += im[lin] * w;
                    }
                }
                accAll += cum;
            }
            volOut[(long long)vk * V * V + (long long)vj * V + vi] += accAll;
        }
    }
}

/* Number of OpenMP threads the timed path will use (1 without OpenMP). */
int cpu_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
