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
 * the back-projection slab gate) is evaluated in FP32 under one "reading" of
 * the reference source: the reference is built by nvcc with no flags
 * (src/CMakeLists.txt:14-16,80-83), and nvcc's default --fmad=true contracts a
 * multiply feeding an add into one fused multiply-add, so the source text
 * alone does not fix the arithmetic.  The readings restated here:
 *   ORC_READING_SOURCE  : no contraction, the source's operation order;
 *   ORC_READING_FIRST   : (default) contraction without reassociation, each
 *                         add of two products fusing its first operand:
 *                         a*x + b*y + c  -> fl(fma(a, x, fl(b*y)) + c),
 *                         A + B + C (+ c)-> fma(C, fma(A, fl(B))) (+ c),
 *                         r*kerScale + kerCenter -> fma(r, kerScale, kerCenter);
 *   ORC_READING_SECOND  : as FIRST with the second product fused instead;
 *   ORC_READING_REASSOC : FIRST plus the nested-FMA reassociation older LLVM
 *                         DAG combiners applied under contraction alone:
 *                         a*x + b*y + c -> fma(a, x, fma(b, y, c)),
 *                         A + B + C + c -> fma(C, fma(A, fma(B, c)));
 *   ORC_READING_INDEX   : only the table index fused (sums as SOURCE).
 * This file must be compiled with -ffp-contract=off and without -ffast-math
 * (oracle/Makefile does so): every fused operation is written as fmaf.
 * Accumulation here is FP64.
 *
 * The mask radius reaches the reference kernels through an int parameter
 * (gpuForwardProject.h:57, gpuBackProject.h:56; gpuGridder.cpp:262,422 pass
 * the float member), so every entry point truncates it toward zero first.
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

enum {
    ORC_READING_SOURCE = 0,
    ORC_READING_FIRST = 1,
    ORC_READING_SECOND = 2,
    ORC_READING_REASSOC = 3,
    ORC_READING_INDEX = 4,
};
static int g_reading = ORC_READING_FIRST;

/* Selects the arithmetic reading (see the header); returns the previous one,
 * or -1 for an unknown value (then nothing changes). */
int orc_set_reading(int reading)
{
    if (reading < ORC_READING_SOURCE || reading > ORC_READING_INDEX) return -1;
    int prev = g_reading;
    g_reading = reading;
    return prev;
}

int orc_get_reading(void) { return g_reading; }

/* The int conversion of the mask radius (gpuForwardProject.h:57). */
static inline float ref_mask_radius(float r)
{
    if (!(r > -2147483648.0f && r < 2147483648.0f)) return r;  /* NaN, huge: no int */
    return (float)(int)r;
}

/* a*x + b*y + c: a pixel's rotated coordinate (gpuForwardProject.cu:53-55) and
 * a back-projection tap's position (gpuBackProject.cu:77,84,91). */
static inline float sum2c(float a, float x, float b, float y, float c)
{
    switch (g_reading) {
    case ORC_READING_FIRST: return fmaf(a, x, b * y) + c;
    case ORC_READING_SECOND: return fmaf(b, y, a * x) + c;
    case ORC_READING_REASSOC: return fmaf(a, x, fmaf(b, y, c));
    default: {
        float p = a * x, q = b * y;
        float s = p + q;
        return s + c;
    }
    }
}

/* x0*a0 + x1*a1 + x2*a2, then + c when addC: a voxel's image-frame coordinate
 * (gpuBackProject.cu:62-64; fk has no constant). */
static inline float sum3c(float x0, float a0, float x1, float a1, float x2, float a2, float c, int addC)
{
    switch (g_reading) {
    case ORC_READING_FIRST: {
        float s = fmaf(x2, a2, fmaf(x0, a0, x1 * a1));
        return addC ? s + c : s;
    }
    case ORC_READING_SECOND: {
        float s = fmaf(x2, a2, fmaf(x1, a1, x0 * a0));
        return addC ? s + c : s;
    }
    case ORC_READING_REASSOC:
        if (addC) return fmaf(x2, a2, fmaf(x0, a0, fmaf(x1, a1, c)));
        return fmaf(x2, a2, fmaf(x0, a0, x1 * a1));
    default: {
        float p = x0 * a0, q = x1 * a1, t = x2 * a2;
        float s = p + q;
        s = s + t;
        return addC ? s + c : s;
    }
    }
}

/* Kernel-table lookup index for a signed offset r (gpuForwardProject.cu:63-67,
 * gpuBackProject.cu:79-82): clamp r to [-convW, convW], scale and offset,
 * round half away from zero, clamp to the table. */
static inline int kb_index(float r, float convW, float kerScale, float kerCenter, int K)
{
    r = fminf(fmaxf(r, -convW), convW);
    float s;
    if (g_reading == ORC_READING_SOURCE) {
        s = r * kerScale;
        s = s + kerCenter;
    } else {
        s = fmaf(r, kerScale, kerCenter);
    }
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
    maskRadius = ref_mask_radius(maskRadius);
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
                    float x = sum2c(nx[0], fi, ny[0], fj, (float)cv);
                    float y = sum2c(nx[1], fi, ny[1], fj, (float)cv);
                    float z = sum2c(nx[2], fi, ny[2], fj, (float)cv);
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
    maskRadius = ref_mask_radius(maskRadius);
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
                float x = sum2c(nx[0], fi, ny[0], fj, (float)cv);
                float y = sum2c(nx[1], fi, ny[1], fj, (float)cv);
                float z = sum2c(nx[2], fi, ny[2], fj, (float)cv);
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
    maskRadius = ref_mask_radius(maskRadius);
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
            const float di = (float)vi - cv, dj = (float)vj - cv, dk = (float)vk - cv;
            float fi = sum3c(di, nx[0], dj, nx[1], dk, nx[2], (float)ci, 1);
            float fj = sum3c(di, ny[0], dj, ny[1], dk, ny[2], (float)ci, 1);
            float fk = sum3c(di, nz[0], dj, nz[1], dk, nz[2], 0.f, 0);
            if (!(fabsf(fk) <= kerHWidth)) continue;
            int I = (int)roundf(fi), J = (int)roundf(fj);
            const float *im = img + (long long)a * M2;
            double cum = 0.0;
            for (int j1 = J - convW; j1 <= J + convW; j1++) {
                for (int i1 = I - convW; i1 <= I + convW; i1++) {
                    const float ti = (float)(i1 - ci), tj = (float)(j1 - ci);
                    float px = sum2c(ti, nx[0], tj, ny[0], (float)cv);
                    float w = ker[kb_index(px - vi, fW, kerScale, kerCenter, K)];
                    float py = sum2c(ti, nx[1], tj, ny[1], (float)cv);
                    w = w * ker[kb_index(py - vj, fW, kerScale, kerCenter, K)];
                    float pz = sum2c(ti, nx[2], tj, ny[2], (float)cv);
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
    maskRadius = ref_mask_radius(maskRadius);
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
                    float x = sum2c(nx[0], fi, ny[0], fj, (float)cv);
                    float y = sum2c(nx[1], fi, ny[1], fj, (float)cv);
                    float z = sum2c(nx[2], fi, ny[2], fj, (float)cv);
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
    maskRadius = ref_mask_radius(maskRadius);
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
                const float di = (float)vi - cv, dj = (float)vj - cv, dk = (float)vk - cv;
                float fi = sum3c(di, nx[0], dj, nx[1], dk, nx[2], (float)ci, 1);
                float fj = sum3c(di, ny[0], dj, ny[1], dk, ny[2], (float)ci, 1);
                float fk = sum3c(di, nz[0], dj, nz[1], dk, nz[2], 0.f, 0);
                if (!(fabsf(fk) <= kerHWidth)) continue;
                int I = (int)roundf(fi), J = (int)roundf(fj);
                const float *im = img + (long long)a * M2;
                float cum = 0.0f;
                for (int j1 = J - convW; j1 <= J + convW; j1++) {
                    for (int i1 = I - convW; i1 <= I + convW; i1++) {
                        const float ti = (float)(i1 - ci), tj = (float)(j1 - ci);
                        float px = sum2c(ti, nx[0], tj, ny[0], (float)cv);
                        float w = ker[kb_index(px - vi, fW, kerScale, kerCenter, K)];
                        float py = sum2c(ti, nx[1], tj, ny[1], (float)cv);
                        w = w * ker[kb_index(py - vj, fW, kerScale, kerCenter, K)];
                        float pz = sum2c(ti, nx[2], tj, ny[2], (float)cv);
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
