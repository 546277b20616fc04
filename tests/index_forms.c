/* Checks the table-index forms the projection kernels use against the
 * reference's index arithmetic as nvcc contracts it (--fmad=true, the default;
 * Projection/gpuBackProject.cu:78-81, gpuForwardProject.cu:65-68):
 *   r = fl(p - v), r = clamp(r, -2, 2), k = roundf(fma(r, 125, 250)),
 *   k = clamp(k, 0, 500).
 * The kernels round with floor(s + 0.5) (one v_cvt_rpi_i32_f32) and replace
 * the clamp of r by a clamp of the index (or a table padded with copies of its
 * end entries).  Each form is checked twice: with floor(s + 0.5) exact and
 * with s + 0.5 first rounded to FP32, so the result does not depend on how the
 * instruction rounds internally.
 *
 *   N: floor(fma(fl(p - v), 125, 250) + 0.5)      (any v; depends on r only)
 *   A: floor(fma(p, 125, 250 - 125 v) + 0.5)      (claimed for v >= 9)
 *
 * Arguments:
 *   lo:hi  every float p in [v - 4.5, v + 4.5] for each integer v in [lo, hi];
 *          prints "lo hi pairs mismA_exact mismA_fp32"
 *   r      every float r in [-4.5, 4.5] through form N;
 *          prints "r count mismN_exact mismN_fp32" */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

static int sat(long k) { return k < 0 ? 0 : (k > 500 ? 500 : (int)k); }

static int ref_from_r(float r)
{
    r = fminf(fmaxf(r, -2.0f), 2.0f);
    volatile float s = fmaf(r, 125.0f, 250.0f);
    return sat(lroundf(s));
}

static int rpi_exact(float s) { return sat((long)floor((double)s + 0.5)); }

static int rpi_fp32(float s)
{
    volatile float t = s + 0.5f;
    return sat((long)floorf(t));
}

int main(int argc, char **argv)
{
    for (int a = 1; a < argc; a++) {
        if (argv[a][0] == 'r' && argv[a][1] == 0) {
            long long n = 0, me = 0, mf = 0;
            for (float r = -4.5f; r <= 4.5f; r = nextafterf(r, INFINITY)) {
                const int k = ref_from_r(r);
                const float s = fmaf(r, 125.0f, 250.0f);
                n++;
                me += rpi_exact(s) != k;
                mf += rpi_fp32(s) != k;
            }
            printf("r %lld %lld %lld\n", n, me, mf);
            continue;
        }
        int lo, hi;
        if (sscanf(argv[a], "%d:%d", &lo, &hi) != 2) return 2;
        long long n = 0, me = 0, mf = 0;
        for (int v = lo; v <= hi; v++) {
            const float end = (float)v + 4.5f;
            const float c = 250.0f - 125.0f * (float)v;
            for (float p = (float)v - 4.5f; p <= end; p = nextafterf(p, INFINITY)) {
                volatile float r = p - (float)v;
                const int k = ref_from_r(r);
                const float s = fmaf(p, 125.0f, c);
                n++;
                me += rpi_exact(s) != k;
                mf += rpi_fp32(s) != k;
            }
        }
        printf("%d %d %lld %lld %lld\n", lo, hi, n, me, mf);
    }
    return 0;
}
