/* Checks the fused table-index forms the back-projection kernel uses against
 * the reference's index arithmetic (Projection/gpuBackProject.cu:77-95):
 *   r = fl(p - v), r = clamp(r, -2, 2), k = roundf(fl(fl(r * 125) + 250)),
 *   k = clamp(k, 0, 500)
 * for every float p in [v - 4.5, v + 4.5] and each integer v given on the
 * command line (ranges lo:hi).  Forms (index saturated to [0, 500], the end
 * entries the padded table repeats):
 *   A: floor(fma(p, 125, 250 - 125 v) + 0.5)            (claimed for v >= 130)
 *   B: floor(fl(fma(p, 125, -125 v) + 250) + 0.5)       (claimed for v >= 9)
 * Prints one line per range: "lo hi pairs mismA mismB". */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

static int sat(long k) { return k < 0 ? 0 : (k > 500 ? 500 : (int)k); }

static int ref_index(float p, int v)
{
    volatile float r = p - (float)v;
    r = fminf(fmaxf(r, -2.0f), 2.0f);
    volatile float a = r * 125.0f;
    volatile float s = a + 250.0f;
    return sat(lroundf(s));
}

static int form_a(float p, int v)
{
    float s = fmaf(p, 125.0f, 250.0f - 125.0f * (float)v);
    return sat((long)floor((double)s + 0.5));
}

static int form_b(float p, int v)
{
    volatile float t = fmaf(p, 125.0f, -125.0f * (float)v);
    volatile float s = t + 250.0f;
    return sat((long)floor((double)s + 0.5));
}

int main(int argc, char **argv)
{
    for (int a = 1; a < argc; a++) {
        int lo, hi;
        if (sscanf(argv[a], "%d:%d", &lo, &hi) != 2) return 2;
        long long n = 0, ma = 0, mb = 0;
        for (int v = lo; v <= hi; v++) {
            const float end = (float)v + 4.5f;
            for (float p = (float)v - 4.5f; p <= end; p = nextafterf(p, INFINITY)) {
                const int k = ref_index(p, v);
                n++;
                ma += form_a(p, v) != k;
                mb += form_b(p, v) != k;
            }
        }
        printf("%d %d %lld %lld %lld\n", lo, hi, n, ma, mb);
    }
    return 0;
}
