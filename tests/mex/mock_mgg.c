/* Test double of the gridder C ABI (include/mgg.h) for the MEX dispatch tests
 * on a machine without a GPU: every call appends one line to a log the test
 * reads back, and mgg_create hands out a fake object. */
#include "../../include/mgg.h"

#include <stdarg.h>
#include <stdio.h>
#include <string.h>

static char g_log[65536];
static size_t g_len;
static int g_fail;  /* status the next mocked call returns (then reset) */
static int g_fake_object;

static int logf_(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    int n = vsnprintf(g_log + g_len, sizeof g_log - g_len, fmt, ap);
    va_end(ap);
    if (n > 0 && g_len + (size_t)n < sizeof g_log) g_len += (size_t)n;
    int rc = g_fail;
    g_fail = 0;
    return rc;
}

const char *mock_log(void) { return g_log; }
void mock_reset(void) { g_len = 0; g_log[0] = 0; g_fail = 0; }
void mock_fail_next(int code) { g_fail = code; }
const char *mgg_last_error(void) { return "mock failure"; }

int mgg_create(int n, int p, float f, int pad, int ngpu, const int *gpus, int runfft, int verbose, mgg_gridder **out)
{
    int rc = logf_("create %d %d %g %d %d %d %d %d\n", n, p, f, pad, ngpu, gpus ? gpus[0] : -1, runfft, verbose);
    if (rc == 0) *out = (mgg_gridder *)&g_fake_object;
    return rc;
}
int mgg_destroy(mgg_gridder *g) { return logf_("destroy %d\n", g == (mgg_gridder *)&g_fake_object); }
#define SETARR(name)                                                                              \
    int name(mgg_gridder *g, float *p, const int *d)                                             \
    {                                                                                             \
        (void)g;                                                                                  \
        return logf_(#name " %g %d %d %d\n", p[0], d[0], d[1], d[2]);                             \
    }
SETARR(mgg_set_volume)
SETARR(mgg_set_images)
SETARR(mgg_set_cas_images)
SETARR(mgg_set_cas_volume)
int mgg_set_coord_axes(mgg_gridder *g, float *p, const int *d) { (void)g; return logf_("mgg_set_coord_axes %g %d %d\n", p[0], d[0], d[1]); }
int mgg_set_kb_table(mgg_gridder *g, float *p, const int *d) { (void)g; return logf_("mgg_set_kb_table %g %d\n", p[0], d[0]); }
int mgg_set_mask_radius(mgg_gridder *g, float r) { (void)g; return logf_("mgg_set_mask_radius %g\n", r); }
#define SETINT(name)                                                                     \
    int name(mgg_gridder *g, int v)                                                     \
    {                                                                                    \
        (void)g;                                                                         \
        return logf_(#name " %d\n", v);                                                  \
    }
SETINT(mgg_set_num_axes)
SETINT(mgg_set_max_axes_to_allocate)
SETINT(mgg_set_num_streams_fp)
SETINT(mgg_set_num_streams_bp)
int mgg_set_gpus(mgg_gridder *g, const int *gpus, int n) { (void)g; return logf_("mgg_set_gpus %d %d\n", n, n ? gpus[n - 1] : -1); }
#define ACTION(name)                                          \
    int name(mgg_gridder *g)                                  \
    {                                                         \
        return logf_(#name " %d\n", g == (mgg_gridder *)&g_fake_object); \
    }
ACTION(mgg_forward_project)
ACTION(mgg_back_project)
ACTION(mgg_cas_volume_to_host)
ACTION(mgg_reconstruct_volume)
