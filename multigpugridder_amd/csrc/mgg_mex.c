/*
 * MEX dispatch layer (include/mgg_mex.h): the reference's Matlab entry points
 * (the src/Mex directory) over the gridder C ABI, in plain C so that it builds and is
 * tested without Matlab.  The MEX files in multigpugridder_amd/mex only
 * convert mxArrays into mgg_mex_arg records and call mgg_mex_call.
 */
#include "../../include/mgg_mex.h"

#include "../../include/mgg.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* The record a Matlab handle points to (Mex/mexFunctionWrapper.h:29-47: a
 * signature, the object's type name and the object). */
typedef struct mgg_mex_handle {
    uint32_t signature;
    char type_name[16];
    mgg_gridder *g;
} mgg_mex_handle;

static const char kTypeName[] = "mgg_gridder";

/* Handles that mexCreateGridder returned and mexDeleteGridder has not freed.
 * A handle is dereferenced only after it is found here, so a stale or
 * arbitrary uint64 is refused instead of read (the reference reads whatever
 * the integer points at, Mex/mexFunctionWrapper.h:70-73).  MEX files run on
 * Matlab's one interpreter thread. */
typedef struct live_node {
    mgg_mex_handle *h;
    struct live_node *next;
} live_node;
static live_node *g_live;

static int live_add(mgg_mex_handle *h)
{
    live_node *n = (live_node *)malloc(sizeof *n);
    if (!n) return 0;
    n->h = h;
    n->next = g_live;
    g_live = n;
    return 1;
}

static int live_has(const mgg_mex_handle *h)
{
    for (const live_node *n = g_live; n; n = n->next)
        if (n->h == h) return 1;
    return 0;
}

static void live_remove(const mgg_mex_handle *h)
{
    for (live_node **pp = &g_live; *pp; pp = &(*pp)->next)
        if ((*pp)->h == h) {
            live_node *d = *pp;
            *pp = d->next;
            free(d);
            return;
        }
}

static int set_msg(char *msg, size_t len, const char *fmt, ...)
{
    if (msg && len) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(msg, len, fmt, ap);
        va_end(ap);
    }
    return 1;
}

static int is_numeric(int cls)
{
    return cls >= MGG_MEX_DOUBLE && cls <= MGG_MEX_LOGICAL;
}

/* Element i of a numeric array as a double (what mxGetScalar reads for i = 0). */
static double elem(const mgg_mex_arg *a, size_t i)
{
    switch (a->cls) {
    case MGG_MEX_DOUBLE: return ((const double *)a->data)[i];
    case MGG_MEX_SINGLE: return ((const float *)a->data)[i];
    case MGG_MEX_INT8: return ((const int8_t *)a->data)[i];
    case MGG_MEX_UINT8: return ((const uint8_t *)a->data)[i];
    case MGG_MEX_INT16: return ((const int16_t *)a->data)[i];
    case MGG_MEX_UINT16: return ((const uint16_t *)a->data)[i];
    case MGG_MEX_INT32: return ((const int32_t *)a->data)[i];
    case MGG_MEX_UINT32: return ((const uint32_t *)a->data)[i];
    case MGG_MEX_INT64: return (double)((const int64_t *)a->data)[i];
    case MGG_MEX_UINT64: return (double)((const uint64_t *)a->data)[i];
    case MGG_MEX_LOGICAL: return ((const uint8_t *)a->data)[i] ? 1.0 : 0.0;
    default: return NAN;
    }
}

static int scalar(const mgg_mex_arg *a, const char *what, double *out, char *msg, size_t len)
{
    if (!is_numeric(a->cls) || a->is_complex || a->numel < 1 || !a->data)
        return set_msg(msg, len, "%s must be a real numeric scalar", what);
    *out = elem(a, 0);
    return 0;
}

/* A real numeric scalar that is an integer in the int range. */
static int int_scalar(const mgg_mex_arg *a, const char *what, int *out, char *msg, size_t len)
{
    double d;
    if (scalar(a, what, &d, msg, len)) return 1;
    if (!(d == floor(d)) || d < -2147483648.0 || d > 2147483647.0)
        return set_msg(msg, len, "%s must be an integer, got %g", what, d);
    *out = (int)d;
    return 0;
}

/* A real numeric vector of integers, copied into out[0..max). */
static int int_vector(const mgg_mex_arg *a, const char *what, int *out, int max, int *n, char *msg, size_t len)
{
    if (!is_numeric(a->cls) || a->is_complex || (a->numel > 0 && !a->data))
        return set_msg(msg, len, "%s must be a real numeric vector", what);
    if (a->numel > (size_t)max) return set_msg(msg, len, "%s has %zu entries, at most %d allowed", what, a->numel, max);
    for (size_t i = 0; i < a->numel; i++) {
        const double d = elem(a, i);
        if (!(d == floor(d)) || d < -2147483648.0 || d > 2147483647.0)
            return set_msg(msg, len, "%s must hold integers, entry %zu is %g", what, i + 1, d);
        out[i] = (int)d;
    }
    *n = (int)a->numel;
    return 0;
}

/* Mex/mexFunctionWrapper.h:64-75: a real uint64 scalar pointing at a handle
 * record with the right signature and type name. */
static mgg_mex_handle *to_handle(const mgg_mex_arg *a, char *msg, size_t len)
{
    if (a->numel != 1 || a->cls != MGG_MEX_UINT64 || a->is_complex || !a->data) {
        set_msg(msg, len, "Input must be a real uint64 scalar.");
        return NULL;
    }
    mgg_mex_handle *h = (mgg_mex_handle *)(uintptr_t)((const uint64_t *)a->data)[0];
    if (!h || !live_has(h) || h->signature != MGG_MEX_HANDLE_SIGNATURE || strcmp(h->type_name, kTypeName) != 0) {
        set_msg(msg, len, "Handle not valid.");
        return NULL;
    }
    return h;
}

static int status(int rc, const char *entry, const char *what, char *msg, size_t len)
{
    if (rc == MGG_OK) return 0;
    return set_msg(msg, len, "%s: %s failed (status %d): %s", entry, what, rc, mgg_last_error());
}

static int create_gridder(int nrhs, const mgg_mex_arg *prhs, uint64_t *handle_out, char *msg, size_t len)
{
    static const char *E = "mexCreateGridder";
    if (nrhs != 8)
        return set_msg(msg, len,
                       "%s: There should be 8 inputs: Volume size, number of coordinate axes, the interpolation "
                       "factor, amount of extra padding, number of GPUs, a vector of GPU device numbers, a flag to "
                       "run the FFTs on the GPU or not, and a flag for verbose output.", E);
    if (!handle_out) return set_msg(msg, len, "%s: no output for the handle", E);
    int n, p, pad, ngpu, runfft, verbose, gpus[64], ng = 0;
    double f, vb;
    if (int_scalar(&prhs[0], "Volume size", &n, msg, len) || int_scalar(&prhs[1], "number of coordinate axes", &p, msg, len) ||
        scalar(&prhs[2], "interpolation factor", &f, msg, len) || int_scalar(&prhs[3], "extra padding", &pad, msg, len) ||
        int_scalar(&prhs[4], "number of GPUs", &ngpu, msg, len) ||
        int_vector(&prhs[5], "GPU device numbers", gpus, 64, &ng, msg, len) ||
        int_scalar(&prhs[6], "RunFFTOnGPU flag", &runfft, msg, len) || scalar(&prhs[7], "verbose flag", &vb, msg, len))
        return 1;
    verbose = vb != 0.0;
    if (ngpu > ng) return set_msg(msg, len, "%s: %d GPUs requested, %d device numbers given", E, ngpu, ng);
    mgg_mex_handle *h = (mgg_mex_handle *)calloc(1, sizeof *h);
    if (!h) return set_msg(msg, len, "%s: out of host memory", E);
    mgg_gridder *g = NULL;
    const int rc = mgg_create(n, p, (float)f, pad, ngpu, ng ? gpus : NULL, runfft, verbose, &g);
    if (rc != MGG_OK) {
        free(h);
        return status(rc, E, "mgg_create", msg, len);
    }
    h->signature = MGG_MEX_HANDLE_SIGNATURE;
    memcpy(h->type_name, kTypeName, sizeof kTypeName);
    h->g = g;
    if (!live_add(h)) {
        mgg_destroy(g);
        free(h);
        return set_msg(msg, len, "%s: out of host memory", E);
    }
    *handle_out = (uint64_t)(uintptr_t)h;
    return 0;
}

/* A single-precision array and its Matlab size vector, padded with trailing
 * ones to `want` entries (Matlab drops trailing singleton dimensions); the
 * array must hold exactly the product of the sizes. */
static int array_and_dims(const mgg_mex_arg *data, const mgg_mex_arg *dims, int want, const char *what, float **ptr,
                          int *d, char *msg, size_t len)
{
    if (data->cls != MGG_MEX_SINGLE || data->is_complex || !data->data)
        return set_msg(msg, len, "%s must be a real single array", what);
    int n = 0;
    if (int_vector(dims, "the size vector", d, 8, &n, msg, len)) return 1;
    if (n < 1 || n > want) return set_msg(msg, len, "%s: size vector has %d entries, expected up to %d", what, n, want);
    for (int i = n; i < want; i++) d[i] = 1;
    double prod = 1;
    for (int i = 0; i < want; i++) {
        if (d[i] < 0) return set_msg(msg, len, "%s: negative size", what);
        prod *= d[i];
    }
    if (prod != (double)data->numel)
        return set_msg(msg, len, "%s holds %zu values, its size vector says %.0f", what, data->numel, prod);
    *ptr = (float *)data->data;
    return 0;
}

static int set_variables(int nrhs, const mgg_mex_arg *prhs, char *msg, size_t len)
{
    static const char *E = "mexSetVariables";
    if (nrhs != 3 && nrhs != 4)
        return set_msg(msg, len, "%s: There should be 3 or 4 inputs: a command, the gridder handle, a value and "
                                 "(for arrays) its size vector.", E);
    if (prhs[0].cls != MGG_MEX_CHAR || !prhs[0].str || strlen(prhs[0].str) >= 64)
        return set_msg(msg, len, "First input should be a command string less than 64 characters long.");
    const char *cmd = prhs[0].str;
    mgg_mex_handle *h = to_handle(&prhs[1], msg, len);
    if (!h) return 1;
    mgg_gridder *g = h->g;
    int iv, d[8];
    double dv;
    float *ptr;
    /* array setters: value + size vector */
    struct { const char *name; int ndims; int (*fn)(mgg_gridder *, float *, const int *); } arrays[] = {
        {"SetVolume", 3, mgg_set_volume},   {"SetCoordAxes", 2, mgg_set_coord_axes},
        {"SetCASVolume", 3, mgg_set_cas_volume}, {"SetImages", 3, mgg_set_images},
        {"SetCASImages", 3, mgg_set_cas_images},
    };
    for (size_t k = 0; k < sizeof arrays / sizeof arrays[0]; k++) {
        if (strcmp(cmd, arrays[k].name)) continue;
        if (nrhs != 4) return set_msg(msg, len, "%s %s: needs the array and its size vector", E, cmd);
        if (array_and_dims(&prhs[2], &prhs[3], arrays[k].ndims, cmd, &ptr, d, msg, len)) return 1;
        return status(arrays[k].fn(g, ptr, d), E, cmd, msg, len);
    }
    if (!strcmp(cmd, "SetKBTable")) {
        /* size(KBTable) is a 1 x K or K x 1 vector: the table length is the product */
        if (nrhs != 4) return set_msg(msg, len, "%s %s: needs the table and its size vector", E, cmd);
        if (array_and_dims(&prhs[2], &prhs[3], 2, cmd, &ptr, d, msg, len)) return 1;
        const int k = d[0] * d[1];
        return status(mgg_set_kb_table(g, ptr, &k), E, cmd, msg, len);
    }
    if (!strcmp(cmd, "SetGPUs")) {
        int gpus[64], ng = 0;
        if (int_vector(&prhs[2], "GPU device numbers", gpus, 64, &ng, msg, len)) return 1;
        if (nrhs == 4) {
            int cnt;
            if (int_scalar(&prhs[3], "number of GPUs", &cnt, msg, len)) return 1;
            if (cnt != ng) return set_msg(msg, len, "%s SetGPUs: %d device numbers, count says %d", E, ng, cnt);
        }
        return status(mgg_set_gpus(g, gpus, ng), E, cmd, msg, len);
    }
    if (!strcmp(cmd, "SetMaskRadius")) {
        if (scalar(&prhs[2], "mask radius", &dv, msg, len)) return 1;
        return status(mgg_set_mask_radius(g, (float)dv), E, cmd, msg, len);
    }
    struct { const char *name; int (*fn)(mgg_gridder *, int); } ints[] = {
        {"SetMaxAxesToAllocate", mgg_set_max_axes_to_allocate},
        {"SetNumberStreamsFP", mgg_set_num_streams_fp},
        {"SetNumberStreamsBP", mgg_set_num_streams_bp},
        {"SetNumAxes", mgg_set_num_axes},
    };
    for (size_t k = 0; k < sizeof ints / sizeof ints[0]; k++) {
        if (strcmp(cmd, ints[k].name)) continue;
        if (int_scalar(&prhs[2], cmd, &iv, msg, len)) return 1;
        return status(ints[k].fn(g, iv), E, cmd, msg, len);
    }
    return set_msg(msg, len, "%s: unknown command '%s'", E, cmd);
}

static int one_handle_action(const char *E, int nrhs, const mgg_mex_arg *prhs, int (*fn)(mgg_gridder *),
                             const char *what, char *msg, size_t len)
{
    if (nrhs != 1) return set_msg(msg, len, "%s: There should be 1 input.", E);
    mgg_mex_handle *h = to_handle(&prhs[0], msg, len);
    if (!h) return 1;
    return status(fn(h->g), E, what, msg, len);
}

static int delete_gridder(int nlhs, int nrhs, const mgg_mex_arg *prhs, char *msg, size_t len)
{
    if (nrhs < 1) return set_msg(msg, len, "mexDeleteGridder: the gridder handle is missing");
    mgg_mex_handle *h = to_handle(&prhs[0], msg, len);
    if (!h) return 1;
    const int rc = mgg_destroy(h->g);
    live_remove(h);  /* a stale copy of the handle is refused from now on */
    h->signature = 0;
    h->g = NULL;
    free(h);
    if (rc != MGG_OK) return status(rc, "mexDeleteGridder", "mgg_destroy", msg, len);
    if (nlhs != 0 || nrhs != 1) {
        set_msg(msg, len, "Delete: Unexpected arguments ignored.");
        return -1;
    }
    return 0;
}

int mgg_mex_call(const char *entry, int nlhs, int nrhs, const mgg_mex_arg *prhs, uint64_t *handle_out, char *msg,
                 size_t msg_len)
{
    if (msg && msg_len) msg[0] = 0;
    if (!entry) return set_msg(msg, msg_len, "no entry point name");
    if (nrhs < 0 || (nrhs > 0 && !prhs)) return set_msg(msg, msg_len, "%s: bad argument list", entry);
    if (!strcmp(entry, "mexCreateGridder")) return create_gridder(nrhs, prhs, handle_out, msg, msg_len);
    if (!strcmp(entry, "mexSetVariables")) return set_variables(nrhs, prhs, msg, msg_len);
    if (!strcmp(entry, "mexMultiGPUForwardProject"))
        return one_handle_action(entry, nrhs, prhs, mgg_forward_project, "ForwardProject", msg, msg_len);
    if (!strcmp(entry, "mexMultiGPUBackProject"))
        return one_handle_action(entry, nrhs, prhs, mgg_back_project, "BackProject", msg, msg_len);
    if (!strcmp(entry, "mexMultiGPUGetVolume"))
        return one_handle_action(entry, nrhs, prhs, mgg_cas_volume_to_host, "CASVolumeToVolume", msg, msg_len);
    if (!strcmp(entry, "mexMultiGPUReconstructVolume"))
        return one_handle_action(entry, nrhs, prhs, mgg_reconstruct_volume, "ReconstructVolume", msg, msg_len);
    if (!strcmp(entry, "mexDeleteGridder")) return delete_gridder(nlhs, nrhs, prhs, msg, msg_len);
    return set_msg(msg, msg_len, "unknown MEX entry point '%s'", entry);
}
