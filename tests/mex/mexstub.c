/* Implementation of the test stand-in mex.h, plus the calls a test uses to
 * build arguments, run a mexFunction as Matlab would (mexErrMsgTxt ends the
 * call), and read the results. */
#include "mex.h"

#include <setjmp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

struct mxArray_tag {
    mxClassID cls;
    int is_complex;
    void *data;
    size_t numel;
    int owns;
    char *text;
};

static jmp_buf g_jmp;
static char g_err[1024], g_warn[1024];
static int g_locks;

mxClassID mxGetClassID(const mxArray *a) { return a->cls; }
int mxIsComplex(const mxArray *a) { return a->is_complex; }
void *mxGetData(const mxArray *a) { return a->data; }
size_t mxGetNumberOfElements(const mxArray *a) { return a->numel; }
char *mxArrayToString(const mxArray *a)
{
    if (a->cls != mxCHAR_CLASS || !a->text) return NULL;
    char *s = (char *)malloc(strlen(a->text) + 1);
    strcpy(s, a->text);
    return s;
}
void mxFree(void *p) { free(p); }
mxArray *mxCreateNumericMatrix(size_t m, size_t n, mxClassID cls, mxComplexity c)
{
    mxArray *a = (mxArray *)calloc(1, sizeof *a);
    size_t es = cls == mxDOUBLE_CLASS || cls == mxINT64_CLASS || cls == mxUINT64_CLASS ? 8 : 4;
    a->cls = cls;
    a->is_complex = c == mxCOMPLEX;
    a->numel = m * n;
    a->data = calloc(m * n ? m * n : 1, es);
    a->owns = 1;
    return a;
}
void mexErrMsgTxt(const char *msg)
{
    snprintf(g_err, sizeof g_err, "%s", msg);
    longjmp(g_jmp, 1);
}
void mexWarnMsgTxt(const char *msg) { snprintf(g_warn, sizeof g_warn, "%s", msg); }
void mexLock(void) { g_locks++; }
void mexUnlock(void) { g_locks--; }

/* ---- test-side calls ---- */
mxArray *stub_array(int cls, size_t numel, void *data, int is_complex)
{
    mxArray *a = (mxArray *)calloc(1, sizeof *a);
    a->cls = (mxClassID)cls;
    a->is_complex = is_complex;
    a->data = data;  /* borrowed, as Matlab hands the MEX file the array's own memory */
    a->numel = numel;
    return a;
}
mxArray *stub_string(const char *s)
{
    mxArray *a = (mxArray *)calloc(1, sizeof *a);
    a->cls = mxCHAR_CLASS;
    a->text = (char *)malloc(strlen(s) + 1);
    strcpy(a->text, s);
    a->numel = strlen(s);
    a->data = a->text;
    return a;
}
void stub_free(mxArray *a)
{
    if (!a) return;
    if (a->owns) free(a->data);
    free(a->text);
    free(a);
}
typedef void (*mexfn)(int, mxArray **, int, const mxArray **);
/* Runs fn; 0 when it returned, 1 when it raised an error (text in err). */
int stub_call(mexfn fn, int nlhs, mxArray **plhs, int nrhs, const mxArray **prhs)
{
    g_err[0] = 0;
    g_warn[0] = 0;
    if (setjmp(g_jmp)) return 1;
    fn(nlhs, plhs, nrhs, prhs);
    return 0;
}
const char *stub_error(void) { return g_err; }
const char *stub_warning(void) { return g_warn; }
int stub_locks(void) { return g_locks; }
uint64_t stub_u64(const mxArray *a) { return *(const uint64_t *)a->data; }
