/*
 * The Matlab side of the MEX entry points: converts mxArray arguments into the
 * mgg_mex_arg records of include/mgg_mex.h, runs the entry point through
 * mgg_mex_call (libMultiGPUGridder.so) and reports its outcome with
 * mexErrMsgTxt / mexWarnMsgTxt, as the reference's MEX files do
 * (Mex/mexCreateGridder.cpp, mexSetVariables.cpp, mexMultiGPU*.cpp,
 * mexDeleteGridder.cpp).  Compiled only where Matlab's mex.h exists (see
 * build_mex.m); every check and all dispatch live in the library.
 */
#ifndef MGG_MEX_ADAPTER_H
#define MGG_MEX_ADAPTER_H

#include "mex.h"
#include "mgg_mex.h"

#include <string.h>

static int mgg_mex_class_of(const mxArray *a)
{
    switch (mxGetClassID(a)) {
    case mxDOUBLE_CLASS: return MGG_MEX_DOUBLE;
    case mxSINGLE_CLASS: return MGG_MEX_SINGLE;
    case mxINT8_CLASS: return MGG_MEX_INT8;
    case mxUINT8_CLASS: return MGG_MEX_UINT8;
    case mxINT16_CLASS: return MGG_MEX_INT16;
    case mxUINT16_CLASS: return MGG_MEX_UINT16;
    case mxINT32_CLASS: return MGG_MEX_INT32;
    case mxUINT32_CLASS: return MGG_MEX_UINT32;
    case mxINT64_CLASS: return MGG_MEX_INT64;
    case mxUINT64_CLASS: return MGG_MEX_UINT64;
    case mxLOGICAL_CLASS: return MGG_MEX_LOGICAL;
    case mxCHAR_CLASS: return MGG_MEX_CHAR;
    default: return MGG_MEX_OTHER;
    }
}

static void mgg_mex_run(const char *entry, int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    enum { kMaxArgs = 16 };
    mgg_mex_arg args[kMaxArgs];
    char *strs[kMaxArgs];
    char msg[1024];
    uint64_t handle = 0;
    int i, rc;
    if (nrhs > kMaxArgs) mexErrMsgTxt("too many input arguments");
    memset(args, 0, sizeof args);
    memset(strs, 0, sizeof strs);
    for (i = 0; i < nrhs; i++) {
        args[i].cls = mgg_mex_class_of(prhs[i]);
        args[i].is_complex = mxIsComplex(prhs[i]) ? 1 : 0;
        args[i].data = mxGetData(prhs[i]);
        args[i].numel = mxGetNumberOfElements(prhs[i]);
        strs[i] = args[i].cls == MGG_MEX_CHAR ? mxArrayToString(prhs[i]) : NULL;
        args[i].str = strs[i];
    }
    rc = mgg_mex_call(entry, nlhs, nrhs, args, &handle, msg, sizeof msg);
    for (i = 0; i < nrhs; i++)
        if (strs[i]) mxFree(strs[i]);
    if (rc > 0) mexErrMsgTxt(msg);
    if (rc < 0) mexWarnMsgTxt(msg);
    if (!strcmp(entry, "mexCreateGridder")) {
        /* the handle keeps the MEX file loaded until mexDeleteGridder
         * (Mex/mexFunctionWrapper.h:53-60) */
        mexLock();
        plhs[0] = mxCreateNumericMatrix(1, 1, mxUINT64_CLASS, mxREAL);
        *(uint64_t *)mxGetData(plhs[0]) = handle;
    } else if (!strcmp(entry, "mexDeleteGridder")) {
        mexUnlock();
    }
}

#endif /* MGG_MEX_ADAPTER_H */
