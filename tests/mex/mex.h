/* Test stand-in for the part of Matlab's MEX API the adapters in
 * multigpugridder_amd/mex use (mxArray, class ids, data access, error and
 * warning reporting, locking).  Implemented by mexstub.c; lets the tests
 * compile the adapters and drive them as Matlab would, without Matlab. */
#ifndef MGG_TEST_MEX_STUB_H
#define MGG_TEST_MEX_STUB_H

#include <stddef.h>
#include <stdint.h>

typedef enum {
    mxUNKNOWN_CLASS = 0, mxCELL_CLASS, mxSTRUCT_CLASS, mxLOGICAL_CLASS, mxCHAR_CLASS, mxVOID_CLASS,
    mxDOUBLE_CLASS, mxSINGLE_CLASS, mxINT8_CLASS, mxUINT8_CLASS, mxINT16_CLASS, mxUINT16_CLASS,
    mxINT32_CLASS, mxUINT32_CLASS, mxINT64_CLASS, mxUINT64_CLASS, mxFUNCTION_CLASS
} mxClassID;
typedef enum { mxREAL = 0, mxCOMPLEX } mxComplexity;
typedef struct mxArray_tag mxArray;

mxClassID mxGetClassID(const mxArray *a);
int mxIsComplex(const mxArray *a);
void *mxGetData(const mxArray *a);
size_t mxGetNumberOfElements(const mxArray *a);
char *mxArrayToString(const mxArray *a);
void mxFree(void *p);
mxArray *mxCreateNumericMatrix(size_t m, size_t n, mxClassID cls, mxComplexity c);
void mexErrMsgTxt(const char *msg);
void mexWarnMsgTxt(const char *msg);
void mexLock(void);
void mexUnlock(void);

#endif
