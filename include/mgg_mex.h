/*
 * MEX dispatch layer: the Matlab entry points of the reference
 * (the src/Mex directory) as plain C over the gridder C ABI (mgg.h).
 *
 * The reference builds one MEX file per entry point, each a mexFunction over
 * a C++ MultiGPUGridder object held through a uint64 "class handle"
 * (Mex/mexFunctionWrapper.h:29-82).  Here all argument checking, command
 * dispatch, handle validation and status-to-message conversion live in this
 * layer, which compiles without Matlab and is exported by
 * libMultiGPUGridder.so; each MEX file (multigpugridder_amd/mex) only
 * converts its mxArray arguments into mgg_mex_arg records and turns a non-zero
 * return into mexErrMsgTxt.
 *
 * Handle convention (Mex/mexFunctionWrapper.h:29-82): mexCreateGridder
 * returns a real uint64 scalar holding the address of a handle record with
 * the signature 0xFF00F0A5 and a type name; every other entry point checks
 * that the argument is a real uint64 scalar naming a live handle (one created
 * and not yet deleted; the reference reads whatever the integer points at),
 * that the signature and the type name match, and fails with "Handle not
 * valid." otherwise.
 */
#ifndef MGG_MEX_H
#define MGG_MEX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MGG_MEX_HANDLE_SIGNATURE 0xFF00F0A5u

/* Matlab array classes the entry points accept (mxClassID subset). */
enum {
    MGG_MEX_OTHER = 0,
    MGG_MEX_DOUBLE,
    MGG_MEX_SINGLE,
    MGG_MEX_INT8,
    MGG_MEX_UINT8,
    MGG_MEX_INT16,
    MGG_MEX_UINT16,
    MGG_MEX_INT32,
    MGG_MEX_UINT32,
    MGG_MEX_INT64,
    MGG_MEX_UINT64,
    MGG_MEX_LOGICAL,
    MGG_MEX_CHAR
};

/* One right-hand-side argument of a mexFunction.  data is mxGetData (the
 * array's own memory, borrowed by the gridder where the reference borrows it);
 * numel is mxGetNumberOfElements; str is the NUL-terminated text of a char
 * array (NULL otherwise). */
typedef struct mgg_mex_arg {
    int cls;          /* MGG_MEX_* */
    int is_complex;   /* mxIsComplex */
    void *data;
    size_t numel;
    const char *str;
} mgg_mex_arg;

/* Runs one MEX entry point by its reference name:
 *   "mexCreateGridder"             (Mex/mexCreateGridder.cpp:12-27; 8 inputs,
 *                                   returns the handle in *handle_out)
 *   "mexSetVariables"              (Mex/mexSetVariables.cpp:27-97; command,
 *                                   handle, value[, dims]; commands SetVolume,
 *                                   SetCoordAxes, SetMaskRadius, SetCASVolume,
 *                                   SetMaxAxesToAllocate, SetNumberStreamsFP,
 *                                   SetNumberStreamsBP, SetImages, SetCASImages,
 *                                   SetKBTable, SetNumAxes and SetGPUs)
 *   "mexMultiGPUForwardProject"    (Mex/mexMultiGPUForwardProject.cpp)
 *   "mexMultiGPUBackProject"       (Mex/mexMultiGPUBackProject.cpp)
 *   "mexMultiGPUGetVolume"         (Mex/mexMultiGPUGetVolume.cpp: the
 *                                   CASVolumeToVolume copy to the CAS-volume array)
 *   "mexMultiGPUReconstructVolume" (an empty stub in the reference,
 *                                   Mex/mexMultiGPUReconstructVolume.cpp:9-14;
 *                                   here mgg_reconstruct_volume)
 *   "mexDeleteGridder"             (Mex/mexDeleteGridder.cpp:6-17)
 * Returns 0 on success, 1 on an error (msg holds the text for mexErrMsgTxt),
 * or -1 on success with a warning (msg holds the text for mexWarnMsgTxt).
 * msg may be NULL when msg_len is 0. */
int mgg_mex_call(const char *entry, int nlhs, int nrhs, const mgg_mex_arg *prhs, uint64_t *handle_out,
                 char *msg, size_t msg_len);

#ifdef __cplusplus
}
#endif

#endif /* MGG_MEX_H */
