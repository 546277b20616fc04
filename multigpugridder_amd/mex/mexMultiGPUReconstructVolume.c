/* mexMultiGPUReconstructVolume: the reference's MEX entry point of the same name, over libMultiGPUGridder.so. */
#include "mgg_mex_adapter.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    mgg_mex_run("mexMultiGPUReconstructVolume", nlhs, plhs, nrhs, prhs);
}
