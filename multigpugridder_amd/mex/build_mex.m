% Builds the seven MEX entry points the reference's MultiGPUGridder_Matlab_Class
% calls (mexCreateGridder, mexSetVariables, mexMultiGPUForwardProject,
% mexMultiGPUBackProject, mexMultiGPUGetVolume, mexMultiGPUReconstructVolume,
% mexDeleteGridder) against libMultiGPUGridder.so.  Run from this folder in
% Matlab after building the library (make -C ../csrc).  The MEX files go to
% this folder; put it on the Matlab path ahead of the reference's Mex folder.
here = fileparts(mfilename('fullpath'));
lib = fullfile(here, '..');
inc = fullfile(here, '..', '..', 'include');
entries = {'mexCreateGridder', 'mexSetVariables', 'mexMultiGPUForwardProject', ...
           'mexMultiGPUBackProject', 'mexMultiGPUGetVolume', ...
           'mexMultiGPUReconstructVolume', 'mexDeleteGridder'};
for k = 1:numel(entries)
    mex('-R2018a', ['-I' inc], ['-L' lib], '-lMultiGPUGridder', ...
        ['LDFLAGS=$LDFLAGS -Wl,-rpath,' lib], '-outdir', here, ...
        fullfile(here, [entries{k} '.c']));
end
