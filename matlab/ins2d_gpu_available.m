function ok = ins2d_gpu_available(options)
% INS2D_GPU_AVAILABLE  The GPU path is used when the case asks for it
% (options.gpu.enable = 1, e.g. main('LDC_unsteady','gpu')) and the MEX
% gateway is on the path, in the pattern of the reference's probe for the
% cg MEX file (spatial/operators/operator_divergence.m:219-234): otherwise
% the MATLAB path runs, with a message.
want = isfield(options, 'gpu') && isfield(options.gpu, 'enable') && options.gpu.enable == 1;
ok = want && exist(['ins2d_mex.' mexext], 'file') == 3;
if want && ~ok
    disp('No ins2d_mex file available; running the MATLAB path');
end
end
