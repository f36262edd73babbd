function [V, p, T] = ins2d_gpu_time_ERK(h, Vn, pn, Tn, tn, dt, options, xin, yin, resident)
% INS2D_GPU_TIME_ERK  One step of unsteady/time_ERK.m on the GPU: the drop-in
% for the call in unsteady/solver_unsteady.m:101-102 (method 20).  The state
% stays on the GPU between steps; with resident = true, Vn and pn are taken
% to be the state the GPU already holds (no upload), and V, p are downloaded
% only when the caller asks for them.
if nargin < 10 || ~resident
    ins2d_mex('set_state', h, Vn, pn);
end
if options.BC.BC_unsteady == 1
    [S, R] = ins2d_gpu_samples(options, xin, yin, tn, dt);
    ins2d_mex('step_unsteady', h, tn, dt, 1, S, R);
else
    ins2d_mex('step', h, tn, dt, 1);
end
if nargout > 0
    [V, p] = ins2d_mex('get_state', h);
end
T = Tn;
end
