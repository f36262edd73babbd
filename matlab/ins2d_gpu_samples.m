function [S, R] = ins2d_gpu_samples(options, xin, yin, tn, dt)
% INS2D_GPU_SAMPLES  Boundary values of one RK step with time-dependent
% boundary conditions (BC_unsteady = 1), in the layout of
% ins2d_mex('step_unsteady', ...): the 12 sample columns at tn, at the stage
% times tn + c_i dt (i = 2..s) and at tn + dt (time_ERK.m:62-64,106-109),
% stacked; R: the 4 rate columns at tn + dt (pressure_additional_solve.m:8-15).
[~, ~, c] = getRKmethod(options.time.RK);
c = c(:)';
S = [];
for ti = [tn, tn + c(2:end) * dt, tn + dt]
    s = ins2d_gpu_bc_samples(options, xin, yin, ti);
    S = [S; vertcat(s{:})]; %#ok<AGROW>
end
r = ins2d_gpu_bc_rates(options, tn + dt);
R = vertcat(r{:});
end
