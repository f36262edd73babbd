function [h, xin, yin] = ins2d_gpu_create(options)
% INS2D_GPU_CREATE  Build the libins2d_hip handle of the case that main.m has
% set up: grid lines (options.grid.x, y), boundary kinds (options.BC), Re,
% the Butcher tableau of getRKmethod(options.time.RK), the pressure-solver
% settings, and the boundary values at t_start sampled on the point sets of
% set_bc_vectors.m:53-71.  The operators of spatial/operators/operator_*.m
% are built inside the library.
%
% options.gpu (optional): fft_symbol (0: the closed form of
% operator_divergence.m:283, 1: the exact symbol of the assembled A, for the
% fourth-order operators), cg_precond (0 Jacobi, 1 IC(0) as cg_matlab.m).

g  = options.grid;
BC = options.BC;
kind = containers.Map({'dir','per','pres','sym'}, {0,1,2,3});
bcu = [kind(BC.u.left) kind(BC.u.right) kind(BC.u.low) kind(BC.u.up)];
bcv = [kind(BC.v.left) kind(BC.v.right) kind(BC.v.low) kind(BC.v.up)];
[A, b, c] = getRKmethod(options.time.RK);
fft_symbol = 0;
cg_precond = 0;
if isfield(options, 'gpu')
    if isfield(options.gpu, 'fft_symbol'), fft_symbol = options.gpu.fft_symbol; end
    if isfield(options.gpu, 'cg_precond'), cg_precond = options.gpu.cg_precond; end
end
s = options.solversettings;
h = ins2d_mex('create', g.x(:), g.y(:), bcu, bcv, options.fluid.Re, options.discretization.order4, ...
              A, b(:), c(:), s.poisson, s.CG_acc, s.CG_maxit, s.p_add_solve, fft_symbol, cg_precond);
[~, ~, ~, ~, ~, ~, xin, yin] = ins2d_mex('grid_info', h);
t = options.time.t_start;
ins2d_mex('set_bc', h, t, ins2d_gpu_bc_samples(options, xin, yin, t));
if BC.BC_unsteady == 1
    % ydM of the initial pressure (pressure_additional_solve.m:8-15)
    ins2d_mex('set_bc_rate', h, ins2d_gpu_bc_rates(options, t));
end
end
