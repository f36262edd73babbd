function R = ins2d_gpu_bc_rates(options, t)
% INS2D_GPU_BC_RATES  Time derivatives of the boundary values on the point
% sets of set_bc_vectors.m:73-78 (dudtLe_i, dudtRi_i, dvdtLo_i, dvdtUp_i) at
% time t: they make ydM of the additional pressure solve.
g = options.grid;
p = options.case.project;
R = {reshape(feval([p '_dudtBC'], g.x(1), g.yp, t, options), [], 1), ...
     reshape(feval([p '_dudtBC'], g.x(end), g.yp, t, options), [], 1), ...
     reshape(feval([p '_dvdtBC'], g.xp, g.y(1), t, options), [], 1), ...
     reshape(feval([p '_dvdtBC'], g.xp, g.y(end), t, options), [], 1)};
end
