function S = ins2d_gpu_bc_samples(options, xin, yin, t)
% INS2D_GPU_BC_SAMPLES  The 12 boundary-value columns of set_bc_vectors.m:53-71
% at time t, in the order ins2d_mex('set_bc', ...) takes them:
% uLo, uUp, uLo_i, uUp_i, uLe_i, uRi_i, vLe, vRi, vLo_i, vUp_i, vLe_i, vRi_i.
g = options.grid;
p = options.case.project;
uBC = @(x, y) reshape(feval([p '_uBC'], x, y, t, options), [], 1);
vBC = @(x, y) reshape(feval([p '_vBC'], x, y, t, options), [], 1);
S = {uBC(g.x, g.y(1)), uBC(g.x, g.y(end)), uBC(xin, g.y(1)), uBC(xin, g.y(end)), ...
     uBC(g.x(1), g.yp), uBC(g.x(end), g.yp), vBC(g.x(1), g.y), vBC(g.x(end), g.y), ...
     vBC(g.xp, g.y(1)), vBC(g.xp, g.y(end)), vBC(g.x(1), yin), vBC(g.x(end), yin)};
end
