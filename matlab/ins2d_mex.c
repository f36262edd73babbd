/*
 * MEX gateway of libins2d_hip: the MATLAB side of the drop-in for
 * unsteady/time_ERK.m (called from unsteady/solver_unsteady.m:101-102) and
 * solvers/pressure/pressure_poisson.m, in the pattern of the reference's own
 * native solver (solvers/pressure/cg.F:31, discovered by
 * spatial/operators/operator_divergence.m:219-234).
 *
 * Build:  mex ins2d_mex.c -I<repo>/include -L<repo>/ins2d_amd -lins2d_hip
 *
 * Usage:  h = ins2d_mex('create', x, y, bcu, bcv, Re, order4, A, b, c, poisson, CG_acc, CG_maxit,
 *                       p_add_solve [, fft_symbol [, cg_precond]])
 *         ins2d_mex('set_bc', h, t, {12 sample columns})      set_bc_vectors.m:53-71 order
 *         ins2d_mex('set_bc_rate', h, {4 rate columns})       set_bc_vectors.m:73-78 order
 *         ins2d_mex('set_state', h, V, p)
 *         ins2d_mex('input_check', h, t, p_initial)
 *         ins2d_mex('step', h, tn, dt, nsteps)                time_ERK.m (steady BC)
 *         ins2d_mex('step_unsteady', h, tn, dt, nsteps, S, R) time_ERK.m (BC_unsteady = 1)
 *         d = ins2d_mex('diagnostics', h, t)                  [maxres maxdiv umom vmom k]
 *         [V, p] = ins2d_mex('get_state', h)
 *         [Nux,Nuy,Nvx,Nvy,Npx,Npy,xin,yin] = ins2d_mex('grid_info', h)
 *         [maxres, F] = ins2d_mex('rhs', h, V, p, t)          F.m
 *         [dp, iter, relres] = ins2d_mex('poisson', h, f)     pressure_poisson.m
 *         ins2d_mex('destroy', h)
 *
 * Errors: non-convergence of PCG raises INS2D:noconv, every other failure
 * INS2D:error (or INS2D:args / INS2D:handle for bad calls), with the text of
 * ins2d_last_error, as cg.F:262 raises a MATLAB error.  Handles live in a
 * table until 'destroy' or until MATLAB clears the MEX file (mexAtExit).
 */
#include <string.h>

#include "mex.h"
#include "ins2d.h"

#define MAXH 64
#define MAXS 8

static ins2d_handle* handles[MAXH];
static int registered = 0;

static void cleanup(void) {
  for (int i = 0; i < MAXH; ++i)
    if (handles[i]) {
      ins2d_destroy(handles[i]);
      handles[i] = NULL;
    }
}

static void need(int nrhs, int n, const char* cmd) {
  if (nrhs < n) mexErrMsgIdAndTxt("INS2D:args", "%s: %d arguments expected", cmd, n - 1);
}

static ins2d_handle* get(const mxArray* a) {
  int i = (int)mxGetScalar(a);
  if (i < 0 || i >= MAXH || !handles[i]) mexErrMsgIdAndTxt("INS2D:handle", "invalid handle");
  return handles[i];
}

static void check(const ins2d_handle* h, int rc) {
  if (rc) mexErrMsgIdAndTxt(rc == INS2D_E_NOCONV ? "INS2D:noconv" : "INS2D:error", "%s", ins2d_last_error(h));
}

static const double* numeric(const mxArray* a, const char* what) {
  const double* p = mxGetPr(a);
  if (!p) mexErrMsgIdAndTxt("INS2D:args", "%s: a double array expected", what);
  return p;
}

static const double* cell_item(const mxArray* a, int k, const char* what) {
  const mxArray* c = mxGetCell(a, k);
  if (!c) mexErrMsgIdAndTxt("INS2D:args", "%s: a cell array expected", what);
  return numeric(c, what);
}

static const double* column(const mxArray* a, size_t n, const char* what) {
  if (mxGetNumberOfElements(a) != n) mexErrMsgIdAndTxt("INS2D:args", "%s: %d values expected", what, (int)n);
  return numeric(a, what);
}

static void create(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  ins2d_case_desc d;
  double At[MAXS * MAXS];
  need(nrhs, 14, "create");
  memset(&d, 0, sizeof d);
  d.Nx = (int)mxGetNumberOfElements(prhs[1]) - 1;
  d.x = numeric(prhs[1], "x");
  d.Ny = (int)mxGetNumberOfElements(prhs[2]) - 1;
  d.y = numeric(prhs[2], "y");
  for (int k = 0; k < 4; ++k) {
    d.bc_u[k] = (int)column(prhs[3], 4, "bcu")[k];
    d.bc_v[k] = (int)column(prhs[4], 4, "bcv")[k];
  }
  d.Re = mxGetScalar(prhs[5]);
  d.order4 = (int)mxGetScalar(prhs[6]);
  d.alfa = 81.0;   /* accumulate_structure.m:94-96 */
  d.beta = 9.0 / 8.0;
  d.rk_stages = (int)mxGetNumberOfElements(prhs[8]);
  if (d.rk_stages < 1 || d.rk_stages > MAXS) mexErrMsgIdAndTxt("INS2D:args", "create: 1 to %d RK stages", MAXS);
  {
    /* getRKmethod.m returns A as a MATLAB (column-major) matrix; the C ABI
       takes it row-major */
    const double* A = column(prhs[7], (size_t)d.rk_stages * d.rk_stages, "A");
    for (int i = 0; i < d.rk_stages; ++i)
      for (int j = 0; j < d.rk_stages; ++j) At[i * d.rk_stages + j] = A[i + d.rk_stages * j];
  }
  d.rk_A = At;
  d.rk_b = numeric(prhs[8], "b");
  d.rk_c = column(prhs[9], (size_t)d.rk_stages, "c");
  d.poisson = (int)mxGetScalar(prhs[10]);
  d.cg_tol = mxGetScalar(prhs[11]);
  d.cg_maxit = (int)mxGetScalar(prhs[12]);
  d.p_add_solve = (int)mxGetScalar(prhs[13]);
  d.fft_symbol = nrhs > 14 ? (int)mxGetScalar(prhs[14]) : 0;
  d.cg_precond = nrhs > 15 ? (int)mxGetScalar(prhs[15]) : 0;
  {
    int slot = 0;
    while (slot < MAXH && handles[slot]) ++slot;
    if (slot == MAXH) mexErrMsgIdAndTxt("INS2D:handle", "create: too many handles");
    check(NULL, ins2d_create(&d, &handles[slot]));
    if (nlhs > 0) plhs[0] = mxCreateDoubleScalar(slot);
  }
}

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  char cmd[32];
  if (!registered) {
    mexAtExit(cleanup);
    registered = 1;
  }
  if (nrhs < 1 || mxGetString(prhs[0], cmd, sizeof cmd)) mexErrMsgIdAndTxt("INS2D:args", "command expected");
  if (!strcmp(cmd, "create")) {
    create(nlhs, plhs, nrhs, prhs);
    return;
  }
  if (!strcmp(cmd, "destroy")) {
    need(nrhs, 2, cmd);
    int i = (int)mxGetScalar(prhs[1]);
    if (i >= 0 && i < MAXH && handles[i]) {
      ins2d_destroy(handles[i]);
      handles[i] = NULL;
    }
    return;
  }
  {
    static const char* known[] = {"set_bc", "set_bc_rate", "set_state", "input_check", "step", "step_unsteady",
                                  "diagnostics", "get_state", "grid_info", "rhs", "poisson"};
    int ok = 0;
    for (size_t k = 0; k < sizeof known / sizeof known[0]; ++k) ok |= !strcmp(cmd, known[k]);
    if (!ok) mexErrMsgIdAndTxt("INS2D:args", "unknown command %s", cmd);
  }
  need(nrhs, 2, cmd);
  ins2d_handle* h = get(prhs[1]);
  int n[7];
  check(h, ins2d_grid_info(h, n, NULL, NULL));
  const size_t NV = (size_t)n[0] * n[1] + (size_t)n[2] * n[3], Np = (size_t)n[4] * n[5];
  if (!strcmp(cmd, "set_bc")) {
    const double* s[12];
    need(nrhs, 4, cmd);
    if (mxGetNumberOfElements(prhs[3]) != 12) mexErrMsgIdAndTxt("INS2D:args", "set_bc: 12 samples expected");
    for (int k = 0; k < 12; ++k) s[k] = cell_item(prhs[3], k, "set_bc");
    check(h, ins2d_set_bc(h, mxGetScalar(prhs[2]), s));
  } else if (!strcmp(cmd, "set_bc_rate")) {
    const double* r[4];
    need(nrhs, 3, cmd);
    if (mxGetNumberOfElements(prhs[2]) != 4) mexErrMsgIdAndTxt("INS2D:args", "set_bc_rate: 4 rates expected");
    for (int k = 0; k < 4; ++k) r[k] = cell_item(prhs[2], k, "set_bc_rate");
    check(h, ins2d_set_bc_rate(h, r));
  } else if (!strcmp(cmd, "set_state")) {
    need(nrhs, 4, cmd);
    check(h, ins2d_set_state(h, column(prhs[2], NV, "V"), column(prhs[3], Np, "p")));
  } else if (!strcmp(cmd, "input_check")) {
    need(nrhs, 4, cmd);
    check(h, ins2d_input_check(h, mxGetScalar(prhs[2]), (int)mxGetScalar(prhs[3])));
  } else if (!strcmp(cmd, "step")) {
    need(nrhs, 5, cmd);
    check(h, ins2d_step(h, mxGetScalar(prhs[2]), mxGetScalar(prhs[3]), (int)mxGetScalar(prhs[4])));
  } else if (!strcmp(cmd, "step_unsteady")) {
    need(nrhs, 7, cmd);
    check(h, ins2d_step_unsteady(h, mxGetScalar(prhs[2]), mxGetScalar(prhs[3]), (int)mxGetScalar(prhs[4]),
                                 numeric(prhs[5], "S"), numeric(prhs[6], "R")));
  } else if (!strcmp(cmd, "diagnostics")) {
    need(nrhs, 3, cmd);
    plhs[0] = mxCreateDoubleMatrix(1, 5, mxREAL);
    check(h, ins2d_diagnostics(h, mxGetScalar(prhs[2]), mxGetPr(plhs[0])));
  } else if (!strcmp(cmd, "get_state")) {
    mxArray* V = mxCreateDoubleMatrix(NV, 1, mxREAL);
    mxArray* p = mxCreateDoubleMatrix(Np, 1, mxREAL);
    check(h, ins2d_get_state(h, mxGetPr(V), mxGetPr(p)));
    plhs[0] = V;
    if (nlhs > 1) plhs[1] = p;
    else mxDestroyArray(p);
  } else if (!strcmp(cmd, "grid_info")) {
    for (int k = 0; k < 6 && k < nlhs; ++k) plhs[k] = mxCreateDoubleScalar(n[k]);
    if (nlhs > 7) {
      plhs[6] = mxCreateDoubleMatrix(n[0], 1, mxREAL);
      plhs[7] = mxCreateDoubleMatrix(n[3], 1, mxREAL);
      check(h, ins2d_grid_info(h, n, mxGetPr(plhs[6]), mxGetPr(plhs[7])));
    }
  } else if (!strcmp(cmd, "rhs")) {
    double mr = 0.0;
    need(nrhs, 5, cmd);
    mxArray* F = mxCreateDoubleMatrix(NV, 1, mxREAL);
    check(h, ins2d_rhs(h, column(prhs[2], NV, "V"), column(prhs[3], Np, "p"), mxGetScalar(prhs[4]), mxGetPr(F),
                       &mr));
    plhs[0] = mxCreateDoubleScalar(mr);
    if (nlhs > 1) plhs[1] = F;
    else mxDestroyArray(F);
  } else if (!strcmp(cmd, "poisson")) {
    int it = 0;
    double rel = 0.0;
    need(nrhs, 3, cmd);
    plhs[0] = mxCreateDoubleMatrix(Np, 1, mxREAL);
    check(h, ins2d_poisson(h, column(prhs[2], Np, "f"), mxGetPr(plhs[0]), &it, &rel));
    if (nlhs > 1) plhs[1] = mxCreateDoubleScalar(it);
    if (nlhs > 2) plhs[2] = mxCreateDoubleScalar(rel);
  } else {
    mexErrMsgIdAndTxt("INS2D:args", "unknown command %s", cmd);
  }
}
