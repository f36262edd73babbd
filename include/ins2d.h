/*
 * libins2d_hip: MI355X (gfx950) implementation of the unsteady explicit
 * Runge-Kutta time stepping of INS2D (energy-conserving staggered finite
 * volumes for 2-D incompressible Navier-Stokes), FP64.
 *
 * C ABI: plain pointers and sizes, host arrays in the reference's layout
 *   V = [u(:); v(:)], u stored Nux_in x Nuy_in and v Nvx_in x Nvy_in,
 *   p Npx x Npy, x index fastest (column-major, operator_mesh.m:417-419).
 * Host arrays are borrowed for the duration of a call and copied; the state,
 * the operators, FFT plans and solver buffers live on the GPU inside the
 * handle.  One HIP stream per handle; a handle is not thread-safe.  Every call
 * returns with its results on the host, except ins2d_step, which leaves the
 * advanced state resident on the device.
 *
 * Each entry point names the reference function it replaces (paths relative
 * to the reference root).
 */
#ifndef INS2D_H
#define INS2D_H

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define INS2D_OK 0
#define INS2D_E_ARG 1          /* invalid argument / unsupported option   */
#define INS2D_E_HIP 2          /* HIP runtime error                       */
#define INS2D_E_NOCONV 3       /* PCG did not converge within maxit       */
#define INS2D_E_UNSUPPORTED 4  /* combination not implemented            */

/* boundary kinds (boundary_conditions.m:17-40) */
#define INS2D_BC_DIR 0
#define INS2D_BC_PER 1
#define INS2D_BC_PRES 2
#define INS2D_BC_SYM 3

/* pressure solvers (pressure_poisson.m:42-104 numbering) */
#define INS2D_POISSON_DIRECT 1 /* direct solve by fast diagonalization of the
                                  tensor-product A (singular all-Neumann A:
                                  the solution of zero area-weighted mean);
                                  fourth order with walls: the dense inverse
                                  of A (at most 8192 cells; singular A: the
                                  zero-mean solution)                        */
#define INS2D_POISSON_PCG_ILU 2 /* CG preconditioned by the unit lower ILU(0)
                                  factor L as M = L L^T (pcg + ilu 'nofill',
                                  pressure_poisson.m:52-64); at most 2^20
                                  cells, one workgroup                        */
#define INS2D_POISSON_CG 3     /* preconditioned CG (cg.F slot)             */
#define INS2D_POISSON_CG_IC 4  /* preconditioned CG (cg_matlab.m slot)      */
#define INS2D_POISSON_FFT 6    /* periodic uniform grids                    */

typedef struct ins2d_handle ins2d_handle;

/* Case description (the data main.m hands to solver_unsteady through
 * options: grid lines from <case>_mesh.m, BC types from <case>_BCtype.m, the
 * parameters of <case>_parameters.m and the Butcher tableau of
 * getRKmethod.m). */
typedef struct {
  int Nx, Ny;                /* pressure cells                                  */
  const double* x;           /* Nx+1 grid lines in x                            */
  const double* y;           /* Ny+1 grid lines in y                            */
  int bc_u[4];               /* u: left, right, low, up                         */
  int bc_v[4];               /* v: left, right, low, up                         */
  double Re;                 /* Reynolds number (1e100 = inviscid)              */
  int order4;                /* 0: second order, 1: fourth order: periodic
                                uniform grids with the FFT solve take the
                                closed-form kernels, any other grid whose
                                ends are Dirichlet or periodic the assembled
                                operators (CSR) with poisson = DIRECT and
                                steady boundary values                          */
  double alfa, beta;         /* fourth-order parameters (81, 9/8)               */
  int rk_stages;             /* s                                               */
  const double* rk_A;        /* s*s, row-major                                  */
  const double* rk_b;        /* s                                               */
  const double* rk_c;        /* s                                               */
  int poisson;               /* INS2D_POISSON_*                                 */
  int fft_symbol;            /* 0: operator_divergence.m:283 closed form
                                (second order, also used by the reference for
                                fourth order); 1: exact symbol of the
                                assembled A = M Om^-1 G                         */
  double cg_tol;             /* relative residual ||r||/||f|| (CG_acc)          */
  int cg_maxit;              /* CG_maxit                                        */
  int p_add_solve;           /* time-dependent boundary values: 1 additional
                                pressure solve per step, 0 the last stage
                                pressure (time_ERK.m:145-158); steady boundary
                                values always take the additional solve
                                (time_ERK.m:160-163)                            */
  int kernels;               /* 0: auto, 1: force generic kron-factored kernels */
  int cg_precond;            /* PCG modes: 0 Jacobi; 1 IC(0) as the reference
                                (make_cholinc.m, cg_matlab.m:135-198), one
                                workgroup with wavefront triangular solves,
                                up to 2^20 cells; 2 Jacobi plus an
                                aggregation coarse-grid correction (single
                                case: ins2d_set_poisson_accel flag 2, set at
                                create); 3 the fast-diagonalization inverse of
                                the pressure matrix (of each instance in an
                                ensemble) */
} ins2d_case_desc;

/* operator construction: replaces operators.m -> operator_mesh/averaging/
 * interpolation/divergence/convection_diffusion (spatial/operators/operator_*.m) and
 * the solver preparation of operator_divergence.m:219-291 */
int ins2d_create(const ins2d_case_desc* desc, ins2d_handle** out);
void ins2d_destroy(ins2d_handle* h);
/* h NULL: the error of the last failed ins2d_create or handle-free call */
const char* ins2d_last_error(const ins2d_handle* h);

/* sizes: out[0..6] = Nux_in, Nuy_in, Nvx_in, Nvy_in, Npx, Npy, kernel path
 * (0 generic or assembled fourth order, 2 periodic second order, 4 periodic
 * fourth order); xin (Nux_in),
 * yin (Nvy_in) are the interior velocity grid lines (operator_mesh.m:185-320) */
int ins2d_grid_info(const ins2d_handle* h, int* out, double* xin, double* yin);

/* boundary samples at time t, in the order set_bc_vectors.m:53-71 takes them:
 * uLo, uUp (x, len Nx+1), uLo_i, uUp_i (xin), uLe_i, uRi_i (yp, len Ny),
 * vLe, vRi (y, len Ny+1), vLo_i, vUp_i (xp, len Nx), vLe_i, vRi_i (yin).
 * Replaces set_bc_vectors.m; with time-dependent boundary values
 * (BC_unsteady = 1) the caller sets them at the time it then evaluates:
 * ins2d_rhs, ins2d_input_check and ins2d_diagnostics use the last set. */
int ins2d_set_bc(ins2d_handle* h, double t, const double* const* samples);

/* state transfer (host arrays, reference layout) */
int ins2d_set_state(ins2d_handle* h, const double* V, const double* p);
int ins2d_get_state(ins2d_handle* h, double* V, double* p);

/* F.m:1-117: F = -conv + diff - (G p + y_p); maxres = max|F| */
int ins2d_rhs(ins2d_handle* h, const double* V, const double* p, double t, double* F, double* maxres);

/* pressure_poisson.m: solve A dp = f (x0 = 0); iters/relres for CG modes */
int ins2d_poisson(ins2d_handle* h, const double* f, double* dp, int* iters, double* relres);

/* input_check.m:92-128 on the resident state: initial projection when
 * max|M V + yM| > 1e-12, then the initial pressure (pressure_additional_solve)
 * when p_initial = 1 */
int ins2d_input_check(ins2d_handle* h, double t, int p_initial);

/* nsteps of time_ERK.m:1-173 (with pressure_additional_solve.m at the end of
 * each step) on the resident state, t = tn, tn + dt, ... (steady boundary
 * values: those of the last ins2d_set_bc).  With cg_precond 3 (and profiling
 * off) each pressure solve runs one PCG iteration whose outcome the device
 * records, and the call waits on the host once; when a solve missed the
 * stopping rule, the call restores the state it started from and redoes its
 * steps with the convergence test read after every iteration (the results
 * are those of that path bit for bit).  When another step of the same call
 * follows (profiling off; the periodic paths with Nx even, a multiple of 128
 * at fourth order, or the generic operators), the F of a step's end
 * velocity is formed once for the additional pressure solve and for the
 * next step's first stage (the reference forms it twice); the results agree
 * with one-step calls to rounding. */
int ins2d_step(ins2d_handle* h, double tn, double dt, int nsteps);

/* time derivatives of the boundary values at the time of the last
 * ins2d_set_bc, in the order set_bc_vectors.m:73-78 takes them: dudtLe_i,
 * dudtRi_i (yp, len Ny), dvdtLo_i, dvdtUp_i (xp, len Nx).  They make ydM,
 * the boundary term of the additional pressure solve of ins2d_input_check
 * (pressure_additional_solve.m:8-21, set_bc_vectors.m:120-146); NULL clears
 * them (ydM = 0, as for steady boundary values). */
int ins2d_set_bc_rate(ins2d_handle* h, const double* const* rates);

/* nsteps of time_ERK.m:1-173 with time-dependent boundary values
 * (BC_unsteady = 1) on the resident state, step n from tn + n dt.  samples:
 * per step s + 1 sample sets, each the 12 arrays of ins2d_set_bc concatenated
 * in that order, at tn + n dt (time_ERK.m:62-64) and at the stage times
 * tn + (n + c_i) dt, i = 2..s, then tn + (n + 1) dt (time_ERK.m:106-109; c_i
 * from rk_c).  rates: per step the 4 arrays of ins2d_set_bc_rate at
 * tn + (n + 1) dt.  Each stage evaluates F with the boundary values at its
 * start (F.m:27-29) and takes a pressure solve (time_ERK.m:121-123); each step
 * ends with the additional solve with ydM (p_add_solve = 1) or the last stage
 * pressure (time_ERK.m:145-158).  Returns with the boundary state of
 * tn + nsteps dt set as by ins2d_set_bc and ins2d_set_bc_rate.  Needs a case
 * with boundaries (INS2D_E_UNSUPPORTED on the fully periodic kernels). */
int ins2d_step_unsteady(ins2d_handle* h, double tn, double dt, int nsteps, const double* samples,
                        const double* rates);

/* check_conservation.m:1-76 + the extra F of process_iteration.m:22 on the
 * resident state: out = {maxres, maxdiv, umom, vmom, k} */
int ins2d_diagnostics(ins2d_handle* h, double t, double* out);

/* last Poisson solve: iterations and relative residual (CG modes) */
int ins2d_solver_stats(const ins2d_handle* h, int* iters, double* relres, long long* total_iters);

/* one FP64 16x16x4 matrix-core product of row-major A (16 x 4) and B
 * (4 x 16) with the operand layout of the direct solve's GEMM; Draw receives
 * each lane's four accumulator values (lane * 4 + r), which pins the result
 * layout the GEMM assumes (lane l, register r = D[l / 16 + 4 r][l % 16]) */
int ins2d_fdm_mfma_probe(const double* A, const double* B, double* Draw);

/* true relative residual ||f - A x|| / ||f|| of a pressure solve */
int ins2d_residual(ins2d_handle* h, const double* f, const double* x, double* relres);

/* the right-hand side of the pressure equation for the resident velocity,
 * f = M Om^-1 (-conv + diff) (pressure_additional_solve.m:19-22 with p = 0):
 * its solution is the pressure consistent with V (second order) */
int ins2d_pressure_rhs(ins2d_handle* h, double t, double* f);

/* kernel timing with HIP events on the handle's stream:
 * enable != 0 starts accumulating; read returns the accumulated milliseconds
 * and launch count of one kernel class: 0 momentum RHS, 1 divergence,
 * 2 Poisson solve, 3 projection of the RK stages (time_ERK.m:78-132);
 * 4 momentum RHS, 5 divergence, 6 Poisson solve of the additional solve
 * (pressure_additional_solve.m:19-25); the kernels of the register-transform
 * FFT solve over all solves: 8 forward row transform, 9 column solve,
 * 10 inverse row transform, 11 inverse row transform adding dp into p */
int ins2d_profile(ins2d_handle* h, int enable);
int ins2d_profile_read(const ins2d_handle* h, int kind, double* ms, long long* count);

/* wait for all work queued on the handle's stream */
int ins2d_sync(ins2d_handle* h);

/* ---- batched ensemble (no reference counterpart: many solver_unsteady runs
 * of uneven size advanced together; second order, PCG pressure).  All
 * instances share the Butcher tableau, cg_tol, cg_maxit, cg_precond (0 Jacobi,
 * 2 two-level with 16 x 16 aggregates, at most 2048 coarse cells per
 * instance, 3 fast-diagonalization) and p_add_solve of descs[0]; each has its
 * own grid, boundary kinds and Re. */
typedef struct ins2d_batch ins2d_batch;
int ins2d_batch_create(const ins2d_case_desc* descs, int n, ins2d_batch** out);
void ins2d_batch_destroy(ins2d_batch* b);
const char* ins2d_batch_last_error(const ins2d_batch* b);
/* sizes[0..5] and interior grid lines of instance k (as ins2d_grid_info) */
int ins2d_batch_info(const ins2d_batch* b, int k, int* sizes, double* xin, double* yin);
/* boundary samples of instance k (same order as ins2d_set_bc) */
int ins2d_batch_set_bc(ins2d_batch* b, int k, double t, const double* const* samples);
int ins2d_batch_set_state(ins2d_batch* b, int k, const double* V, const double* p);
int ins2d_batch_get_state(ins2d_batch* b, int k, double* V, double* p);
/* input_check.m:92-128 for every instance */
int ins2d_batch_input_check(ins2d_batch* b, int p_initial);
/* nsteps of time_ERK.m for every instance, instance k with time step dt[k] */
int ins2d_batch_step(ins2d_batch* b, const double* dt, int nsteps);
/* pick the workgroup shape of the stage kernels (velocity tiles) by timing
 * the momentum right-hand side over all instances for a few candidates on the
 * resident state; returns the chosen shape (no reference counterpart) */
int ins2d_batch_autotune(ins2d_batch* b, int* tx, int* ty);
/* set that workgroup shape directly: tx * ty = 256 with tx in {16, 32, 64,
 * 128, 256} (no reference counterpart) */
int ins2d_batch_set_tile(ins2d_batch* b, int tx, int ty);
/* accumulated PCG iterations over all instances and solves, the largest
 * per-solve iteration count, total pressure cells */
int ins2d_batch_stats(const ins2d_batch* b, long long* total_iters, int* max_iters, long long* cells);

/* ---- learned Poisson warm start (no reference counterpart).  A U-shaped
 * CNN (4 levels, 16/32/64/64 channels, 3x3 convolutions, BF16 weights and
 * activations, FP32 accumulation on the matrix cores) maps a Poisson
 * right-hand side f on an H x W pressure grid (x fastest, H, W multiples of 8)
 * and optionally a coarse-grid solution xc (ins2d_poisson_coarse) to
 * p0 = xc + out_scale * rms(f) * net(f / rms(f), xc / (out_scale rms(f))),
 * the start vector of PCG (xc NULL: zero).  Parameters are one flat FP32
 * vector (ins2d_nn_num_params entries). */
typedef struct ins2d_nn ins2d_nn;
int ins2d_nn_create(int H, int W, ins2d_nn** out);
void ins2d_nn_destroy(ins2d_nn* n);
const char* ins2d_nn_last_error(const ins2d_nn* n);
long ins2d_nn_num_params(const ins2d_nn* n);
int ins2d_nn_set_params(ins2d_nn* n, const float* params, double out_scale);
int ins2d_nn_get_params(ins2d_nn* n, float* params);
/* He-normal weights from a seed, zero biases, fresh optimizer state */
int ins2d_nn_init(ins2d_nn* n, unsigned long long seed, double out_scale);
/* host arrays of H*W doubles (xc optional) */
int ins2d_nn_forward(ins2d_nn* n, const double* f, const double* xc, double* p);
/* device arrays, queued on the given HIP stream (NULL: the null stream); xc
 * optional and may be p itself.  Inference form: 14 kernel launches, the input
 * scaling, pooling, upsampling and the output layer fused into the
 * convolutions */
int ins2d_nn_forward_dev(ins2d_nn* n, const double* f, const double* xc, double* p, void* stream);
/* the same map through the training forward (input tensor, pooled and
 * upsampled tensors and the 1x1 output layer stored separately); equal to the
 * fused form up to FP32 summation order */
int ins2d_nn_forward_unfused(ins2d_nn* n, const double* f, const double* xc, double* p);
int ins2d_nn_forward_unfused_dev(ins2d_nn* n, const double* f, const double* xc, double* p, void* stream);
/* mean-squared-error loss of net(f_b / rms(f_b), xc_b / (out_scale rms(f_b)))
 * against (p_b - xc_b) / (out_scale rms(f_b)) over a batch of device arrays
 * (xc NULL, or an entry NULL: zero); gradients into the gradient vector;
 * mode bit 0: also one Adam step (lr, beta 0.9 / 0.999); bit 1: the relative
 * loss sum (y - t)^2 / sum t^2 per pair instead of the mean square (pairs whose
 * targets differ in scale weigh alike) */
int ins2d_nn_train_step(ins2d_nn* n, const double* const* f, const double* const* p, const double* const* xc,
                        int batch, float lr, int mode, double* loss_out, void* stream);
int ins2d_nn_get_grad(ins2d_nn* n, float* grad);
/* training pairs held by the network object in device memory: reserve
 * room for count pairs, upload pair k from host arrays (xc NULL: zero), and
 * train on the pairs idx[0..batch) (as ins2d_nn_train_step) */
int ins2d_nn_data_reserve(ins2d_nn* n, int count);
int ins2d_nn_data_set(ins2d_nn* n, int k, const double* f, const double* p, const double* xc);
int ins2d_nn_train(ins2d_nn* n, const int* idx, int batch, float lr, int mode, double* loss_out);
/* the device array of pair k: which 0 f, 1 p, 2 xc (NULL when out of
 * range); pairs can be written there directly (ins2d_pressure_rhs_dev,
 * ins2d_poisson_dev, ins2d_poisson_coarse_dev) */
double* ins2d_nn_data_ptr(ins2d_nn* n, int k, int which);
/* copy pair k to host arrays (any of them NULL: skipped) */
int ins2d_nn_data_get(ins2d_nn* n, int k, double* f, double* p, double* xc);
/* pair dst = a pair_i / rms(f_i) + b pair_j / rms(f_j) on all three arrays
 * (A is linear, so this is again an exact pair) */
int ins2d_nn_data_combine(ins2d_nn* n, int dst, int i, int j, double a, double b);
/* one convolution layer's kernels on host FP32 arrays (inputs, weights and
 * gradients rounded to BF16 where the kernels read BF16): out = act(conv(x)
 * + bias); din, dw, db = the input, weight and bias gradients for the output
 * gradient g.  Layouts as the network's: x (H, W, C) channels innermost,
 * w[t][cout][ca + cb] with t = ks * (dy + r) + (dx + r). */
int ins2d_nn_layer_check(int H, int W, int ca, int cb, int cout, int ks, int relu, const float* xa, const float* xb,
                         const float* w, const float* bias, const float* g, float* out, float* din, float* dw,
                         float* db);
/* one 16x16x16 BF16 matrix-core product D = A B of row-major 16x16 FP32
 * matrices rounded to BF16 (checks the operand layout the kernels use) */
int ins2d_nn_mfma_probe(const float* A, const float* B, float* D);
/* H (rows, y) and W (columns, x) the network was created for */
int ins2d_nn_shape(const ins2d_nn* n, int* H, int* W);

/* ---- accelerated PCG for the pressure modes 3/4 (no reference
 * counterpart).  flags bit 0: start from the network's p0 (nn on the handle's
 * pressure grid); bit 1: two-level preconditioner, Jacobi plus an aggregation
 * coarse-grid correction over agg x agg cells solved exactly (agg 0:
 * automatic); bit 2 (with bits 0 and 1): the network takes the coarse-grid
 * solution P Ac^-1 P^T f as its second input and predicts its correction;
 * bit 3 (with bit 1, without bit 0): start from the coarse-grid solution
 * alone; bits 8-15: damped Jacobi sweep pairs after a bit-2 start (0: 4).  A
 * start x0 is kept only when it is closer to the solution than x0 = 0 in the
 * (-A)-norm that CG minimises (x0 (f + f - A x0) < 0), else the solve falls
 * back to x0 = 0.  The stopping rule stays ||r|| <= cg_tol ||f|| (cg.F:247);
 * ins2d_solver_stats then reports the true relative residual
 * ||f - A x|| / ||f||.  flags 0 restores the plain PCG.  The network is
 * borrowed and must outlive its use. */
int ins2d_set_poisson_accel(ins2d_handle* h, ins2d_nn* nn, int flags, int agg);
/* device-array forms (HIP device pointers; each call returns after its
 * work): the pressure right-hand side of the resident velocity (as
 * ins2d_pressure_rhs), a pressure solve (as ins2d_poisson) and the coarse-grid
 * solution (as ins2d_poisson_coarse).  For generating training pairs in
 * device memory. */
int ins2d_pressure_rhs_dev(ins2d_handle* h, double t, double* f);
int ins2d_poisson_dev(ins2d_handle* h, const double* f, double* dp);
int ins2d_poisson_coarse_dev(ins2d_handle* h, const double* f, double* xc);
/* xc = P Ac^-1 P^T f, the coarse-grid solution of the two-level
 * preconditioner (needs bit 1 of the ins2d_set_poisson_accel flags); host
 * arrays */
int ins2d_poisson_coarse(ins2d_handle* h, const double* f, double* xc);
/* number of warm-start fallbacks so far, aggregate size in use */
int ins2d_poisson_accel_stats(const ins2d_handle* h, long long* fallbacks, int* agg);

/* ---- the periodic FFT pressure solve without a handle and without
 * allocation.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (work captured into HIP graphs, tight
 * serving loops, harnesses that forbid hipMalloc).  Solves A dp = f on the
 * periodic uniform Nx x Ny grid with spacings hx, hy as ins2d_poisson does
 * for a handle created with poisson = INS2D_POISSON_FFT and the same order4
 * and fft_symbol, with f and dp device arrays (x fastest) and the work queued
 * on the given HIP stream (NULL: the null stream); it returns once the work
 * is queued.  Nx and Ny must be powers of two from 16 to 4096 with the half
 * spectrum (Nx/2 + 1) * Ny at most 525312 complex values (e.g. 1024 x 1024 or
 * 4096 x 256): larger grids return INS2D_E_ARG.  The scratch memory (about
 * 8.7 MB per GPU, held until the process ends) is one fixed device array
 * that every call on a GPU shares: calls that overlap on one GPU race.  The
 * error text of a failed call is ins2d_last_error(NULL). */
int ins2d_poisson_fft_noalloc(int Nx, int Ny, double hx, double hy, int order4, int fft_symbol, const double* f,
                              double* dp, void* stream);

#ifdef __cplusplus
}
#endif
#endif
