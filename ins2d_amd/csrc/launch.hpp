// Host launch wrappers of the kernels in kernels.hip.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <vector>

#include "device.hpp"

namespace ins2d {

// right-hand-side modes
enum : int { RHS_WITH_P = 1, RHS_SCALE_OM = 2 };

// Kernel timestamps for profiling: while g_ktimer is set (on the calling
// thread), the kernels of a launcher go out through launch_timed, which
// records the first kernel's own start timestamp into `start` and each
// kernel's end timestamp into `stop` (hipExtLaunchKernelGGL), so that the
// elapsed time between the two events is the launcher's kernel time without
// the cost of separate event commands on the stream.
struct KernelTimer {
  hipEvent_t start = nullptr, stop = nullptr;
  int launches = 0;
};
extern thread_local KernelTimer* g_ktimer;
// the same for the three kernels of the register-transform FFT solve, one
// timer each: 0 forward row transform, 1 column solve, 2 inverse row
// transform, 3 inverse row transform that adds dp into its output
extern thread_local KernelTimer* g_fft_timer[4];

template <typename F, typename... Args>
inline void launch_timed(F kernel, dim3 grid, dim3 block, unsigned shmem, hipStream_t st, Args... args) {
  if (KernelTimer* t = g_ktimer) {
    hipExtLaunchKernelGGL(kernel, grid, block, shmem, st, t->launches == 0 ? t->start : nullptr, t->stop, 0u, args...);
    ++t->launches;
  } else {
    hipLaunchKernelGGL(kernel, grid, block, shmem, st, args...);
  }
}

struct CGStateH {
  double* s;
  int* flag;
  double* part;
  int nblk;
  int maxit;
  double* e = nullptr;   // 4 + nblk doubles for launch_cg_iter2 (null: not used)
};

// missed[0] += 1 when the PCG's convergence test has not passed or it
// stopped at maxit (one thread)
void launch_cgx_missed(hipStream_t st, const CGStateH& S, int* missed);

void launch_rhs_generic(hipStream_t st, const DComp* dCu, const DComp* dCv, int nux, int nuy, int nvx, int nvy,
                        const double* u, const double* v, const double* p, int mode, double* out_u, double* out_v,
                        const StageCombine& scu, const StageCombine& scv, double* vt_u, double* vt_v);
// the same over split tiles (interior closed form, boundary ring generic)
// rows per thread of launch_rhs_split: its tiles are built with this
// rows_per_thread (split_tiles)
int rhs_split_rows();
void launch_rhs_split(hipStream_t st, const DComp* dCu, const DComp* dCv, const RTile* tiles, int ntiles,
                      const double* u, const double* v, const double* p, int mode, double* out_u, double* out_v,
                      const StageCombine& scu, const StageCombine& scv, double* vt_u, double* vt_v);
void launch_div_generic(hipStream_t st, const DComp* dCu, const DComp* dCv, int npx, int npy, const double* A,
                        const double* B, double ca, double cb, double ycoef, double oscale, int nu, double* f);
// tV = k1 - Om^-1 G p at the nu u and nv v points (generic operators)
void launch_sub_grad_generic(hipStream_t st, const DComp* dCu, const DComp* dCv, int nu, int nv, const double* k1,
                             const double* p, double* tV);
void launch_proj_generic(hipStream_t st, const DComp* dCu, const DComp* dCv, int nux, int nuy, int nvx, int nvy,
                         int nu, const double* Vn, const double* Vt, const double* dp, double dt, double c, double* V);
// out2 (optional; rhs_per_dual_ok shapes, 16-byte aligned arrays): also
// (F - G p) / Omega from the same F, as a RHS_WITH_P | RHS_SCALE_OM launch
// would form it
void launch_rhs_per(hipStream_t st, int order, int Nx, int Ny, double hx, double hy, double invRe, const double* u,
                    const double* v, const double* p, int mode, double* out, const StageCombine& sc, double* vt,
                    double* out2 = nullptr);
bool rhs_per_dual_ok(int order, int Nx);
void launch_div_per(hipStream_t st, int order, int Nx, int Ny, double hx, double hy, const double* A, const double* B,
                    double ca, double cb, double oscale, double* f);
void launch_proj_per(hipStream_t st, int order, int Nx, int Ny, double hx, double hy, const double* Vn,
                     const double* Vt, const double* dp, double dt, double c, double* V);
void launch_fft_poisson(hipStream_t st, const FFTPlan& P, const double* f, double* dp);
// the register-transform solve (fft_register_lengths only) with its ends
// fused: A non-null: the right-hand side is the periodic divergence of
// order 2 (oscale (hy dx + hx dy)(ca A + cb B)) or 4 (that of
// launch_div_per, fft_fused_div4 lengths only) of the velocity A, B (B may
// be null), computed in the first transform (f unused), else f;
// acc: out += dp, else out = dp
void launch_fft_poisson_fused(hipStream_t st, const FFTPlan& P, const double* f, double hx, double hy,
                              const double* A, const double* B, double ca, double cb, double oscale, double* out,
                              bool acc, int order = 2);
// both lengths take the register transforms (no row-major spectra buffer)
bool fft_register_lengths(int Nx, int Ny);
// the lengths whose forward row transform computes the fourth-order
// divergence itself (4096-point rows)
bool fft_fused_div4(int Nx, int Ny);
void launch_lap(hipStream_t st, const DLap& L, const double* p, double* q);
// the early exit of pressure_poisson.m:32-38: max |f| as 64 partial maxima
// (part: 64 doubles), then x = 0 when their maximum is below tol
void launch_absmax_part(hipStream_t st, const double* f, size_t n, double* part);
void launch_zero_if_small(hipStream_t st, const double* part, double tol, double* x, size_t n);
void launch_cg_single(hipStream_t st, const DLap& L, const double* f, double* x, double* r, double* z, double* p,
                      double* q, double tol2, int maxit, int x0_zero, int* out_it, double* out_rel);
void launch_cg_init(hipStream_t st, const DLap& L, const CGStateH& S, const double* f, const double* x, double* r,
                    double* z, double* p, int x0_zero, double tol2);
// PCG with a preconditioner applied by the caller between the phases:
// 0 r = f (+ f.f); [z = M^-1 r]; 1 r.z, p = z, initial scalars; per
// iteration: 2 q = A p, alpha, x and r, convergence test; [z = M^-1 r];
// 3 r.z, beta, p = z + beta p
void launch_cgx(hipStream_t st, const DLap& L, const CGStateH& S, int phase, const double* a, double* b, double* c,
                double* d, double tol2);
void launch_cg_iter(hipStream_t st, const DLap& L, const CGStateH& S, double* x, double* r, double* z, double* p,
                    double* q);
// the same iteration in two launches (k_cg2_apply, k_cg2_update; S.e set
// before launch_cg_init); p is double-buffered: the iteration reads p_in and
// leaves the new direction in p_out.  The stopping rule of an iteration is
// applied by the next iteration's first launch, or by launch_cg2_check after
// the last iteration of a chunk.  The scalars are launch_cg_iter's, with
// A p formed as A z + beta A p_old (equal up to rounding).
void launch_cg_iter2(hipStream_t st, const DLap& L, const CGStateH& S, double* x, double* r, double* z,
                     const double* p_in, double* p_out, double* q);
void launch_cg2_check(hipStream_t st, const CGStateH& S);
void launch_axpy(hipStream_t st, size_t n, double a, const double* x, double* y);

// fast-diagonalization direct Poisson solve (fdm.hip; factors from
// build_fdm_host in fdm_host.cpp)
struct FDMPlan {
  int nx, ny, nx0, ny0;
  const double *Vx, *VxT, *Vy, *VyT, *lx, *ly;
};
// dp = A^-1 f by fast diagonalization; the launches do nothing while *skip
// (optional device flag) is set; with emax (ceil(Npy / 16) doubles) dp = 0
// when max |f| < 1e-14 (the early exit of pressure_poisson.m:21-39); acc
// (with emax only): dp += A^-1 f, dp unchanged on the early exit
void launch_fdm_poisson(hipStream_t st, const FDMPlan& F, const double* f, double* dp, double* w1, double* w2,
                        const int* skip = nullptr, double* emax = nullptr, bool acc = false);

// accelerated PCG (pcg2.hip): optional warm start and two-level preconditioner
constexpr int kTLMaxCoarse = 2048;   // coarse unknowns solved in one workgroup
struct TLCoarse {
  int ncx, ncy, agg, nx0, ny0;
  const double *Vx, *Vy, *lx, *ly;   // coarse fast-diagonalization factors
  double* rc;                        // aggregate sums of r (ncx * ncy)
  double* ec;                        // coarse correction
};
struct TLPlan {
  DLap L;
  TLCoarse C;
  double* s;      // 8 scalars
  int* flag;      // 4 flags
  double* part;   // 3 per aggregate block
  int maxit;
  int coarse;
  double* en;     // 1 per aggregate block (energy partials of a start vector)
};
// IC(0)-preconditioned CG in one workgroup, triangular solves by
// anti-diagonal wavefronts (pcg2.hip); factor from build_ic0_host
struct DIC {
  const double *D, *cw, *cs, *cwx, *cwy;   // cwx / cwy null when not periodic
  int nx, ny;
  // 0: M = (D + L) D^-1 (D + L^T), the IC(0) factor of make_cholinc.m;
  // 1: M = L1 L1^T with L1 = (D + L) D^-1 the unit lower ILU(0) factor
  // (pressure_poisson.m:52-64: pcg(-A, -f, tol, maxit, L, L'))
  int lu;
};
void launch_cg_ic0(hipStream_t st, const DLap& L, const DIC& P, const double* f, double* x, double* r, double* z,
                   double* p, double* q, double tol2, int maxit, int x0_zero, int* out_it, double* out_rel);
void launch_tl_init(hipStream_t st, const TLPlan& P, const double* f, const double* x, double* r, double* pdir,
                    int x0_zero, double tol2);
void launch_tl_iter(hipStream_t st, const TLPlan& P, double* x, double* r, double* pdir, double* q);
void launch_tl_true_residual(hipStream_t st, const TLPlan& P, const double* f, const double* x, double* r);
// x += P Ac^-1 P^T (f - A x) (needs the coarse factors; r: work array)
void launch_tl_coarse_correct(hipStream_t st, const TLPlan& P, const double* f, double* x, double* r);
// x = P Ac^-1 P^T f, the coarse-grid solution (r: work array)
void launch_tl_coarse_solution(hipStream_t st, const TLPlan& P, const double* f, double* x, double* r);
// part[b] = sum over aggregate b of x (f + r), r = f - A x (one value per
// aggregate; their sum is 2 phi(x), phi(x) = -x^T A x / 2 + f^T x)
void launch_tl_energy(hipStream_t st, const TLPlan& P, const double* f, const double* x, const double* r,
                      double* part);
// 2 * sweeps damped Jacobi sweeps on x (tmp: work array)
void launch_tl_smooth(hipStream_t st, const TLPlan& P, const double* f, double* x, double* tmp, int sweeps,
                      double omega);
// fourth-order operators on grids with walls (w4.hip)
void launch_w4_rhs(hipStream_t st, const W4Dev& D, const double* V, const double* p, int mode, double* out,
                   const StageCombine& sc, double* vt);
void launch_w4_div(hipStream_t st, const W4Dev& D, const double* A, const double* B, double ca, double cb,
                   double ycoef, double oscale, double* f);
void launch_w4_proj(hipStream_t st, const W4Dev& D, const double* Vn, const double* Vt, const double* dp, double dt,
                    double c, double* V);
// dp = K f
void launch_w4_solve(hipStream_t st, const W4Dev& D, const double* f, double* dp);
// a (n x n, symmetric positive definite) <- -a^-1 + shift, by n Gauss-Jordan
// sweeps (rowk, colk: n doubles of work space each)
void w4_invert(hipStream_t st, double* a, int n, double* rowk, double* colk, double shift);
void launch_reduce_stats(hipStream_t st, const double* x, int nx, int ny, const double* wx, const double* wy,
                         double* part, int nblk);

}  // namespace ins2d
