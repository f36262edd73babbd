// Accelerated preconditioned CG for the pressure equation (no reference
// counterpart): the reference's Jacobi/IC PCG (cg.F:158-280) with two
// optional additions, both keeping its stopping rule ||r||^2 / ||f||^2 <= tol^2
// (cg.F:247), so the answer is judged exactly as the plain solver's:
//  * a start vector x0 from the learned network (the caller fills x before
//    the solve), smoothed by damped Jacobi sweeps, with a fallback to x0 = 0
//    when ||f - A x0|| > ||f||;
//  * a two-level preconditioner z = D^-1 r + P Ac^-1 P^T r, with P the 0/1
//    aggregation over agg x agg blocks of cells and Ac = P^T A P the Galerkin
//    coarse operator, solved exactly by fast diagonalization inside one
//    workgroup (Ac is again a sum of two Kronecker products on a tensor grid).
//    The preconditioner is symmetric and fixed, so standard PCG applies.
//
// Every kernel over cells uses one workgroup per aggregate (blockDim = agg x
// agg), so the restriction P^T r is a workgroup sum fused into the update.
#include <hip/hip_runtime.h>

#include "launch.hpp"
#include "stencil.hpp"

namespace ins2d {

namespace {

// s: 0 rz, 1 ff, 2 alpha, 3 beta, 4 rel^2, 5 tol^2, 6 rr0 (initial), 7 true rel^2
// flag: 0 done, 1 iterations, 2 failed
struct TL {
  DLap L;
  TLCoarse C;
  double* s;
  int* flag;
  double* part;   // 3 per aggregate block
  int maxit;
  int coarse;     // 1: two-level preconditioner
};

__device__ __forceinline__ int blk_id() { return blockIdx.x + gridDim.x * blockIdx.y; }

// r = f - A x (or f), partials ff, rr, r.D^-1 r, aggregate sums of r
__global__ void k_tl_init(TL T, const double* __restrict__ f, const double* __restrict__ x, double* __restrict__ r,
                          int x0_zero) {
  __shared__ double red[16];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  double fi = 0.0, ri = 0.0, di = 0.0;
  if (i < T.L.nx && j < T.L.ny) {
    const size_t idx = (size_t)i + (size_t)T.L.nx * j;
    fi = f[idx];
    ri = x0_zero ? fi : fi - lap_at(T.L, x, i, j);
    r[idx] = ri;
    di = T.L.dinv[idx];
  }
  const double ff = block_sum(fi * fi, red);
  const double rr = block_sum(ri * ri, red);
  const double rd = block_sum(ri * di * ri, red);
  const double rs = block_sum(ri, red);
  if (lin_tid() == 0) {
    const int b = blk_id();
    T.part[3 * b] = ff;
    T.part[3 * b + 1] = rr;
    T.part[3 * b + 2] = rd;
    T.C.rc[b] = rs;
  }
}

// q = A p, partial p.q
__global__ void k_tl_apply(TL T, const double* __restrict__ p, double* __restrict__ q) {
  __shared__ double red[16];
  if (T.flag[0]) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  double pq = 0.0;
  if (i < T.L.nx && j < T.L.ny) {
    const size_t idx = (size_t)i + (size_t)T.L.nx * j;
    const double a = lap_at(T.L, p, i, j);
    q[idx] = a;
    pq = p[idx] * a;
  }
  const double s = block_sum(pq, red);
  if (lin_tid() == 0) T.part[3 * blk_id()] = s;
}

__global__ void k_tl_alpha(TL T, int nblk) {
  __shared__ double red[16];
  if (T.flag[0]) return;
  double a = 0.0;
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) a += T.part[3 * b];
  const double pq = block_sum(a, red);
  if (threadIdx.x == 0) T.s[2] = T.s[0] / pq;
}

// x += alpha p, r -= alpha q; partials rr, r.D^-1 r; aggregate sums of r
__global__ void k_tl_update(TL T, const double* __restrict__ p, const double* __restrict__ q, double* __restrict__ x,
                            double* __restrict__ r) {
  __shared__ double red[16];
  if (T.flag[0]) return;
  const double alpha = T.s[2];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  double ri = 0.0, di = 0.0;
  if (i < T.L.nx && j < T.L.ny) {
    const size_t idx = (size_t)i + (size_t)T.L.nx * j;
    x[idx] += alpha * p[idx];
    ri = r[idx] - alpha * q[idx];
    r[idx] = ri;
    di = T.L.dinv[idx];
  }
  const double rr = block_sum(ri * ri, red);
  const double rd = block_sum(ri * di * ri, red);
  const double rs = block_sum(ri, red);
  if (lin_tid() == 0) {
    const int b = blk_id();
    T.part[3 * b + 1] = rr;
    T.part[3 * b + 2] = rd;
    T.C.rc[b] = rs;
  }
}

// one workgroup: convergence test, coarse solve ec = Ac^-1 rc, rz and beta.
// init: first call after k_tl_init (sets ff, rr0, beta = 0)
__global__ void __launch_bounds__(1024) k_tl_coarse(TL T, int nblk, int init, double tol2) {
  __shared__ double red[16];
  __shared__ double A[kTLMaxCoarse];
  __shared__ double B[kTLMaxCoarse];
  __shared__ double bc[2];
  __shared__ int done;
  if (!init && T.flag[0]) return;
  double ff = 0.0, rr = 0.0, rd = 0.0;
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
    if (init) ff += T.part[3 * b];
    rr += T.part[3 * b + 1];
    rd += T.part[3 * b + 2];
  }
  ff = block_sum(ff, red);
  rr = block_sum(rr, red);
  rd = block_sum(rd, red);
  if (threadIdx.x == 0) {
    int d = 0;
    if (init) {
      T.s[1] = ff;
      T.s[6] = rr;
      T.s[4] = rr / fmax(ff, 1e-300);
      T.s[5] = tol2;
      T.flag[1] = 0;
      T.flag[2] = 0;
      // cg.F:222: an already tiny right-hand side returns x0
      d = (ff <= tol2) ? 1 : 0;
    } else {
      const int it = T.flag[1] + 1;
      T.flag[1] = it;
      const double rel = rr / T.s[1];
      T.s[4] = rel;
      if (rel <= T.s[5]) {
        d = 1;
      } else if (it >= T.maxit) {
        d = 1;
        T.flag[2] = 1;
      }
    }
    T.flag[0] = d;
    done = d;
    bc[0] = rd;
  }
  __syncthreads();
  if (done) return;
  double rz = bc[0];
  if (T.coarse) {
    const TLCoarse& C = T.C;
    const int m = C.ncx, n = C.ncy, mn = m * n;
    for (int e = threadIdx.x; e < mn; e += blockDim.x) A[e] = C.rc[e];
    __syncthreads();
    // B = Vx^T A
    for (int e = threadIdx.x; e < mn; e += blockDim.x) {
      const int a = e % m, b = e / m;
      double s = 0.0;
      for (int k = 0; k < m; ++k) s += C.Vx[k + (size_t)m * a] * A[k + m * b];
      B[e] = s;
    }
    __syncthreads();
    // A = -(B Vy) / (lx + ly), null mode dropped
    for (int e = threadIdx.x; e < mn; e += blockDim.x) {
      const int a = e % m, b = e / m;
      double s = 0.0;
      for (int k = 0; k < n; ++k) s += B[a + m * k] * C.Vy[k + (size_t)n * b];
      A[e] = (a == C.nx0 && b == C.ny0) ? 0.0 : -s / (C.lx[a] + C.ly[b]);
    }
    __syncthreads();
    // B = Vx A
    for (int e = threadIdx.x; e < mn; e += blockDim.x) {
      const int a = e % m, b = e / m;
      double s = 0.0;
      for (int k = 0; k < m; ++k) s += C.Vx[a + (size_t)m * k] * A[k + m * b];
      B[e] = s;
    }
    __syncthreads();
    // ec = B Vy^T; rz += rc . ec
    double dot = 0.0;
    for (int e = threadIdx.x; e < mn; e += blockDim.x) {
      const int a = e % m, b = e / m;
      double s = 0.0;
      for (int k = 0; k < n; ++k) s += B[a + m * k] * C.Vy[b + (size_t)n * k];
      C.ec[e] = s;
      dot += C.rc[e] * s;
    }
    dot = block_sum(dot, red);
    if (threadIdx.x == 0) bc[1] = dot;
    __syncthreads();
    rz += bc[1];
  }
  if (threadIdx.x == 0) {
    if (init) {
      T.s[3] = 0.0;
    } else {
      T.s[3] = rz / T.s[0];
    }
    T.s[0] = rz;
  }
}

// p = D^-1 r + ec[aggregate] + beta p
__global__ void k_tl_dir(TL T, const double* __restrict__ r, double* __restrict__ p) {
  if (T.flag[0]) return;
  const double beta = T.s[3];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= T.L.nx || j >= T.L.ny) return;
  const size_t idx = (size_t)i + (size_t)T.L.nx * j;
  double z = T.L.dinv[idx] * r[idx];
  if (T.coarse) z += T.C.ec[blk_id()];
  p[idx] = z + beta * p[idx];
}

// damped Jacobi sweep xo = xi + omega D^-1 (f - A xi): removes the rough
// part of a start vector's error, which dominates its residual
__global__ void k_tl_jacobi(TL T, const double* __restrict__ f, const double* __restrict__ xi, double* __restrict__ xo,
                            double omega) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= T.L.nx || j >= T.L.ny) return;
  const size_t idx = (size_t)i + (size_t)T.L.nx * j;
  xo[idx] = xi[idx] + omega * T.L.dinv[idx] * (f[idx] - lap_at(T.L, xi, i, j));
}

// x += ec[aggregate] (coarse-grid correction of a start vector)
__global__ void k_tl_addcoarse(TL T, double* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= T.L.nx || j >= T.L.ny) return;
  x[(size_t)i + (size_t)T.L.nx * j] += T.C.ec[blk_id()];
}

// true relative residual of the final x: s[7] = ||f - A x||^2 / ||f||^2
__global__ void k_tl_true(TL T, int nblk) {
  __shared__ double red[16];
  double ff = 0.0, rr = 0.0;
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
    ff += T.part[3 * b];
    rr += T.part[3 * b + 1];
  }
  ff = block_sum(ff, red);
  rr = block_sum(rr, red);
  if (threadIdx.x == 0) T.s[7] = ff > 0.0 ? rr / ff : 0.0;
}

TL make(const TLPlan& P) {
  TL T;
  T.L = P.L;
  T.C = P.C;
  T.s = P.s;
  T.flag = P.flag;
  T.part = P.part;
  T.maxit = P.maxit;
  T.coarse = P.coarse;
  return T;
}

dim3 grid_of(const TLPlan& P) { return dim3(P.C.ncx, P.C.ncy); }
dim3 block_of(const TLPlan& P) { return dim3(P.C.agg, P.C.agg); }

}  // namespace

void launch_tl_init(hipStream_t st, const TLPlan& P, const double* f, const double* x, double* r, double* pdir,
                    int x0_zero, double tol2) {
  TL T = make(P);
  const int nblk = P.C.ncx * P.C.ncy;
  hipLaunchKernelGGL(k_tl_init, grid_of(P), block_of(P), 0, st, T, f, x, r, x0_zero);
  hipLaunchKernelGGL(k_tl_coarse, dim3(1), dim3(1024), 0, st, T, nblk, 1, tol2);
  hipLaunchKernelGGL(k_tl_dir, grid_of(P), block_of(P), 0, st, T, r, pdir);
}

void launch_tl_iter(hipStream_t st, const TLPlan& P, double* x, double* r, double* pdir, double* q) {
  TL T = make(P);
  const int nblk = P.C.ncx * P.C.ncy;
  hipLaunchKernelGGL(k_tl_apply, grid_of(P), block_of(P), 0, st, T, pdir, q);
  hipLaunchKernelGGL(k_tl_alpha, dim3(1), dim3(1024), 0, st, T, nblk);
  hipLaunchKernelGGL(k_tl_update, grid_of(P), block_of(P), 0, st, T, pdir, q, x, r);
  hipLaunchKernelGGL(k_tl_coarse, dim3(1), dim3(1024), 0, st, T, nblk, 0, 0.0);
  hipLaunchKernelGGL(k_tl_dir, grid_of(P), block_of(P), 0, st, T, r, pdir);
}

void launch_tl_smooth(hipStream_t st, const TLPlan& P, const double* f, double* x, double* tmp, int sweeps,
                      double omega) {
  TL T = make(P);
  for (int k = 0; k < sweeps; ++k) {
    hipLaunchKernelGGL(k_tl_jacobi, grid_of(P), block_of(P), 0, st, T, f, x, tmp, omega);
    hipLaunchKernelGGL(k_tl_jacobi, grid_of(P), block_of(P), 0, st, T, f, tmp, x, omega);
  }
}

void launch_tl_coarse_correct(hipStream_t st, const TLPlan& P, const double* f, double* x, double* r) {
  TL T = make(P);
  const int nblk = P.C.ncx * P.C.ncy;
  hipLaunchKernelGGL(k_tl_init, grid_of(P), block_of(P), 0, st, T, f, x, r, 0);
  hipLaunchKernelGGL(k_tl_coarse, dim3(1), dim3(1024), 0, st, T, nblk, 1, -1.0);
  hipLaunchKernelGGL(k_tl_addcoarse, grid_of(P), block_of(P), 0, st, T, x);
}

void launch_tl_true_residual(hipStream_t st, const TLPlan& P, const double* f, const double* x, double* r) {
  TL T = make(P);
  const int nblk = P.C.ncx * P.C.ncy;
  hipLaunchKernelGGL(k_tl_init, grid_of(P), block_of(P), 0, st, T, f, x, r, 0);
  hipLaunchKernelGGL(k_tl_true, dim3(1), dim3(1024), 0, st, T, nblk);
}

}  // namespace ins2d
