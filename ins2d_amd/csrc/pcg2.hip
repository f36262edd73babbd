// Accelerated preconditioned CG for the pressure equation (no reference
// counterpart): the reference's Jacobi/IC PCG (cg.F:158-280) with two
// optional additions, both keeping its stopping rule ||r||^2 / ||f||^2 <= tol^2
// (cg.F:247), so the answer is judged exactly as the plain solver's:
//  * a start vector x0 from the learned network (the caller fills x before
//    the solve), smoothed by damped Jacobi sweeps, with a fallback to x0 = 0
//    unless x0 is closer to the solution than 0 in the (-A)-norm
//    (x0 . (f + r0) < 0, k_tl_energy);
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

// x = ec[aggregate] (the coarse-grid solution P Ac^-1 P^T f after an init
// with x0 = 0)
__global__ void k_tl_setcoarse(TL T, double* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= T.L.nx || j >= T.L.ny) return;
  x[(size_t)i + (size_t)T.L.nx * j] = T.C.ec[blk_id()];
}

// per-aggregate partials of x . (f + r) with r = f - A x: twice the energy
// phi(x) = -x^T A x / 2 + f^T x, whose difference to phi(A^-1 f) is half the
// squared (-A)-norm of the error; phi(0) = 0
__global__ void k_tl_energy(TL T, const double* __restrict__ f, const double* __restrict__ x,
                            const double* __restrict__ r, double* __restrict__ part) {
  __shared__ double red[16];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  double e = 0.0;
  if (i < T.L.nx && j < T.L.ny) {
    const size_t idx = (size_t)i + (size_t)T.L.nx * j;
    e = x[idx] * (f[idx] + r[idx]);
  }
  e = block_sum(e, red);
  if (lin_tid() == 0) part[blk_id()] = e;
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

// z = M^-1 r for M = (D + L) D^-1 (D + L^T) (cg_matlab.m:135-198): forward
// (D + L) q = r, then (I + D^-1 L^T) z = q in place; cell (i, j) depends on
// its west/south (forward) or east/north (backward) neighbours and on the
// periodic wrap cells, all on earlier anti-diagonals i + j in sweep order.
// P.lu: M = (D + L) D^-2 (D + L^T), the backward sweep solves
// (I + D^-1 L^T) z = D q
__device__ void ic0_apply(const DIC& P, const double* __restrict__ r, double* __restrict__ z) {
  const int nx = P.nx, ny = P.ny, T = blockDim.x;
  const int ns = nx + ny - 1;
  for (int s = 0; s < ns; ++s) {
    const int ilo = max(0, s - (ny - 1)), ihi = min(s, nx - 1);
    for (int i = ilo + (int)threadIdx.x; i <= ihi; i += T) {
      const int j = s - i;
      const size_t c = (size_t)i + (size_t)nx * j;
      double acc = r[c];
      if (i > 0) acc -= P.cw[c] * z[c - 1];
      if (j > 0) acc -= P.cs[c] * z[c - nx];
      if (P.cwx && i == nx - 1) acc -= P.cwx[j] * z[(size_t)nx * j];
      if (P.cwy && j == ny - 1) acc -= P.cwy[i] * z[i];
      z[c] = acc / P.D[c];
    }
    __syncthreads();
  }
  for (int s = ns - 1; s >= 0; --s) {
    const int ilo = max(0, s - (ny - 1)), ihi = min(s, nx - 1);
    for (int i = ilo + (int)threadIdx.x; i <= ihi; i += T) {
      const int j = s - i;
      const size_t c = (size_t)i + (size_t)nx * j;
      double acc = 0.0;
      if (i < nx - 1) acc += P.cw[c + 1] * z[c + 1];
      if (j < ny - 1) acc += P.cs[c + nx] * z[c + nx];
      if (P.cwx && i == 0) acc += P.cwx[j] * z[c + nx - 1];
      if (P.cwy && j == 0) acc += P.cwy[i] * z[(size_t)i + (size_t)nx * (ny - 1)];
      z[c] = (P.lu ? P.D[c] * z[c] : z[c]) - acc / P.D[c];
    }
    __syncthreads();
  }
}

// IC(0)-PCG in one workgroup, the reference's iteration (cg_matlab.m:52-100):
// stop when ||r||^2 / ||f||^2 <= tol^2, checked before the preconditioner
__global__ void __launch_bounds__(1024) k_cg_ic0(DLap L, DIC P, const double* __restrict__ f, double* __restrict__ x,
                                                double* __restrict__ r, double* __restrict__ z, double* __restrict__ p,
                                                double* __restrict__ q, double tol2, int maxit, int x0_zero,
                                                int* __restrict__ out_it, double* __restrict__ out_rel) {
  __shared__ double red[16];
  __shared__ double bc[2];
  const int n = L.nx * L.ny;
  const int T = blockDim.x;
  double ff = 0.0;
  for (int idx = threadIdx.x; idx < n; idx += T) {
    const int i = idx % L.nx, j = idx / L.nx;
    r[idx] = x0_zero ? f[idx] : f[idx] - lap_at(L, x, i, j);
    if (x0_zero) x[idx] = 0.0;
    ff += f[idx] * f[idx];
  }
  ff = block_sum(ff, red);
  if (threadIdx.x == 0) bc[0] = ff;
  __syncthreads();
  ff = bc[0];
  if (ff <= tol2) {
    if (threadIdx.x == 0) { *out_it = 0; *out_rel = 0.0; }
    return;
  }
  ic0_apply(P, r, z);
  double rz = 0.0;
  for (int idx = threadIdx.x; idx < n; idx += T) {
    p[idx] = z[idx];
    rz += r[idx] * z[idx];
  }
  rz = block_sum(rz, red);
  if (threadIdx.x == 0) bc[0] = rz;
  __syncthreads();
  rz = bc[0];
  int it = 0;
  double rel = 1.0;
  while (true) {
    double pq = 0.0;
    for (int idx = threadIdx.x; idx < n; idx += T) {
      const int i = idx % L.nx, j = idx / L.nx;
      const double a = lap_at(L, p, i, j);
      q[idx] = a;
      pq += p[idx] * a;
    }
    pq = block_sum(pq, red);
    if (threadIdx.x == 0) bc[0] = pq;
    __syncthreads();
    const double alpha = rz / bc[0];
    double rr = 0.0;
    for (int idx = threadIdx.x; idx < n; idx += T) {
      x[idx] += alpha * p[idx];
      const double rn = r[idx] - alpha * q[idx];
      r[idx] = rn;
      rr += rn * rn;
    }
    rr = block_sum(rr, red);
    if (threadIdx.x == 0) bc[0] = rr;
    __syncthreads();
    ++it;
    rel = bc[0] / ff;
    if (rel <= tol2 || it >= maxit) break;
    ic0_apply(P, r, z);
    double rzn = 0.0;
    for (int idx = threadIdx.x; idx < n; idx += T) rzn += r[idx] * z[idx];
    rzn = block_sum(rzn, red);
    if (threadIdx.x == 0) bc[1] = rzn;
    __syncthreads();
    const double beta = bc[1] / rz;
    rz = bc[1];
    for (int idx = threadIdx.x; idx < n; idx += T) p[idx] = z[idx] + beta * p[idx];
    __syncthreads();
  }
  if (threadIdx.x == 0) { *out_it = it; *out_rel = rel; }
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

void launch_cg_ic0(hipStream_t st, const DLap& L, const DIC& P, const double* f, double* x, double* r, double* z,
                   double* p, double* q, double tol2, int maxit, int x0_zero, int* out_it, double* out_rel) {
  hipLaunchKernelGGL(k_cg_ic0, dim3(1), dim3(1024), 0, st, L, P, f, x, r, z, p, q, tol2, maxit, x0_zero, out_it,
                     out_rel);
}

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

void launch_tl_coarse_solution(hipStream_t st, const TLPlan& P, const double* f, double* x, double* r) {
  TL T = make(P);
  const int nblk = P.C.ncx * P.C.ncy;
  hipLaunchKernelGGL(k_tl_init, grid_of(P), block_of(P), 0, st, T, f, (const double*)nullptr, r, 1);
  hipLaunchKernelGGL(k_tl_coarse, dim3(1), dim3(1024), 0, st, T, nblk, 1, -1.0);
  hipLaunchKernelGGL(k_tl_setcoarse, grid_of(P), block_of(P), 0, st, T, x);
}

void launch_tl_energy(hipStream_t st, const TLPlan& P, const double* f, const double* x, const double* r,
                      double* part) {
  TL T = make(P);
  hipLaunchKernelGGL(k_tl_energy, grid_of(P), block_of(P), 0, st, T, f, x, r, part);
}

void launch_tl_true_residual(hipStream_t st, const TLPlan& P, const double* f, const double* x, double* r) {
  TL T = make(P);
  const int nblk = P.C.ncx * P.C.ncy;
  hipLaunchKernelGGL(k_tl_init, grid_of(P), block_of(P), 0, st, T, f, x, r, 0);
  hipLaunchKernelGGL(k_tl_true, dim3(1), dim3(1024), 0, st, T, nblk);
}

}  // namespace ins2d
