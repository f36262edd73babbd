// Direct pressure solve on tensor-product grids by fast diagonalization
// (Lynch, Rice & Thomas, Numer. Math. 6, 1964).
//
// The second-order pressure matrix of a rectangular INS2D grid is a sum of
// two Kronecker products (operator_divergence.m:54-167 builds A = M Om^-1 G
// from 1-D factors):
//     A = -(Hy (x) Lx + Ly (x) Hx),  Hx = diag(hx), Hy = diag(hy),
//     Lx = Mx diag(1/gx) Mx^T,  Ly = My diag(1/gy) My^T   (symmetric PSD).
// With the generalized eigen-decompositions Lx Vx = Hx Vx Lamx, Vx^T Hx Vx = I
// (and the same in y),
//     A^-1 f = -(Vy (x) Vx) diag(1 / (lamx_i + lamy_j)) (Vy (x) Vx)^T f,
// i.e. with F the Npx x Npy array of f (x fastest):
//     P = Vx [ -(Vx^T F Vy) ./ (lamx_i + lamy_j) ] Vy^T,
// four dense FP64 GEMMs on the GPU.  When A is singular (no 'pres' boundary)
// the mode with lamx = lamy = 0 (the constant) is dropped, which returns the
// solution with zero area-weighted mean; the reference's direct solve of
// such a system fixes an arbitrary constant (pressure_poisson.m:12 with
// main.m:34 silencing the singular-matrix warning).
//
// The 1-D eigenproblems are solved once on the host (fdm_host.cpp).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <vector>

#include "../../include/ins2d.h"
#include "dgemm.hpp"
#include "launch.hpp"
#include "ops1d.hpp"

namespace ins2d {

namespace {

// skip: optional device flag (a converged PCG); the launch does nothing
// while it is set.  The early exit of pressure_poisson.m:21-39 (dp = 0 when
// max |f| < tol) rides on the first and last launches of a direct solve:
// EXIT 1, the blocks of the first row tile also record the largest |B| of
// their 32 columns (all K rows; B = f), taken from the operand values the
// tile loads, in emax[blockIdx.y]; EXIT 2, every
// block reads the ne partial maxima first and, when they are all below tol,
// stores zeros instead of its product (EPI 2, which adds the product into
// C: leaves C as it is).
template <int EPI, int EXIT, int TN>
__global__ void __launch_bounds__(64 * kGemmWaves) k_dgemm(int M, int N, int K, const double* __restrict__ A, int lda,
                                                          const double* __restrict__ B, int ldb,
                                                          double* __restrict__ C, int ldc,
                                                          const double* __restrict__ lx,
                                                          const double* __restrict__ ly, int nx0, int ny0,
                                                          const int* __restrict__ skip, double* __restrict__ emax,
                                                          int ne, double tol) {
  __shared__ GemmPart part;
  __shared__ double wmax[kGemmWaves];
  __shared__ int tiny;
  if (skip && *skip) return;
  const int m0 = blockIdx.x * 32, n0 = blockIdx.y * TN;
  if (EXIT == 2) {
    if (threadIdx.x < 64) {
      double m = 0.0;
      for (int e = threadIdx.x; e < ne; e += 64) m = fmax(m, emax[e]);
      for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
      if (threadIdx.x == 0) tiny = m < tol;
    }
    __syncthreads();
    if (tiny) {
      if (EPI == 2) return;   // adds zero
      for (int e = threadIdx.x; e < 32 * TN; e += blockDim.x) {
        const int gm = m0 + (e & 31), gn = n0 + (e >> 5);
        if (gm < M && gn < N) C[(size_t)gm + (size_t)ldc * gn] = 0.0;
      }
      return;
    }
  }
  double m = 0.0;
  dgemm_tile<EPI, EXIT == 1, TN>(M, N, K, A, lda, B, ldb, C, ldc, lx, ly, nx0, ny0, m0, n0, part, &m);
  if (EXIT == 1 && blockIdx.x == 0) {
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
      double b = 0.0;
      for (int w = 0; w < kGemmWaves; ++w) b = fmax(b, wmax[w]);
      emax[blockIdx.y] = b;
    }
  }
}

// 32 x 16 tiles when 32 x 32 ones would give fewer than 128 workgroups (the
// small grids: LDC 64^2 17.9 against 20.5 us per solve; at BFS 1024 x 256,
// 256 workgroups of 32 x 32, the wider tiles: 62.9 against 69.9 us).  The
// first and last launches of a solve share M and N, so they agree on the
// number of early-exit partials.
template <int EPI, int EXIT = 0>
void gemm(hipStream_t st, const int* skip, int M, int N, int K, const double* A, int lda, const double* B, int ldb,
          double* C, int ldc, const double* lx = nullptr, const double* ly = nullptr, int nx0 = -1, int ny0 = -1,
          double* emax = nullptr, double tol = 0.0) {
  if (((M + 31) / 32) * ((N + 31) / 32) < 128)
    hipLaunchKernelGGL((k_dgemm<EPI, EXIT, 16>), dim3((M + 31) / 32, (N + 15) / 16), dim3(64 * kGemmWaves), 0, st, M,
                       N, K, A, lda, B, ldb, C, ldc, lx, ly, nx0, ny0, skip, emax, (N + 15) / 16, tol);
  else
    hipLaunchKernelGGL((k_dgemm<EPI, EXIT, 32>), dim3((M + 31) / 32, (N + 31) / 32), dim3(64 * kGemmWaves), 0, st, M,
                       N, K, A, lda, B, ldb, C, ldc, lx, ly, nx0, ny0, skip, emax, (N + 31) / 32, tol);
}

// one v_mfma_f64_16x16x4f64 with the operand layout the GEMM uses; the raw
// accumulator of every lane is written out (Draw[lane * 4 + r])
__global__ void k_mfma64_probe(const double* __restrict__ A, const double* __restrict__ B, double* __restrict__ Draw) {
  const int l = threadIdx.x;
  const double a = A[(l & 15) * 4 + (l >> 4)];    // A row-major 16 x 4
  const double b = B[(l >> 4) * 16 + (l & 15)];   // B row-major 4 x 16
  double4_t acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, double4_t{0.0, 0.0, 0.0, 0.0}, 0, 0, 0);
  for (int r = 0; r < 4; ++r) Draw[l * 4 + r] = acc[r];
}

}  // namespace

int fdm_mfma_probe(const double* A, const double* B, double* Draw) {
  double *dA = nullptr, *dB = nullptr, *dD = nullptr;
  if (hipMalloc(&dA, 64 * 8) != hipSuccess || hipMalloc(&dB, 64 * 8) != hipSuccess || hipMalloc(&dD, 256 * 8) != hipSuccess)
    return 2;
  (void)hipMemcpy(dA, A, 64 * 8, hipMemcpyHostToDevice);
  (void)hipMemcpy(dB, B, 64 * 8, hipMemcpyHostToDevice);
  hipLaunchKernelGGL(k_mfma64_probe, dim3(1), dim3(64), 0, nullptr, dA, dB, dD);
  hipError_t e = hipMemcpy(Draw, dD, 256 * 8, hipMemcpyDeviceToHost);
  (void)hipFree(dA);
  (void)hipFree(dB);
  (void)hipFree(dD);
  return e == hipSuccess ? 0 : 2;
}

// dp = A^-1 f; w1, w2: Npx*Npy work arrays; nothing while *skip is set;
// emax (ceil(Npy / 32) doubles, optional): dp = 0 when max |f| < 1e-14
void launch_fdm_poisson(hipStream_t st, const FDMPlan& F, const double* f, double* dp, double* w1, double* w2,
                        const int* skip, double* emax, bool acc) {
  const int nx = F.nx, ny = F.ny;
  if (emax)
    gemm<0, 1>(st, skip, nx, ny, nx, F.VxT, nx, f, nx, w1, nx, nullptr, nullptr, -1, -1, emax);   // Vx^T F, max |f|
  else
    gemm<0>(st, skip, nx, ny, nx, F.VxT, nx, f, nx, w1, nx);                                       // Vx^T F
  gemm<1>(st, skip, nx, ny, ny, w1, nx, F.Vy, ny, w2, nx, F.lx, F.ly, F.nx0, F.ny0);   // (.) Vy, scaled
  gemm<0>(st, skip, nx, ny, nx, F.Vx, nx, w2, nx, w1, nx);                            // Vx (.)
  if (emax && acc)
    gemm<2, 2>(st, skip, nx, ny, ny, w1, nx, F.VyT, ny, dp, nx, nullptr, nullptr, -1, -1, emax, 1e-14);   // += (.) Vy^T
  else if (emax)
    gemm<0, 2>(st, skip, nx, ny, ny, w1, nx, F.VyT, ny, dp, nx, nullptr, nullptr, -1, -1, emax, 1e-14);   // (.) Vy^T
  else if (acc)
    throw std::runtime_error("fast-diagonalization solve: accumulation needs the early-exit partials");
  else
    gemm<0>(st, skip, nx, ny, ny, w1, nx, F.VyT, ny, dp, nx);                                             // (.) Vy^T
}

}  // namespace ins2d

extern "C" int ins2d_fdm_mfma_probe(const double* A, const double* B, double* Draw) {
  if (!A || !B || !Draw) return INS2D_E_ARG;
  return ins2d::fdm_mfma_probe(A, B, Draw);
}
