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

#include "launch.hpp"
#include "ops1d.hpp"

namespace ins2d {

namespace {

// C = A B (column-major; M x K times K x N), 64 x 64 tile per workgroup,
// 4 x 4 outputs per thread; EPI 1 scales C_ij by -1 / (lx_i + ly_j) and zeroes
// the null mode (nx0, ny0)
template <int EPI>
__global__ void __launch_bounds__(256) k_dgemm(int M, int N, int K, const double* __restrict__ A, int lda,
                                               const double* __restrict__ B, int ldb, double* __restrict__ C, int ldc,
                                               const double* __restrict__ lx, const double* __restrict__ ly, int nx0,
                                               int ny0) {
  __shared__ double As[16][64];
  __shared__ double Bs[16][65];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int m0 = blockIdx.x * 64, n0 = blockIdx.y * 64;
  double acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
  for (int k0 = 0; k0 < K; k0 += 16) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int e = threadIdx.x + 256 * r;
      const int mm = e & 63, kk = e >> 6;
      const int gm = m0 + mm, gk = k0 + kk;
      As[kk][mm] = (gm < M && gk < K) ? A[(size_t)gm + (size_t)lda * gk] : 0.0;
      const int kb = e & 15, nn = e >> 4;
      const int gk2 = k0 + kb, gn = n0 + nn;
      Bs[kb][nn] = (gk2 < K && gn < N) ? B[(size_t)gk2 + (size_t)ldb * gn] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) {
      double a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[kk][tx + 16 * i];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = Bs[kk][ty + 16 * j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = fma(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int gn = n0 + ty + 16 * j;
    if (gn >= N) continue;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int gm = m0 + tx + 16 * i;
      if (gm >= M) continue;
      double v = acc[i][j];
      if (EPI == 1) v = (gm == nx0 && gn == ny0) ? 0.0 : -v / (lx[gm] + ly[gn]);
      C[(size_t)gm + (size_t)ldc * gn] = v;
    }
  }
}

template <int EPI>
void gemm(hipStream_t st, int M, int N, int K, const double* A, int lda, const double* B, int ldb, double* C, int ldc,
          const double* lx = nullptr, const double* ly = nullptr, int nx0 = -1, int ny0 = -1) {
  hipLaunchKernelGGL(k_dgemm<EPI>, dim3((M + 63) / 64, (N + 63) / 64), dim3(256), 0, st, M, N, K, A, lda, B, ldb, C,
                     ldc, lx, ly, nx0, ny0);
}

}  // namespace

// dp = A^-1 f; w1, w2: Npx*Npy work arrays
void launch_fdm_poisson(hipStream_t st, const FDMPlan& F, const double* f, double* dp, double* w1, double* w2) {
  const int nx = F.nx, ny = F.ny;
  gemm<0>(st, nx, ny, nx, F.VxT, nx, f, nx, w1, nx);                           // Vx^T F
  gemm<1>(st, nx, ny, ny, w1, nx, F.Vy, ny, w2, nx, F.lx, F.ly, F.nx0, F.ny0);   // (.) Vy, scaled
  gemm<0>(st, nx, ny, nx, F.Vx, nx, w2, nx, w1, nx);                            // Vx (.)
  gemm<0>(st, nx, ny, ny, w1, nx, F.VyT, ny, dp, nx);                           // (.) Vy^T
}

}  // namespace ins2d
