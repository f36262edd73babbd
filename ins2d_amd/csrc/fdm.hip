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
#include "launch.hpp"
#include "ops1d.hpp"

namespace ins2d {

namespace {

// C = A B (column-major; M x K times K x N) on the FP64 matrix cores
// (v_mfma_f64_16x16x4f64).  A workgroup owns a 32 x 32 tile of C; its
// kWaves waves split K (wave w takes the k-steps w, w + kWaves, ...), each
// keeping four independent 16 x 16 accumulators (2 x 2 tiles) so that
// consecutive matrix instructions do not wait on each other.  Operands come
// straight from memory (L2), kDepth k-steps ahead of their use (a register
// ring), which covers the L2 latency with one or two waves per SIMD; the
// partial tiles are summed through LDS.  EPI 1 scales C_ij by
// -1 / (lx_i + ly_j) and zeroes the null mode (nx0, ny0).
// Operand layout of the 16x16x4 instruction: lane l holds A[l % 16][l / 16]
// and B[l / 16][l % 16]; the result lane l, register r holds
// D[l / 16 + 4 r][l % 16] (pinned by a GPU test through ins2d_fdm_mfma_probe).
typedef double double4_t __attribute__((ext_vector_type(4)));
constexpr int kWaves = 8, kDepth = 4;

template <int EPI>
__global__ void __launch_bounds__(64 * kWaves) k_dgemm(int M, int N, int K, const double* __restrict__ A, int lda,
                                                      const double* __restrict__ B, int ldb, double* __restrict__ C,
                                                      int ldc, const double* __restrict__ lx,
                                                      const double* __restrict__ ly, int nx0, int ny0) {
  __shared__ double part[kWaves][32][33];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m0 = blockIdx.x * 32, n0 = blockIdx.y * 32;
  const int lr = lane & 15, lk = lane >> 4;
  const int ma = m0 + lr, mb = m0 + 16 + lr, na = n0 + lr, nb = n0 + 16 + lr;
  const bool ina = ma < M, inb = mb < M, jna = na < N, jnb = nb < N;
  const double* Aa = A + (ina ? ma : 0);
  const double* Ab = A + (inb ? mb : 0);
  const double* Ba = B + (size_t)ldb * (jna ? na : 0);
  const double* Bb = B + (size_t)ldb * (jnb ? nb : 0);
  double4_t acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = double4_t{0.0, 0.0, 0.0, 0.0};
  constexpr int kStep = 4 * kWaves;   // k advance per step of one wave
  double ra[kDepth], rb[kDepth], rc[kDepth], rd[kDepth];
  auto ld = [&](int k, int d) {
    const bool kin = k < K;
    ra[d] = (kin && ina) ? Aa[(size_t)lda * k] : 0.0;
    rb[d] = (kin && inb) ? Ab[(size_t)lda * k] : 0.0;
    rc[d] = (kin && jna) ? Ba[k] : 0.0;
    rd[d] = (kin && jnb) ? Bb[k] : 0.0;
  };
  const int kw = 4 * wave + lk;   // this lane's k in the wave's first step
#pragma unroll
  for (int d = 0; d < kDepth; ++d) ld(kw + d * kStep, d);
  for (int k0 = 4 * wave; k0 < K; k0 += kDepth * kStep) {
#pragma unroll
    for (int d = 0; d < kDepth; ++d) {
      if (k0 + d * kStep < K) {
        acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(ra[d], rc[d], acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(ra[d], rd[d], acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(rb[d], rc[d], acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(rb[d], rd[d], acc[1][1], 0, 0, 0);
      }
      ld(k0 - 4 * wave + kw + (d + kDepth) * kStep, d);
    }
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) part[wave][16 * a + lk + 4 * r][16 * b + lr] = acc[a][b][r];
  __syncthreads();
  for (int e = threadIdx.x; e < 32 * 32; e += 64 * kWaves) {
    const int mm = e & 31, nn = e >> 5;
    const int gm = m0 + mm, gn = n0 + nn;
    if (gm >= M || gn >= N) continue;
    double v = 0.0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) v += part[w][mm][nn];
    if (EPI == 1) v = (gm == nx0 && gn == ny0) ? 0.0 : -v / (lx[gm] + ly[gn]);
    C[(size_t)gm + (size_t)ldc * gn] = v;
  }
}

template <int EPI>
void gemm(hipStream_t st, int M, int N, int K, const double* A, int lda, const double* B, int ldb, double* C, int ldc,
          const double* lx = nullptr, const double* ly = nullptr, int nx0 = -1, int ny0 = -1) {
  hipLaunchKernelGGL(k_dgemm<EPI>, dim3((M + 31) / 32, (N + 31) / 32), dim3(64 * kWaves), 0, st, M, N, K, A, lda, B,
                     ldb, C, ldc, lx, ly, nx0, ny0);
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

// dp = A^-1 f; w1, w2: Npx*Npy work arrays
void launch_fdm_poisson(hipStream_t st, const FDMPlan& F, const double* f, double* dp, double* w1, double* w2) {
  const int nx = F.nx, ny = F.ny;
  gemm<0>(st, nx, ny, nx, F.VxT, nx, f, nx, w1, nx);                           // Vx^T F
  gemm<1>(st, nx, ny, ny, w1, nx, F.Vy, ny, w2, nx, F.lx, F.ly, F.nx0, F.ny0);   // (.) Vy, scaled
  gemm<0>(st, nx, ny, nx, F.Vx, nx, w2, nx, w1, nx);                            // Vx (.)
  gemm<0>(st, nx, ny, ny, w1, nx, F.VyT, ny, dp, nx);                           // (.) Vy^T
}

}  // namespace ins2d

extern "C" int ins2d_fdm_mfma_probe(const double* A, const double* B, double* Draw) {
  if (!A || !B || !Draw) return INS2D_E_ARG;
  return ins2d::fdm_mfma_probe(A, B, Draw);
}
