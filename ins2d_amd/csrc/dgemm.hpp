// Dense FP64 GEMM tile on the matrix cores, shared by the single-case
// fast-diagonalization solve (fdm.hip) and the batched one (batch.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace ins2d {

// C = A B (column-major; M x K times K x N) on the FP64 matrix cores
// (v_mfma_f64_16x16x4f64).  A workgroup of 64 kGemmWaves threads owns the
// 32 x 32 tile (m0, n0) of C; its waves split K (wave w takes the k-steps
// w, w + kGemmWaves, ...), each keeping four independent 16 x 16 accumulators
// (2 x 2 tiles) so that consecutive matrix instructions do not wait on each
// other.  Operands come straight from memory (L2), kGemmDepth k-steps ahead of
// their use (a register ring), which covers the L2 latency with one or two
// waves per SIMD; the partial tiles are summed through LDS (part).  EPI 1
// scales C_ij by -1 / (lx_i + ly_j) and zeroes the null mode (nx0, ny0).
// Operand layout of the 16x16x4 instruction: lane l holds A[l % 16][l / 16]
// and B[l / 16][l % 16]; the result lane l, register r holds
// D[l / 16 + 4 r][l % 16] (pinned by a GPU test through ins2d_fdm_mfma_probe).
typedef double double4_t __attribute__((ext_vector_type(4)));
constexpr int kGemmWaves = 8, kGemmDepth = 4;
typedef double GemmPart[kGemmWaves][32][33];

template <int EPI>
__device__ __forceinline__ void dgemm_tile(int M, int N, int K, const double* __restrict__ A, int lda,
                                           const double* __restrict__ B, int ldb, double* __restrict__ C, int ldc,
                                           const double* __restrict__ lx, const double* __restrict__ ly, int nx0,
                                           int ny0, int m0, int n0, GemmPart& part) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
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
  constexpr int kStep = 4 * kGemmWaves;   // k advance per step of one wave
  double ra[kGemmDepth], rb[kGemmDepth], rc[kGemmDepth], rd[kGemmDepth];
  auto ld = [&](int k, int d) {
    const bool kin = k < K;
    ra[d] = (kin && ina) ? Aa[(size_t)lda * k] : 0.0;
    rb[d] = (kin && inb) ? Ab[(size_t)lda * k] : 0.0;
    rc[d] = (kin && jna) ? Ba[k] : 0.0;
    rd[d] = (kin && jnb) ? Bb[k] : 0.0;
  };
  const int kw = 4 * wave + lk;   // this lane's k in the wave's first step
#pragma unroll
  for (int d = 0; d < kGemmDepth; ++d) ld(kw + d * kStep, d);
  for (int k0 = 4 * wave; k0 < K; k0 += kGemmDepth * kStep) {
#pragma unroll
    for (int d = 0; d < kGemmDepth; ++d) {
      if (k0 + d * kStep < K) {
        acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(ra[d], rc[d], acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(ra[d], rd[d], acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(rb[d], rc[d], acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(rb[d], rd[d], acc[1][1], 0, 0, 0);
      }
      ld(k0 - 4 * wave + kw + (d + kGemmDepth) * kStep, d);
    }
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) part[wave][16 * a + lk + 4 * r][16 * b + lr] = acc[a][b][r];
  __syncthreads();
  for (int e = threadIdx.x; e < 32 * 32; e += 64 * kGemmWaves) {
    const int mm = e & 31, nn = e >> 5;
    const int gm = m0 + mm, gn = n0 + nn;
    if (gm >= M || gn >= N) continue;
    double v = 0.0;
#pragma unroll
    for (int w = 0; w < kGemmWaves; ++w) v += part[w][mm][nn];
    if (EPI == 1) v = (gm == nx0 && gn == ny0) ? 0.0 : -v / (lx[gm] + ly[gn]);
    C[(size_t)gm + (size_t)ldc * gn] = v;
  }
}

}  // namespace ins2d
