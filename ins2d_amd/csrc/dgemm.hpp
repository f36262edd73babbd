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
// scales C_ij by -1 / (lx_i + ly_j) and zeroes the null mode (nx0, ny0);
// EPI 2 adds the product into C.
// Operand layout of the 16x16x4 instruction: lane l holds A[l % 16][l / 16]
// and B[l / 16][l % 16]; the result lane l, register r holds
// D[l / 16 + 4 r][l % 16] (pinned by a GPU test through ins2d_fdm_mfma_probe).
typedef double double4_t __attribute__((ext_vector_type(4)));
constexpr int kGemmWaves = 8, kGemmDepth = 4;
typedef double GemmPart[kGemmWaves][32][33];

// BMAX: also the largest |B| among the operand values this thread loads
// (all of the tile's 32 columns over the block's threads) into *bmax
// TN: the tile's width, 32 or 16 columns (one 16-column block: half the
// accumulators, twice the workgroups for a given C)
template <int EPI, bool BMAX = false, int TN = 32>
__device__ __forceinline__ void dgemm_tile(int M, int N, int K, const double* __restrict__ A, int lda,
                                           const double* __restrict__ B, int ldb, double* __restrict__ C, int ldc,
                                           const double* __restrict__ lx, const double* __restrict__ ly, int nx0,
                                           int ny0, int m0, int n0, GemmPart& part, double* bmax = nullptr) {
  static_assert(TN == 32 || TN == 16, "tile width 32 or 16");
  constexpr int NB = TN / 16;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int ma = m0 + lr, mb = m0 + 16 + lr, na = n0 + lr, nb = n0 + 16 + lr;
  const bool ina = ma < M, inb = mb < M, jna = na < N, jnb = NB == 2 && nb < N;
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
  double bm = 0.0;
  auto ld = [&](int k, int d) {
    const bool kin = k < K;
    ra[d] = (kin && ina) ? Aa[(size_t)lda * k] : 0.0;
    rb[d] = (kin && inb) ? Ab[(size_t)lda * k] : 0.0;
    rc[d] = (kin && jna) ? Ba[k] : 0.0;
    rd[d] = (NB == 2 && kin && jnb) ? Bb[k] : 0.0;
    if (BMAX) bm = fmax(bm, fmax(fabs(rc[d]), fabs(rd[d])));
  };
  const int kw = 4 * wave + lk;   // this lane's k in the wave's first step
#pragma unroll
  for (int d = 0; d < kGemmDepth; ++d) ld(kw + d * kStep, d);
  for (int k0 = 4 * wave; k0 < K; k0 += kGemmDepth * kStep) {
#pragma unroll
    for (int d = 0; d < kGemmDepth; ++d) {
      if (k0 + d * kStep < K) {
        acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(ra[d], rc[d], acc[0][0], 0, 0, 0);
        if (NB == 2) acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(ra[d], rd[d], acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(rb[d], rc[d], acc[1][0], 0, 0, 0);
        if (NB == 2) acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(rb[d], rd[d], acc[1][1], 0, 0, 0);
      }
      ld(k0 - 4 * wave + kw + (d + kGemmDepth) * kStep, d);
    }
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) part[wave][16 * a + lk + 4 * r][16 * b + lr] = acc[a][b][r];
  __syncthreads();
  for (int e = threadIdx.x; e < 32 * TN; e += 64 * kGemmWaves) {
    const int mm = e & 31, nn = e >> 5;
    const int gm = m0 + mm, gn = n0 + nn;
    if (gm >= M || gn >= N) continue;
    double v = 0.0;
#pragma unroll
    for (int w = 0; w < kGemmWaves; ++w) v += part[w][mm][nn];
    if (EPI == 1) v = (gm == nx0 && gn == ny0) ? 0.0 : -v / (lx[gm] + ly[gn]);
    if (EPI == 2) C[(size_t)gm + (size_t)ldc * gn] += v;
    else C[(size_t)gm + (size_t)ldc * gn] = v;
  }
  if (BMAX) *bmax = bm;
}

// C = A B as above, LDS-tiled: a workgroup of 256 threads (4 waves) owns the
// 64 x 64 tile (m0, n0) of C, each wave a 32 x 32 quarter (2 x 2 matrix-core
// tiles).  K advances in stages of 16: every thread loads 4 values of the
// 64 x 16 slice of A (a wave reads 64 consecutive rows of one column) and 4
// of the 16 x 64 slice of B (a wave reads 16 consecutive k of 4 columns),
// writes them to the LDS buffer the next stage reads, and the matrix-core
// operands come from LDS (double-buffered: one barrier per stage).  Per
// matrix instruction a wave issues half a memory load instead of one, and the
// loads cover whole cache-line runs; the result tile is transposed through
// LDS so that the stores of C run along its columns.  EPI as in dgemm_tile.
constexpr int kG64 = 64, kG64K = 16;
union Gemm64Smem {
  struct {
    double a[2][kG64K][kG64];       // a[buf][k][m]
    double b[2][kG64][kG64K + 1];   // b[buf][n][k]
  } s;
  double c[kG64][kG64 + 1];         // c[n][m], the result tile
};

template <int EPI>
__device__ __forceinline__ void dgemm_tile64(int M, int N, int K, const double* __restrict__ A, int lda,
                                             const double* __restrict__ B, int ldb, double* __restrict__ C, int ldc,
                                             const double* __restrict__ lx, const double* __restrict__ ly, int nx0,
                                             int ny0, int m0, int n0, Gemm64Smem& sm) {
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
  // this thread's global load positions: A (row am, k rows ak + 4 r),
  // B (k bk, columns bn + 16 r)
  const int am = t & 63, ak = t >> 6;
  const int bk = t & 15, bn = t >> 4;
  const bool arow = m0 + am < M;
  const double* Ap = A + (arow ? m0 + am : 0);
  double ra[4], rb[4];
  auto load = [&](int k0) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int k = k0 + ak + 4 * r;
      ra[r] = (arow && k < K) ? Ap[(size_t)lda * k] : 0.0;
      const int n = n0 + bn + 16 * r, kb = k0 + bk;
      rb[r] = (n < N && kb < K) ? B[(size_t)ldb * n + kb] : 0.0;
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      sm.s.a[buf][ak + 4 * r][am] = ra[r];
      sm.s.b[buf][bn + 16 * r][bk] = rb[r];
    }
  };
  double4_t acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = double4_t{0.0, 0.0, 0.0, 0.0};
  const int nst = (K + kG64K - 1) / kG64K;
  // a wave whose quarter lies wholly outside C (edge tiles of matrices whose
  // sides are not multiples of 64) only loads and stores: its matrix
  // instructions are skipped
  const bool active = m0 + wm < M && n0 + wn < N;
  load(0);
  store(0);
  __syncthreads();
  for (int st = 0; st < nst; ++st) {
    const int buf = st & 1;
    if (st + 1 < nst) load((st + 1) * kG64K);
    if (active) {
#pragma unroll
      for (int kk = 0; kk < kG64K; kk += 4) {
        const double a0 = sm.s.a[buf][kk + lk][wm + lr], a1 = sm.s.a[buf][kk + lk][wm + 16 + lr];
        const double b0 = sm.s.b[buf][wn + lr][kk + lk], b1 = sm.s.b[buf][wn + 16 + lr][kk + lk];
        acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
      }
    }
    if (st + 1 < nst) store(buf ^ 1);
    __syncthreads();
  }
  // result lane l, register r: row l / 16 + 4 r, column l % 16 of its tile
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) sm.c[wn + 16 * b + lr][wm + 16 * a + lk + 4 * r] = acc[a][b][r];
  __syncthreads();
  const int gm = m0 + am;
  if (gm >= M) return;
#pragma unroll 4
  for (int r = 0; r < kG64 / 4; ++r) {
    const int nn = ak + 4 * r, gn = n0 + nn;
    if (gn >= N) break;
    double v = sm.c[nn][am];
    if (EPI == 1) v = (gm == nx0 && gn == ny0) ? 0.0 : -v / (lx[gm] + ly[gn]);
    C[(size_t)gm + (size_t)ldc * gn] = v;
  }
}

}  // namespace ins2d
