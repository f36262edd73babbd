// Fourth-order operators on grids with walls (operators assembled by
// ops4.cpp), applied in CSR form:
//  * face products w = (If V + yI) .* (Af V + yA) of the eight convective
//    terms, then the momentum right-hand side
//      F = -Cw w + Diff V + yD (- G p) (/ Om)
//    with the stage combination of time_ERK.m fused in its epilogue;
//  * divergence M (ca A + cb B) + ycoef yM and projection Vn + dt (Vt - c Om^-1 G dp);
//  * the pressure solve as a product with the dense inverse of the pressure
//    matrix (pressure_poisson.m:42-47, poisson = 1), inverted once at set-up
//    by Gauss-Jordan sweeps on the device.
#include "device.hpp"
#include "launch.hpp"

namespace ins2d {

namespace {

__device__ __forceinline__ double csr_row(const DCsr& A, int r, const double* __restrict__ x) {
  double s = 0.0;
  for (int k = A.rp[r]; k < A.rp[r + 1]; ++k) s += A.v[k] * x[A.ci[k]];
  return s;
}

__global__ void __launch_bounds__(256) k_w4_faces(W4Dev D, const double* __restrict__ V, double* __restrict__ w) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= D.Nf) return;
  const double fi = D.yI[r] + csr_row(D.If, r, V);
  const double fa = D.yA[r] + csr_row(D.Af, r, V);
  w[r] = fi * fa;
}

__global__ void __launch_bounds__(256) k_w4_rhs(W4Dev D, const double* __restrict__ V, const double* __restrict__ p,
                                               const double* __restrict__ w, int mode, double* __restrict__ out,
                                               StageCombine sc, double* __restrict__ vtemp) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= D.NV) return;
  double acc = -csr_row(D.Cw, r, w) + (csr_row(D.Diff, r, V) + D.yD[r]);
  if (mode & RHS_WITH_P) acc -= csr_row(D.G, r, p);
  if (mode & RHS_SCALE_OM) acc = acc / D.Om[r];
  if (out) out[r] = acc;
  if (vtemp) {
    double vt = sc.aself * acc;
    for (int q = 0; q < sc.n; ++q) vt += sc.a[q] * sc.k[q][r];
    vtemp[r] = vt;
  }
}

__global__ void __launch_bounds__(256) k_w4_div(W4Dev D, const double* __restrict__ A, const double* __restrict__ B,
                                               double ca, double cb, double ycoef, double oscale,
                                               double* __restrict__ f) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= D.Np) return;
  double s = 0.0;
  for (int k = D.M.rp[r]; k < D.M.rp[r + 1]; ++k) {
    const int c = D.M.ci[k];
    s += D.M.v[k] * (B ? ca * A[c] + cb * B[c] : ca * A[c]);
  }
  f[r] = oscale * (s + ycoef * D.yM[r]);
}

__global__ void __launch_bounds__(256) k_w4_proj(W4Dev D, const double* __restrict__ Vn,
                                                const double* __restrict__ Vt, const double* __restrict__ dp,
                                                double dt, double c, double* __restrict__ V) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= D.NV) return;
  const double g = csr_row(D.G, r, dp) / D.Om[r];
  const double vt = Vt ? Vt[r] : 0.0;
  V[r] = Vn[r] + dt * (vt - c * g);
}

// dp = K f: one wavefront per row of the row-major n x n inverse
__global__ void __launch_bounds__(256) k_w4_gemv(const double* __restrict__ K, int n, const double* __restrict__ f,
                                                double* __restrict__ dp) {
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= n) return;   // whole wavefronts leave together
  const double* k = K + (size_t)row * n;
  double s = 0.0;
  for (int j = lane; j < n; j += 64) s += k[j] * f[j];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (lane == 0) dp[row] = s;
}

// one Gauss-Jordan sweep on pivot k (in place, no pivoting: the swept
// matrix is symmetric positive definite): save row and column k first, then
//   a_kk <- 1/p, a_kj <- a_kj/p, a_ik <- -a_ik/p, a_ij <- a_ij - a_ik a_kj/p
// so that after the sweeps over all k the array holds the inverse
__global__ void __launch_bounds__(256) k_gj_save(const double* __restrict__ a, int n, int k, double* __restrict__ rowk,
                                                double* __restrict__ colk) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  rowk[i] = a[(size_t)k * n + i];
  colk[i] = a[(size_t)i * n + k];
}

__global__ void __launch_bounds__(256) k_gj_sweep(double* __restrict__ a, int n, int k,
                                                 const double* __restrict__ rowk, const double* __restrict__ colk) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= n) return;
  const double rp = 1.0 / rowk[k];
  double& x = a[(size_t)i * n + j];
  if (i == k) x = (j == k) ? rp : rowk[j] * rp;
  else if (j == k) x = -colk[i] * rp;
  else x = x - colk[i] * (rowk[j] * rp);
}

// K = -C^-1 + s (every entry), C^-1 in place of C
__global__ void __launch_bounds__(256) k_gj_finish(double* __restrict__ a, size_t nn, double s) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < nn) a[i] = -a[i] + s;
}

}  // namespace

void launch_w4_rhs(hipStream_t st, const W4Dev& D, const double* V, const double* p, int mode, double* out,
                   const StageCombine& sc, double* vt) {
  launch_timed(k_w4_faces, dim3((D.Nf + 255) / 256), dim3(256), 0, st, D, V, D.w);
  launch_timed(k_w4_rhs, dim3((D.NV + 255) / 256), dim3(256), 0, st, D, V, p, (const double*)D.w, mode, out, sc, vt);
}

void launch_w4_div(hipStream_t st, const W4Dev& D, const double* A, const double* B, double ca, double cb,
                   double ycoef, double oscale, double* f) {
  launch_timed(k_w4_div, dim3((D.Np + 255) / 256), dim3(256), 0, st, D, A, B, ca, cb, ycoef, oscale, f);
}

void launch_w4_proj(hipStream_t st, const W4Dev& D, const double* Vn, const double* Vt, const double* dp, double dt,
                    double c, double* V) {
  launch_timed(k_w4_proj, dim3((D.NV + 255) / 256), dim3(256), 0, st, D, Vn, Vt, dp, dt, c, V);
}

void launch_w4_solve(hipStream_t st, const W4Dev& D, const double* f, double* dp) {
  launch_timed(k_w4_gemv, dim3((D.Np + 3) / 4), dim3(256), 0, st, (const double*)D.K, D.Np, f, dp);
}

void w4_invert(hipStream_t st, double* a, int n, double* rowk, double* colk, double shift) {
  const dim3 b(256), g1((n + 255) / 256), g2((n + 255) / 256, n);
  for (int k = 0; k < n; ++k) {
    hipLaunchKernelGGL(k_gj_save, g1, b, 0, st, (const double*)a, n, k, rowk, colk);
    hipLaunchKernelGGL(k_gj_sweep, g2, b, 0, st, a, n, k, (const double*)rowk, (const double*)colk);
  }
  const size_t nn = (size_t)n * n;
  hipLaunchKernelGGL(k_gj_finish, dim3((unsigned)((nn + 255) / 256)), b, 0, st, a, nn, shift);
}

}  // namespace ins2d
