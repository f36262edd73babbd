// Device-side point operators shared by the single-case and batched kernels.
#pragma once

#include "device.hpp"
#include "launch.hpp"

namespace ins2d {

// ===========================================================================
// kron-factored helpers
// ===========================================================================
__device__ __forceinline__ double kron_at(const DKron& K, const double* __restrict__ in, int f, int g) {
  const int KY = K.Y.K, KX = K.X.K, ld = K.X.ncols;
  const int* cy = K.Y.col + (size_t)g * KY;
  const double* wy = K.Y.val + (size_t)g * KY;
  const int* cx = K.X.col + (size_t)f * KX;
  const double* wx = K.X.val + (size_t)f * KX;
  double acc = 0.0;
  for (int a = 0; a < KY; ++a) {
    const double* row = in + (size_t)cy[a] * ld;
    double s = 0.0;
    for (int b = 0; b < KX; ++b) s += wx[b] * row[cx[b]];
    acc += wy[a] * s;
  }
  return acc;
}

// kron applied to (ca * A + cb * B) without forming the sum
__device__ __forceinline__ double kron_at2(const DKron& K, const double* __restrict__ A, double ca,
                                           const double* __restrict__ B, double cb, int f, int g) {
  const int KY = K.Y.K, KX = K.X.K, ld = K.X.ncols;
  const int* cy = K.Y.col + (size_t)g * KY;
  const double* wy = K.Y.val + (size_t)g * KY;
  const int* cx = K.X.col + (size_t)f * KX;
  const double* wx = K.X.val + (size_t)f * KX;
  double acc = 0.0;
  for (int a = 0; a < KY; ++a) {
    const size_t ro = (size_t)cy[a] * ld;
    double s = 0.0;
    if (B) {
      for (int b = 0; b < KX; ++b) s += wx[b] * (ca * A[ro + cx[b]] + cb * B[ro + cx[b]]);
    } else {
      for (int b = 0; b < KX; ++b) s += wx[b] * (ca * A[ro + cx[b]]);
    }
    acc += wy[a] * s;
  }
  return acc;
}

__device__ __forceinline__ double yv_at(const DYV& Y, int f, int g) {
  double s = 0.0;
  for (int t = 0; t < Y.nt; ++t) s += Y.ax[t][f] * Y.ay[t][g];
  return s;
}

// momentum right-hand side at (i, j) of component COMP
//   out = [ -sum conv + sum diff  (- G p if WITH_P) ] (/ volume if SCALE_OM)
// and, for an RK stage, vtemp = aself * out + sum_j a_j k_j at the same point.
// ``koff`` offsets the k_j pointers (batched layouts).
template <int COMP>
__device__ __forceinline__ void rhs_point(const DComp& C, const double* __restrict__ u, const double* __restrict__ v,
                                          const double* __restrict__ p, int mode, double* __restrict__ out,
                                          const StageCombine& sc, double* __restrict__ vtemp, size_t koff, int i,
                                          int j) {
  const double* own = COMP == 0 ? u : v;
  double acc = 0.0;
  for (int t = 0; t < C.nconv; ++t) {
    const DConv& T = C.conv[t];
    const double* cf = T.icomp == 0 ? u : v;
    const int KDY = T.D.Y.K, KDX = T.D.X.K;
    double s = 0.0;
    for (int a = 0; a < KDY; ++a) {
      const int g = T.D.Y.col[(size_t)j * KDY + a];
      const double dy = T.D.Y.val[(size_t)j * KDY + a];
      if (dy == 0.0) continue;
      double sx = 0.0;
      for (int b = 0; b < KDX; ++b) {
        const int f = T.D.X.col[(size_t)i * KDX + b];
        const double dx = T.D.X.val[(size_t)i * KDX + b];
        if (dx == 0.0) continue;
        const double fi = kron_at(T.I, cf, f, g) + yv_at(T.yI, f, g);
        const double fa = kron_at(T.A, own, f, g) + yv_at(T.yA, f, g);
        sx += dx * (fi * fa);
      }
      s += dy * sx;
    }
    acc -= T.scale * s;
  }
  for (int t = 0; t < C.ndiff; ++t) {
    const DDiff& T = C.diff[t];
    const int KDY = T.D.Y.K, KDX = T.D.X.K;
    double s = 0.0;
    for (int a = 0; a < KDY; ++a) {
      const int g = T.D.Y.col[(size_t)j * KDY + a];
      const double dy = T.D.Y.val[(size_t)j * KDY + a];
      if (dy == 0.0) continue;
      double sx = 0.0;
      for (int b = 0; b < KDX; ++b) {
        const int f = T.D.X.col[(size_t)i * KDX + b];
        const double dx = T.D.X.val[(size_t)i * KDX + b];
        if (dx == 0.0) continue;
        sx += dx * (T.sscale * kron_at(T.S, own, f, g) + yv_at(T.yS, f, g));
      }
      s += dy * sx;
    }
    acc += s;
  }
  if (mode & RHS_WITH_P) acc -= kron_at(C.G, p, i, j);
  if (mode & RHS_SCALE_OM) acc = acc / (C.ox[i] * C.oy[j]);
  const size_t idx = (size_t)i + (size_t)C.nx * j;
  if (out) out[idx] = acc;
  if (vtemp) {
    double vt = sc.aself * acc;
    for (int q = 0; q < sc.n; ++q) vt += sc.a[q] * sc.k[q][koff + idx];
    vtemp[idx] = vt;
  }
}

// divergence at pressure cell (i, j): Mu (ca A_u + cb B_u) + Mv (...) + ycoef yM
__device__ __forceinline__ double div_point(const DComp& Cu, const DComp& Cv, const double* __restrict__ A,
                                            const double* __restrict__ B, double ca, double cb, double ycoef, int nu,
                                            int i, int j) {
  double s = kron_at2(Cu.M, A, ca, B, cb, i, j);
  s += kron_at2(Cv.M, A + nu, ca, B ? B + nu : nullptr, cb, i, j);
  if (ycoef != 0.0) s += ycoef * (yv_at(Cu.yM, i, j) + yv_at(Cv.yM, i, j));
  return s;
}

// projected velocity at (i, j): Vn + dt (Vtemp - c Om^-1 G dp)
__device__ __forceinline__ double proj_point(const DComp& C, const double* __restrict__ Vn,
                                             const double* __restrict__ Vt, const double* __restrict__ dp, double dt,
                                             double c, int i, int j) {
  const size_t idx = (size_t)i + (size_t)C.nx * j;
  const double g = kron_at(C.G, dp, i, j) / (C.ox[i] * C.oy[j]);
  const double vt = Vt ? Vt[idx] : 0.0;
  return Vn[idx] + dt * (vt - c * g);
}

__device__ __forceinline__ double lap_at(const DLap& L, const double* __restrict__ p, int i, int j) {
  return kron_at(L.L1, p, i, j) + kron_at(L.L2, p, i, j);
}

// block reductions over all threads of a (possibly 2-D) block; the result is
// valid on the block's first thread
__device__ __forceinline__ int lin_tid() { return threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z); }
__device__ __forceinline__ int block_threads() { return blockDim.x * blockDim.y * blockDim.z; }

__device__ __forceinline__ double block_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int tid = lin_tid();
  const int lane = tid & 63, wid = tid >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (tid == 0)
    for (int w = 0; w < (block_threads() + 63) / 64; ++w) s += red[w];
  return s;
}
__device__ __forceinline__ double block_max(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  const int tid = lin_tid();
  const int lane = tid & 63, wid = tid >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (tid == 0)
    for (int w = 0; w < (block_threads() + 63) / 64; ++w) s = fmax(s, red[w]);
  return s;
}

}  // namespace ins2d
