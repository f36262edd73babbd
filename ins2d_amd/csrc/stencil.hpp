// Device-side point operators shared by the single-case and batched kernels.
#pragma once

#include "device.hpp"
#include "launch.hpp"
#include "uniform.hpp"

namespace ins2d {

// ===========================================================================
// kron-factored helpers
// ===========================================================================
// Pointers read from the operator records are generic pointers to the
// compiler, which then loads through them with flat instructions; every array
// here lives in device global memory, so the loads below go through the
// global address space instead.
template <class T>
__device__ __forceinline__ const __attribute__((address_space(1))) T* glb(const T* p) {
  return (const __attribute__((address_space(1))) T*)p;
}
// A record that no kernel writes, at a wave-uniform address (the operator
// records of the batched kernels, reached through the instance table): copied
// word by word through the constant address space, so its fields arrive by
// scalar loads instead of one flat load per lane and field (a plain copy
// would bind the record to a generic reference and load it through flat
// instructions).  The host pass of the compiler sees the plain copy.
template <class T>
__device__ __forceinline__ T cst(const T* p) {
#if defined(__HIP_DEVICE_COMPILE__)
  static_assert(sizeof(T) % 8 == 0, "record size");
  T v;
  const auto* src = (const __attribute__((address_space(4))) unsigned long long*)p;
  auto* dst = reinterpret_cast<unsigned long long*>(&v);
#pragma unroll
  for (unsigned k = 0; k < sizeof(T) / 8; ++k) dst[k] = src[k];
  return v;
#else
  return *p;
#endif
}
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

__device__ __forceinline__ double yv_at(const DYV& Y, int f, int g) {
  double s = 0.0;
#pragma unroll
  for (int t = 0; t < 4; ++t)
    if (t < Y.nt && f >= Y.x0[t] && f < Y.x1[t] && g >= Y.y0[t] && g < Y.y1[t]) s += glb(Y.ax[t])[f] * glb(Y.ay[t])[g];
  return s;
}

// ---------------------------------------------------------------------------
// Second-order operators of the generic path: every 1-D factor is stored with
// exactly two (column, weight) entries per row (zero-padded by the host), so
// the index and data loads below are unrolled and independent of each other
// and issue together instead of as a chain of loop-carried loads.  A factor
// whose rows hold one entry (kmax 1: identities, diagonals) skips the second
// entry on a uniform branch.
// ---------------------------------------------------------------------------
struct Row2 {
  int2 c;
  double2 w;
};
__device__ __forceinline__ Row2 row2(const DOp1& O, int r) {
  const long long c = glb(reinterpret_cast<const long long*>(O.col))[r];
  const auto w = glb(O.val) + 2 * r;
  Row2 q;
  q.c.x = (int)c;
  q.c.y = (int)(c >> 32);
  q.w.x = w[0];
  q.w.y = w[1];
  return q;
}

// (kron(Y, X) in)(f, g)
__device__ __forceinline__ double kron2_at(const DKron& K, const double* __restrict__ in, int f, int g) {
  const Row2 y = row2(K.Y, g), x = row2(K.X, f);
  // 32-bit element offsets (every array here holds fewer than 2^31 values)
  const unsigned ld = K.X.ncols;
  const bool x2 = K.X.kmax > 1;
  const auto a = glb(in);
  const unsigned r0 = (unsigned)y.c.x * ld;
  double s0 = x.w.x * a[r0 + (unsigned)x.c.x];
  if (x2) s0 += x.w.y * a[r0 + (unsigned)x.c.y];
  double acc = y.w.x * s0;
  if (K.Y.kmax > 1) {
    const unsigned r1 = (unsigned)y.c.y * ld;
    double s1 = x.w.x * a[r1 + (unsigned)x.c.x];
    if (x2) s1 += x.w.y * a[r1 + (unsigned)x.c.y];
    acc += y.w.y * s1;
  }
  return acc;
}

// kron applied to (ca * A + cb * B) without forming the sum
__device__ __forceinline__ double kron2_at2(const DKron& K, const double* __restrict__ A, double ca,
                                            const double* __restrict__ B, double cb, int f, int g) {
  const Row2 y = row2(K.Y, g), x = row2(K.X, f);
  const size_t ld = K.X.ncols;
  const bool x2 = K.X.kmax > 1;
  auto w = [&](size_t o) { return B ? ca * glb(A)[o] + cb * glb(B)[o] : ca * glb(A)[o]; };
  const size_t r0 = (size_t)y.c.x * ld;
  double s0 = x.w.x * w(r0 + x.c.x);
  if (x2) s0 += x.w.y * w(r0 + x.c.y);
  double acc = y.w.x * s0;
  if (K.Y.kmax > 1) {
    const size_t r1 = (size_t)y.c.y * ld;
    double s1 = x.w.x * w(r1 + x.c.x);
    if (x2) s1 += x.w.y * w(r1 + x.c.y);
    acc += y.w.y * s1;
  }
  return acc;
}

// Branch-free forms of kron2_at / kron2_at2: both stored entries of every row
// are used (the host pads a row's second entry with the row's first column
// and weight 0, so the padded term adds an exact zero), and with B as a
// template flag there is no branch left: the loads of several points issue
// together instead of waiting behind each point's branches.
__device__ __forceinline__ double kron2_at_nb(const DKron& K, const double* __restrict__ in, int f, int g) {
  const Row2 y = row2(K.Y, g), x = row2(K.X, f);
  const unsigned ld = K.X.ncols;
  const auto a = glb(in);
  const unsigned r0 = (unsigned)y.c.x * ld, r1 = (unsigned)y.c.y * ld;
  double s0 = x.w.x * a[r0 + (unsigned)x.c.x];
  s0 += x.w.y * a[r0 + (unsigned)x.c.y];
  double s1 = x.w.x * a[r1 + (unsigned)x.c.x];
  s1 += x.w.y * a[r1 + (unsigned)x.c.y];
  double acc = y.w.x * s0;
  acc += y.w.y * s1;
  return acc;
}
template <bool HASB>
__device__ __forceinline__ double kron2_at2_nb(const DKron& K, const double* __restrict__ A, double ca,
                                               const double* __restrict__ B, double cb, int f, int g) {
  const Row2 y = row2(K.Y, g), x = row2(K.X, f);
  const unsigned ld = K.X.ncols;
  auto w = [&](unsigned o) { return HASB ? ca * glb(A)[o] + cb * glb(B)[o] : ca * glb(A)[o]; };
  const unsigned r0 = (unsigned)y.c.x * ld, r1 = (unsigned)y.c.y * ld;
  double s0 = x.w.x * w(r0 + (unsigned)x.c.x);
  s0 += x.w.y * w(r0 + (unsigned)x.c.y);
  double s1 = x.w.x * w(r1 + (unsigned)x.c.x);
  s1 += x.w.y * w(r1 + (unsigned)x.c.y);
  double acc = y.w.x * s0;
  acc += y.w.y * s1;
  return acc;
}

// sum over the (up to 2 x 2) entries of D's rows i (x) and j (y) of
// D(i, f) D(j, g) face(f, g)
template <class FACE>
__device__ __forceinline__ double apply_d2(const DKron& D, int i, int j, FACE&& face) {
  const Row2 dy = row2(D.Y, j), dx = row2(D.X, i);
  const bool x2 = D.X.kmax > 1;
  double s0 = dx.w.x * face(dx.c.x, dy.c.x);
  if (x2) s0 += dx.w.y * face(dx.c.y, dy.c.x);
  double acc = dy.w.x * s0;
  if (D.Y.kmax > 1) {
    double s1 = dx.w.x * face(dx.c.x, dy.c.y);
    if (x2) s1 += dx.w.y * face(dx.c.y, dy.c.y);
    acc += dy.w.y * s1;
  }
  return acc;
}

// momentum right-hand side at (i, j) of component COMP
//   out = [ -sum conv + sum diff  (- G p if WITH_P) ] (/ volume if SCALE_OM)
// and, for an RK stage, vtemp = aself * out + sum_j a_j k_j at the same point.
// ``koff`` offsets the k_j pointers (batched layouts).
// conv term:  D [ (I c + yI) .* (A own + yA) ]   (convection.m:198-212)
// diff term:  D [ sscale S own + yS ]            (diffusion.m:29-30 with
//             Diffu = Dux (Su_ux / Re) + ..., operator_convection_diffusion.m:460)
template <int SEG>
__device__ __forceinline__ int seg_min(int x) {
  for (int o = SEG / 2; o > 0; o >>= 1) x = min(x, __shfl_xor(x, o, SEG));
  return x;
}
template <int SEG>
__device__ __forceinline__ int seg_max(int x) {
  for (int o = SEG / 2; o > 0; o >>= 1) x = max(x, __shfl_xor(x, o, SEG));
  return x;
}

// (D face)(i, j).  SEG > 0: the wave's lanes form segments of SEG lanes
// (16, 32 or 64), each holding consecutive points of one row, every lane
// calling.  A D that differences in x (one y entry per row) then has each
// face of a segment computed once: lane L of the segment computes face
// fmin + L (and fmin + SEG + L when the segment's faces span more than SEG),
// and every point takes its two faces by shuffles, instead of each face being
// computed by both points that read it.  Other D shapes, and spans of 2 SEG
// or more (periodic wrap), take the direct form.
template <int SEG, class FACE>
__device__ __forceinline__ double apply_d2s(const DKron& D, int i, int j, FACE&& face) {
  if (SEG > 0 && D.Y.kmax == 1) {
    const Row2 dx = row2(D.X, i);
    const bool x2 = D.X.kmax > 1;
    const int fmin = seg_min<SEG>(x2 ? min(dx.c.x, dx.c.y) : dx.c.x);
    const int fmax = seg_max<SEG>(x2 ? max(dx.c.x, dx.c.y) : dx.c.x);
    // segments of one wave may differ here: the branch is per lane, but every
    // lane of a segment takes the same side
    if (fmax - fmin < 2 * SEG) {
      const Row2 dy = row2(D.Y, j);
      const bool two = fmax - fmin >= SEG;
      const int lane = __lane_id() & (SEG - 1);
      double Fa = 0.0, Fb = 0.0;
      if (fmin + lane <= fmax) Fa = face(fmin + lane, dy.c.x);
      if (two && fmin + SEG + lane <= fmax) Fb = face(fmin + SEG + lane, dy.c.x);
      auto pick = [&](int k) {
        const double a = __shfl(Fa, k & (SEG - 1), SEG);
        const double b = __shfl(Fb, k & (SEG - 1), SEG);
        return k < SEG ? a : b;
      };
      double s = dx.w.x * pick(dx.c.x - fmin);
      if (x2) s += dx.w.y * pick(dx.c.y - fmin);
      return dy.w.x * s;
    }
  }
  return apply_d2(D, i, j, face);
}

// -sum conv + sum diff at (i, j); BND: with the boundary terms (yI, yA, yS)
template <bool BND, int SHARE>
__device__ __forceinline__ double rhs_terms(const DComp& C, const double* __restrict__ u,
                                            const double* __restrict__ v, const double* __restrict__ own, int i,
                                            int j) {
  double acc = 0.0;
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    if (t < C.nconv) {
      const DConv& T = C.conv[t];
      const double* cf = T.icomp == 0 ? u : v;
      acc -= T.scale * apply_d2s<SHARE>(T.D, i, j, [&](int f, int g) {
        double fi = kron2_at(T.I, cf, f, g);
        double fa = kron2_at(T.A, own, f, g);
        if (BND) {
          fi += yv_at(T.yI, f, g);
          fa += yv_at(T.yA, f, g);
        }
        return fi * fa;
      });
    }
  }
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    if (t < C.ndiff) {
      const DDiff& T = C.diff[t];
      acc += apply_d2s<SHARE>(T.D, i, j, [&](int f, int g) {
        const double s = T.sscale * kron2_at(T.S, own, f, g);
        return BND ? s + yv_at(T.yS, f, g) : s;
      });
    }
  }
  return acc;
}

// -G p, 1/volume, and the stores of the right-hand side and the stage
// combination at (i, j)
__device__ __forceinline__ void rhs_store(const DComp& C, const double* __restrict__ p, int mode,
                                          double* __restrict__ out, const StageCombine& sc,
                                          double* __restrict__ vtemp, size_t koff, int i, int j, double acc) {
  if (mode & RHS_WITH_P) acc -= kron2_at(C.G, p, i, j);
  if (mode & RHS_SCALE_OM) acc = acc / (glb(C.ox)[i] * glb(C.oy)[j]);
  const size_t idx = (size_t)i + (size_t)C.nx * j;
  if (out) out[idx] = acc;
  if (vtemp) {
    double vt = sc.aself * acc;
    for (int q = 0; q < sc.n; ++q) vt += sc.a[q] * glb(sc.k[q])[koff + idx];
    vtemp[idx] = vt;
  }
}

// SEG (16, 32 or 64): the block is SEG wide, so every SEG consecutive lanes
// hold consecutive points of one row, and every lane calls, also those past
// the row's end (i >= nx, which compute shared faces and store nothing):
// x-direction faces are shared within segments; with SEG = 64, j and
// everything read through the y factors is also wave-uniform
template <int COMP, int SEG = 0>
__device__ __forceinline__ void rhs_point(const DComp& C, const double* __restrict__ u, const double* __restrict__ v,
                                          const double* __restrict__ p, int mode, double* __restrict__ out,
                                          const StageCombine& sc, double* __restrict__ vtemp, size_t koff, int i,
                                          int j) {
  const bool ok = i < C.nx;
  if (SEG == 64) j = __builtin_amdgcn_readfirstlane(j);
  if (SEG > 0) i = ok ? i : C.nx - 1;
  const double* own = COMP == 0 ? u : v;
  // interior points (no boundary term reaches their faces) take the path
  // without the boundary terms; whole waves of interior points skip the
  // other (with shared faces the wave decides together: a face outside every
  // box reads zero there)
  double acc;
  const bool bnd = (glb(C.bmx)[i] & glb(C.bmy)[j]) != 0;
  // interior points of uniform grids take the closed form (a segment that
  // shares faces decides together: its lanes either all take it or none)
  bool closed = false;
  if (C.uni.fast) {
    closed = glb(C.uni.fast)[(size_t)i + (size_t)C.nx * j] && !bnd;
    if (SEG > 0) closed = seg_min<SEG>(closed ? 1 : 0) == 1;
  }
  if (closed) {
    const UniForm& g = C.uni;
    const auto gu = glb(u), gv = glb(v);
    auto U = [&](int a, int b) { return gu[(unsigned)(a - g.ou) + (unsigned)g.nux * (unsigned)b]; };
    auto V = [&](int a, int b) { return gv[(unsigned)a + (unsigned)g.nvx * (unsigned)(b - g.ov)]; };
    acc = COMP == 0 ? uni2_u(g, U, V, i + g.ou, j) : uni2_v(g, U, V, i, j + g.ov);
  } else if (SEG > 0 ? __any(bnd) : bnd) {
    acc = rhs_terms<true, SEG>(C, u, v, own, i, j);
  } else {
    acc = rhs_terms<false, SEG>(C, u, v, own, i, j);
  }
  if (SEG > 0 && !ok) return;
  rhs_store(C, p, mode, out, sc, vtemp, koff, i, j, acc);
}

// a point of the interior of a uniform grid (uni.fast set, no boundary term):
// the closed form only
template <int COMP>
__device__ __forceinline__ void rhs_closed_point(const DComp& C, const double* __restrict__ u,
                                                 const double* __restrict__ v, const double* __restrict__ p, int mode,
                                                 double* __restrict__ out, const StageCombine& sc,
                                                 double* __restrict__ vtemp, size_t koff, int i, int j) {
  const UniForm& g = C.uni;
  const auto gu = glb(u), gv = glb(v);
  auto U = [&](int a, int b) { return gu[(unsigned)(a - g.ou) + (unsigned)g.nux * (unsigned)b]; };
  auto V = [&](int a, int b) { return gv[(unsigned)a + (unsigned)g.nvx * (unsigned)(b - g.ov)]; };
  const double acc = COMP == 0 ? uni2_u(g, U, V, i + g.ou, j) : uni2_v(g, U, V, i, j + g.ov);
  rhs_store(C, p, mode, out, sc, vtemp, koff, i, j, acc);
}

// divergence at pressure cell (i, j): Mu (ca A_u + cb B_u) + Mv (...) + ycoef yM
__device__ __forceinline__ double div_point(const DComp& Cu, const DComp& Cv, const double* __restrict__ A,
                                            const double* __restrict__ B, double ca, double cb, double ycoef, int nu,
                                            int i, int j) {
  double s = kron2_at2(Cu.M, A, ca, B, cb, i, j);
  s += kron2_at2(Cv.M, A + nu, ca, B ? B + nu : nullptr, cb, i, j);
  if (ycoef != 0.0) s += ycoef * (yv_at(Cu.yM, i, j) + yv_at(Cv.yM, i, j));
  return s;
}

// projected velocity at (i, j): Vn + dt (Vtemp - c Om^-1 G dp)
__device__ __forceinline__ double proj_point(const DComp& C, const double* __restrict__ Vn,
                                             const double* __restrict__ Vt, const double* __restrict__ dp, double dt,
                                             double c, int i, int j) {
  const size_t idx = (size_t)i + (size_t)C.nx * j;
  const double g = kron2_at(C.G, dp, i, j) / (C.ox[i] * C.oy[j]);
  const double vt = Vt ? Vt[idx] : 0.0;
  return Vn[idx] + dt * (vt - c * g);
}

__device__ __forceinline__ double lap_at(const DLap& L, const double* __restrict__ p, int i, int j) {
  return kron_at(L.L1, p, i, j) + kron_at(L.L2, p, i, j);
}
// the same for the factor widths of the pressure Laplacian,
// kron(diagonal, tridiagonal) + kron(tridiagonal, diagonal): every index and
// weight load is independent of the others, so they issue together (the
// generic form loads an entry's index, then its value, one entry at a time)
__device__ __forceinline__ double lap_tri_at(const DLap& L, const double* __restrict__ p, int i, int j) {
  const auto a = glb(p);
  const unsigned ld = (unsigned)L.L1.X.ncols;
  const unsigned i3 = 3u * (unsigned)i, j3 = 3u * (unsigned)j;
  const auto c1x = glb(L.L1.X.col), c2y = glb(L.L2.Y.col);
  const auto w1x = glb(L.L1.X.val), w2y = glb(L.L2.Y.val);
  const unsigned r1 = (unsigned)glb(L.L1.Y.col)[j] * ld;
  const double w1y = glb(L.L1.Y.val)[j];
  const unsigned c2 = (unsigned)glb(L.L2.X.col)[i];
  const double w2 = glb(L.L2.X.val)[i];
  double s1 = w1x[i3] * a[r1 + (unsigned)c1x[i3]];
  s1 += w1x[i3 + 1] * a[r1 + (unsigned)c1x[i3 + 1]];
  s1 += w1x[i3 + 2] * a[r1 + (unsigned)c1x[i3 + 2]];
  double s2 = w2y[j3] * a[(unsigned)c2y[j3] * ld + c2];
  s2 += w2y[j3 + 1] * a[(unsigned)c2y[j3 + 1] * ld + c2];
  s2 += w2y[j3 + 2] * a[(unsigned)c2y[j3 + 2] * ld + c2];
  return w1y * s1 + w2 * s2;
}
__device__ __forceinline__ bool lap_is_tri(const DLap& L) {
  return L.L1.Y.K == 1 && L.L1.X.K == 3 && L.L2.Y.K == 3 && L.L2.X.K == 1;
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
