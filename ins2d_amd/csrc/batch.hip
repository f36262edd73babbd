// Batched ensemble: many independent cavity/channel instances of uneven size
// and Reynolds number advanced together.  Every phase of an RK stage is one
// launch over a work list of 64x4 tiles that spans all instances; the
// pressure solve is a batched PCG with per-instance scalars (converged
// instances skip their tiles), run in chunks of iterations that are captured
// once into a hipGraph and replayed.  Preconditioners: Jacobi, Jacobi plus an
// aggregation coarse grid, or the fast-diagonalization inverse of each
// instance's pressure matrix (batched FP64 GEMMs on the matrix cores).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../../include/ins2d.h"
#include "bcmask.hpp"
#include "uniform.hpp"
#include "device.hpp"
#include "dgemm.hpp"
#include "launch.hpp"
#include "ops1d.hpp"
#include "stencil.hpp"

namespace ins2d {

struct BInst {
  const DComp* cu;
  const DComp* cv;
  DLap lap;
  int nux, nuy, nvx, nvy, npx, npy;
  long voff, poff;
  int nu;
  int pt0, pt1;   // range of pressure tiles of this instance
  // two-level preconditioner: aggregate tiles [at0, at1) of kAgg x kAgg cells,
  // coarse grid ncx x ncy at offset coff of rc / ec, fast-diagonalization
  // factors of the Galerkin coarse operator
  int at0, at1, ncx, ncy, nx0, ny0;
  long coff;
  const double *cVx, *cVy, *clx, *cly;
  // fast-diagonalization preconditioner: the factors of this instance's
  // pressure matrix (Vx, Vx^T, Vy, Vy^T, eigenvalues, null mode)
  const double *fVx, *fVxT, *fVy, *fVyT, *flx, *fly;
  int fnx0, fny0;
};
struct Tile {
  int inst, comp, i0, j0;
};


constexpr int TX = 64, TY = 4;
constexpr int kAgg = 16;            // aggregate side of the two-level preconditioner
constexpr int kBMaxCoarse = 2048;   // coarse unknowns per instance

// ------------------------------------------------------------------ kernels
// right-hand side over the split tiles, as two kernels: the closed form over
// the interior tiles (k_b_rhs_closed, small) and the generic operators over
// the ring (k_b_rhs_split); point q of the block is (i0 + q mod w, j0 + q div w)
__global__ void __launch_bounds__(256) k_b_rhs_closed(const BInst* __restrict__ I, const RTile* __restrict__ tiles,
                                                     const double* __restrict__ V, const double* __restrict__ P,
                                                     int mode, double* __restrict__ out, StageCombine sc,
                                                     double* __restrict__ vtemp) {
  const RTile t = tiles[blockIdx.x];
  const int q = threadIdx.x;
  if (q >= t.w * t.h) return;
  const int i = t.i0 + q % t.w, j = t.j0 + q / t.w;
  const BInst& B = I[t.inst];
  const double* u = V + B.voff;
  const double* v = u + B.nu;
  const double* p = P + B.poff;
  if ((t.comp & 1) == 0) {
    const long o = B.voff;
    rhs_closed_point<0>(cst(B.cu), u, v, p, mode, out ? out + o : nullptr, sc, vtemp ? vtemp + o : nullptr, o, i, j);
  } else {
    const long o = B.voff + B.nu;
    rhs_closed_point<1>(cst(B.cv), u, v, p, mode, out ? out + o : nullptr, sc, vtemp ? vtemp + o : nullptr, o, i, j);
  }
}

__global__ void __launch_bounds__(256) k_b_rhs_split(const BInst* __restrict__ I, const RTile* __restrict__ tiles,
                                                    const double* __restrict__ V, const double* __restrict__ P,
                                                    int mode, double* __restrict__ out, StageCombine sc,
                                                    double* __restrict__ vtemp) {
  const RTile t = tiles[blockIdx.x];
  const int q = threadIdx.x;
  if (q >= t.w * t.h) return;
  const int i = t.i0 + q % t.w, j = t.j0 + q / t.w;
  const BInst& B = I[t.inst];
  const double* u = V + B.voff;
  const double* v = u + B.nu;
  const double* p = P + B.poff;
  const bool closed = (t.comp & 2) != 0;
  if ((t.comp & 1) == 0) {
    const DComp& C = *B.cu;
    const long o = B.voff;
    double* ou = out ? out + o : nullptr;
    double* vu = vtemp ? vtemp + o : nullptr;
    if (closed && !(glb(C.bmx)[i] & glb(C.bmy)[j]))
      rhs_closed_point<0>(C, u, v, p, mode, ou, sc, vu, o, i, j);
    else
      rhs_point<0, 0>(C, u, v, p, mode, ou, sc, vu, o, i, j);
  } else {
    const DComp& C = *B.cv;
    const long o = B.voff + B.nu;
    double* ov = out ? out + o : nullptr;
    double* vv = vtemp ? vtemp + o : nullptr;
    if (closed && !(glb(C.bmx)[i] & glb(C.bmy)[j]))
      rhs_closed_point<1>(C, u, v, p, mode, ov, sc, vv, o, i, j);
    else
      rhs_point<1, 0>(C, u, v, p, mode, ov, sc, vv, o, i, j);
  }
}

// SEG: 0, or the segment width of the face sharing (rhs_point): 64 for
// tiles a multiple of 64 wide, else the tile width (16 or 32)
template <int SEG>
__global__ void __launch_bounds__(256) k_b_rhs(const BInst* __restrict__ I, const Tile* __restrict__ tiles,
                                              const double* __restrict__ V, const double* __restrict__ P, int mode,
                                              double* __restrict__ out, StageCombine sc,
                                              double* __restrict__ vtemp) {
  const Tile t = tiles[blockIdx.x];
  const BInst& B = I[t.inst];
  const int i = t.i0 + threadIdx.x, j = t.j0 + threadIdx.y;
  const double* u = V + B.voff;
  const double* v = u + B.nu;
  const double* p = P + B.poff;
  // with SEG, lanes past the row's end stay (they take part in the face
  // sharing of their segment); rows past the end leave as whole segments
  if (t.comp == 0) {
    if ((!SEG && i >= B.nux) || j >= B.nuy) return;
    rhs_point<0, SEG>(*B.cu, u, v, p, mode, out ? out + B.voff : nullptr, sc,
                           vtemp ? vtemp + B.voff : nullptr, B.voff, i, j);
  } else {
    if ((!SEG && i >= B.nvx) || j >= B.nvy) return;
    const long o = B.voff + B.nu;
    rhs_point<1, SEG>(*B.cv, u, v, p, mode, out ? out + o : nullptr, sc, vtemp ? vtemp + o : nullptr, o, i, j);
  }
}

// The divergence and the projection run over tall tiles (TX x kTallRows; the
// block is TX x TY, each thread takes the rows j, j + TY, ...): the chain of
// dependent loads that starts every block (tile, instance, operator records)
// is then paid once per kTallRows / TY rows.  The operator records arrive by
// scalar loads (cst), the per-point arithmetic is branch-free (kron2_at_nb,
// optional operands as template flags) and rows past the component's end are
// clamped to its last row (valid addresses) and not stored, so the loads of
// all rows of a thread issue together.
constexpr int kTallRows = 16;
constexpr int kTallR = kTallRows / TY;

// f = oscale * (M (ca A + cb B) + ycoef yM); with dt != null: ca = ycoef = 1/dt_k
template <bool HASB>
__global__ void __launch_bounds__(256) k_b_div(const BInst* __restrict__ I, const Tile* __restrict__ tiles,
                                              const double* __restrict__ A, const double* __restrict__ Bv,
                                              const double* __restrict__ dt, double ca, double cb, double ycoef,
                                              double oscale, double* __restrict__ f) {
  const Tile t = tiles[blockIdx.x];
  const BInst& B = I[t.inst];
  const int i = t.i0 + threadIdx.x;
  if (i >= B.npx) return;
  double a = ca, y = ycoef;
  if (dt) { a = 1.0 / dt[t.inst]; y = a; }
  const DComp Cu = cst(B.cu), Cv = cst(B.cv);
  const double* Au = A + B.voff;
  const double* Bu = HASB ? Bv + B.voff : nullptr;
  double s[kTallR];
#pragma unroll
  for (int r = 0; r < kTallR; ++r) {
    const int j = min(t.j0 + (int)threadIdx.y + TY * r, B.npy - 1);
    s[r] = kron2_at2_nb<HASB>(Cu.M, Au, a, Bu, cb, i, j);
    s[r] += kron2_at2_nb<HASB>(Cv.M, Au + B.nu, a, HASB ? Bu + B.nu : nullptr, cb, i, j);
  }
  if (y != 0.0) {
#pragma unroll
    for (int r = 0; r < kTallR; ++r) {
      const int j = min(t.j0 + (int)threadIdx.y + TY * r, B.npy - 1);
      s[r] += y * (yv_at(Cu.yM, i, j) + yv_at(Cv.yM, i, j));
    }
  }
#pragma unroll
  for (int r = 0; r < kTallR; ++r) {
    const int j = t.j0 + (int)threadIdx.y + TY * r;
    if (j < B.npy) f[B.poff + (long)i + (long)B.npx * j] = oscale * s[r];
  }
}

// V = Vn + dt_k (Vt - c Om^-1 G dp)  (dt null: dt_k = 1; without Vt: Vt = 0)
template <bool HASVT>
__global__ void __launch_bounds__(256) k_b_proj(const BInst* __restrict__ I, const Tile* __restrict__ tiles,
                                               const double* __restrict__ Vn, const double* __restrict__ Vt,
                                               const double* __restrict__ dp, const double* __restrict__ dt, double c,
                                               double* __restrict__ V) {
  const Tile t = tiles[blockIdx.x];
  const BInst& B = I[t.inst];
  const int i = t.i0 + threadIdx.x;
  const int nx = t.comp == 0 ? B.nux : B.nvx, ny = t.comp == 0 ? B.nuy : B.nvy;
  if (i >= nx) return;
  const double d = dt ? dt[t.inst] : 1.0;
  const double* pp = dp + B.poff;
  const DComp C = cst(t.comp == 0 ? B.cu : B.cv);
  const long o = t.comp == 0 ? B.voff : B.voff + B.nu;
  const auto vn = glb(Vn + o);
  double w[kTallR];
#pragma unroll
  for (int r = 0; r < kTallR; ++r) {
    const int j = min(t.j0 + (int)threadIdx.y + TY * r, ny - 1);
    const unsigned idx = (unsigned)i + (unsigned)nx * (unsigned)j;
    const double g = kron2_at_nb(C.G, pp, i, j) / (glb(C.ox)[i] * glb(C.oy)[j]);
    const double vt = HASVT ? glb(Vt + o)[idx] : 0.0;
    w[r] = vn[idx] + d * (vt - c * g);
  }
#pragma unroll
  for (int r = 0; r < kTallR; ++r) {
    const int j = t.j0 + (int)threadIdx.y + TY * r;
    if (j < ny) V[o + (long)i + (long)nx * j] = w[r];
  }
}

// ---- batched Jacobi-PCG; per-instance state: s[k*8 + {0 rz, 1 ff, 2 alpha,
// 3 beta, 4 rel^2, 5 tol^2}], f[k*4 + {0 done, 1 iterations, 2 failed}]
struct BCG {
  double* s;
  int* f;
  double* part;   // 2 per pressure tile
  int* nactive;
  int maxit;
};

__global__ void __launch_bounds__(256) k_b_cg_init(const BInst* __restrict__ I, const Tile* __restrict__ tiles,
                                                  BCG S, const double* __restrict__ f, double* __restrict__ x,
                                                  double* __restrict__ r, double* __restrict__ z,
                                                  double* __restrict__ p) {
  __shared__ double red[4];
  const Tile t = tiles[blockIdx.x];
  const BInst& B = I[t.inst];
  const int tid = threadIdx.x + TX * threadIdx.y;
  const int i = t.i0 + threadIdx.x, j = t.j0 + threadIdx.y;
  double ff = 0.0, rz = 0.0;
  if (i < B.npx && j < B.npy) {
    const long k = B.poff + (long)i + (long)B.npx * j;
    const double rn = f[k];
    x[k] = 0.0;
    r[k] = rn;
    const double zn = B.lap.dinv[(long)i + (long)B.npx * j] * rn;
    z[k] = zn;
    p[k] = zn;
    ff = rn * rn;
    rz = rn * zn;
  }
  (void)tid;
  const double a = block_sum(ff, red);
  const double b = block_sum(rz, red);
  if (threadIdx.x == 0 && threadIdx.y == 0) { S.part[2 * blockIdx.x] = a; S.part[2 * blockIdx.x + 1] = b; }
}

__global__ void k_b_cg_init_fin(const BInst* __restrict__ I, BCG S, double tol2) {
  __shared__ double red[4];
  const int k = blockIdx.x;
  const BInst& B = I[k];
  double a = 0.0, b = 0.0;
  for (int t = B.pt0 + threadIdx.x; t < B.pt1; t += blockDim.x) { a += S.part[2 * t]; b += S.part[2 * t + 1]; }
  const double ff = block_sum(a, red);
  const double rz = block_sum(b, red);
  if (threadIdx.x == 0) {
    S.s[8 * k + 0] = rz;
    S.s[8 * k + 1] = ff;
    S.s[8 * k + 4] = 1.0;
    S.s[8 * k + 5] = tol2;
    S.f[4 * k + 1] = 0;
    S.f[4 * k + 2] = 0;
    const int done = (ff <= tol2) ? 1 : 0;   // cg.F:222
    S.f[4 * k + 0] = done;
    if (done) atomicSub(S.nactive, 1);
  }
}

__global__ void __launch_bounds__(256) k_b_cg_apply(const BInst* __restrict__ I, const Tile* __restrict__ tiles,
                                                   BCG S, const double* __restrict__ p, double* __restrict__ q) {
  __shared__ double red[4];
  const Tile t = tiles[blockIdx.x];
  if (S.f[4 * t.inst]) return;
  const BInst& B = I[t.inst];
  const int i = t.i0 + threadIdx.x, j = t.j0 + threadIdx.y;
  double acc = 0.0;
  if (i < B.npx && j < B.npy) {
    const long k = B.poff + (long)i + (long)B.npx * j;
    const double a = lap_at(B.lap, p + B.poff, i, j);
    q[k] = a;
    acc = p[k] * a;
  }
  const double s = block_sum(acc, red);
  if (threadIdx.x == 0 && threadIdx.y == 0) S.part[2 * blockIdx.x] = s;
}

// Tall-tile forms of the per-tile reductions: a TX x kTallRows tile covers
// kTallR of the TX x TY pressure tiles whose partial sums the per-instance
// kernels add up over [pt0, pt1).  It writes its sum into the slot of the
// first of them and zeros into the others, so those sums are unchanged.
__device__ __forceinline__ void tall_partial(const BInst& B, const Tile& t, double v, double* __restrict__ part,
                                             int comp) {
  const int ntx = (B.npx + TX - 1) / TX;
  const int s0 = B.pt0 + (t.j0 / TY) * ntx + t.i0 / TX;
  part[2 * s0 + comp] = v;
  for (int m = 1; m < kTallR; ++m)
    if (t.j0 + m * TY < B.npy) part[2 * (s0 + m * ntx) + comp] = 0.0;
}

// k_b_cg_apply over tall tiles, every instance's Laplacian in the widths of
// lap_tri_at: the x factors of the column loaded once for all its rows
__global__ void __launch_bounds__(256) k_b_cg_apply_tall(const BInst* __restrict__ I, const Tile* __restrict__ tiles,
                                                        BCG S, const double* __restrict__ p, double* __restrict__ q) {
  __shared__ double red[4];
  const Tile t = tiles[blockIdx.x];
  if (S.f[4 * t.inst]) return;
  const BInst& B = I[t.inst];
  const int i = t.i0 + threadIdx.x;
  const int npx = B.npx, npy = B.npy;
  double acc = 0.0;
  if (i < npx) {
    const DLap L = cst(&B.lap);
    const auto a = glb(p + B.poff);
    const unsigned ld = (unsigned)L.L1.X.ncols;
    const unsigned i3 = 3u * (unsigned)i;
    const auto c1x = glb(L.L1.X.col), c2y = glb(L.L2.Y.col);
    const auto w1x = glb(L.L1.X.val), w2y = glb(L.L2.Y.val);
    const unsigned x0 = (unsigned)c1x[i3], x1 = (unsigned)c1x[i3 + 1], x2 = (unsigned)c1x[i3 + 2];
    const double u0 = w1x[i3], u1 = w1x[i3 + 1], u2 = w1x[i3 + 2];
    const unsigned c2 = (unsigned)glb(L.L2.X.col)[i];
    const double w2 = glb(L.L2.X.val)[i];
    double qv[kTallR], pv[kTallR];
#pragma unroll
    for (int r = 0; r < kTallR; ++r) {
      const int j = min(t.j0 + (int)threadIdx.y + TY * r, npy - 1);
      const unsigned j3 = 3u * (unsigned)j;
      const unsigned r1 = (unsigned)glb(L.L1.Y.col)[j] * ld;
      const double w1y = glb(L.L1.Y.val)[j];
      double s1 = u0 * a[r1 + x0];
      s1 += u1 * a[r1 + x1];
      s1 += u2 * a[r1 + x2];
      double s2 = w2y[j3] * a[(unsigned)c2y[j3] * ld + c2];
      s2 += w2y[j3 + 1] * a[(unsigned)c2y[j3 + 1] * ld + c2];
      s2 += w2y[j3 + 2] * a[(unsigned)c2y[j3 + 2] * ld + c2];
      qv[r] = w1y * s1 + w2 * s2;
      pv[r] = a[(unsigned)j * ld + (unsigned)i];
    }
#pragma unroll
    for (int r = 0; r < kTallR; ++r) {
      const int j = t.j0 + (int)threadIdx.y + TY * r;
      if (j < npy) {
        q[B.poff + (long)i + (long)npx * j] = qv[r];
        acc += pv[r] * qv[r];
      }
    }
  }
  const double s = block_sum(acc, red);
  if (threadIdx.x == 0 && threadIdx.y == 0) tall_partial(B, t, s, S.part, 0);
}

__global__ void k_b_cg_alpha(const BInst* __restrict__ I, BCG S) {
  __shared__ double red[4];
  const int k = blockIdx.x;
  if (S.f[4 * k]) return;
  const BInst& B = I[k];
  double a = 0.0;
  for (int t = B.pt0 + threadIdx.x; t < B.pt1; t += blockDim.x) a += S.part[2 * t];
  const double pq = block_sum(a, red);
  if (threadIdx.x == 0) S.s[8 * k + 2] = S.s[8 * k + 0] / pq;
}

__global__ void __launch_bounds__(256) k_b_cg_update(const BInst* __restrict__ I, const Tile* __restrict__ tiles,
                                                    BCG S, const double* __restrict__ p,
                                                    const double* __restrict__ q, double* __restrict__ x,
                                                    double* __restrict__ r, double* __restrict__ z) {
  __shared__ double red[4];
  const Tile t = tiles[blockIdx.x];
  if (S.f[4 * t.inst]) return;
  const BInst& B = I[t.inst];
  const int i = t.i0 + threadIdx.x, j = t.j0 + threadIdx.y;
  const double alpha = S.s[8 * t.inst + 2];
  double rr = 0.0, rz = 0.0;
  if (i < B.npx && j < B.npy) {
    const long k = B.poff + (long)i + (long)B.npx * j;
    x[k] += alpha * p[k];
    const double rn = r[k] - alpha * q[k];
    r[k] = rn;
    const double zn = B.lap.dinv[(long)i + (long)B.npx * j] * rn;
    z[k] = zn;
    rr = rn * rn;
    rz = rn * zn;
  }
  const double a = block_sum(rr, red);
  const double b = block_sum(rz, red);
  if (threadIdx.x == 0 && threadIdx.y == 0) { S.part[2 * blockIdx.x] = a; S.part[2 * blockIdx.x + 1] = b; }
}

__global__ void k_b_cg_beta(const BInst* __restrict__ I, BCG S) {
  __shared__ double red[4];
  const int k = blockIdx.x;
  if (S.f[4 * k]) return;
  const BInst& B = I[k];
  double a = 0.0, b = 0.0;
  for (int t = B.pt0 + threadIdx.x; t < B.pt1; t += blockDim.x) { a += S.part[2 * t]; b += S.part[2 * t + 1]; }
  const double rr = block_sum(a, red);
  const double rz = block_sum(b, red);
  if (threadIdx.x == 0) {
    const int it = S.f[4 * k + 1] + 1;
    S.f[4 * k + 1] = it;
    const double rel = rr / S.s[8 * k + 1];
    S.s[8 * k + 4] = rel;
    int done = 0;
    if (rel <= S.s[8 * k + 5]) {
      done = 1;
    } else if (it >= S.maxit) {
      done = 1;
      S.f[4 * k + 2] = 1;
    }
    S.s[8 * k + 3] = rz / S.s[8 * k + 0];
    S.s[8 * k + 0] = rz;
    if (done) {
      S.f[4 * k + 0] = 1;
      atomicSub(S.nactive, 1);
    }
  }
}

__global__ void __launch_bounds__(256) k_b_cg_dir(const BInst* __restrict__ I, const Tile* __restrict__ tiles, BCG S,
                                                 const double* __restrict__ z, double* __restrict__ p) {
  const Tile t = tiles[blockIdx.x];
  if (S.f[4 * t.inst]) return;
  const BInst& B = I[t.inst];
  const int i = t.i0 + threadIdx.x, j = t.j0 + threadIdx.y;
  if (i >= B.npx || j >= B.npy) return;
  const long k = B.poff + (long)i + (long)B.npx * j;
  p[k] = z[k] + S.s[8 * t.inst + 3] * p[k];
}

// ---- batched two-level PCG: z = D^-1 r + P Ac^-1 P^T r per instance (P the
// aggregation over kAgg x kAgg tiles, Ac = P^T A P solved by fast
// diagonalization in one workgroup per instance); partials 3 per tile
// {p.q or f.f, r.r, r.D^-1 r}; same per-instance scalars as the Jacobi PCG
struct BTL {
  double* part;
  double* rc;
  double* ec;
};

__device__ __forceinline__ int agg_index(const BInst& B, const Tile& t) {
  return (t.i0 / kAgg) + B.ncx * (t.j0 / kAgg);
}

__global__ void __launch_bounds__(256) k_b2_init(const BInst* __restrict__ I, const Tile* __restrict__ tiles, BCG S,
                                                BTL T, const double* __restrict__ f, double* __restrict__ x,
                                                double* __restrict__ r) {
  __shared__ double red[4];
  const Tile t = tiles[blockIdx.x];
  const BInst& B = I[t.inst];
  const int i = t.i0 + threadIdx.x, j = t.j0 + threadIdx.y;
  double fi = 0.0, di = 0.0;
  if (i < B.npx && j < B.npy) {
    const long k = B.poff + (long)i + (long)B.npx * j;
    fi = f[k];
    x[k] = 0.0;
    r[k] = fi;
    di = B.lap.dinv[(long)i + (long)B.npx * j];
  }
  const double ff = block_sum(fi * fi, red);
  const double rd = block_sum(fi * di * fi, red);
  const double rs = block_sum(fi, red);
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    T.part[3 * blockIdx.x] = ff;
    T.part[3 * blockIdx.x + 1] = ff;
    T.part[3 * blockIdx.x + 2] = rd;
    T.rc[B.coff + agg_index(B, t)] = rs;
  }
  (void)S;
}

template <bool TRI>
__global__ void __launch_bounds__(256) k_b2_apply(const BInst* __restrict__ I, const Tile* __restrict__ tiles, BCG S,
                                                 BTL T, const double* __restrict__ p, double* __restrict__ q) {
  __shared__ double red[4];
  const Tile t = tiles[blockIdx.x];
  if (S.f[4 * t.inst]) return;
  const BInst& B = I[t.inst];
  const int i = t.i0 + threadIdx.x, j = t.j0 + threadIdx.y;
  double acc = 0.0;
  if (i < B.npx && j < B.npy) {
    const long k = B.poff + (long)i + (long)B.npx * j;
    const double a = TRI ? lap_tri_at(cst(&B.lap), p + B.poff, i, j) : lap_at(B.lap, p + B.poff, i, j);
    q[k] = a;
    acc = p[k] * a;
  }
  const double s = block_sum(acc, red);
  if (threadIdx.x == 0 && threadIdx.y == 0) T.part[3 * blockIdx.x] = s;
}

__global__ void k_b2_alpha(const BInst* __restrict__ I, BCG S, BTL T) {
  __shared__ double red[4];
  const int k = blockIdx.x;
  if (S.f[4 * k]) return;
  const BInst& B = I[k];
  double a = 0.0;
  for (int t = B.at0 + threadIdx.x; t < B.at1; t += blockDim.x) a += T.part[3 * t];
  const double pq = block_sum(a, red);
  if (threadIdx.x == 0) S.s[8 * k + 2] = S.s[8 * k + 0] / pq;
}

__global__ void __launch_bounds__(256) k_b2_update(const BInst* __restrict__ I, const Tile* __restrict__ tiles, BCG S,
                                                  BTL T, const double* __restrict__ p, const double* __restrict__ q,
                                                  double* __restrict__ x, double* __restrict__ r) {
  __shared__ double red[4];
  const Tile t = tiles[blockIdx.x];
  if (S.f[4 * t.inst]) return;
  const BInst& B = I[t.inst];
  const int i = t.i0 + threadIdx.x, j = t.j0 + threadIdx.y;
  const double alpha = S.s[8 * t.inst + 2];
  double ri = 0.0, di = 0.0;
  if (i < B.npx && j < B.npy) {
    const long k = B.poff + (long)i + (long)B.npx * j;
    x[k] += alpha * p[k];
    ri = r[k] - alpha * q[k];
    r[k] = ri;
    di = B.lap.dinv[(long)i + (long)B.npx * j];
  }
  const double rr = block_sum(ri * ri, red);
  const double rd = block_sum(ri * di * ri, red);
  const double rs = block_sum(ri, red);
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    T.part[3 * blockIdx.x + 1] = rr;
    T.part[3 * blockIdx.x + 2] = rd;
    T.rc[B.coff + agg_index(B, t)] = rs;
  }
}

// one workgroup per instance: convergence test (cg.F:247), coarse solve,
// r.z and beta; init: after k_b2_init (ff, beta = 0)
__global__ void __launch_bounds__(256) k_b2_coarse(const BInst* __restrict__ I, BCG S, BTL T, int init,
                                                  double tol2) {
  __shared__ double red[4];
  __shared__ double Ab[kBMaxCoarse];
  __shared__ double Bb[kBMaxCoarse];
  __shared__ double bc[2];
  __shared__ int done;
  const int k = blockIdx.x;
  if (S.f[4 * k]) return;
  const BInst& B = I[k];
  double rr = 0.0, rd = 0.0;
  for (int t = B.at0 + threadIdx.x; t < B.at1; t += blockDim.x) {
    rr += T.part[3 * t + 1];
    rd += T.part[3 * t + 2];
  }
  rr = block_sum(rr, red);
  rd = block_sum(rd, red);
  if (threadIdx.x == 0) {
    int d = 0;
    if (init) {
      S.s[8 * k + 1] = rr;   // ||f||^2 (r = f)
      S.s[8 * k + 4] = 1.0;
      S.s[8 * k + 5] = tol2;
      S.f[4 * k + 1] = 0;
      S.f[4 * k + 2] = 0;
      d = (rr <= tol2) ? 1 : 0;   // cg.F:222
    } else {
      const int it = S.f[4 * k + 1] + 1;
      S.f[4 * k + 1] = it;
      const double rel = rr / S.s[8 * k + 1];
      S.s[8 * k + 4] = rel;
      if (rel <= S.s[8 * k + 5]) {
        d = 1;
      } else if (it >= S.maxit) {
        d = 1;
        S.f[4 * k + 2] = 1;
      }
    }
    if (d) {
      S.f[4 * k + 0] = 1;
      atomicSub(S.nactive, 1);
    }
    done = d;
    bc[0] = rd;
  }
  __syncthreads();
  if (done) return;
  double rz = bc[0];
  const int m = B.ncx, n = B.ncy, mn = m * n;
  const double* rc = T.rc + B.coff;
  double* ec = T.ec + B.coff;
  for (int e = threadIdx.x; e < mn; e += blockDim.x) Ab[e] = rc[e];
  __syncthreads();
  for (int e = threadIdx.x; e < mn; e += blockDim.x) {
    const int a = e % m, b2 = e / m;
    double s = 0.0;
    for (int q = 0; q < m; ++q) s += B.cVx[q + (long)m * a] * Ab[q + m * b2];
    Bb[e] = s;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < mn; e += blockDim.x) {
    const int a = e % m, b2 = e / m;
    double s = 0.0;
    for (int q = 0; q < n; ++q) s += Bb[a + m * q] * B.cVy[q + (long)n * b2];
    Ab[e] = (a == B.nx0 && b2 == B.ny0) ? 0.0 : -s / (B.clx[a] + B.cly[b2]);
  }
  __syncthreads();
  for (int e = threadIdx.x; e < mn; e += blockDim.x) {
    const int a = e % m, b2 = e / m;
    double s = 0.0;
    for (int q = 0; q < m; ++q) s += B.cVx[a + (long)m * q] * Ab[q + m * b2];
    Bb[e] = s;
  }
  __syncthreads();
  double dot = 0.0;
  for (int e = threadIdx.x; e < mn; e += blockDim.x) {
    const int a = e % m, b2 = e / m;
    double s = 0.0;
    for (int q = 0; q < n; ++q) s += Bb[a + m * q] * B.cVy[b2 + (long)n * q];
    ec[e] = s;
    dot += rc[e] * s;
  }
  dot = block_sum(dot, red);
  if (threadIdx.x == 0) {
    rz += dot;
    S.s[8 * k + 3] = init ? 0.0 : rz / S.s[8 * k + 0];
    S.s[8 * k + 0] = rz;
  }
}

__global__ void __launch_bounds__(256) k_b2_dir(const BInst* __restrict__ I, const Tile* __restrict__ tiles, BCG S,
                                               BTL T, const double* __restrict__ r, double* __restrict__ p) {
  const Tile t = tiles[blockIdx.x];
  if (S.f[4 * t.inst]) return;
  const BInst& B = I[t.inst];
  const int i = t.i0 + threadIdx.x, j = t.j0 + threadIdx.y;
  if (i >= B.npx || j >= B.npy) return;
  const long k = B.poff + (long)i + (long)B.npx * j;
  const double z = B.lap.dinv[(long)i + (long)B.npx * j] * r[k] + T.ec[B.coff + agg_index(B, t)];
  p[k] = z + S.s[8 * t.inst + 3] * p[k];
}

// ---- batched PCG with the fast-diagonalization preconditioner: z = A^-1 r
// by the four GEMMs of the direct solve (fdm.hip), per instance, over a work
// list of 64 x 64 tiles of the Npx x Npy pressure arrays; the per-instance
// scalars and the convergence test (cg.F:247) are the Jacobi PCG's, so the
// preconditioner (the exact inverse on these tensor-product grids, up to
// rounding) changes only how many iterations reach the tolerance.
// PH 0: dst = Vx^T src; 1: dst = -(src Vy) ./ (lx + ly); 2: dst = Vx src;
// 3: dst = src Vy^T
template <int PH>
__global__ void __launch_bounds__(256) k_bf_gemm(const BInst* __restrict__ I, const Tile* __restrict__ tiles, BCG S,
                                                const double* __restrict__ src, double* __restrict__ dst) {
  __shared__ Gemm64Smem sm;
  const Tile t = tiles[blockIdx.x];
  if (S.f[4 * t.inst]) return;
  const BInst& B = I[t.inst];
  const int m = B.npx, n = B.npy;
  const double* s = src + B.poff;
  double* d = dst + B.poff;
  if (PH == 0) dgemm_tile64<0>(m, n, m, B.fVxT, m, s, m, d, m, nullptr, nullptr, -1, -1, t.i0, t.j0, sm);
  if (PH == 1) dgemm_tile64<1>(m, n, n, s, m, B.fVy, n, d, m, B.flx, B.fly, B.fnx0, B.fny0, t.i0, t.j0, sm);
  if (PH == 2) dgemm_tile64<0>(m, n, m, B.fVx, m, s, m, d, m, nullptr, nullptr, -1, -1, t.i0, t.j0, sm);
  if (PH == 3) dgemm_tile64<0>(m, n, n, s, m, B.fVyT, n, d, m, nullptr, nullptr, -1, -1, t.i0, t.j0, sm);
}

// x = 0, r = f; partial f.f per pressure tile
__global__ void __launch_bounds__(256) k_bf_init(const BInst* __restrict__ I, const Tile* __restrict__ tiles, BCG S,
                                                const double* __restrict__ f, double* __restrict__ x,
                                                double* __restrict__ r) {
  __shared__ double red[4];
  const Tile t = tiles[blockIdx.x];
  const BInst& B = I[t.inst];
  const int i = t.i0 + threadIdx.x, j = t.j0 + threadIdx.y;
  double ff = 0.0;
  if (i < B.npx && j < B.npy) {
    const long k = B.poff + (long)i + (long)B.npx * j;
    const double rn = f[k];
    x[k] = 0.0;
    r[k] = rn;
    ff = rn * rn;
  }
  const double a = block_sum(ff, red);
  if (threadIdx.x == 0 && threadIdx.y == 0) S.part[2 * blockIdx.x] = a;
}

// partial r.z per pressure tile (active instances)
__global__ void __launch_bounds__(256) k_bf_rz(const BInst* __restrict__ I, const Tile* __restrict__ tiles, BCG S,
                                              const double* __restrict__ r, const double* __restrict__ z) {
  __shared__ double red[4];
  const Tile t = tiles[blockIdx.x];
  if (S.f[4 * t.inst]) return;
  const BInst& B = I[t.inst];
  const int i = t.i0 + threadIdx.x, j = t.j0 + threadIdx.y;
  double rz = 0.0;
  if (i < B.npx && j < B.npy) {
    const long k = B.poff + (long)i + (long)B.npx * j;
    rz = r[k] * z[k];
  }
  const double a = block_sum(rz, red);
  if (threadIdx.x == 0 && threadIdx.y == 0) S.part[2 * blockIdx.x + 1] = a;
}

// k_bf_rz over tall tiles
__global__ void __launch_bounds__(256) k_bf_rz_tall(const BInst* __restrict__ I, const Tile* __restrict__ tiles, BCG S,
                                                   const double* __restrict__ r, const double* __restrict__ z) {
  __shared__ double red[4];
  const Tile t = tiles[blockIdx.x];
  if (S.f[4 * t.inst]) return;
  const BInst& B = I[t.inst];
  const int i = t.i0 + threadIdx.x;
  const int npx = B.npx, npy = B.npy;
  double rz = 0.0;
  if (i < npx) {
    const auto rr = glb(r + B.poff), zz = glb(z + B.poff);
    double a[kTallR], b[kTallR];
#pragma unroll
    for (int q = 0; q < kTallR; ++q) {
      const int j = min(t.j0 + (int)threadIdx.y + TY * q, npy - 1);
      const unsigned k = (unsigned)i + (unsigned)npx * (unsigned)j;
      a[q] = rr[k];
      b[q] = zz[k];
    }
#pragma unroll
    for (int q = 0; q < kTallR; ++q)
      if (t.j0 + (int)threadIdx.y + TY * q < npy) rz += a[q] * b[q];
  }
  const double s = block_sum(rz, red);
  if (threadIdx.x == 0 && threadIdx.y == 0) tall_partial(B, t, s, S.part, 1);
}

// x += alpha p, r -= alpha q; partial r.r per pressure tile
__global__ void __launch_bounds__(256) k_bf_update(const BInst* __restrict__ I, const Tile* __restrict__ tiles, BCG S,
                                                  const double* __restrict__ p, const double* __restrict__ q,
                                                  double* __restrict__ x, double* __restrict__ r) {
  __shared__ double red[4];
  const Tile t = tiles[blockIdx.x];
  if (S.f[4 * t.inst]) return;
  const BInst& B = I[t.inst];
  const int i = t.i0 + threadIdx.x, j = t.j0 + threadIdx.y;
  const double alpha = S.s[8 * t.inst + 2];
  double rr = 0.0;
  if (i < B.npx && j < B.npy) {
    const long k = B.poff + (long)i + (long)B.npx * j;
    x[k] += alpha * p[k];
    const double rn = r[k] - alpha * q[k];
    r[k] = rn;
    rr = rn * rn;
  }
  const double a = block_sum(rr, red);
  if (threadIdx.x == 0 && threadIdx.y == 0) S.part[2 * blockIdx.x] = a;
}

// per instance: iteration count and the convergence test of cg.F:247 on
// ||r||^2 (partials 2t), before the preconditioner is applied again
__global__ void k_bf_test(const BInst* __restrict__ I, BCG S) {
  __shared__ double red[4];
  const int k = blockIdx.x;
  if (S.f[4 * k]) return;
  const BInst& B = I[k];
  double a = 0.0;
  for (int t = B.pt0 + threadIdx.x; t < B.pt1; t += blockDim.x) a += S.part[2 * t];
  const double rr = block_sum(a, red);
  if (threadIdx.x == 0) {
    const int it = S.f[4 * k + 1] + 1;
    S.f[4 * k + 1] = it;
    const double rel = rr / S.s[8 * k + 1];
    S.s[8 * k + 4] = rel;
    int done = 0;
    if (rel <= S.s[8 * k + 5]) {
      done = 1;
    } else if (it >= S.maxit) {
      done = 1;
      S.f[4 * k + 2] = 1;
    }
    if (done) {
      S.f[4 * k + 0] = 1;
      atomicSub(S.nactive, 1);
    }
  }
}

// per instance: beta = (r.z)_new / (r.z)_old (partials 2t + 1)
__global__ void k_bf_beta(const BInst* __restrict__ I, BCG S) {
  __shared__ double red[4];
  const int k = blockIdx.x;
  if (S.f[4 * k]) return;
  const BInst& B = I[k];
  double b = 0.0;
  for (int t = B.pt0 + threadIdx.x; t < B.pt1; t += blockDim.x) b += S.part[2 * t + 1];
  const double rz = block_sum(b, red);
  if (threadIdx.x == 0) {
    S.s[8 * k + 3] = rz / S.s[8 * k + 0];
    S.s[8 * k + 0] = rz;
  }
}

// first direction p = z (active instances)
__global__ void __launch_bounds__(256) k_bf_dir0(const BInst* __restrict__ I, const Tile* __restrict__ tiles, BCG S,
                                                const double* __restrict__ z, double* __restrict__ p) {
  const Tile t = tiles[blockIdx.x];
  if (S.f[4 * t.inst]) return;
  const BInst& B = I[t.inst];
  const int i = t.i0 + threadIdx.x, j = t.j0 + threadIdx.y;
  if (i >= B.npx || j >= B.npy) return;
  const long k = B.poff + (long)i + (long)B.npx * j;
  p[k] = z[k];
}

__global__ void k_b_set_int(int* x, int v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *x = v;
}

__global__ void k_b_clear_flags(int* f, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) f[4 * k] = 0;
}

// end of a device-iterated solve: per instance the iteration count into the
// running totals, and a failure when the stopping rule was not met (or the
// iteration limit reached) within the solve's iterations
__global__ void k_bf_stats(BCG S, int n, unsigned long long* __restrict__ it_sum, int* __restrict__ it_max,
                           int* __restrict__ nfail) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const int it = S.f[4 * k + 1];
    atomicAdd(it_sum, (unsigned long long)it);
    atomicMax(it_max, it);
    if (!S.f[4 * k] || S.f[4 * k + 2]) atomicAdd(nfail, 1);
  }
}

}  // namespace ins2d

using namespace ins2d;

namespace {

struct BErr : std::runtime_error {
  int code;
  BErr(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define BHIP(x)                                                                                     \
  do {                                                                                              \
    hipError_t e_ = (x);                                                                            \
    if (e_ != hipSuccess) throw BErr(INS2D_E_HIP, std::string(#x ": ") + hipGetErrorString(e_));   \
  } while (0)

// single device buffer holding every operator array of the ensemble
struct Arena {
  std::vector<char> host;
  char* dev = nullptr;
  size_t cap = 0;
  bool counting = true;
  size_t used = 0;
  template <typename T>
  T* push(const std::vector<T>& v) {
    const size_t bytes = std::max<size_t>(v.size(), 1) * sizeof(T);
    used = (used + 255) & ~size_t(255);
    size_t off = used;
    used += bytes;
    if (counting) return nullptr;
    if (used > cap) throw std::runtime_error("arena overflow");
    if (!v.empty()) std::memcpy(host.data() + off, v.data(), v.size() * sizeof(T));
    return (T*)(dev + off);
  }
};

DOp1 a_op(Arena& A, const HOp1& o) {
  DOp1 d{};
  d.nrows = o.nrows; d.ncols = o.ncols; d.K = o.K; d.kmax = o.kmax;
  d.col = A.push(o.col);
  d.val = A.push(o.val);
  return d;
}
DKron a_kron(Arena& A, const HKron& k) { return DKron{a_op(A, k.X), a_op(A, k.Y)}; }

}  // namespace

struct ins2d_batch {
  std::string err;
  hipStream_t st = nullptr;
  int n = 0;
  std::vector<std::unique_ptr<Plan>> plans;
  std::vector<BCSamples> samp;
  std::vector<DComp> hcomp;   // 2 per instance, host copies with device pointers
  DComp* dcomp = nullptr;
  std::vector<BInst> hinst;
  BInst* dinst = nullptr;
  std::vector<Tile> pt;   // pressure tiles, TX x TY
  Tile* dpt = nullptr;
  // velocity and pressure areas in tall tiles (TX x kTallRows) for the
  // divergence and the projection
  std::vector<Tile> uvt_tall, pt_tall;
  Tile *duv_tall = nullptr, *dpt_tall = nullptr;
  // momentum right-hand side: members with a closed form over split tiles
  // (rs), the others over tiles of the tuned shape tx x ty (rg)
  std::vector<RTile> rs;
  RTile* drs = nullptr;
  size_t rs_cap = 0, nclosed = 0;
  // no closed-form point reads a boundary term (checked against the boundary
  // masks whenever they change): the interior tiles may run the closed-form
  // kernel, which reads no masks
  bool closed_ok = false;
  std::vector<Tile> rg;
  Tile* drg = nullptr;
  size_t rg_cap = 0;
  Arena ops;                  // operator arrays
  std::vector<void*> allocs, yv_allocs;
  long NVt = 0, Npt = 0;
  double *V = nullptr, *Vb = nullptr, *P = nullptr, *Vt = nullptr, *F = nullptr, *dP = nullptr, *tV = nullptr;
  std::vector<double*> kv;
  double *x = nullptr, *r = nullptr, *z = nullptr, *pp = nullptr, *q = nullptr;
  double* dt = nullptr;
  BCG cg{};
  double tol = 1e-10;
  int maxit = 1000;
  int s = 0;
  std::vector<double> At, ct;
  std::vector<bool> keep;
  hipGraphExec_t cg_chunk = nullptr;
  hipGraph_t cg_graph = nullptr;
  // fast-diagonalization PCG: the tail of an iteration (cg_chunk holds its head)
  hipGraphExec_t cg_tail = nullptr;
  hipGraph_t cg_tail_graph = nullptr;
  // fast-diagonalization PCG with device-side iterations: a whole RK step
  // (every stage and solve) captured as one graph per parity of the state
  // buffers (the step swaps V and Vb), replayed without host checks;
  // iteration totals and failures accumulate on the device
  bool device_iter = false;
  int fixed_iters = 1;
  double *Vsave = nullptr, *Psave = nullptr;   // state at the start of a call (for the fallback)
  hipGraphExec_t step_exec[2] = {nullptr, nullptr};
  hipGraph_t step_graph[2] = {nullptr, nullptr};
  const double* step_V[2] = {nullptr, nullptr};
  unsigned long long* d_it_sum = nullptr;
  int* d_it_max = nullptr;   // d_it_max[0]: largest count, d_it_max[1]: failures
  unsigned long long h_it_sum = 0;
  // stage-kernel tile shape (velocity tiles), chosen by ins2d_batch_autotune
  int tx = TX, ty = TY;
  // two-level PCG (cg_precond = 2)
  bool two_level = false;
  // fast-diagonalization PCG (cg_precond = 3): per-instance factors, the
  // 64 x 64 GEMM tile list, two work arrays
  bool fdm = false;
  bool lap_tri = false;   // every instance's Laplacian in the widths of lap_tri_at
  std::vector<FDMHost> fh;
  // closed-form maps of the momentum kernels (uniform members), per member
  std::vector<UniForm> ug;
  std::vector<UniArrays> ua;
  std::vector<std::vector<unsigned char>> ufu, ufv;
  std::vector<Tile> gt;
  // the GEMM tiles in two launch orders, longest inner dimension first (npx
  // for the products with Vx, npy for those with Vy): the blocks of the
  // largest members start first and the small ones fill in at the end
  Tile *dgtx = nullptr, *dgty = nullptr;
  double *gw1 = nullptr, *gw2 = nullptr;
  std::vector<Tile> at;
  Tile* dat = nullptr;
  BTL tl{};
  int chunk = 16;
  long long total_it = 0;
  int max_it = 0;
  std::vector<int> last_it;
  bool bc_dirty = false;
};

namespace {

template <typename T>
T* balloc(ins2d_batch* b, size_t n) {
  void* p = nullptr;
  BHIP(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
  b->allocs.push_back(p);
  return (T*)p;
}

// the captured steps hold the addresses of the tile lists and boundary
// vectors: dropped whenever those are rebuilt
void drop_step_graphs(ins2d_batch* b) {
  for (int k = 0; k < 2; ++k) {
    if (b->step_exec[k]) (void)hipGraphExecDestroy(b->step_exec[k]);
    if (b->step_graph[k]) (void)hipGraphDestroy(b->step_graph[k]);
    b->step_exec[k] = nullptr;
    b->step_graph[k] = nullptr;
    b->step_V[k] = nullptr;
  }
}

void b_upload_bc(ins2d_batch* b) {
  drop_step_graphs(b);
  for (void* p : b->yv_allocs) (void)hipFree(p);
  b->yv_allocs.clear();
  // size, then fill one buffer with every boundary term of every instance
  std::vector<std::pair<DYV*, std::vector<std::vector<double>>>> pending;
  for (int k = 0; k < b->n; ++k) {
    const Plan& P = *b->plans[k];
    for (int c = 0; c < 2; ++c) {
      const HComp& H = c == 0 ? P.cu : P.cv;
      DComp& D = b->hcomp[2 * k + c];
      auto add = [&](const HYVec& Y, DYV& dy) {
        std::vector<std::vector<double>> ax, ay;
        eval_yvec(Y, b->samp[k], ax, ay);
        if (ax.size() > 4) throw BErr(INS2D_E_ARG, "too many boundary terms");
        dy.nt = (int)ax.size();
        std::vector<std::vector<double>> v;
        for (size_t t = 0; t < ax.size(); ++t) { v.push_back(ax[t]); v.push_back(ay[t]); }
        pending.push_back({&dy, v});
      };
      for (int t = 0; t < D.nconv; ++t) { add(H.conv[t].yI, D.conv[t].yI); add(H.conv[t].yA, D.conv[t].yA); }
      for (int t = 0; t < D.ndiff; ++t) add(H.diff[t].yS, D.diff[t].yS);
      add(H.yM, D.yM);
    }
  }
  size_t total = 0;
  for (auto& pe : pending)
    for (auto& v : pe.second) total += v.size();
  double* dbuf = nullptr;
  BHIP(hipMalloc(&dbuf, std::max<size_t>(total, 1) * sizeof(double)));
  b->yv_allocs.push_back(dbuf);
  std::vector<double> hbuf(std::max<size_t>(total, 1));
  size_t off = 0;
  for (auto& pe : pending) {
    DYV& dy = *pe.first;
    for (int t = 0; t < 4; ++t) {
      dy.ax[t] = nullptr;
      dy.ay[t] = nullptr;
      dy.x0[t] = dy.x1[t] = dy.y0[t] = dy.y1[t] = 0;
    }
    for (int t = 0; t < dy.nt; ++t) {
      const auto& ax = pe.second[2 * t];
      const auto& ay = pe.second[2 * t + 1];
      nonzero_range(ax.data(), ax.size(), dy.x0[t], dy.x1[t]);
      nonzero_range(ay.data(), ay.size(), dy.y0[t], dy.y1[t]);
      std::copy(ax.begin(), ax.end(), hbuf.begin() + off);
      dy.ax[t] = dbuf + off;
      off += ax.size();
      std::copy(ay.begin(), ay.end(), hbuf.begin() + off);
      dy.ay[t] = dbuf + off;
      off += ay.size();
    }
  }
  BHIP(hipMemcpy(dbuf, hbuf.data(), std::max<size_t>(total, 1) * sizeof(double), hipMemcpyHostToDevice));
  // boundary masks of every instance and component, in one buffer
  std::vector<unsigned> masks;
  std::vector<size_t> moff;
  for (int k = 0; k < b->n; ++k)
    for (int c = 0; c < 2; ++c) {
      std::vector<unsigned> mx, my;
      bc_masks(c == 0 ? b->plans[k]->cu : b->plans[k]->cv, b->hcomp[2 * k + c], mx, my);
      moff.push_back(masks.size());
      masks.insert(masks.end(), mx.begin(), mx.end());
      moff.push_back(masks.size());
      masks.insert(masks.end(), my.begin(), my.end());
    }
  unsigned* dmask = nullptr;
  BHIP(hipMalloc(&dmask, std::max<size_t>(masks.size(), 1) * sizeof(unsigned)));
  b->yv_allocs.push_back(dmask);
  BHIP(hipMemcpy(dmask, masks.data(), masks.size() * sizeof(unsigned), hipMemcpyHostToDevice));
  for (size_t q = 0; q < b->hcomp.size(); ++q) {
    b->hcomp[q].bmx = dmask + moff[2 * q];
    b->hcomp[q].bmy = dmask + moff[2 * q + 1];
  }
  // the closed-form kernel reads no boundary masks: allowed only while no
  // point of a closed-form map reads a boundary term
  b->closed_ok = true;
  for (int k = 0; k < b->n && b->closed_ok; ++k)
    for (int c = 0; c < 2 && b->closed_ok; ++c) {
      const std::vector<unsigned char>& f = c == 0 ? b->ufu[k] : b->ufv[k];
      if (f.empty()) continue;
      const int nx = b->hcomp[2 * k + c].nx, ny = b->hcomp[2 * k + c].ny;
      const unsigned* mx = masks.data() + moff[2 * (2 * k + c)];
      const unsigned* my = masks.data() + moff[2 * (2 * k + c) + 1];
      for (int j = 0; j < ny && b->closed_ok; ++j)
        for (int i = 0; i < nx; ++i)
          if (f[(size_t)i + (size_t)nx * j] && (mx[i] & my[j])) {
            b->closed_ok = false;
            break;
          }
    }
  BHIP(hipMemcpy(b->dcomp, b->hcomp.data(), b->hcomp.size() * sizeof(DComp), hipMemcpyHostToDevice));
}

void build_ops(ins2d_batch* b, Arena& A) {
  for (int k = 0; k < b->n; ++k) {
    const Plan& P = *b->plans[k];
    for (int c = 0; c < 2; ++c) {
      const HComp& H = c == 0 ? P.cu : P.cv;
      DComp& D = b->hcomp[2 * k + c];
      D.nx = H.nx; D.ny = H.ny;
      D.nconv = (int)H.conv.size();
      D.ndiff = (int)H.diff.size();
      for (int t = 0; t < D.nconv; ++t) {
        D.conv[t].D = a_kron(A, H.conv[t].D);
        D.conv[t].I = a_kron(A, H.conv[t].I);
        D.conv[t].A = a_kron(A, H.conv[t].A);
        D.conv[t].icomp = H.conv[t].icomp;
        D.conv[t].scale = H.conv[t].scale;
      }
      for (int t = 0; t < D.ndiff; ++t) {
        D.diff[t].D = a_kron(A, H.diff[t].D[0]);
        D.diff[t].S = a_kron(A, H.diff[t].S[0]);
        D.diff[t].sscale = H.diff[t].sscale[0];
      }
      D.ox = A.push(H.ox);
      D.oy = A.push(H.oy);
      D.G = a_kron(A, H.G);
      D.M = a_kron(A, H.M);
      const std::vector<unsigned char>& fm = c == 0 ? b->ufu[k] : b->ufv[k];
      D.uni = b->ug[k];
      D.uni.fast = fm.empty() ? nullptr : A.push(fm);
      if (D.uni.ahx) {   // stretched member: the spacing arrays in the arena
        const UniArrays& S = b->ua[k];
        D.uni.ahx = A.push(S.hx); D.uni.ahy = A.push(S.hy); D.uni.agx = A.push(S.gx); D.uni.agy = A.push(S.gy);
        D.uni.arhx = A.push(S.rhx); D.uni.arhy = A.push(S.rhy); D.uni.argx = A.push(S.rgx); D.uni.argy = A.push(S.rgy);
      }
    }
    BInst& I = b->hinst[k];
    I.lap.L1 = a_kron(A, P.L1);
    I.lap.L2 = a_kron(A, P.L2);
    std::vector<double> dinv(P.Adiag.size());
    for (size_t q = 0; q < dinv.size(); ++q) dinv[q] = 1.0 / P.Adiag[q];
    I.lap.dinv = A.push(dinv);
    I.lap.nx = P.Npx;
    I.lap.ny = P.Npy;
    const DLap& Lh = I.lap;
    const bool tri = Lh.L1.Y.K == 1 && Lh.L1.X.K == 3 && Lh.L2.Y.K == 3 && Lh.L2.X.K == 1;
    b->lap_tri = (k == 0 ? true : b->lap_tri) && tri;
    if (b->two_level) {
      FDMHost C;
      build_coarse_host(P, kAgg, C);
      if ((long)C.nx * C.ny > kBMaxCoarse) throw BErr(INS2D_E_UNSUPPORTED, "instance too large for the coarse solve");
      I.ncx = C.nx;
      I.ncy = C.ny;
      I.nx0 = C.nx0;
      I.ny0 = C.ny0;
      I.cVx = A.push(C.Vx);
      I.cVy = A.push(C.Vy);
      I.clx = A.push(C.lx);
      I.cly = A.push(C.ly);
    }
    if (b->fdm) {
      const FDMHost& F = b->fh[k];
      I.fVx = A.push(F.Vx);
      I.fVxT = A.push(F.VxT);
      I.fVy = A.push(F.Vy);
      I.fVyT = A.push(F.VyT);
      I.flx = A.push(F.lx);
      I.fly = A.push(F.ly);
      I.fnx0 = F.nx0;
      I.fny0 = F.ny0;
    }
  }
}

// the fast-diagonalization factors of every instance, on host threads
// the closed-form maps of every member (uniform_maps), on host threads
void build_uniform_maps(ins2d_batch* b) {
  b->ug.assign(b->n, UniForm{});
  b->ua.assign(b->n, UniArrays{});
  b->ufu.assign(b->n, {});
  b->ufv.assign(b->n, {});
  const int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> pool;
  std::vector<std::string> errs(nt);
  for (int w = 0; w < nt; ++w)
    pool.emplace_back([b, w, nt, &errs] {
      try {
        for (int k = w; k < b->n; k += nt) uniform_maps(*b->plans[k], b->ug[k], b->ua[k], b->ufu[k], b->ufv[k]);
      } catch (const std::exception& e) {
        errs[w] = e.what();
      }
    });
  for (auto& th : pool) th.join();
  for (auto& e : errs)
    if (!e.empty()) throw std::runtime_error("closed-form maps: " + e);
}

void build_fdm_factors(ins2d_batch* b) {
  b->fh.assign(b->n, FDMHost{});
  const int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> pool;
  std::vector<std::string> errs(nt);
  for (int w = 0; w < nt; ++w)
    pool.emplace_back([b, w, nt, &errs] {
      try {
        for (int k = w; k < b->n; k += nt) build_fdm_host(*b->plans[k], b->fh[k]);
      } catch (const std::exception& e) {
        errs[w] = e.what();
      }
    });
  for (auto& th : pool) th.join();
  for (auto& e : errs)
    if (!e.empty()) throw std::runtime_error("fast-diagonalization factors: " + e);
}

// z = A^-1 r for every active instance (four batched GEMM phases)
void fdm_apply(ins2d_batch* b, const double* r, double* z) {
  const dim3 g((unsigned)b->gt.size()), t(256);
  hipLaunchKernelGGL(k_bf_gemm<0>, g, t, 0, b->st, b->dinst, b->dgtx, b->cg, r, b->gw1);
  hipLaunchKernelGGL(k_bf_gemm<1>, g, t, 0, b->st, b->dinst, b->dgty, b->cg, b->gw1, b->gw2);
  hipLaunchKernelGGL(k_bf_gemm<2>, g, t, 0, b->st, b->dinst, b->dgtx, b->cg, b->gw2, b->gw1);
  hipLaunchKernelGGL(k_bf_gemm<3>, g, t, 0, b->st, b->dinst, b->dgty, b->cg, b->gw1, z);
}

// split tiles of member k (split_tiles for both components); false, with
// nothing appended, when either component has no closed-form rectangle
bool append_split_tiles(const BInst& I, int k, const std::vector<unsigned char>& fu,
                        const std::vector<unsigned char>& fv, int tx, int ty, std::vector<RTile>& out) {
  (void)ty;
  std::vector<RTile> add;
  if (!split_tiles(k, 0, I.nux, I.nuy, fu, tx, add) || !split_tiles(k, 1, I.nvx, I.nvy, fv, tx, add)) return false;
  out.insert(out.end(), add.begin(), add.end());
  return true;
}

// the right-hand-side tile lists for interior tiles of tx x ty points: split
// tiles of the members with a closed form, tx x ty tiles of the others
void build_rhs_lists(ins2d_batch* b, int tx, int ty, std::vector<RTile>& rs, std::vector<Tile>& rg);

void append_velocity_tiles(const BInst& I, int k, int tx, int ty, std::vector<Tile>& out) {
  for (int c = 0; c < 2; ++c) {
    const int nx = c == 0 ? I.nux : I.nvx, ny = c == 0 ? I.nuy : I.nvy;
    for (int j = 0; j < ny; j += ty)
      for (int i = 0; i < nx; i += tx) out.push_back(Tile{k, c, i, j});
  }
}

// the momentum right-hand side over a tile list of tx x ty threads (face
// sharing within rows when tx is 16, 32 or a multiple of 64)
void rhs_tiles(ins2d_batch* b, const Tile* tiles, size_t ntiles, int tx, int ty, const double* V, const double* P,
               int mode, double* out, const StageCombine& sc, double* vt) {
  const dim3 g((unsigned)ntiles), t(tx, ty);
  if (tx % 64 == 0)
    hipLaunchKernelGGL(k_b_rhs<64>, g, t, 0, b->st, b->dinst, tiles, V, P, mode, out, sc, vt);
  else
    hipLaunchKernelGGL(k_b_rhs<0>, g, t, 0, b->st, b->dinst, tiles, V, P, mode, out, sc, vt);
}

// Launch order of a tile list: blocks reach the 8 XCDs round-robin (block i
// on XCD i mod 8), so XCD x is given the contiguous chunk of n / 8 tiles
// (one more for x < n mod 8) starting at x (n / 8) + min(x, n mod 8):
// neighbouring tiles of a member, whose halos share cache lines, then run
// on one XCD.  A permutation of [begin, end).
template <class T>
void xcd_order(std::vector<T>& v, size_t begin, size_t end) {
  const size_t n = end - begin;
  if (n < 16) return;
  const size_t q = n / 8, r = n % 8;
  std::vector<T> out(n);
  for (size_t i = 0; i < n; ++i) {
    const size_t x = i & 7;
    out[i] = v[begin + x * q + std::min(x, r) + (i >> 3)];
  }
  std::copy(out.begin(), out.end(), v.begin() + begin);
}

void build_rhs_lists(ins2d_batch* b, int tx, int ty, std::vector<RTile>& rs, std::vector<Tile>& rg) {
  rs.clear();
  rg.clear();
  for (int k = 0; k < b->n; ++k)
    if (!append_split_tiles(b->hinst[k], k, b->ufu[k], b->ufv[k], tx, ty, rs))
      append_velocity_tiles(b->hinst[k], k, tx, ty, rg);
  const size_t nc = (size_t)order_split_tiles(rs);
  xcd_order(rs, 0, nc);
  xcd_order(rs, nc, rs.size());
  xcd_order(rg, 0, rg.size());
}

size_t closed_count(const std::vector<RTile>& rs) {
  size_t n = 0;
  for (const RTile& r : rs) n += (r.comp & 2) ? 1 : 0;
  return n;
}

void rhs_lists_launch(ins2d_batch* b, const RTile* drs, size_t nclosed, size_t nrs, const Tile* drg, size_t nrg,
                      int tx, int ty, const double* V, const double* P, int mode, double* out, const StageCombine& sc,
                      double* vt) {
  if (nclosed)
    hipLaunchKernelGGL(k_b_rhs_closed, dim3((unsigned)nclosed), dim3(256), 0, b->st, b->dinst, drs, V, P, mode, out,
                       sc, vt);
  if (nrs > nclosed)
    hipLaunchKernelGGL(k_b_rhs_split, dim3((unsigned)(nrs - nclosed)), dim3(256), 0, b->st, b->dinst, drs + nclosed, V,
                       P, mode, out, sc, vt);
  if (nrg) rhs_tiles(b, drg, nrg, tx, ty, V, P, mode, out, sc, vt);
}

// upload rs and rg into the device lists (grown as needed)
void upload_rhs_lists(ins2d_batch* b, std::vector<RTile>& rs, std::vector<Tile>& rg) {
  drop_step_graphs(b);
  if (!b->drs || rs.size() > b->rs_cap) {
    b->drs = balloc<RTile>(b, rs.size());
    b->rs_cap = rs.size();
  }
  if (!b->drg || rg.size() > b->rg_cap) {
    b->drg = balloc<Tile>(b, rg.size());
    b->rg_cap = rg.size();
  }
  if (!rs.empty()) BHIP(hipMemcpy(b->drs, rs.data(), rs.size() * sizeof(RTile), hipMemcpyHostToDevice));
  if (!rg.empty()) BHIP(hipMemcpy(b->drg, rg.data(), rg.size() * sizeof(Tile), hipMemcpyHostToDevice));
  b->rs.swap(rs);
  b->rg.swap(rg);
  b->nclosed = closed_count(b->rs);
}

void launch_tiles_rhs(ins2d_batch* b, const double* V, const double* P, int mode, double* out, const StageCombine& sc,
                      double* vt) {
  rhs_lists_launch(b, b->drs, b->closed_ok ? b->nclosed : 0, b->rs.size(), b->drg, b->rg.size(), b->tx, b->ty, V, P,
                   mode, out, sc, vt);
}
void launch_tiles_div(ins2d_batch* b, const double* A, const double* B, const double* dt, double ca, double cb,
                      double yc, double os, double* f) {
  const dim3 g((unsigned)b->pt_tall.size()), t(TX, TY);
  if (B)
    hipLaunchKernelGGL(k_b_div<true>, g, t, 0, b->st, b->dinst, b->dpt_tall, A, B, dt, ca, cb, yc, os, f);
  else
    hipLaunchKernelGGL(k_b_div<false>, g, t, 0, b->st, b->dinst, b->dpt_tall, A, B, dt, ca, cb, yc, os, f);
}
void launch_tiles_proj(ins2d_batch* b, const double* Vn, const double* Vt, const double* dp, const double* dt,
                       double c, double* V) {
  const dim3 g((unsigned)b->uvt_tall.size()), t(TX, TY);
  if (Vt)
    hipLaunchKernelGGL(k_b_proj<true>, g, t, 0, b->st, b->dinst, b->duv_tall, Vn, Vt, dp, dt, c, V);
  else
    hipLaunchKernelGGL(k_b_proj<false>, g, t, 0, b->st, b->dinst, b->duv_tall, Vn, Vt, dp, dt, c, V);
}

// fast-diagonalization PCG iteration, head: q = A p, alpha, x and r, the
// convergence test (converged instances stop here)
// q = A p with partials p.q: over tall tiles when every instance's Laplacian
// has the widths of lap_tri_at, else over the pressure tiles (g, tiles)
void launch_cg_apply(ins2d_batch* b, dim3 g, dim3 t, unsigned sh, hipStream_t st, const BInst* I, const Tile* tiles,
                     BCG S, const double* p, double* q) {
  if (b->lap_tri)
    hipLaunchKernelGGL(k_b_cg_apply_tall, dim3((unsigned)b->pt_tall.size()), t, sh, st, I, b->dpt_tall, S, p, q);
  else hipLaunchKernelGGL(k_b_cg_apply, g, t, sh, st, I, tiles, S, p, q);
}
// partials r.z over tall tiles
void launch_bf_rz(ins2d_batch* b, const double* r, const double* z) {
  hipLaunchKernelGGL(k_bf_rz_tall, dim3((unsigned)b->pt_tall.size()), dim3(TX, TY), 0, b->st, b->dinst, b->dpt_tall,
                     b->cg, r, z);
}

void fdm_iter_head(ins2d_batch* b) {
  const unsigned np = (unsigned)b->pt.size();
  launch_cg_apply(b, dim3(np), dim3(TX, TY), 0, b->st, b->dinst, b->dpt, b->cg, b->pp, b->q);
  hipLaunchKernelGGL(k_b_cg_alpha, dim3(b->n), dim3(256), 0, b->st, b->dinst, b->cg);
  hipLaunchKernelGGL(k_bf_update, dim3(np), dim3(TX, TY), 0, b->st, b->dinst, b->dpt, b->cg, b->pp, b->q, b->x,
                     b->r);
  hipLaunchKernelGGL(k_bf_test, dim3(b->n), dim3(256), 0, b->st, b->dinst, b->cg);
}

// tail: z = A^-1 r for the instances still iterating, r.z, beta, direction
void fdm_iter_tail(ins2d_batch* b) {
  const unsigned np = (unsigned)b->pt.size();
  fdm_apply(b, b->r, b->z);
  launch_bf_rz(b, b->r, b->z);
  hipLaunchKernelGGL(k_bf_beta, dim3(b->n), dim3(256), 0, b->st, b->dinst, b->cg);
  hipLaunchKernelGGL(k_b_cg_dir, dim3(np), dim3(TX, TY), 0, b->st, b->dinst, b->dpt, b->cg, b->z, b->pp);
}

// one iteration of the Jacobi or two-level PCG (replayed in chunks)
void cg_iter_launches(ins2d_batch* b) {
  if (b->two_level) {
    const unsigned na = (unsigned)b->at.size();
    const dim3 tb(kAgg, kAgg);
    if (b->lap_tri) hipLaunchKernelGGL(k_b2_apply<true>, dim3(na), tb, 0, b->st, b->dinst, b->dat, b->cg, b->tl, b->pp, b->q);
    else hipLaunchKernelGGL(k_b2_apply<false>, dim3(na), tb, 0, b->st, b->dinst, b->dat, b->cg, b->tl, b->pp, b->q);
    hipLaunchKernelGGL(k_b2_alpha, dim3(b->n), dim3(256), 0, b->st, b->dinst, b->cg, b->tl);
    hipLaunchKernelGGL(k_b2_update, dim3(na), tb, 0, b->st, b->dinst, b->dat, b->cg, b->tl, b->pp, b->q, b->x, b->r);
    hipLaunchKernelGGL(k_b2_coarse, dim3(b->n), dim3(256), 0, b->st, b->dinst, b->cg, b->tl, 0, 0.0);
    hipLaunchKernelGGL(k_b2_dir, dim3(na), tb, 0, b->st, b->dinst, b->dat, b->cg, b->tl, b->r, b->pp);
    return;
  }
  const unsigned np = (unsigned)b->pt.size();
  launch_cg_apply(b, dim3(np), dim3(TX, TY), 0, b->st, b->dinst, b->dpt, b->cg, b->pp, b->q);
  hipLaunchKernelGGL(k_b_cg_alpha, dim3(b->n), dim3(256), 0, b->st, b->dinst, b->cg);
  hipLaunchKernelGGL(k_b_cg_update, dim3(np), dim3(TX, TY), 0, b->st, b->dinst, b->dpt, b->cg, b->pp, b->q, b->x,
                     b->r, b->z);
  hipLaunchKernelGGL(k_b_cg_beta, dim3(b->n), dim3(256), 0, b->st, b->dinst, b->cg);
  hipLaunchKernelGGL(k_b_cg_dir, dim3(np), dim3(TX, TY), 0, b->st, b->dinst, b->dpt, b->cg, b->z, b->pp);
}

// batched A dp = f for every instance (x0 = 0); result in b->x
void b_poisson(ins2d_batch* b, const double* f) {
  const unsigned np = (unsigned)b->pt.size();
  hipLaunchKernelGGL(k_b_set_int, dim3(1), dim3(1), 0, b->st, b->cg.nactive, b->n);
  if (b->fdm) {
    // flags cleared first: the GEMM phases skip finished instances
    hipLaunchKernelGGL(k_b_clear_flags, dim3((b->n + 255) / 256), dim3(256), 0, b->st, b->cg.f, b->n);
    hipLaunchKernelGGL(k_bf_init, dim3(np), dim3(TX, TY), 0, b->st, b->dinst, b->dpt, b->cg, f, b->x, b->r);
    fdm_apply(b, b->r, b->z);
    launch_bf_rz(b, b->r, b->z);
    hipLaunchKernelGGL(k_b_cg_init_fin, dim3(b->n), dim3(256), 0, b->st, b->dinst, b->cg, b->tol * b->tol);
    hipLaunchKernelGGL(k_bf_dir0, dim3(np), dim3(TX, TY), 0, b->st, b->dinst, b->dpt, b->cg, b->z, b->pp);
  } else if (b->two_level) {
    const unsigned na = (unsigned)b->at.size();
    const dim3 tb(kAgg, kAgg);
    // per-instance flags cleared first: k_b2_coarse skips finished instances
    hipLaunchKernelGGL(k_b_clear_flags, dim3((b->n + 255) / 256), dim3(256), 0, b->st, b->cg.f, b->n);
    hipLaunchKernelGGL(k_b2_init, dim3(na), tb, 0, b->st, b->dinst, b->dat, b->cg, b->tl, f, b->x, b->r);
    hipLaunchKernelGGL(k_b2_coarse, dim3(b->n), dim3(256), 0, b->st, b->dinst, b->cg, b->tl, 1, b->tol * b->tol);
    hipLaunchKernelGGL(k_b2_dir, dim3(na), tb, 0, b->st, b->dinst, b->dat, b->cg, b->tl, b->r, b->pp);
  } else {
    hipLaunchKernelGGL(k_b_cg_init, dim3(np), dim3(TX, TY), 0, b->st, b->dinst, b->dpt, b->cg, f, b->x, b->r, b->z,
                       b->pp);
    hipLaunchKernelGGL(k_b_cg_init_fin, dim3(b->n), dim3(256), 0, b->st, b->dinst, b->cg, b->tol * b->tol);
  }
  if (b->fdm && b->device_iter) {
    // a fixed number of iterations on the device (the exact inverse as
    // preconditioner meets the stopping rule in one): instances that have
    // converged skip the rest; no host check, so the solve can be captured
    for (int it = 0; it < b->fixed_iters; ++it) {
      fdm_iter_head(b);
      if (it + 1 < b->fixed_iters) fdm_iter_tail(b);
    }
    hipLaunchKernelGGL(k_bf_stats, dim3((b->n + 255) / 256), dim3(256), 0, b->st, b->cg, b->n, b->d_it_sum,
                       b->d_it_max, b->d_it_max + 1);
    return;
  }
  int active = b->n;
  if (b->fdm) {
    // with the exact inverse as preconditioner one iteration is the rule:
    // the host checks after the head of each iteration, so a converged
    // solve ends before the tail (whose launches would find every instance
    // done); the tail runs only for another iteration
    if (!b->cg_chunk) {
      BHIP(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
      fdm_iter_head(b);
      BHIP(hipStreamEndCapture(b->st, &b->cg_graph));
      BHIP(hipGraphInstantiate(&b->cg_chunk, b->cg_graph, nullptr, nullptr, 0));
      BHIP(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
      fdm_iter_tail(b);
      BHIP(hipStreamEndCapture(b->st, &b->cg_tail_graph));
      BHIP(hipGraphInstantiate(&b->cg_tail, b->cg_tail_graph, nullptr, nullptr, 0));
    }
    while (true) {
      BHIP(hipGraphLaunch(b->cg_chunk, b->st));
      BHIP(hipMemcpyAsync(&active, b->cg.nactive, sizeof(int), hipMemcpyDeviceToHost, b->st));
      BHIP(hipStreamSynchronize(b->st));
      if (active <= 0) break;
      BHIP(hipGraphLaunch(b->cg_tail, b->st));
    }
  } else {
    if (!b->cg_chunk) {
      BHIP(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
      for (int c = 0; c < b->chunk; ++c) cg_iter_launches(b);
      BHIP(hipStreamEndCapture(b->st, &b->cg_graph));
      BHIP(hipGraphInstantiate(&b->cg_chunk, b->cg_graph, nullptr, nullptr, 0));
    }
    while (true) {
      BHIP(hipGraphLaunch(b->cg_chunk, b->st));
      BHIP(hipMemcpyAsync(&active, b->cg.nactive, sizeof(int), hipMemcpyDeviceToHost, b->st));
      BHIP(hipStreamSynchronize(b->st));
      if (active <= 0) break;
    }
  }
  std::vector<int> fl(4 * (size_t)b->n);
  BHIP(hipMemcpy(fl.data(), b->cg.f, fl.size() * sizeof(int), hipMemcpyDeviceToHost));
  b->last_it.assign(b->n, 0);
  for (int k = 0; k < b->n; ++k) {
    b->last_it[k] = fl[4 * k + 1];
    b->total_it += fl[4 * k + 1];
    b->max_it = std::max(b->max_it, fl[4 * k + 1]);
    if (fl[4 * k + 2]) throw BErr(INS2D_E_NOCONV, "CG did not converge in maxit iterations (instance " + std::to_string(k) + ")");
  }
}

void b_additional(ins2d_batch* b, const double* V) {
  launch_tiles_rhs(b, V, b->P, RHS_WITH_P | RHS_SCALE_OM, b->tV, StageCombine{0, 0.0, {0}, {nullptr}}, nullptr);
  launch_tiles_div(b, b->tV, nullptr, nullptr, 1.0, 0.0, 0.0, 1.0, b->F);
  b_poisson(b, b->F);
  launch_axpy(b->st, (size_t)b->Npt, 1.0, b->x, b->P);
}

void b_step(ins2d_batch* b) {
  double* Vn = b->V;
  double* Vs = b->Vb;
  for (int i = 0; i < b->s; ++i) {
    const double* Vsrc = i == 0 ? Vn : Vs;
    StageCombine sc{};
    sc.aself = b->At[i * b->s + i];
    for (int j = 0; j < i; ++j) {
      const double a = b->At[i * b->s + j];
      if (a != 0.0) { sc.a[sc.n] = a; sc.k[sc.n] = b->kv[j]; ++sc.n; }
    }
    launch_tiles_rhs(b, Vsrc, b->P, RHS_SCALE_OM, b->keep[i] ? b->kv[i] : nullptr, sc, b->Vt);
    const double c = b->ct[i];
    const double* dpp = b->P;
    if (i > 0) {
      launch_tiles_div(b, Vn, b->Vt, b->dt, 0.0, 1.0, 0.0, 1.0 / c, b->F);
      b_poisson(b, b->F);
      dpp = b->x;
    }
    launch_tiles_proj(b, Vn, b->Vt, dpp, b->dt, c, Vs);
  }
  std::swap(b->V, b->Vb);
  // steady boundary values: always the additional solve (time_ERK.m:160-163)
  b_additional(b, b->V);
}

int bfail(ins2d_batch* b, const std::exception& e, int code) {
  if (b) b->err = e.what();
  return code;
}

#define B_TRY try {
#define B_CATCH(b)                                                  \
  }                                                                 \
  catch (const BErr& e) { return bfail(b, e, e.code); }             \
  catch (const std::exception& e) { return bfail(b, e, INS2D_E_ARG); }

std::string g_batch_err;

}  // namespace

extern "C" {

const char* ins2d_batch_last_error(const ins2d_batch* b) { return b ? b->err.c_str() : g_batch_err.c_str(); }

int ins2d_batch_create(const ins2d_case_desc* descs, int n, ins2d_batch** out) {
  if (!descs || n < 1 || !out) return INS2D_E_ARG;
  *out = nullptr;
  auto b = std::make_unique<ins2d_batch>();
  try {
    b->n = n;
    const ins2d_case_desc& d0 = descs[0];
    if (d0.rk_stages < 1 || d0.rk_stages > 8) throw BErr(INS2D_E_ARG, "bad tableau");
    b->s = d0.rk_stages;
    b->tol = d0.cg_tol;
    b->maxit = d0.cg_maxit;
    for (int k = 0; k < n; ++k) {
      const ins2d_case_desc& d = descs[k];
      if (d.order4) throw BErr(INS2D_E_UNSUPPORTED, "the ensemble runs the second-order operators");
      if (d.poisson == INS2D_POISSON_FFT) throw BErr(INS2D_E_UNSUPPORTED, "the ensemble solves the pressure by PCG");
      if (d.rk_stages != b->s) throw BErr(INS2D_E_ARG, "all instances share one tableau");
      for (int q = 0; q < b->s * b->s; ++q)
        if (d.rk_A[q] != d0.rk_A[q]) throw BErr(INS2D_E_ARG, "all instances share one tableau");
      CaseDesc cd;
      cd.Nx = d.Nx; cd.Ny = d.Ny;
      cd.x.assign(d.x, d.x + d.Nx + 1);
      cd.y.assign(d.y, d.y + d.Ny + 1);
      cd.u_left = d.bc_u[0]; cd.u_right = d.bc_u[1]; cd.u_low = d.bc_u[2]; cd.u_up = d.bc_u[3];
      cd.v_left = d.bc_v[0]; cd.v_right = d.bc_v[1]; cd.v_low = d.bc_v[2]; cd.v_up = d.bc_v[3];
      cd.Re = d.Re;
      b->plans.emplace_back(new Plan());
      build_plan(cd, *b->plans.back());
    }
    if (d0.cg_precond == 2) b->two_level = true;
    else if (d0.cg_precond == 3) b->fdm = true;
    else if (d0.cg_precond != 0)
      throw BErr(INS2D_E_UNSUPPORTED, "the ensemble runs Jacobi, two-level or fast-diagonalization PCG");
    if (b->fdm) build_fdm_factors(b.get());
    build_uniform_maps(b.get());
    BHIP(hipStreamCreate(&b->st));
    b->hcomp.assign(2 * (size_t)n, DComp{});
    b->hinst.assign(n, BInst{});
    b->samp.resize(n);
    // operator arena: size, allocate, fill, upload
    Arena A;
    build_ops(b.get(), A);
    A.cap = A.used + 256;
    A.counting = false;
    A.used = 0;
    A.host.assign(A.cap, 0);
    BHIP(hipMalloc((void**)&A.dev, A.cap));
    b->allocs.push_back(A.dev);
    build_ops(b.get(), A);
    BHIP(hipMemcpy(A.dev, A.host.data(), A.cap, hipMemcpyHostToDevice));
    // layout and tiles
    long voff = 0, poff = 0, coff = 0;
    for (int k = 0; k < n; ++k) {
      const Plan& P = *b->plans[k];
      BInst& I = b->hinst[k];
      I.nux = P.Nux_in; I.nuy = P.Nuy_in; I.nvx = P.Nvx_in; I.nvy = P.Nvy_in; I.npx = P.Npx; I.npy = P.Npy;
      I.nu = P.Nu;
      I.voff = voff;
      I.poff = poff;
      voff += P.NV;
      poff += P.Np;
      append_velocity_tiles(I, k, TX, kTallRows, b->uvt_tall);
      for (int j = 0; j < I.npy; j += kTallRows)
        for (int i = 0; i < I.npx; i += TX) b->pt_tall.push_back(Tile{k, 2, i, j});
      I.pt0 = (int)b->pt.size();
      for (int j = 0; j < I.npy; j += TY)
        for (int i = 0; i < I.npx; i += TX) b->pt.push_back(Tile{k, 2, i, j});
      I.pt1 = (int)b->pt.size();
      I.at0 = (int)b->at.size();
      I.coff = coff;
      if (b->two_level) {
        for (int j = 0; j < I.npy; j += kAgg)
          for (int i = 0; i < I.npx; i += kAgg) b->at.push_back(Tile{k, 2, i, j});
        coff += (long)I.ncx * I.ncy;
      }
      I.at1 = (int)b->at.size();
      if (b->fdm)
        for (int j = 0; j < I.npy; j += kG64)
          for (int i = 0; i < I.npx; i += kG64) b->gt.push_back(Tile{k, 2, i, j});
      // samples default to zero
      BCSamples& S = b->samp[k];
      for (int s = 0; s < S_NSRC; ++s) S.s[s].assign(sample_len(P, s), 0.0);
    }
    b->NVt = voff;
    b->Npt = poff;
    b->dcomp = balloc<DComp>(b.get(), 2 * (size_t)n);
    for (int k = 0; k < n; ++k) {
      b->hinst[k].cu = b->dcomp + 2 * k;
      b->hinst[k].cv = b->dcomp + 2 * k + 1;
    }
    b_upload_bc(b.get());
    b->dinst = balloc<BInst>(b.get(), n);
    BHIP(hipMemcpy(b->dinst, b->hinst.data(), n * sizeof(BInst), hipMemcpyHostToDevice));
    {
      std::vector<RTile> rs;
      std::vector<Tile> rg;
      build_rhs_lists(b.get(), b->tx, b->ty, rs, rg);
      upload_rhs_lists(b.get(), rs, rg);
    }
    b->dpt = balloc<Tile>(b.get(), b->pt.size());
    BHIP(hipMemcpy(b->dpt, b->pt.data(), b->pt.size() * sizeof(Tile), hipMemcpyHostToDevice));
    xcd_order(b->uvt_tall, 0, b->uvt_tall.size());
    xcd_order(b->pt_tall, 0, b->pt_tall.size());
    b->duv_tall = balloc<Tile>(b.get(), b->uvt_tall.size());
    BHIP(hipMemcpy(b->duv_tall, b->uvt_tall.data(), b->uvt_tall.size() * sizeof(Tile), hipMemcpyHostToDevice));
    b->dpt_tall = balloc<Tile>(b.get(), b->pt_tall.size());
    BHIP(hipMemcpy(b->dpt_tall, b->pt_tall.data(), b->pt_tall.size() * sizeof(Tile), hipMemcpyHostToDevice));
    if (b->two_level) {
      b->dat = balloc<Tile>(b.get(), b->at.size());
      BHIP(hipMemcpy(b->dat, b->at.data(), b->at.size() * sizeof(Tile), hipMemcpyHostToDevice));
      b->tl.part = balloc<double>(b.get(), 3 * b->at.size());
      b->tl.rc = balloc<double>(b.get(), (size_t)coff);
      b->tl.ec = balloc<double>(b.get(), (size_t)coff);
    }
    if (b->fdm) {
      for (int ph = 0; ph < 2; ++ph) {
        std::vector<Tile> o = b->gt;
        std::stable_sort(o.begin(), o.end(), [&](const Tile& x, const Tile& y) {
          const BInst &X = b->hinst[x.inst], &Y = b->hinst[y.inst];
          return ph == 0 ? X.npx > Y.npx : X.npy > Y.npy;
        });
        Tile*& d = ph == 0 ? b->dgtx : b->dgty;
        d = balloc<Tile>(b.get(), o.size());
        BHIP(hipMemcpy(d, o.data(), o.size() * sizeof(Tile), hipMemcpyHostToDevice));
      }
      b->gw1 = balloc<double>(b.get(), b->Npt);
      b->gw2 = balloc<double>(b.get(), b->Npt);
    }
    // tableau (time_ERK.m:29,32)
    const int s = b->s;
    b->At.assign((size_t)s * s, 0.0);
    b->ct.assign(s, 0.0);
    for (int i = 0; i < s; ++i) {
      for (int j = 0; j < s; ++j) b->At[i * s + j] = (i < s - 1) ? d0.rk_A[(i + 1) * s + j] : d0.rk_b[j];
      b->ct[i] = (i < s - 1) ? d0.rk_c[i + 1] : 1.0;
    }
    b->keep.assign(s, false);
    for (int j = 0; j < s; ++j)
      for (int i = j + 1; i < s; ++i)
        if (b->At[i * s + j] != 0.0) b->keep[j] = true;
    // state and work buffers
    b->V = balloc<double>(b.get(), b->NVt);
    b->Vb = balloc<double>(b.get(), b->NVt);
    b->Vt = balloc<double>(b.get(), b->NVt);
    b->tV = balloc<double>(b.get(), b->NVt);
    b->kv.assign(s, nullptr);
    for (int j = 0; j < s; ++j)
      if (b->keep[j]) b->kv[j] = balloc<double>(b.get(), b->NVt);
    b->P = balloc<double>(b.get(), b->Npt);
    b->F = balloc<double>(b.get(), b->Npt);
    b->x = balloc<double>(b.get(), b->Npt);
    b->r = balloc<double>(b.get(), b->Npt);
    b->z = balloc<double>(b.get(), b->Npt);
    b->pp = balloc<double>(b.get(), b->Npt);
    b->q = balloc<double>(b.get(), b->Npt);
    b->dt = balloc<double>(b.get(), n);
    b->cg.s = balloc<double>(b.get(), 8 * (size_t)n);
    b->cg.f = balloc<int>(b.get(), 4 * (size_t)n);
    b->cg.part = balloc<double>(b.get(), 2 * b->pt.size());
    b->cg.nactive = balloc<int>(b.get(), 1);
    b->cg.maxit = b->maxit;
    b->d_it_sum = balloc<unsigned long long>(b.get(), 1);
    b->d_it_max = balloc<int>(b.get(), 2);
    BHIP(hipMemset(b->d_it_sum, 0, sizeof(unsigned long long)));
    BHIP(hipMemset(b->d_it_max, 0, 2 * sizeof(int)));
    b->fixed_iters = 1;
    if (b->fdm) {
      b->Vsave = balloc<double>(b.get(), b->NVt);
      b->Psave = balloc<double>(b.get(), b->Npt);
    }
    BHIP(hipMemset(b->V, 0, sizeof(double) * b->NVt));
    BHIP(hipMemset(b->P, 0, sizeof(double) * b->Npt));
    BHIP(hipDeviceSynchronize());
  } catch (const BErr& e) {
    g_batch_err = e.what();
    for (void* p : b->allocs) (void)hipFree(p);
    for (void* p : b->yv_allocs) (void)hipFree(p);
    b->allocs.clear();
    b->yv_allocs.clear();
    if (b->st) (void)hipStreamDestroy(b->st);
    b->st = nullptr;
    return e.code;
  } catch (const std::exception& e) {
    g_batch_err = e.what();
    for (void* p : b->allocs) (void)hipFree(p);
    for (void* p : b->yv_allocs) (void)hipFree(p);
    b->allocs.clear();
    b->yv_allocs.clear();
    if (b->st) (void)hipStreamDestroy(b->st);
    b->st = nullptr;
    return INS2D_E_ARG;
  }
  *out = b.release();
  return INS2D_OK;
}

// time the stage kernel (the momentum right-hand side over every velocity
// point) for a few interior tile shapes on the resident state and keep the
// fastest (closed-form interior tiles of the members with a closed form,
// generic tiles of the others)
int ins2d_batch_autotune(ins2d_batch* b, int* tx, int* ty) {
  if (!b) return INS2D_E_ARG;
  B_TRY
  // 256 points each (the kernels are compiled for at most 256 threads)
  static const int cand[][2] = {{64, 4}, {32, 8}, {128, 2}, {16, 16}, {256, 1}};
  hipEvent_t e0, e1;
  BHIP(hipEventCreate(&e0));
  BHIP(hipEventCreate(&e1));
  float best = 1e30f;
  int bx = b->tx, by = b->ty;
  const StageCombine none{0, 0.0, {0}, {nullptr}};
  for (const auto& c : cand) {
    std::vector<RTile> rs;
    std::vector<Tile> rg;
    build_rhs_lists(b, c[0], c[1], rs, rg);
    RTile* d1 = nullptr;
    Tile* d2 = nullptr;
    BHIP(hipMalloc(&d1, std::max<size_t>(rs.size(), 1) * sizeof(RTile)));
    BHIP(hipMalloc(&d2, std::max<size_t>(rg.size(), 1) * sizeof(Tile)));
    if (!rs.empty()) BHIP(hipMemcpy(d1, rs.data(), rs.size() * sizeof(RTile), hipMemcpyHostToDevice));
    if (!rg.empty()) BHIP(hipMemcpy(d2, rg.data(), rg.size() * sizeof(Tile), hipMemcpyHostToDevice));
    float tot = 0.f;
    for (int rep = 0; rep < 4; ++rep) {
      BHIP(hipEventRecord(e0, b->st));
      rhs_lists_launch(b, d1, b->closed_ok ? closed_count(rs) : 0, rs.size(), d2, rg.size(), c[0], c[1], b->V, b->P,
                       RHS_SCALE_OM, b->tV, none, nullptr);
      BHIP(hipEventRecord(e1, b->st));
      BHIP(hipEventSynchronize(e1));
      float ms = 0.f;
      BHIP(hipEventElapsedTime(&ms, e0, e1));
      if (rep > 0) tot += ms;   // the first launch warms caches
    }
    BHIP(hipFree(d1));
    BHIP(hipFree(d2));
    if (tot < best) {
      best = tot;
      bx = c[0];
      by = c[1];
    }
  }
  BHIP(hipEventDestroy(e0));
  BHIP(hipEventDestroy(e1));
  std::vector<RTile> rs;
  std::vector<Tile> rg;
  build_rhs_lists(b, bx, by, rs, rg);
  upload_rhs_lists(b, rs, rg);
  b->tx = bx;
  b->ty = by;
  if (tx) *tx = bx;
  if (ty) *ty = by;
  return INS2D_OK;
  B_CATCH(b)
}

int ins2d_batch_set_tile(ins2d_batch* b, int tx, int ty) {
  if (!b || tx * ty != 256 || !(tx == 16 || tx == 32 || tx == 64 || tx == 128 || tx == 256)) return INS2D_E_ARG;
  B_TRY
  BHIP(hipStreamSynchronize(b->st));
  std::vector<RTile> rs;
  std::vector<Tile> rg;
  build_rhs_lists(b, tx, ty, rs, rg);
  upload_rhs_lists(b, rs, rg);
  b->tx = tx;
  b->ty = ty;
  return INS2D_OK;
  B_CATCH(b)
}

void ins2d_batch_destroy(ins2d_batch* b) {
  if (!b) return;
  if (b->st) (void)hipStreamSynchronize(b->st);
  if (b->cg_chunk) (void)hipGraphExecDestroy(b->cg_chunk);
  if (b->cg_graph) (void)hipGraphDestroy(b->cg_graph);
  if (b->cg_tail) (void)hipGraphExecDestroy(b->cg_tail);
  if (b->cg_tail_graph) (void)hipGraphDestroy(b->cg_tail_graph);
  drop_step_graphs(b);
  for (void* p : b->allocs) (void)hipFree(p);
  for (void* p : b->yv_allocs) (void)hipFree(p);
  if (b->st) (void)hipStreamDestroy(b->st);
  delete b;
}

int ins2d_batch_info(const ins2d_batch* b, int k, int* sizes, double* xin, double* yin) {
  if (!b || k < 0 || k >= b->n || !sizes) return INS2D_E_ARG;
  const Plan& P = *b->plans[k];
  sizes[0] = P.Nux_in; sizes[1] = P.Nuy_in; sizes[2] = P.Nvx_in; sizes[3] = P.Nvy_in; sizes[4] = P.Npx;
  sizes[5] = P.Npy;
  if (xin) std::copy(P.xin.begin(), P.xin.end(), xin);
  if (yin) std::copy(P.yin.begin(), P.yin.end(), yin);
  return INS2D_OK;
}

int ins2d_batch_set_bc(ins2d_batch* b, int k, double t, const double* const* samples) {
  if (!b || k < 0 || k >= b->n || !samples) return INS2D_E_ARG;
  (void)t;
  B_TRY
  BCSamples& S = b->samp[k];
  for (int s = 0; s < S_NSRC; ++s) {
    if (!samples[s]) throw BErr(INS2D_E_ARG, "missing boundary sample");
    std::copy(samples[s], samples[s] + S.s[s].size(), S.s[s].begin());
  }
  b->bc_dirty = true;
  return INS2D_OK;
  B_CATCH(b)
}

int ins2d_batch_set_state(ins2d_batch* b, int k, const double* V, const double* p) {
  if (!b || k < 0 || k >= b->n) return INS2D_E_ARG;
  B_TRY
  const BInst& I = b->hinst[k];
  const Plan& P = *b->plans[k];
  if (V) BHIP(hipMemcpy(b->V + I.voff, V, sizeof(double) * P.NV, hipMemcpyHostToDevice));
  if (p) BHIP(hipMemcpy(b->P + I.poff, p, sizeof(double) * P.Np, hipMemcpyHostToDevice));
  return INS2D_OK;
  B_CATCH(b)
}

int ins2d_batch_get_state(ins2d_batch* b, int k, double* V, double* p) {
  if (!b || k < 0 || k >= b->n) return INS2D_E_ARG;
  B_TRY
  BHIP(hipStreamSynchronize(b->st));
  const BInst& I = b->hinst[k];
  const Plan& P = *b->plans[k];
  if (V) BHIP(hipMemcpy(V, b->V + I.voff, sizeof(double) * P.NV, hipMemcpyDeviceToHost));
  if (p) BHIP(hipMemcpy(p, b->P + I.poff, sizeof(double) * P.Np, hipMemcpyDeviceToHost));
  return INS2D_OK;
  B_CATCH(b)
}

int ins2d_batch_input_check(ins2d_batch* b, int p_initial) {
  if (!b) return INS2D_E_ARG;
  B_TRY
  BHIP(hipStreamSynchronize(b->st));
  b_upload_bc(b);
  b->bc_dirty = false;
  // initial projection of every instance (input_check.m:104-107); an
  // instance that is already divergence-free gets dp ~ 0
  launch_tiles_div(b, b->V, nullptr, nullptr, 1.0, 0.0, 1.0, 1.0, b->F);
  b_poisson(b, b->F);
  BHIP(hipMemcpyAsync(b->Vb, b->V, sizeof(double) * b->NVt, hipMemcpyDeviceToDevice, b->st));
  launch_tiles_proj(b, b->Vb, nullptr, b->x, nullptr, 1.0, b->V);
  if (p_initial) b_additional(b, b->V);
  BHIP(hipStreamSynchronize(b->st));
  return INS2D_OK;
  B_CATCH(b)
}

int ins2d_batch_step(ins2d_batch* b, const double* dt, int nsteps) {
  if (!b || !dt || nsteps < 0) return INS2D_E_ARG;
  B_TRY
  for (int k = 0; k < b->n; ++k)
    if (!(dt[k] > 0)) throw BErr(INS2D_E_ARG, "time steps must be positive");
  if (b->bc_dirty) {
    BHIP(hipStreamSynchronize(b->st));
    b_upload_bc(b);
    b->bc_dirty = false;
  }
  BHIP(hipMemcpyAsync(b->dt, dt, sizeof(double) * b->n, hipMemcpyHostToDevice, b->st));
  if (!b->fdm) {
    for (int n = 0; n < nsteps; ++n) b_step(b);
    BHIP(hipStreamSynchronize(b->st));
    return INS2D_OK;
  }
  // fast-diagonalization PCG: each step is one graph replay (captured on
  // first use for each parity of the state buffers) with one PCG iteration
  // per solve on the device, and one host wait at the end.  Should any solve
  // miss the stopping rule in that iteration, the call starts over from the
  // saved state on the host-checked path (iterations up to cg_maxit).
  double* V_start = b->V;
  BHIP(hipMemcpyAsync(b->Vsave, b->V, sizeof(double) * b->NVt, hipMemcpyDeviceToDevice, b->st));
  BHIP(hipMemcpyAsync(b->Psave, b->P, sizeof(double) * b->Npt, hipMemcpyDeviceToDevice, b->st));
  for (int n = 0; n < nsteps; ++n) {
    int par = -1;
    for (int k = 0; k < 2; ++k)
      if (b->step_exec[k] && b->step_V[k] == b->V) par = k;
    if (par < 0) {
      par = b->step_exec[0] ? 1 : 0;
      b->device_iter = true;
      BHIP(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
      double* V0 = b->V;
      b_step(b);   // swaps V and Vb
      std::swap(b->V, b->Vb);
      BHIP(hipStreamEndCapture(b->st, &b->step_graph[par]));
      b->device_iter = false;
      BHIP(hipGraphInstantiate(&b->step_exec[par], b->step_graph[par], nullptr, nullptr, 0));
      b->step_V[par] = V0;
    }
    BHIP(hipGraphLaunch(b->step_exec[par], b->st));
    std::swap(b->V, b->Vb);
  }
  unsigned long long sum = 0;
  int mx[2] = {0, 0};
  BHIP(hipMemcpyAsync(&sum, b->d_it_sum, sizeof(sum), hipMemcpyDeviceToHost, b->st));
  BHIP(hipMemcpyAsync(mx, b->d_it_max, sizeof(mx), hipMemcpyDeviceToHost, b->st));
  BHIP(hipStreamSynchronize(b->st));
  if (!mx[1]) {
    b->total_it += (long long)(sum - b->h_it_sum);
    b->h_it_sum = sum;
    b->max_it = std::max(b->max_it, mx[0]);
    return INS2D_OK;
  }
  // fallback: the counts of the discarded attempt are dropped
  b->h_it_sum = sum;
  BHIP(hipMemsetAsync(b->d_it_max + 1, 0, sizeof(int), b->st));
  if (b->V != V_start) std::swap(b->V, b->Vb);
  BHIP(hipMemcpyAsync(b->V, b->Vsave, sizeof(double) * b->NVt, hipMemcpyDeviceToDevice, b->st));
  BHIP(hipMemcpyAsync(b->P, b->Psave, sizeof(double) * b->Npt, hipMemcpyDeviceToDevice, b->st));
  for (int n = 0; n < nsteps; ++n) b_step(b);
  BHIP(hipStreamSynchronize(b->st));
  return INS2D_OK;
  B_CATCH(b)
}

int ins2d_batch_stats(const ins2d_batch* b, long long* total_iters, int* max_iters, long long* cells) {
  if (!b) return INS2D_E_ARG;
  if (total_iters) *total_iters = b->total_it;
  if (max_iters) *max_iters = b->max_it;
  if (cells) *cells = b->Npt;
  return INS2D_OK;
}

}  // extern "C"
