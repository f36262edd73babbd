// HIP kernels of libins2d_hip (gfx950, FP64).
//
//  * generic kron-factored stencils (any boundary combination, non-uniform
//    grids, second order): momentum right-hand side with the stage
//    combination fused in its epilogue, divergence, projection, Laplacian;
//  * periodic uniform stencils (second and fourth order) with the boundary
//    closures resolved at compile time;
//  * hand-written FFT Poisson solve (register-butterfly Stockham passes for
//    power-of-two lengths, whose row transforms store and load the half
//    spectra transposed for the column solve, so three kernels per solve;
//    direct DFT plus explicit transposes for other lengths);
//  * Jacobi-preconditioned CG (multi-kernel, device-resident scalars, and a
//    one-workgroup-per-system form for small systems);
//  * reductions for the per-step diagnostics.
#include "device.hpp"
#include "launch.hpp"
#include "stencil.hpp"

#include <cstdint>
#include <stdexcept>

namespace ins2d {

// ===========================================================================
// generic momentum right-hand side
//   out = [ -sum conv + sum diff  (- G p if WITH_P) ] (/ volume if SCALE_OM)
// and, for an RK stage, vtemp = aself * out + sum_j a_j k_j (same point)
// ===========================================================================
template <int COMP>
__global__ void __launch_bounds__(256) k_rhs_generic(const DComp* __restrict__ Cp, const double* __restrict__ u,
                                                    const double* __restrict__ v, const double* __restrict__ p,
                                                    int mode, double* __restrict__ out, StageCombine sc,
                                                    double* __restrict__ vtemp) {
  const DComp& C = *Cp;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (j >= C.ny) return;   // whole waves (64-wide blocks); i >= nx lanes share faces
  rhs_point<COMP, 64>(C, u, v, p, mode, out, sc, vtemp, 0, i, j);
}

// the same over split tiles: the closed form over the interior tiles, the
// generic operators over the ring (and at any point a boundary term reaches);
// thread q of a block takes the points (i0 + q mod w, j0 + q div w + r P),
// r < kRhsRows, P = 256 / w the rows of one pass (tiles of up to kRhsRows P
// rows: the tile record and operator loads, and the launch's waves, shared
// by kRhsRows points).  One launch: on single
// grids a separate launch for the interior tiles measured slower at every
// size up to 1024^2.
constexpr int kRhsRows = 4;
__global__ void __launch_bounds__(256) k_rhs_split(const DComp* __restrict__ Cu, const DComp* __restrict__ Cv,
                                                  const RTile* __restrict__ tiles, const double* __restrict__ u,
                                                  const double* __restrict__ v, const double* __restrict__ p, int mode,
                                                  double* __restrict__ out_u, double* __restrict__ out_v,
                                                  StageCombine scu, StageCombine scv, double* __restrict__ vt_u,
                                                  double* __restrict__ vt_v) {
  const RTile t = tiles[blockIdx.x];
  const int q = threadIdx.x;
  const int pass = 256 / t.w;   // rows per pass of the block
  if (q >= t.w * pass) return;
  const int i = t.i0 + q % t.w;
  const bool closed = (t.comp & 2) != 0;
#pragma unroll
  for (int r = 0; r < kRhsRows; ++r) {
    const int jj = q / t.w + r * pass;
    if (jj >= t.h) break;
    const int j = t.j0 + jj;
    if ((t.comp & 1) == 0) {
      const DComp& C = *Cu;
      if (closed && !(glb(C.bmx)[i] & glb(C.bmy)[j]))
        rhs_closed_point<0>(C, u, v, p, mode, out_u, scu, vt_u, 0, i, j);
      else
        rhs_point<0, 0>(C, u, v, p, mode, out_u, scu, vt_u, 0, i, j);
    } else {
      const DComp& C = *Cv;
      if (closed && !(glb(C.bmx)[i] & glb(C.bmy)[j]))
        rhs_closed_point<1>(C, u, v, p, mode, out_v, scv, vt_v, 0, i, j);
      else
        rhs_point<1, 0>(C, u, v, p, mode, out_v, scv, vt_v, 0, i, j);
    }
  }
}

// f = oscale * ( Mu (ca A_u + cb B_u) + Mv (ca A_v + cb B_v) + ycoef yM )
__global__ void __launch_bounds__(256) k_div_generic(const DComp* __restrict__ Cu, const DComp* __restrict__ Cv,
                                                    int npx, int npy, const double* __restrict__ A,
                                                    const double* __restrict__ B, double ca, double cb,
                                                    double ycoef, double oscale, int nu, double* __restrict__ f) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= npx || j >= npy) return;
  f[(size_t)i + (size_t)npx * j] = oscale * div_point(*Cu, *Cv, A, B, ca, cb, ycoef, nu, i, j);
}

// V = Vn + dt * (Vtemp - c * Om^-1 G dp)      (time_ERK.m:132)
template <int COMP>
__global__ void __launch_bounds__(256) k_proj_generic(const DComp* __restrict__ Cp, const double* __restrict__ Vn,
                                                     const double* __restrict__ Vt, const double* __restrict__ dp,
                                                     double dt, double c, double* __restrict__ V) {
  const DComp& C = *Cp;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= C.nx || j >= C.ny) return;
  V[(size_t)i + (size_t)C.nx * j] = proj_point(C, Vn, Vt, dp, dt, c, i, j);
}

// tV = k1 - Om^-1 G p over the velocity points of both components (u then
// v, nu of them u): the additional pressure solve's right-hand side
// (F - G p) / Omega from the first-stage k1 = F / Omega of the same velocity
__global__ void __launch_bounds__(256) k_sub_grad_generic(const DComp* __restrict__ Cu, const DComp* __restrict__ Cv,
                                                         int nu, int nv, const double* __restrict__ k1,
                                                         const double* __restrict__ p, double* __restrict__ tV) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nu + nv) return;
  const bool isv = e >= nu;
  const DComp& C = isv ? *Cv : *Cu;
  const int q = isv ? e - nu : e;
  const int i = q % C.nx, j = q / C.nx;
  tV[e] = k1[e] - kron2_at(C.G, p, i, j) / (glb(C.ox)[i] * glb(C.oy)[j]);
}

// ===========================================================================
// periodic uniform stencils (second and fourth order)
// ===========================================================================
// rhx = 1 / hx, rhy = 1 / hy, rom = 1 / Omega (hx hy, 72 hx hy at fourth
// order): the kernels multiply by these instead of dividing (an FP64
// division costs about ten instructions, one of them quarter rate)
struct PerGrid {
  int Nx, Ny;
  double hx, hy, invRe;
  double rhx, rhy, rom;
};

__device__ __forceinline__ int wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// Block coordinates with the blocks of each XCD on one band of block rows.
// Blocks reach the 8 XCDs round-robin in launch order, so with the plain
// order the left and right neighbours of a block sit on other XCDs and the
// cache lines their halos share are fetched into two L2s; here XCD x takes
// block rows [x band, (x + 1) band), all columns, when the row count divides
// by 8 (else the launch order).  A bijection of the grid.
template <bool BAND>
__device__ __forceinline__ int2 block_xy() {
  if (!BAND || (gridDim.y & 7)) return make_int2(blockIdx.x, blockIdx.y);
  const int gx = gridDim.x;
  const int b = blockIdx.x + gx * blockIdx.y;
  const int x = b & 7, l = b >> 3, band = gridDim.y >> 3;
  return make_int2(l % gx, x * band + l / gx);
}

// second order, one thread computes both components at (i, j)
// (interior formulas of the operators in operator_averaging.m,
// operator_interpolation.m and operator_convection_diffusion.m specialised to
// all-periodic uniform grids)
// the point formula from its 12 stencil values: u and v at (i, j) and its
// four neighbours, v(i-1, j+1) and u(i+1, j-1)
__device__ __forceinline__ void rhs_per2_vals(const PerGrid& G, double u0, double ue, double uw, double un, double us,
                                              double v0, double ve, double vw, double vn, double vs, double v_wn,
                                              double u_se, double& fu, double& fv) {
  const double hx = G.hx, hy = G.hy;
  // u-momentum
  const double Ue = 0.5 * (u0 + ue), Uw = 0.5 * (uw + u0);
  const double du2dx = (hy * Ue) * Ue - (hy * Uw) * Uw;
  const double Vn = 0.5 * hx * (v_wn + vn), Vs = 0.5 * hx * (vw + v0);
  const double duvdy = Vn * (0.5 * (u0 + un)) - Vs * (0.5 * (us + u0));
  const double d2u = hy * ((ue - u0) * G.rhx - (u0 - uw) * G.rhx) + hx * ((un - u0) * G.rhy - (u0 - us) * G.rhy);
  fu = -(du2dx + duvdy) + G.invRe * d2u;
  // v-momentum
  const double Ue2 = 0.5 * hy * (u_se + ue), Uw2 = 0.5 * hy * (us + u0);
  const double duvdx = Ue2 * (0.5 * (v0 + ve)) - Uw2 * (0.5 * (vw + v0));
  const double Vn2 = 0.5 * (v0 + vn), Vs2 = 0.5 * (vs + v0);
  const double dv2dy = (hx * Vn2) * Vn2 - (hx * Vs2) * Vs2;
  const double d2v = hy * ((ve - v0) * G.rhx - (v0 - vw) * G.rhx) + hx * ((vn - v0) * G.rhy - (v0 - vs) * G.rhy);
  fv = -(duvdx + dv2dy) + G.invRe * d2v;
}

__device__ __forceinline__ void rhs_per2(const PerGrid& G, const double* __restrict__ u, const double* __restrict__ v,
                                         int i, int j, double& fu, double& fv) {
  const int Nx = G.Nx;
  const int im = wrap(i - 1, Nx), ip = wrap(i + 1, Nx);
  const int jm = wrap(j - 1, G.Ny), jp = wrap(j + 1, G.Ny);
  const size_t r0 = (size_t)Nx * j, rm = (size_t)Nx * jm, rp = (size_t)Nx * jp;
  rhs_per2_vals(G, u[i + r0], u[ip + r0], u[im + r0], u[i + rp], u[i + rm], v[i + r0], v[ip + r0], v[im + r0],
                v[i + rp], v[i + rm], v[im + rp], u[ip + rm], fu, fv);
}

// the points (i, j) and (i+1, j), i even (Nx even): of their 24 stencil
// values 18 are distinct, the pairs at i, i+1 loaded 16 bytes at a time
__device__ __forceinline__ void rhs_per2_pair(const PerGrid& G, const double* __restrict__ u,
                                              const double* __restrict__ v, int i, int j, double2& fu, double2& fv) {
  const int Nx = G.Nx;
  const int im = wrap(i - 1, Nx), i2 = wrap(i + 2, Nx);
  const int jm = wrap(j - 1, G.Ny), jp = wrap(j + 1, G.Ny);
  const size_t r0 = (size_t)Nx * j, rm = (size_t)Nx * jm, rp = (size_t)Nx * jp;
  const double2 uc = *(const double2*)(u + i + r0), vc = *(const double2*)(v + i + r0);
  const double2 un = *(const double2*)(u + i + rp), vn = *(const double2*)(v + i + rp);
  const double2 us = *(const double2*)(u + i + rm), vs = *(const double2*)(v + i + rm);
  const double uw = u[im + r0], ue2 = u[i2 + r0], vw = v[im + r0], ve2 = v[i2 + r0];
  const double vwn = v[im + rp], use2 = u[i2 + rm];
  rhs_per2_vals(G, uc.x, uc.y, uw, un.x, us.x, vc.x, vc.y, vw, vn.x, vs.x, vwn, us.y, fu.x, fv.x);
  rhs_per2_vals(G, uc.y, ue2, uc.x, un.y, us.y, vc.y, ve2, vc.x, vn.y, vs.y, vn.x, use2, fu.y, fv.y);
}

// fourth order: accessor of u(i + a, j + b), v(i + a, j + b), |a|, |b| <= 3,
// from the rows j-3 .. j+3 staged in LDS (pointers already at this thread's
// column)
struct RingAcc4 {
  const double* ur[7];
  const double* vr[7];
  __device__ __forceinline__ double U(int a, int b) const { return ur[b + 3][a]; }
  __device__ __forceinline__ double V(int a, int b) const { return vr[b + 3][a]; }
};

template <int ORDER>
__device__ __forceinline__ void grad_per(const PerGrid& G, const double* __restrict__ p, int i, int j, double& gx,
                                         double& gy) {
  const int Nx = G.Nx, Ny = G.Ny;
  const size_t r0 = (size_t)Nx * j;
  const int im = wrap(i - 1, Nx);
  const int jm = wrap(j - 1, Ny);
  const double p0 = p[i + r0];
  if (ORDER == 2) {
    gx = G.hy * (p0 - p[im + r0]);
    gy = G.hx * (p0 - p[i + (size_t)Nx * jm]);
  } else {
    const int ip = wrap(i + 1, Nx), imm = wrap(i - 2, Nx);
    const int jp = wrap(j + 1, Ny), jmm = wrap(j - 2, Ny);
    gx = 81.0 * G.hy * (p0 - p[im + r0]) - 3.0 * G.hy * (p[ip + r0] - p[imm + r0]);
    gy = 81.0 * G.hx * (p0 - p[i + (size_t)Nx * jm]) - 3.0 * G.hx * (p[i + (size_t)Nx * jp] - p[i + (size_t)Nx * jmm]);
  }
}

// second order, one thread per cell (fourth order: k_rhs_per4s)
template <int ORDER, bool BAND>
__global__ void __launch_bounds__(256) k_rhs_per(PerGrid G, const double* __restrict__ u, const double* __restrict__ v,
                                                const double* __restrict__ p, int mode, double* __restrict__ out,
                                                StageCombine sc, double* __restrict__ vtemp) {
  static_assert(ORDER == 2, "the fourth-order right-hand side is k_rhs_per4s");
  const int2 bxy = block_xy<BAND>();
  const int i = bxy.x * blockDim.x + threadIdx.x;
  const int j = bxy.y * blockDim.y + threadIdx.y;
  if (i >= G.Nx || j >= G.Ny) return;
  double fu, fv;
  rhs_per2(G, u, v, i, j, fu, fv);
  if (mode & RHS_WITH_P) {
    double gx, gy;
    grad_per<ORDER>(G, p, i, j, gx, gy);
    fu -= gx;
    fv -= gy;
  }
  if (mode & RHS_SCALE_OM) {
    fu *= G.rom;
    fv *= G.rom;
  }
  const size_t N = (size_t)G.Nx * G.Ny;
  const size_t idx = (size_t)i + (size_t)G.Nx * j;
  if (out) { out[idx] = fu; out[N + idx] = fv; }
  if (vtemp) {
    double a = sc.aself * fu, b = sc.aself * fv;
    for (int q = 0; q < sc.n; ++q) { a += sc.a[q] * sc.k[q][idx]; b += sc.a[q] * sc.k[q][N + idx]; }
    vtemp[idx] = a;
    vtemp[N + idx] = b;
  }
}

// second order, two cells (i, i+1) per thread (even Nx): the shared stencil
// values loaded once, 16-byte loads and stores
// DUAL: also (F - G p) / Omega into out2, as a RHS_WITH_P | RHS_SCALE_OM
// launch forms it (the F of a step's end velocity shared by the additional
// pressure solve and the next step's first stage)
template <bool BAND, bool DUAL>
__global__ void __launch_bounds__(256) k_rhs_per2x(PerGrid G, const double* __restrict__ u,
                                                  const double* __restrict__ v, const double* __restrict__ p,
                                                  int mode, double* __restrict__ out, StageCombine sc,
                                                  double* __restrict__ vtemp, double* __restrict__ out2) {
  const int2 bxy = block_xy<BAND>();
  const int i = 2 * (bxy.x * blockDim.x + threadIdx.x);
  const int j = bxy.y * blockDim.y + threadIdx.y;
  if (i >= G.Nx || j >= G.Ny) return;
  double2 fu, fv;
  rhs_per2_pair(G, u, v, i, j, fu, fv);
  if constexpr (DUAL) {
    double gx0, gy0, gx1, gy1;
    grad_per<2>(G, p, i, j, gx0, gy0);
    grad_per<2>(G, p, i + 1, j, gx1, gy1);
    const size_t n2 = (size_t)G.Nx * G.Ny, id2 = (size_t)i + (size_t)G.Nx * j;
    *(double2*)(out2 + id2) = make_double2((fu.x - gx0) * G.rom, (fu.y - gx1) * G.rom);
    *(double2*)(out2 + n2 + id2) = make_double2((fv.x - gy0) * G.rom, (fv.y - gy1) * G.rom);
  }
  if (mode & RHS_WITH_P) {
    double gx, gy;
    grad_per<2>(G, p, i, j, gx, gy);
    fu.x -= gx;
    fv.x -= gy;
    grad_per<2>(G, p, i + 1, j, gx, gy);
    fu.y -= gx;
    fv.y -= gy;
  }
  if (mode & RHS_SCALE_OM) {
    fu.x *= G.rom; fu.y *= G.rom;
    fv.x *= G.rom; fv.y *= G.rom;
  }
  const size_t N = (size_t)G.Nx * G.Ny;
  const size_t idx = (size_t)i + (size_t)G.Nx * j;
  if (out) {
    *(double2*)(out + idx) = fu;
    *(double2*)(out + N + idx) = fv;
  }
  if (vtemp) {
    double2 a = make_double2(sc.aself * fu.x, sc.aself * fu.y), b = make_double2(sc.aself * fv.x, sc.aself * fv.y);
    for (int q = 0; q < sc.n; ++q) {
      const double2 ka = *(const double2*)(sc.k[q] + idx), kb = *(const double2*)(sc.k[q] + N + idx);
      a.x += sc.a[q] * ka.x; a.y += sc.a[q] * ka.y;
      b.x += sc.a[q] * kb.x; b.y += sc.a[q] * kb.y;
    }
    *(double2*)(vtemp + idx) = a;
    *(double2*)(vtemp + N + idx) = b;
  }
}

// fourth order (alfa = 81, beta = 9/8: alfa * (1h operators) - (3h
// operators), convection.m:38-43, operator_convection_diffusion.m:463-470),
// row-streaming: one wave per 64-wide column strip of H rows; the rows
// j-3 .. j+3 of u and v live in an LDS ring of 7 rows (+-3 halo in x), so
// every value is read from memory once per strip (plus the 6-row start-up and
// the 6-column halo).  The y-direction face fluxes are streamed: each row
// computes only the fluxes of its newest faces (1h faces j+1, 3h faces j+3,
// diffusive faces j+3/2) and takes the older ones from registers, where the
// rows above left them; the x-direction terms are computed per row.
constexpr int kStripW = 64;
template <bool BAND>
__global__ void __launch_bounds__(64) k_rhs_per4s(PerGrid G, const double* __restrict__ u,
                                                 const double* __restrict__ v, const double* __restrict__ p, int mode,
                                                 double* __restrict__ out, StageCombine sc,
                                                 double* __restrict__ vtemp, int H) {
  __shared__ double su[7][kStripW + 6];
  __shared__ double sv[7][kStripW + 6];
  const int tx = threadIdx.x;
  const int Nx = G.Nx, Ny = G.Ny;
  const int2 bxy = block_xy<BAND>();   // strips of one XCD side by side
  const int x0 = bxy.x * kStripW, j0 = bxy.y * H;
  const int i = x0 + tx;
  // LDS column c holds x = x0 - 3 + c (periodic)
  const int xa = ((x0 - 3 + tx) % Nx + Nx) % Nx;
  const int xb = tx < 6 ? (x0 + kStripW - 3 + tx) % Nx : 0;
  // a row is fetched into registers one iteration ahead (its memory latency
  // overlaps the previous row's arithmetic) and stored to the ring after
  double pu = 0.0, pv = 0.0, pu2 = 0.0, pv2 = 0.0;
  auto fetch_row = [&](int y) {
    const size_t r = (size_t)Nx * wrap(y, Ny);
    pu = u[r + xa];
    pv = v[r + xa];
    if (tx < 6) {
      pu2 = u[r + xb];
      pv2 = v[r + xb];
    }
  };
  auto store_row = [&](int slot) {
    su[slot][tx] = pu;
    sv[slot][tx] = pv;
    if (tx < 6) {
      su[slot][kStripW + tx] = pu2;
      sv[slot][kStripW + tx] = pv2;
    }
  };
  const double hx = G.hx, hy = G.hy;
  const double w1 = 9.0 / 16.0, w2 = -1.0 / 16.0;   // beta/2, (1-beta)/2
  const double r72x = G.rhx * (1.0 / 72.0), r72y = G.rhy * (1.0 / 72.0);
  RingAcc4 acc;
  // rows j-3 .. j+3 of the ring (s0: slot of row j - 3)
  auto set_ring = [&](int s0) {
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      const int sl = s0 + k < 7 ? s0 + k : s0 + k - 7;
      acc.ur[k] = &su[sl][tx + 3];
      acc.vr[k] = &sv[sl][tx + 3];
    }
  };
#define U(a, b) acc.U(a, b)
#define V(a, b) acc.V(a, b)
  // y-direction fluxes at the face j + d of the current row j
  auto fy1u = [&](int d) {   // 1h u-flux v * u on grid line y_{j+d}
    return (hx * (w2 * V(-2, d) + w1 * V(-1, d) + w1 * V(0, d) + w2 * V(1, d))) * (0.5 * (U(0, d - 1) + U(0, d)));
  };
  auto fy3u = [&](int d) {   // 3h u-flux on grid line y_{j+d-1}
    return (3.0 * hx * (w2 * V(-2, d - 1) + w1 * V(-1, d - 1) + w1 * V(0, d - 1) + w2 * V(1, d - 1))) *
           (0.5 * (U(0, d - 3) + U(0, d)));
  };
  auto pya = [&](int d) {    // du/dy between rows j+d and j+d+1
    return (81.0 * (U(0, d + 1) - U(0, d)) - 3.0 * (U(0, d + 2) - U(0, d - 1))) * r72y;
  };
  auto gy1v = [&](int d) {   // 1h v-flux v * v
    return (hx * (w2 * V(0, d - 2) + w1 * V(0, d - 1) + w1 * V(0, d) + w2 * V(0, d + 1))) *
           (0.5 * (V(0, d - 1) + V(0, d)));
  };
  auto gy3v = [&](int d) {   // 3h v-flux
    return (3.0 * hx * (w2 * V(0, d - 3) + w1 * V(0, d - 2) + w1 * V(0, d - 1) + w2 * V(0, d))) *
           (0.5 * (V(0, d - 3) + V(0, d)));
  };
  auto qya = [&](int d) {    // dv/dy between rows j+d and j+d+1
    return (81.0 * (V(0, d + 1) - V(0, d)) - 3.0 * (V(0, d + 2) - V(0, d - 1))) * r72y;
  };
  for (int k = 0; k < 6; ++k) {
    fetch_row(j0 - 3 + k);
    store_row(k);
  }
  fetch_row(j0 + 3);
  __syncthreads();
  // the fluxes of the faces at and above row j0 (rows j0-3 .. j0+2 of the ring)
  set_ring(0);
  double F1 = fy1u(0);                                 // face j
  double F30 = fy3u(0), F31 = fy3u(1), F32 = fy3u(2);   // faces j, j+1, j+2
  double Pm2 = pya(-2), Pm1 = pya(-1), P0 = pya(0);      // j-3/2 .. j+1/2
  double G1 = gy1v(0);
  double G30 = gy3v(0), G31 = gy3v(1), G32 = gy3v(2);
  double Qm2 = qya(-2), Qm1 = qya(-1), Q0 = qya(0);
  const size_t N = (size_t)Nx * Ny;
  const int jend = min(j0 + H, Ny);
  int s0 = 0;   // ring slot of row j - 3
  for (int j = j0; j < jend; ++j) {
    store_row(s0 == 0 ? 6 : s0 - 1);   // row j + 3
    __syncthreads();
    if (j + 1 < jend) fetch_row(j + 4);
    set_ring(s0);
    // ---------------- u-momentum ----------------
    auto Ix1 = [&](int f) { return hy * (w2 * U(f - 2, 0) + w1 * U(f - 1, 0) + w1 * U(f, 0) + w2 * U(f + 1, 0)); };
    auto Ax1 = [&](int f) { return 0.5 * (U(f - 1, 0) + U(f, 0)); };
    const double du2dx1 = Ix1(1) * Ax1(1) - Ix1(0) * Ax1(0);
    auto Ix3 = [&](int r) { return 3.0 * hy * (w2 * U(r - 3, 0) + w1 * U(r - 2, 0) + w1 * U(r - 1, 0) + w2 * U(r, 0)); };
    auto Ax3 = [&](int r) { return 0.5 * (U(r - 3, 0) + U(r, 0)); };
    const double du2dx3 = Ix3(3) * Ax3(3) - Ix3(0) * Ax3(0);
    const double F1n = fy1u(1), F3n = fy3u(3), Pn = pya(1);
    const double duvdy1 = F1n - F1;
    const double duvdy3 = F3n - F30;
    const double convu = 81.0 * (du2dx1 + duvdy1) - (du2dx3 + duvdy3);
    auto Px = [&](int k) { return (81.0 * (U(k + 1, 0) - U(k, 0)) - 3.0 * (U(k + 2, 0) - U(k - 1, 0))) * r72x; };
    const double d2u = hy * (81.0 * (Px(0) - Px(-1)) - 3.0 * (Px(1) - Px(-2))) +
                       hx * (81.0 * (P0 - Pm1) - 3.0 * (Pn - Pm2));
    double fu = -convu + G.invRe * d2u;
    // ---------------- v-momentum ----------------
    auto Uf1 = [&](int f) { return hy * (w2 * U(f, -2) + w1 * U(f, -1) + w1 * U(f, 0) + w2 * U(f, 1)); };
    auto Bx1 = [&](int f) { return 0.5 * (V(f - 1, 0) + V(f, 0)); };
    const double duvdx1 = Uf1(1) * Bx1(1) - Uf1(0) * Bx1(0);
    auto Uf3 = [&](int f) { return 3.0 * hy * (w2 * U(f - 1, -2) + w1 * U(f - 1, -1) + w1 * U(f - 1, 0) + w2 * U(f - 1, 1)); };
    auto Bx3 = [&](int f) { return 0.5 * (V(f - 3, 0) + V(f, 0)); };
    const double duvdx3 = Uf3(3) * Bx3(3) - Uf3(0) * Bx3(0);
    const double G1n = gy1v(1), G3n = gy3v(3), Qn = qya(1);
    const double dv2dy1 = G1n - G1;
    const double dv2dy3 = G3n - G30;
    const double convv = 81.0 * (duvdx1 + dv2dy1) - (duvdx3 + dv2dy3);
    auto Qx = [&](int k) { return (81.0 * (V(k + 1, 0) - V(k, 0)) - 3.0 * (V(k + 2, 0) - V(k - 1, 0))) * r72x; };
    const double d2v = hy * (81.0 * (Qx(0) - Qx(-1)) - 3.0 * (Qx(1) - Qx(-2))) +
                       hx * (81.0 * (Q0 - Qm1) - 3.0 * (Qn - Qm2));
    double fv = -convv + G.invRe * d2v;
    // advance the flux queues to the next row
    F1 = F1n;
    F30 = F31; F31 = F32; F32 = F3n;
    Pm2 = Pm1; Pm1 = P0; P0 = Pn;
    G1 = G1n;
    G30 = G31; G31 = G32; G32 = G3n;
    Qm2 = Qm1; Qm1 = Q0; Q0 = Qn;
    if (i < Nx) {
      if (mode & RHS_WITH_P) {
        double gx, gy;
        grad_per<4>(G, p, i, j, gx, gy);
        fu -= gx;
        fv -= gy;
      }
      if (mode & RHS_SCALE_OM) {
        fu *= G.rom;
        fv *= G.rom;
      }
      const size_t idx = (size_t)i + (size_t)Nx * j;
      if (out) { out[idx] = fu; out[N + idx] = fv; }
      if (vtemp) {
        double a = sc.aself * fu, b = sc.aself * fv;
        for (int q = 0; q < sc.n; ++q) { a += sc.a[q] * sc.k[q][idx]; b += sc.a[q] * sc.k[q][N + idx]; }
        vtemp[idx] = a;
        vtemp[N + idx] = b;
      }
    }
    __syncthreads();
    s0 = s0 == 6 ? 0 : s0 + 1;
  }
#undef U
#undef V
}

// k_rhs_per4s with two cells per thread (columns x0 + 2 tx and x0 + 2 tx + 1
// of a 128-wide strip; Nx a multiple of 128, fields 16-byte aligned): the
// x-direction terms the two cells share are formed once (the same
// expressions over the same ring values), the row fetch and the stores are
// 16 bytes per field and lane
constexpr int kStripW2 = 128;
template <bool BAND, bool DUAL>
__global__ void __launch_bounds__(64) k_rhs_per4s2(PerGrid G, const double* __restrict__ u,
                                                 const double* __restrict__ v, const double* __restrict__ p, int mode,
                                                 double* __restrict__ out, StageCombine sc,
                                                 double* __restrict__ vtemp, int H, double* __restrict__ out2) {
  __shared__ double su[7][kStripW2 + 6];
  __shared__ double sv[7][kStripW2 + 6];
  const int tx = threadIdx.x;
  const int Nx = G.Nx, Ny = G.Ny;
  const int2 bxy = block_xy<BAND>();   // strips of one XCD side by side
  const int x0 = bxy.x * kStripW2, j0 = bxy.y * H;
  const int i0 = x0 + 2 * tx;   // this lane's cells i0 and i0 + 1
  // LDS column c holds x = x0 - 3 + c (periodic): the lane's pair at
  // 3 + 2 tx, the halo columns 0..2 and 131..133 from lanes 0..5
  const int xh = tx < 3 ? ((x0 - 3 + tx) % Nx + Nx) % Nx : (tx < 6 ? (x0 + kStripW2 + tx - 3) % Nx : 0);
  const int ch = tx < 3 ? tx : kStripW2 + tx;
  // a row is fetched into registers one iteration ahead (its memory latency
  // overlaps the previous row's arithmetic) and stored to the ring after
  double2 pu = make_double2(0.0, 0.0), pv = make_double2(0.0, 0.0);
  double pu2 = 0.0, pv2 = 0.0;
  auto fetch_row = [&](int y) {
    const size_t r = (size_t)Nx * wrap(y, Ny);
    pu = *(const double2*)(u + r + i0);
    pv = *(const double2*)(v + r + i0);
    if (tx < 6) {
      pu2 = u[r + xh];
      pv2 = v[r + xh];
    }
  };
  auto store_row = [&](int slot) {
    su[slot][3 + 2 * tx] = pu.x;
    su[slot][4 + 2 * tx] = pu.y;
    sv[slot][3 + 2 * tx] = pv.x;
    sv[slot][4 + 2 * tx] = pv.y;
    if (tx < 6) {
      su[slot][ch] = pu2;
      sv[slot][ch] = pv2;
    }
  };
  const double hx = G.hx, hy = G.hy;
  const double w1 = 9.0 / 16.0, w2 = -1.0 / 16.0;   // beta/2, (1-beta)/2
  const double r72x = G.rhx * (1.0 / 72.0), r72y = G.rhy * (1.0 / 72.0);
  RingAcc4 acc;
  int cc = 0;   // the cell of the pair the stencil lambdas take (column offset)
  // rows j-3 .. j+3 of the ring (s0: slot of row j - 3)
  auto set_ring = [&](int s0) {
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      const int sl = s0 + k < 7 ? s0 + k : s0 + k - 7;
      acc.ur[k] = &su[sl][2 * tx + 3];
      acc.vr[k] = &sv[sl][2 * tx + 3];
    }
  };
#define U(a, b) acc.U((a) + cc, b)
#define V(a, b) acc.V((a) + cc, b)
  // y-direction fluxes at the face j + d of the current row j
  auto fy1u = [&](int d) {   // 1h u-flux v * u on grid line y_{j+d}
    return (hx * (w2 * V(-2, d) + w1 * V(-1, d) + w1 * V(0, d) + w2 * V(1, d))) * (0.5 * (U(0, d - 1) + U(0, d)));
  };
  auto fy3u = [&](int d) {   // 3h u-flux on grid line y_{j+d-1}
    return (3.0 * hx * (w2 * V(-2, d - 1) + w1 * V(-1, d - 1) + w1 * V(0, d - 1) + w2 * V(1, d - 1))) *
           (0.5 * (U(0, d - 3) + U(0, d)));
  };
  auto pya = [&](int d) {    // du/dy between rows j+d and j+d+1
    return (81.0 * (U(0, d + 1) - U(0, d)) - 3.0 * (U(0, d + 2) - U(0, d - 1))) * r72y;
  };
  auto gy1v = [&](int d) {   // 1h v-flux v * v
    return (hx * (w2 * V(0, d - 2) + w1 * V(0, d - 1) + w1 * V(0, d) + w2 * V(0, d + 1))) *
           (0.5 * (V(0, d - 1) + V(0, d)));
  };
  auto gy3v = [&](int d) {   // 3h v-flux
    return (3.0 * hx * (w2 * V(0, d - 3) + w1 * V(0, d - 2) + w1 * V(0, d - 1) + w2 * V(0, d))) *
           (0.5 * (V(0, d - 3) + V(0, d)));
  };
  auto qya = [&](int d) {    // dv/dy between rows j+d and j+d+1
    return (81.0 * (V(0, d + 1) - V(0, d)) - 3.0 * (V(0, d + 2) - V(0, d - 1))) * r72y;
  };
  for (int k = 0; k < 6; ++k) {
    fetch_row(j0 - 3 + k);
    store_row(k);
  }
  fetch_row(j0 + 3);
  __syncthreads();
  // the fluxes of the faces at and above row j0 (rows j0-3 .. j0+2 of the ring)
  set_ring(0);
  double F1[2], F30[2], F31[2], F32[2], Pm2[2], Pm1[2], P0[2];
  double G1[2], G30[2], G31[2], G32[2], Qm2[2], Qm1[2], Q0[2];
#pragma unroll
  for (cc = 0; cc < 2; ++cc) {
    F1[cc] = fy1u(0);                                             // face j
    F30[cc] = fy3u(0); F31[cc] = fy3u(1); F32[cc] = fy3u(2);      // faces j, j+1, j+2
    Pm2[cc] = pya(-2); Pm1[cc] = pya(-1); P0[cc] = pya(0);        // j-3/2 .. j+1/2
    G1[cc] = gy1v(0);
    G30[cc] = gy3v(0); G31[cc] = gy3v(1); G32[cc] = gy3v(2);
    Qm2[cc] = qya(-2); Qm1[cc] = qya(-1); Q0[cc] = qya(0);
  }
  const size_t N = (size_t)Nx * Ny;
  const int jend = min(j0 + H, Ny);
  int s0 = 0;   // ring slot of row j - 3
  for (int j = j0; j < jend; ++j) {
    store_row(s0 == 0 ? 6 : s0 - 1);   // row j + 3
    __syncthreads();
    if (j + 1 < jend) fetch_row(j + 4);
    set_ring(s0);
    double fuc[2], fvc[2];
#pragma unroll
    for (cc = 0; cc < 2; ++cc) {
      // ---------------- u-momentum ----------------
      auto Ix1 = [&](int f) { return hy * (w2 * U(f - 2, 0) + w1 * U(f - 1, 0) + w1 * U(f, 0) + w2 * U(f + 1, 0)); };
      auto Ax1 = [&](int f) { return 0.5 * (U(f - 1, 0) + U(f, 0)); };
      const double du2dx1 = Ix1(1) * Ax1(1) - Ix1(0) * Ax1(0);
      auto Ix3 = [&](int r) { return 3.0 * hy * (w2 * U(r - 3, 0) + w1 * U(r - 2, 0) + w1 * U(r - 1, 0) + w2 * U(r, 0)); };
      auto Ax3 = [&](int r) { return 0.5 * (U(r - 3, 0) + U(r, 0)); };
      const double du2dx3 = Ix3(3) * Ax3(3) - Ix3(0) * Ax3(0);
      const double F1n = fy1u(1), F3n = fy3u(3), Pn = pya(1);
      const double duvdy1 = F1n - F1[cc];
      const double duvdy3 = F3n - F30[cc];
      const double convu = 81.0 * (du2dx1 + duvdy1) - (du2dx3 + duvdy3);
      auto Px = [&](int k) { return (81.0 * (U(k + 1, 0) - U(k, 0)) - 3.0 * (U(k + 2, 0) - U(k - 1, 0))) * r72x; };
      const double d2u = hy * (81.0 * (Px(0) - Px(-1)) - 3.0 * (Px(1) - Px(-2))) +
                         hx * (81.0 * (P0[cc] - Pm1[cc]) - 3.0 * (Pn - Pm2[cc]));
      double fu = -convu + G.invRe * d2u;
      // ---------------- v-momentum ----------------
      auto Uf1 = [&](int f) { return hy * (w2 * U(f, -2) + w1 * U(f, -1) + w1 * U(f, 0) + w2 * U(f, 1)); };
      auto Bx1 = [&](int f) { return 0.5 * (V(f - 1, 0) + V(f, 0)); };
      const double duvdx1 = Uf1(1) * Bx1(1) - Uf1(0) * Bx1(0);
      auto Uf3 = [&](int f) { return 3.0 * hy * (w2 * U(f - 1, -2) + w1 * U(f - 1, -1) + w1 * U(f - 1, 0) + w2 * U(f - 1, 1)); };
      auto Bx3 = [&](int f) { return 0.5 * (V(f - 3, 0) + V(f, 0)); };
      const double duvdx3 = Uf3(3) * Bx3(3) - Uf3(0) * Bx3(0);
      const double G1n = gy1v(1), G3n = gy3v(3), Qn = qya(1);
      const double dv2dy1 = G1n - G1[cc];
      const double dv2dy3 = G3n - G30[cc];
      const double convv = 81.0 * (duvdx1 + dv2dy1) - (duvdx3 + dv2dy3);
      auto Qx = [&](int k) { return (81.0 * (V(k + 1, 0) - V(k, 0)) - 3.0 * (V(k + 2, 0) - V(k - 1, 0))) * r72x; };
      const double d2v = hy * (81.0 * (Qx(0) - Qx(-1)) - 3.0 * (Qx(1) - Qx(-2))) +
                         hx * (81.0 * (Q0[cc] - Qm1[cc]) - 3.0 * (Qn - Qm2[cc]));
      double fv = -convv + G.invRe * d2v;
      // advance the flux queues to the next row
      F1[cc] = F1n;
      F30[cc] = F31[cc]; F31[cc] = F32[cc]; F32[cc] = F3n;
      Pm2[cc] = Pm1[cc]; Pm1[cc] = P0[cc]; P0[cc] = Pn;
      G1[cc] = G1n;
      G30[cc] = G31[cc]; G31[cc] = G32[cc]; G32[cc] = G3n;
      Qm2[cc] = Qm1[cc]; Qm1[cc] = Q0[cc]; Q0[cc] = Qn;
      fuc[cc] = fu;
      fvc[cc] = fv;
    }
    {
      if constexpr (DUAL) {
        // also (F - G p) / Omega, as a RHS_WITH_P | RHS_SCALE_OM launch forms it
        const size_t idx = (size_t)i0 + (size_t)Nx * j;
        double gxc[2], gyc[2];
#pragma unroll
        for (int c = 0; c < 2; ++c) grad_per<4>(G, p, i0 + c, j, gxc[c], gyc[c]);
        *(double2*)(out2 + idx) = make_double2((fuc[0] - gxc[0]) * G.rom, (fuc[1] - gxc[1]) * G.rom);
        *(double2*)(out2 + N + idx) = make_double2((fvc[0] - gyc[0]) * G.rom, (fvc[1] - gyc[1]) * G.rom);
      }
      if (mode & RHS_WITH_P) {
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          double gx, gy;
          grad_per<4>(G, p, i0 + c, j, gx, gy);
          fuc[c] -= gx;
          fvc[c] -= gy;
        }
      }
      if (mode & RHS_SCALE_OM) {
        fuc[0] *= G.rom; fuc[1] *= G.rom;
        fvc[0] *= G.rom; fvc[1] *= G.rom;
      }
      const size_t idx = (size_t)i0 + (size_t)Nx * j;
      if (out) {
        *(double2*)(out + idx) = make_double2(fuc[0], fuc[1]);
        *(double2*)(out + N + idx) = make_double2(fvc[0], fvc[1]);
      }
      if (vtemp) {
        double2 a = make_double2(sc.aself * fuc[0], sc.aself * fuc[1]);
        double2 b = make_double2(sc.aself * fvc[0], sc.aself * fvc[1]);
        for (int q = 0; q < sc.n; ++q) {
          const double2 ka = *(const double2*)(sc.k[q] + idx), kb = *(const double2*)(sc.k[q] + N + idx);
          a.x += sc.a[q] * ka.x; a.y += sc.a[q] * ka.y;
          b.x += sc.a[q] * kb.x; b.y += sc.a[q] * kb.y;
        }
        *(double2*)(vtemp + idx) = a;
        *(double2*)(vtemp + N + idx) = b;
      }
    }
    __syncthreads();
    s0 = s0 == 6 ? 0 : s0 + 1;
  }
#undef U
#undef V
}

// second order (fourth order: k_div_per4s)
template <int ORDER, bool BAND>
__global__ void __launch_bounds__(256) k_div_per(PerGrid G, const double* __restrict__ A, const double* __restrict__ B,
                                                double ca, double cb, double oscale, double* __restrict__ f) {
  static_assert(ORDER == 2, "the fourth-order divergence is k_div_per4s");
  const int2 bxy = block_xy<BAND>();
  const int i = bxy.x * blockDim.x + threadIdx.x;
  const int j = bxy.y * blockDim.y + threadIdx.y;
  if (i >= G.Nx || j >= G.Ny) return;
  const int Nx = G.Nx, Ny = G.Ny;
  const size_t N = (size_t)Nx * Ny;
  auto W = [&](size_t k) { return B ? ca * A[k] + cb * B[k] : ca * A[k]; };
  const int ip = wrap(i + 1, Nx), jp = wrap(j + 1, Ny);
  const size_t r0 = (size_t)Nx * j;
  const double s = G.hy * (W(ip + r0) - W(i + r0)) + G.hx * (W(N + i + (size_t)Nx * jp) - W(N + i + r0));
  f[(size_t)i + r0] = oscale * s;
}

// fourth-order divergence, row-streaming like k_rhs_per4s (strips of H
// rows): u of row j with its x halo (-1 .. +2) staged in LDS (two row
// buffers, so one barrier per row); the v column values of rows j-1 .. j+2 of
// this thread's column carried in registers from row to row.  The raw A and B
// values of the next row are loaded one row ahead and combined when stored.
template <bool BAND>
__global__ void __launch_bounds__(64) k_div_per4s(PerGrid G, const double* __restrict__ A,
                                                 const double* __restrict__ B, double ca, double cb, double oscale,
                                                 int H, double* __restrict__ f) {
  __shared__ double su[2][kStripW + 3];
  const int tx = threadIdx.x;
  const int Nx = G.Nx, Ny = G.Ny;
  const size_t N = (size_t)Nx * Ny;
  const int2 bxy = block_xy<BAND>();   // strips of one XCD side by side
  const int x0 = bxy.x * kStripW, j0 = bxy.y * H;
  const int i = x0 + tx;
  // LDS column c holds x = x0 - 1 + c (periodic); c = tx and, for tx < 3, 64 + tx
  const int xa = ((x0 - 1 + tx) % Nx + Nx) % Nx;
  const int xb = tx < 3 ? (x0 + kStripW - 1 + tx) % Nx : xa;
  const int iv = i < Nx ? i : 0;
  const bool hasB = B != nullptr;
  auto vrow = [&](int y) { return N + (size_t)iv + (size_t)Nx * wrap(y, Ny); };
  auto W = [&](size_t k) { return hasB ? ca * A[k] + cb * B[k] : ca * A[k]; };
  double vm1 = W(vrow(j0 - 1)), v0 = W(vrow(j0)), vp1 = W(vrow(j0 + 1)), vp2 = W(vrow(j0 + 2));
  size_t r = (size_t)Nx * j0;
  size_t kv = vrow(j0 + 3);
  double ua = A[r + xa], uc = A[r + xb], va = A[kv];
  double ub = hasB ? B[r + xa] : 0.0, ud = hasB ? B[r + xb] : 0.0, vb = hasB ? B[kv] : 0.0;
  const int jend = min(j0 + H, Ny);
  int s = 0;
  for (int j = j0; j < jend; ++j) {
    su[s][tx] = hasB ? ca * ua + cb * ub : ca * ua;
    if (tx < 3) su[s][kStripW + tx] = hasB ? ca * uc + cb * ud : ca * uc;
    const double vnext = hasB ? ca * va + cb * vb : ca * va;
    if (j + 1 < jend) {
      const size_t rn = r + Nx;
      kv = vrow(j + 4);
      ua = A[rn + xa];
      uc = A[rn + xb];
      va = A[kv];
      if (hasB) {
        ub = B[rn + xa];
        ud = B[rn + xb];
        vb = B[kv];
      }
    }
    __syncthreads();
    if (i < Nx) {
      const double um1 = su[s][tx], u0 = su[s][tx + 1], up1 = su[s][tx + 2], up2 = su[s][tx + 3];
      const double sdiv = 81.0 * G.hy * (up1 - u0) - 3.0 * G.hy * (up2 - um1) + 81.0 * G.hx * (vp1 - v0) -
                          3.0 * G.hx * (vp2 - vm1);
      f[(size_t)i + r] = oscale * sdiv;
    }
    vm1 = v0;
    v0 = vp1;
    vp1 = vp2;
    vp2 = vnext;
    r += Nx;
    s ^= 1;
  }
}

template <int ORDER, bool BAND>
__global__ void __launch_bounds__(256) k_proj_per(PerGrid G, const double* __restrict__ Vn, const double* __restrict__ Vt,
                                                 const double* __restrict__ dp, double dt, double c,
                                                 double* __restrict__ V) {
  const int2 bxy = block_xy<BAND>();
  const int i = bxy.x * blockDim.x + threadIdx.x;
  const int j = bxy.y * blockDim.y + threadIdx.y;
  if (i >= G.Nx || j >= G.Ny) return;
  const size_t N = (size_t)G.Nx * G.Ny;
  const size_t idx = (size_t)i + (size_t)G.Nx * j;
  double gx, gy;
  grad_per<ORDER>(G, dp, i, j, gx, gy);
  const double tu = Vt ? Vt[idx] : 0.0, tv = Vt ? Vt[N + idx] : 0.0;
  V[idx] = Vn[idx] + dt * (tu - c * (gx * G.rom));
  V[N + idx] = Vn[N + idx] + dt * (tv - c * (gy * G.rom));
}

// k_proj_per<4> with two cells (i, i+1) per thread (Nx even, arrays 16-byte
// aligned): the five dp values of the pair's row and the three of each other
// row loaded 16 bytes at a time; the same arithmetic per cell as grad_per
// (at second order the pair form measured no faster: 1024^2, 9.9 against
// 9.9 us)
template <bool BAND>
__global__ void __launch_bounds__(256) k_proj_per4x(PerGrid G, const double* __restrict__ Vn,
                                                   const double* __restrict__ Vt, const double* __restrict__ dp,
                                                   double dt, double c, double* __restrict__ V) {
  const int2 bxy = block_xy<BAND>();
  const int i = 2 * (bxy.x * blockDim.x + threadIdx.x);
  const int j = bxy.y * blockDim.y + threadIdx.y;
  const int Nx = G.Nx, Ny = G.Ny;
  if (i >= Nx || j >= Ny) return;
  const size_t N = (size_t)Nx * Ny, r0 = (size_t)Nx * j;
  const size_t rm = (size_t)Nx * wrap(j - 1, Ny), rp = (size_t)Nx * wrap(j + 1, Ny), rmm = (size_t)Nx * wrap(j - 2, Ny);
  auto ld2 = [](const double* a) { return *(const double2*)a; };
  const double2 pw = ld2(dp + r0 + wrap(i - 2, Nx)), pc = ld2(dp + r0 + i), pe = ld2(dp + r0 + wrap(i + 2, Nx));
  const double2 pm = ld2(dp + rm + i), pp = ld2(dp + rp + i), pmm = ld2(dp + rmm + i);
  // cell i: p(i-2), p(i-1), p(i), p(i+1) = pw.x, pw.y, pc.x, pc.y; cell i+1: pw.y, pc.x, pc.y, pe.x
  const double gx0 = 81.0 * G.hy * (pc.x - pw.y) - 3.0 * G.hy * (pc.y - pw.x);
  const double gx1 = 81.0 * G.hy * (pc.y - pc.x) - 3.0 * G.hy * (pe.x - pw.y);
  const double gy0 = 81.0 * G.hx * (pc.x - pm.x) - 3.0 * G.hx * (pp.x - pmm.x);
  const double gy1 = 81.0 * G.hx * (pc.y - pm.y) - 3.0 * G.hx * (pp.y - pmm.y);
  const size_t idx = r0 + i;
  const double2 un = ld2(Vn + idx), vn = ld2(Vn + N + idx);
  const double2 tu = Vt ? ld2(Vt + idx) : make_double2(0.0, 0.0), tv = Vt ? ld2(Vt + N + idx) : make_double2(0.0, 0.0);
  *(double2*)(V + idx) = make_double2(un.x + dt * (tu.x - c * (gx0 * G.rom)), un.y + dt * (tu.y - c * (gx1 * G.rom)));
  *(double2*)(V + N + idx) =
      make_double2(vn.x + dt * (tv.x - c * (gy0 * G.rom)), vn.y + dt * (tv.y - c * (gy1 * G.rom)));
}

// ===========================================================================
// FFT Poisson solve (periodic, uniform):  A dp = f  <=>  -Ahat dphat = fhat
// (solvers/pressure/pressure_poisson.m:92-104)
// ===========================================================================
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 conjd(double2 a) { return make_double2(a.x, -a.y); }

// in-place FFT of z[0..N) in LDS; INV selects exp(+2 pi i k n / N) (unscaled).
// Stockham autosort radix-4 passes (plus one radix-2 pass when log2 N is odd);
// each thread owns at most 4 radix-4 (8 radix-2) butterflies per pass, held
// in registers between the LDS read and the LDS write (N <= 16 * blockDim).
// C: radix-4 butterflies per thread and pass (2 C in the radix-2 tail);
// the launch must give N <= 4 C blockDim
template <bool INV, int C>
__device__ void lds_fft(double2* z, int N, int log2N, const double2* __restrict__ tw) {
  const int T = blockDim.x;
  const int tid = threadIdx.x;
  int Ns = 1;
  const int passes4 = log2N / 2;
  for (int pass = 0; pass < passes4; ++pass) {
    const int Q = N / 4;
    const int step = N / (Ns * 4);
    double2 o0[C], o1[C], o2[C], o3[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int jq = tid + c * T;
      if (jq < Q) {
        const int k = jq % Ns;
        double2 a0 = z[jq], a1 = z[jq + Q], a2 = z[jq + 2 * Q], a3 = z[jq + 3 * Q];
        double2 t1 = tw[k * step], t2 = tw[2 * k * step], t3 = tw[3 * k * step];
        if (INV) { t1 = conjd(t1); t2 = conjd(t2); t3 = conjd(t3); }
        a1 = cmul(a1, t1); a2 = cmul(a2, t2); a3 = cmul(a3, t3);
        const double2 s02 = cadd(a0, a2), d02 = csub(a0, a2), s13 = cadd(a1, a3), d13 = csub(a1, a3);
        // -i * d13 (forward) or +i * d13 (inverse)
        const double2 jd13 = INV ? make_double2(-d13.y, d13.x) : make_double2(d13.y, -d13.x);
        o0[c] = cadd(s02, s13);
        o1[c] = cadd(d02, jd13);
        o2[c] = csub(s02, s13);
        o3[c] = csub(d02, jd13);
      }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int jq = tid + c * T;
      if (jq < Q) {
        const int k = jq % Ns;
        const int o = (jq / Ns) * Ns * 4 + k;
        z[o] = o0[c];
        z[o + Ns] = o1[c];
        z[o + 2 * Ns] = o2[c];
        z[o + 3 * Ns] = o3[c];
      }
    }
    __syncthreads();
    Ns *= 4;
  }
  if (log2N & 1) {
    const int Q = N / 2;
    const int step = N / (Ns * 2);
    double2 o0[2 * C], o1[2 * C];
#pragma unroll
    for (int c = 0; c < 2 * C; ++c) {
      const int jq = tid + c * T;
      if (jq < Q) {
        const int k = jq % Ns;
        double2 a0 = z[jq], a1 = z[jq + Q];
        double2 t1 = tw[k * step];
        if (INV) t1 = conjd(t1);
        a1 = cmul(a1, t1);
        o0[c] = cadd(a0, a1);
        o1[c] = csub(a0, a1);
      }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 2 * C; ++c) {
      const int jq = tid + c * T;
      if (jq < Q) {
        const int k = jq % Ns;
        const int o = (jq / Ns) * Ns * 2 + k;
        z[o] = o0[c];
        z[o + Ns] = o1[c];
      }
    }
    __syncthreads();
  }
}

// direct DFT for lengths that are not powers of two (O(N^2), N <= 4096)
template <bool INV>
__device__ void lds_dft(double2* z, double2* out, int N, const double2* __restrict__ tw) {
  for (int k = threadIdx.x; k < N; k += blockDim.x) {
    double2 acc = make_double2(0.0, 0.0);
    int m = 0;
    for (int n = 0; n < N; ++n) {
      double2 t = tw[m];
      if (INV) t = conjd(t);
      acc = cadd(acc, cmul(z[n], t));
      m += k;
      if (m >= N) m -= N;
    }
    out[k] = acc;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < N; k += blockDim.x) z[k] = out[k];
  __syncthreads();
}

template <bool INV, int C>
__device__ __forceinline__ void lds_transform(double2* z, double2* scratch, int N, bool pow2, int log2N,
                                              const double2* __restrict__ tw) {
  if (pow2) lds_fft<INV, C>(z, N, log2N, tw);
  else lds_dft<INV>(z, scratch, N, tw);
}

// ---------------------------------------------------------------------------
// Power-of-two FFTs (16 <= N <= 4096) with register butterflies.  A transform
// runs on T = max(64, N / 16) threads; each Stockham pass of radix R (16, and
// one pass of 2, 4 or 8 for the rest of log2 N) gives every thread 16 / R
// butterflies held in registers.  The first pass takes its inputs straight
// from the caller (global memory or registers) and the last pass hands its
// outputs straight back, so only the passes in between go through LDS.  The
// LDS index is swizzled, n ^ ((n >> 4) & 15), which keeps the stride-R writes
// of the early passes and the unit-stride reads free of bank conflicts.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int swz(int n) { return n ^ ((n >> 4) & 15); }

// the early exit of pressure_poisson.m:21-31 (FFT branch): dp = 0 when
// max |fhat| < 1e-14, fhat = fft2(f) (unnormalized).  The maximum of |fhat|^2
// is kept in kFmaxSlots slots (block b of the column solve raises slot
// b mod kFmaxSlots): one address raised by every block serialised its
// atomics (4.8 us per solve at 1024^2).
constexpr double kFftExit2 = 1e-28;   // (1e-14)^2, compared with |fhat|^2
constexpr int kFmaxSlots = 64;
__device__ __forceinline__ void fmax2_reset(unsigned long long* m) {
  if (blockIdx.x == 0 && threadIdx.x < kFmaxSlots) m[threadIdx.x] = 0ull;
}
// block maximum of m2 (>= 0) raised into its slot (one atomic per block);
// every thread of the block must call it
__device__ __forceinline__ void fmax2_raise(unsigned long long* m, double m2) {
  __shared__ double wmax[16];
  for (int o = 32; o > 0; o >>= 1) m2 = fmax(m2, __shfl_xor(m2, o));
  const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  if ((threadIdx.x & 63) == 0) wmax[w] = m2;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = 0.0;
    for (int k = 0; k < nw; ++k) b = fmax(b, wmax[k]);
    atomicMax(m + (blockIdx.x & (kFmaxSlots - 1)), (unsigned long long)__double_as_longlong(b));
  }
}
// every wave reads the slots (one per lane) and reduces them itself: no
// barrier; all 64 lanes of the wave must call it
__device__ __forceinline__ bool fmax2_tiny(const unsigned long long* m) {
  static_assert(kFmaxSlots == 64, "one slot per lane");
  double v = __longlong_as_double((long long)m[threadIdx.x & 63]);
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v < kFftExit2;
}

// x * exp(-+2 pi i m / 16) for a compile-time m in 0 .. 9 (sign + for INV)
template <bool INV>
__device__ __forceinline__ double2 mul_w16(double2 x, int m) {
  constexpr double c1 = 0.92387953251128675613;   // cos(pi / 8)
  constexpr double s1 = 0.38268343236508977173;   // sin(pi / 8)
  constexpr double r2 = 0.70710678118654752440;   // cos(pi / 4)
  double wr, wi;
  switch (m) {
    case 0: return x;
    case 4: return INV ? make_double2(-x.y, x.x) : make_double2(x.y, -x.x);
    case 8: return make_double2(-x.x, -x.y);
    case 1: wr = c1; wi = -s1; break;
    case 2: wr = r2; wi = -r2; break;
    case 3: wr = s1; wi = -c1; break;
    case 5: wr = -s1; wi = -c1; break;
    case 6: wr = -r2; wi = -r2; break;
    case 7: wr = -c1; wi = -s1; break;
    default: wr = -c1; wi = s1; break;   // 9
  }
  if (INV) wi = -wi;
  return make_double2(x.x * wr - x.y * wi, x.x * wi + x.y * wr);
}

template <bool INV>
__device__ __forceinline__ void dft4(double2& a0, double2& a1, double2& a2, double2& a3) {
  const double2 s02 = cadd(a0, a2), d02 = csub(a0, a2), s13 = cadd(a1, a3), d13 = csub(a1, a3);
  const double2 jd13 = INV ? make_double2(-d13.y, d13.x) : make_double2(d13.y, -d13.x);   // -+ i d13
  a0 = cadd(s02, s13);
  a1 = cadd(d02, jd13);
  a2 = csub(s02, s13);
  a3 = csub(d02, jd13);
}

// R-point DFT of x[0 .. R) in place, natural order in and out
template <int R, bool INV>
__device__ __forceinline__ void dft_reg(double2* x) {
  if (R == 2) {
    const double2 a = x[0], b = x[1];
    x[0] = cadd(a, b);
    x[1] = csub(a, b);
  } else if (R == 4) {
    dft4<INV>(x[0], x[1], x[2], x[3]);
  } else if (R == 8) {
    // n = 2 n1 + n2, k = k1 + 4 k2
    double2 e0 = x[0], e1 = x[2], e2 = x[4], e3 = x[6], o0 = x[1], o1 = x[3], o2 = x[5], o3 = x[7];
    dft4<INV>(e0, e1, e2, e3);
    dft4<INV>(o0, o1, o2, o3);
    o1 = mul_w16<INV>(o1, 2);
    o2 = mul_w16<INV>(o2, 4);
    o3 = mul_w16<INV>(o3, 6);
    x[0] = cadd(e0, o0); x[4] = csub(e0, o0);
    x[1] = cadd(e1, o1); x[5] = csub(e1, o1);
    x[2] = cadd(e2, o2); x[6] = csub(e2, o2);
    x[3] = cadd(e3, o3); x[7] = csub(e3, o3);
  } else {
    // n = 4 n1 + n2, k = k1 + 4 k2: DFT4 over n1, twiddle W16^(n2 k1), DFT4 over n2
    double2 y[16];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
      double2 a0 = x[n2], a1 = x[4 + n2], a2 = x[8 + n2], a3 = x[12 + n2];
      dft4<INV>(a0, a1, a2, a3);
      y[4 * n2 + 0] = a0;
      y[4 * n2 + 1] = mul_w16<INV>(a1, n2);
      y[4 * n2 + 2] = mul_w16<INV>(a2, 2 * n2);
      y[4 * n2 + 3] = mul_w16<INV>(a3, 3 * n2);
    }
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      double2 b0 = y[k1], b1 = y[4 + k1], b2 = y[8 + k1], b3 = y[12 + k1];
      dft4<INV>(b0, b1, b2, b3);
      x[k1] = b0;
      x[k1 + 4] = b1;
      x[k1 + 8] = b2;
      x[k1 + 12] = b3;
    }
  }
}

// One radix-R Stockham pass of an N-point transform (Ns: product of the
// earlier radices).  Thread tid owns the butterflies jq = tid + c T
// (c < 16 / R, jq < N / R); ld(s, n) gives input element n for register slot
// s = c R + m, st(s, n, x) takes output element n.  TO_LDS: st writes LDS, so
// the pass waits for every thread's loads before storing and after.
template <int E, int R, bool INV, bool TO_LDS, class LD, class ST>
__device__ __forceinline__ void fft_pass(int N, int Ns, const double2* __restrict__ tw, LD& ld, ST& st) {
  const int T = blockDim.x, tid = threadIdx.x;
  const int Q = N / R, step = N / (Ns * R);
  double2 a[E];
#pragma unroll
  for (int c = 0; c < E / R; ++c) {
    const int jq = tid + c * T;
    if (jq < Q) {
#pragma unroll
      for (int m = 0; m < R; ++m) a[c * R + m] = ld(c * R + m, jq + m * Q);
      if (Ns > 1) {
        // twiddles W^(m ks), m < R: W^ks, W^2ks, W^4ks, W^8ks from the
        // table, the others as products of these (at most three deep)
        const int ks = (jq & (Ns - 1)) * step;
        double2 t[R];
#pragma unroll
        for (int b = 1; b < R; b *= 2) {
          t[b] = tw[b * ks];
          if (INV) t[b] = conjd(t[b]);
        }
#pragma unroll
        for (int m = 3; m < R; ++m)
          if (m & (m - 1)) t[m] = cmul(t[m & (m - 1)], t[m & -m]);
#pragma unroll
        for (int m = 1; m < R; ++m) a[c * R + m] = cmul(a[c * R + m], t[m]);
      }
      dft_reg<R, INV>(a + c * R);
    }
  }
  if (TO_LDS) __syncthreads();
#pragma unroll
  for (int c = 0; c < E / R; ++c) {
    const int jq = tid + c * T;
    if (jq < Q) {
      const int k = jq & (Ns - 1);
      const int o = (jq - k) * R + k;   // (jq / Ns) Ns R + k
#pragma unroll
      for (int m = 0; m < R; ++m) st(c * R + m, o + m * Ns, a[c * R + m]);
    }
  }
  if (TO_LDS) __syncthreads();
}

template <int E, bool INV, bool TO_LDS, class LD, class ST>
__device__ __forceinline__ void fft_pass_r(int Rr, int N, int Ns, const double2* __restrict__ tw, LD& ld, ST& st) {
  if (Rr == 2) fft_pass<E, 2, INV, TO_LDS>(N, Ns, tw, ld, st);
  else if (Rr == 4) fft_pass<E, 4, INV, TO_LDS>(N, Ns, tw, ld, st);
  if constexpr (E >= 8)
    if (Rr == 8) fft_pass<E, 8, INV, TO_LDS>(N, Ns, tw, ld, st);
  if constexpr (E >= 16)
    if (Rr == 16) fft_pass<E, 16, INV, TO_LDS>(N, Ns, tw, ld, st);
}

// N-point transform, N = 2^log2N in [E, 4096], z: N-element LDS work array,
// max(64, N / E) threads.  Pass radices: the remainder 2^(log2N mod log2 E)
// first, then radix E (forward); reversed for INV, so that the forward
// transform's last pass and the inverse transform's first pass use the same
// register slots for the same elements.  OUT_LDS: the caller's st writes LDS.
template <int E, bool INV, bool OUT_LDS, class LD, class ST>
__device__ __forceinline__ void fft_reg(double2* z, int N, int log2N, const double2* __restrict__ tw, LD& ld,
                                        ST& st) {
  auto ldz = [&](int, int n) { return z[swz(n)]; };
  auto stz = [&](int, int n, double2 x) { z[swz(n)] = x; };
  constexpr int lgE = E == 16 ? 4 : (E == 8 ? 3 : 2);
  const int rem = log2N % lgE;
  const int P = log2N / lgE + (rem ? 1 : 0);
  int Ns = 1;
  for (int p = 0; p < P; ++p) {
    const bool remp = rem && (INV ? p == P - 1 : p == 0);
    const int R = remp ? (1 << rem) : E;
    const bool first = p == 0, last = p == P - 1;
    if (first && last) fft_pass_r<E, INV, OUT_LDS>(R, N, Ns, tw, ld, st);
    else if (first) fft_pass_r<E, INV, true>(R, N, Ns, tw, ld, stz);
    else if (last) fft_pass_r<E, INV, OUT_LDS>(R, N, Ns, tw, ldz, st);
    else fft_pass_r<E, INV, true>(R, N, Ns, tw, ldz, stz);
    Ns *= R;
  }
}

template <int E>
constexpr int fft_block(int) { return E == 16 ? 256 : (E == 8 ? 512 : 1024); }

// Work item of block b out of nb.  Blocks go to the 8 XCDs round-robin
// (b mod 8); XCD x gets the contiguous range of nb / 8 items (one more for
// x < nb mod 8) starting at x (nb / 8) + min(x, nb mod 8), so that the
// transposed stores (loads) of neighbouring items, which share cache lines,
// meet in the same L2.  A bijection of [0, nb).
__device__ __forceinline__ int xcd_row_pair(int b, int nb) {
  const int x = b & 7, q = nb >> 3, r = nb & 7;
  return x * q + min(x, r) + (b >> 3);
}

// Right-hand side of a forward row transform taken straight from the
// velocity: the second-order periodic divergence of k_div_per,
// f = oscale (hy (W_u(i+1, j) - W_u(i, j)) + hx (W_v(i, j+1) - W_v(i, j))),
// W = ca A + cb B (B null: ca A), u at A[i + Nx j], v at A[Nx Ny + i + Nx j]
struct DivSrc {
  const double* A;
  const double* B;
  double ca, cb, oscale, hx, hy;
  int pair;    // A, B 16-byte aligned: two elements per thread, 16-byte loads
  int order;   // 2, or 4: the fourth-order divergence of k_div_per4s
};

// rows y0 = 2b and y1 = 2b+1 packed as one complex row, half spectra out
// (power-of-two Nx in [E, 4096], at least max(64, Nx / E) threads: the
// passes leave any beyond Nx / E idle).  TR: the spectra
// are stored transposed, H[k Ny + y] (Nh x Ny), ready for the column solve.
// DIV: the rows are the divergence of D (computed into LDS before the
// transform, f is not read), else the rows of f.
template <int E, bool TR, bool DIV>
__global__ void __launch_bounds__(fft_block<E>(0)) k_fftr_r2c_rows(int Nx, int Ny, int Nh, int log2N,
                                                       const double2* __restrict__ tw, const double* __restrict__ f,
                                                       DivSrc D, double2* __restrict__ H,
                                                       unsigned long long* __restrict__ fmax2) {
  extern __shared__ double2 z[];
  fmax2_reset(fmax2);
  const int y0 = 2 * (TR ? xcd_row_pair(blockIdx.x, gridDim.x) : (int)blockIdx.x), y1 = y0 + 1;
  const bool has1 = y1 < Ny;
  auto st = [&](int, int n, double2 x) { z[swz(n)] = x; };
  if constexpr (DIV) {
    // rows ya = y0 + 1 and yb = y0 + 2 (periodic): ya = y1 when has1; the
    // v values of row ya enter both divergences.  The two rows are staged
    // in LDS first (computed inside the first pass, the 14 loads per
    // element of all its slots at once exceed the register budget), and
    // transformed in place
    const int ya = y0 + 1 == Ny ? 0 : y0 + 1, yb = ya + 1 == Ny ? 0 : ya + 1;
    const double* __restrict__ A = D.A;
    const double* __restrict__ B = D.B;
    const size_t N = (size_t)Nx * Ny;
    const size_t r0 = (size_t)y0 * Nx, ra = (size_t)ya * Nx, rb = (size_t)yb * Nx;
    auto W = [&](size_t k) { return B ? D.ca * A[k] + D.cb * B[k] : D.ca * A[k]; };
    auto W2 = [&](size_t k) {   // elements k, k + 1 (k even)
      const double2 a = *(const double2*)(A + k);
      if (!B) return make_double2(D.ca * a.x, D.ca * a.y);
      const double2 b = *(const double2*)(B + k);
      return make_double2(D.ca * a.x + D.cb * b.x, D.ca * a.y + D.cb * b.y);
    };
    for (int n = 2 * threadIdx.x; D.pair && n < Nx; n += 2 * blockDim.x) {
      const int n2 = (n + 2) & (Nx - 1);
      const double2 u0 = W2(n + r0), ua = W2(n + ra), v0 = W2(N + n + r0), va = W2(N + n + ra);
      const double u0e = W(n2 + r0);
      const double s00 = D.hy * (u0.y - u0.x) + D.hx * (va.x - v0.x);
      const double s01 = D.hy * (u0e - u0.y) + D.hx * (va.y - v0.y);
      double s10 = 0.0, s11 = 0.0;
      if (has1) {
        const double2 vb = W2(N + n + rb);
        const double uae = W(n2 + ra);
        s10 = D.hy * (ua.y - ua.x) + D.hx * (vb.x - va.x);
        s11 = D.hy * (uae - ua.y) + D.hx * (vb.y - va.y);
      }
      z[swz(n)] = make_double2(D.oscale * s00, D.oscale * s10);
      z[swz(n + 1)] = make_double2(D.oscale * s01, D.oscale * s11);
    }
    for (int n = threadIdx.x; !D.pair && n < Nx; n += blockDim.x) {
      const int np = (n + 1) & (Nx - 1);
      const double va = W(N + n + ra);
      const double s0 = D.hy * (W(np + r0) - W(n + r0)) + D.hx * (va - W(N + n + r0));
      const double s1 = has1 ? D.hy * (W(np + ra) - W(n + ra)) + D.hx * (W(N + n + rb) - va) : 0.0;
      z[swz(n)] = make_double2(D.oscale * s0, D.oscale * s1);
    }
    __syncthreads();
    auto ld = [&](int, int n) { return z[swz(n)]; };
    fft_reg<E, false, true>(z, Nx, log2N, tw, ld, st);
  } else {
    const double* f0 = f + (size_t)y0 * Nx;
    const double* f1 = f + (size_t)(has1 ? y1 : y0) * Nx;
    auto ld = [&](int, int n) { return make_double2(f0[n], has1 ? f1[n] : 0.0); };
    fft_reg<E, false, true>(z, Nx, log2N, tw, ld, st);
  }
  for (int k = threadIdx.x; k < Nh; k += blockDim.x) {
    const double2 a = z[swz(k)], b = conjd(z[swz((Nx - k) & (Nx - 1))]);
    const double2 h0 = make_double2(0.5 * (a.x + b.x), 0.5 * (a.y + b.y));
    // (a - b) / (2i) = -i/2 (a - b)
    const double2 h1 = make_double2(0.5 * (a.y - b.y), -0.5 * (a.x - b.x));
    if (TR) {
      double2* o = H + (size_t)k * Ny + y0;
      o[0] = h0;
      if (has1) o[1] = h1;
    } else {
      H[(size_t)y0 * Nh + k] = h0;
      if (has1) H[(size_t)y1 * Nh + k] = h1;
    }
  }
}

// inverse of k_fftr_r2c_rows: the two half spectra as one complex row
// a + i b, transformed and split into the real rows y0 and y1
// (TR: the spectra are read transposed, H[k Ny + y]; ACC: the rows are
// added into out instead of stored)
template <int E, bool TR, bool ACC>
__global__ void __launch_bounds__(fft_block<E>(0)) k_fftr_c2r_rows(int Nx, int Ny, int Nh, int log2N,
                                                       const double2* __restrict__ tw, const double2* __restrict__ H,
                                                       double* __restrict__ out,
                                                       const unsigned long long* __restrict__ fmax2) {
  extern __shared__ double2 z[];
  const int y0 = 2 * (TR ? xcd_row_pair(blockIdx.x, gridDim.x) : (int)blockIdx.x), y1 = y0 + 1;
  const bool has1 = y1 < Ny;
  if (fmax2_tiny(fmax2)) {
    // dp = 0: nothing to add, or zero rows
    if (!ACC)
      for (int n = threadIdx.x; n < Nx; n += blockDim.x) {
        out[(size_t)y0 * Nx + n] = 0.0;
        if (has1) out[(size_t)y1 * Nx + n] = 0.0;
      }
    return;
  }
  const size_t ks = TR ? (size_t)Ny : 1;   // stride of k
  const double2* H0 = H + (TR ? (size_t)y0 : (size_t)y0 * Nh);
  const double2* H1 = H + (TR ? (size_t)(has1 ? y1 : y0) : (size_t)(has1 ? y1 : y0) * Nh);
  double* o0 = out + (size_t)y0 * Nx;
  double* o1 = out + (size_t)(has1 ? y1 : y0) * Nx;
  // ACC: the values the last pass adds to, loaded up front into the
  // register slots of its stores (slot s = c R + m takes element
  // tid + c T + m Q, R the last pass's radix, Q = Nx / R: the remainder
  // radix when log2 Nx is not a multiple of log2 E, else E), so that the
  // loads overlap the transform instead of waiting behind the stores
  double pre0[E], pre1[E];
  if constexpr (ACC) {
    constexpr int lgE = E == 16 ? 4 : (E == 8 ? 3 : 2);
    const int rem = log2N % lgE;
    const int R = rem ? (1 << rem) : E, Q = Nx / R, T = blockDim.x;
#pragma unroll
    for (int sl = 0; sl < E; ++sl) {
      const int c = sl / R, m = sl % R, jq = threadIdx.x + c * T;
      const bool ok = c < E / R && jq < Q;
      pre0[sl] = ok ? o0[jq + m * Q] : 0.0;
      pre1[sl] = ok ? o1[jq + m * Q] : 0.0;
    }
  }
  // the full row a + i b, staged in LDS (the half spectra read once)
  for (int k = threadIdx.x; k < Nh; k += blockDim.x) {
    double2 a = H0[k * ks], b = has1 ? H1[k * ks] : make_double2(0.0, 0.0);
    if (k == 0 || 2 * k == Nx) { a.y = 0.0; b.y = 0.0; }
    z[swz(k)] = make_double2(a.x - b.y, a.y + b.x);
    if (k > 0 && 2 * k < Nx) z[swz(Nx - k)] = make_double2(a.x + b.y, b.x - a.y);   // conj(a) + i conj(b)
  }
  __syncthreads();
  auto ld = [&](int, int n) { return z[swz(n)]; };
  auto st = [&](int sl, int n, double2 x) {
    o0[n] = ACC ? pre0[sl] + x.x : x.x;
    if (has1) o1[n] = ACC ? pre1[sl] + x.y : x.y;
  };
  fft_reg<E, true, false>(z, Nx, log2N, tw, ld, st);
}

// P passes of radix E (N = E^P points) on data in registers: thread tid
// (< N / E, the active threads) owns one butterfly per pass, slots m < E
// holding elements tid + m N/E on input and on output (the last pass's
// output index is again tid + m N/E).  Consecutive passes exchange through
// an N-double LDS array, real parts first, then imaginary parts (half the LDS
// of exchanging complex values, at two more barriers per exchange).  Every
// thread of the block must call it.
template <int E, bool INV>
__device__ __forceinline__ void fft_split_passes(double2 (&a)[E], double* zr, int N, int P,
                                                 const double2* __restrict__ tw) {
  const int tid = threadIdx.x, Q = N / E;
  const bool act = tid < Q;
  int Ns = 1;
  for (int p = 0; p < P; ++p) {
    if (act) {
      if (Ns > 1) {
        const int ks = (tid & (Ns - 1)) * (N / (Ns * E));
        double2 t[E];
#pragma unroll
        for (int b = 1; b < E; b *= 2) {
          t[b] = tw[b * ks];
          if (INV) t[b] = conjd(t[b]);
        }
#pragma unroll
        for (int m = 3; m < E; ++m)
          if (m & (m - 1)) t[m] = cmul(t[m & (m - 1)], t[m & -m]);
#pragma unroll
        for (int m = 1; m < E; ++m) a[m] = cmul(a[m], t[m]);
      }
      dft_reg<E, INV>(a);
    }
    if (p == P - 1) break;
    const int k = tid & (Ns - 1);
    const int o = (tid - k) * E + k;   // output element of slot m: o + m Ns
    __syncthreads();   // the previous exchange's reads are done
    if (act) {
#pragma unroll
      for (int m = 0; m < E; ++m) zr[swz(o + m * Ns)] = a[m].x;
    }
    __syncthreads();
    if (act) {
#pragma unroll
      for (int m = 0; m < E; ++m) a[m].x = zr[swz(tid + m * Q)];
    }
    __syncthreads();
    if (act) {
#pragma unroll
      for (int m = 0; m < E; ++m) zr[swz(o + m * Ns)] = a[m].y;
    }
    __syncthreads();
    if (act) {
#pragma unroll
      for (int m = 0; m < E; ++m) a[m].y = zr[swz(tid + m * Q)];
    }
    Ns *= E;
  }
}

// k_fftr_cols_solve for Ny = E^P: forward transform, division by the
// symbol and inverse transform on one register array per thread, N / E
// threads per column, N doubles of LDS (32 KB at 4096 instead of 64 KB)
template <int E>
__global__ void __launch_bounds__(4096 / E) k_fftr_cols_solve_s(int Nx, int Ny, int P, const double2* __restrict__ tw,
                                                             const double* __restrict__ ax, const double* __restrict__ ay,
                                                             double c, double2* __restrict__ HT,
                                                             unsigned long long* __restrict__ fmax2) {
  extern __shared__ double zr[];
  const int kx = blockIdx.x, tid = threadIdx.x, Q = Ny / E;
  const bool act = tid < Q;
  double2* row = HT + (size_t)kx * Ny;
  double2 a[E];
  if (act) {
#pragma unroll
    for (int m = 0; m < E; ++m) a[m] = row[tid + m * Q];
  }
  fft_split_passes<E, false>(a, zr, Ny, P, tw);
  double m2 = 0.0;
  if (act) {
    const double scale = 1.0 / ((double)Nx * (double)Ny);
    const double axk = ax[kx];
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int ky = tid + m * Q;
      m2 = fmax(m2, a[m].x * a[m].x + a[m].y * a[m].y);
      double Ah = c * (axk + ay[ky]);
      if (kx == 0 && ky == 0) Ah = 1.0;
      const double sc = -scale / Ah;
      a[m] = make_double2(a[m].x * sc, a[m].y * sc);
    }
  }
  fmax2_raise(fmax2, m2);
  fft_split_passes<E, true>(a, zr, Ny, P, tw);
  if (act) {
#pragma unroll
    for (int m = 0; m < E; ++m) row[tid + m * Q] = a[m];
  }
}

// k_fftr_r2c_rows (transposed spectra, rows of f) for Nx = E^P with the split
// exchange: the rows y0 = 2b, y1 = 2b+1 packed as one complex row, loaded
// straight into the registers of the first pass; after the transform the
// partner element Nx - k of each k is fetched through LDS (real parts, then
// imaginary parts) to split the two half spectra.  Nx / E threads, Nx doubles
// of LDS.
// the two real rows y0, y1 transformed as one complex row (element
// k = tid + m Q in a[m]) split into their half spectra, stored transposed
// (H[k Ny + y]); zr: the Nx-double LDS array of the split exchange
template <int E>
__device__ __forceinline__ void r2c_split_store(double2 (&a)[E], double* zr, int Nx, int Ny, int Nh, int y0, bool has1,
                                                double2* __restrict__ H) {
  const int tid = threadIdx.x, Q = Nx / E;
  const bool act = tid < Q;
  // the partner Nx - k (mod Nx) of element k through LDS
  double2 b[E];
  __syncthreads();
  if (act) {
#pragma unroll
    for (int m = 0; m < E; ++m) zr[swz(tid + m * Q)] = a[m].x;
  }
  __syncthreads();
  if (act) {
#pragma unroll
    for (int m = 0; m < E; ++m) b[m].x = zr[swz((Nx - tid - m * Q) & (Nx - 1))];
  }
  __syncthreads();
  if (act) {
#pragma unroll
    for (int m = 0; m < E; ++m) zr[swz(tid + m * Q)] = a[m].y;
  }
  __syncthreads();
  if (act) {
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int k = tid + m * Q;
      b[m].y = -zr[swz((Nx - k) & (Nx - 1))];   // b = conj(z[Nx - k])
      if (k < Nh) {
        const double2 h0 = make_double2(0.5 * (a[m].x + b[m].x), 0.5 * (a[m].y + b[m].y));
        // (a - b) / (2i) = -i/2 (a - b)
        const double2 h1 = make_double2(0.5 * (a[m].y - b[m].y), -0.5 * (a[m].x - b[m].x));
        double2* o = H + (size_t)k * Ny + y0;
        o[0] = h0;
        if (has1) o[1] = h1;
      }
    }
  }
}

template <int E>
__global__ void __launch_bounds__(4096 / E) k_fftr_r2c_rows_s(int Nx, int Ny, int Nh, int P,
                                                           const double2* __restrict__ tw, const double* __restrict__ f,
                                                           double2* __restrict__ H,
                                                           unsigned long long* __restrict__ fmax2) {
  extern __shared__ double zr[];
  fmax2_reset(fmax2);
  const int y0 = 2 * xcd_row_pair(blockIdx.x, gridDim.x), y1 = y0 + 1;
  const bool has1 = y1 < Ny;
  const int tid = threadIdx.x, Q = Nx / E;
  const bool act = tid < Q;
  const double* f0 = f + (size_t)y0 * Nx;
  const double* f1 = f + (size_t)(has1 ? y1 : y0) * Nx;
  double2 a[E];
  if (act) {
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int n = tid + m * Q;
      a[m] = make_double2(f0[n], has1 ? f1[n] : 0.0);
    }
  }
  fft_split_passes<E, false>(a, zr, Nx, P, tw);
  r2c_split_store<E>(a, zr, Nx, Ny, Nh, y0, has1, H);
}

// k_fftr_r2c_rows_s with its rows the fourth-order periodic divergence of
// k_div_per4s, f = oscale (81 hy (W_u(x+1) - W_u(x)) - 3 hy (W_u(x+2) -
// W_u(x-1)) + 81 hx (W_v(y+1) - W_v(y)) - 3 hx (W_v(y+2) - W_v(y-1))),
// W = ca A + cb B, taken straight from the velocity (f is not stored), in
// the same arithmetic order: u of each of the two rows staged through the
// LDS array for its x halo, v of the rows y0-1 .. y0+3 loaded per element
template <int E, bool HASB>
__global__ void __launch_bounds__(4096 / E, 3) k_fftr_r2c_rows_sdiv4(int Nx, int Ny, int Nh, int P,
                                                                  const double2* __restrict__ tw, DivSrc D,
                                                                  double2* __restrict__ H,
                                                                  unsigned long long* __restrict__ fmax2) {
  extern __shared__ double zr[];
  fmax2_reset(fmax2);
  const int y0 = 2 * xcd_row_pair(blockIdx.x, gridDim.x), y1 = y0 + 1;
  const bool has1 = y1 < Ny;
  const int tid = threadIdx.x, Q = Nx / E;
  const bool act = tid < Q;
  const double* __restrict__ A = D.A;
  const double* __restrict__ B = D.B;
  const double ca = D.ca, cb = D.cb, hx = D.hx, hy = D.hy;
  auto W = [&](size_t k) {
    if constexpr (HASB) return ca * A[k] + cb * B[k];
    else return ca * A[k];
  };
  double2 a[E];
  // the u terms of rows y0 and y1
  for (int r = 0; r < 2; ++r) {
    const size_t row = (size_t)Nx * (r == 0 || !has1 ? y0 : y1);
    if (act) {
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int n = tid + m * Q;
        zr[n] = W(row + n);
      }
    }
    __syncthreads();
    if (act) {
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int n = tid + m * Q;
        const double um1 = zr[(n - 1) & (Nx - 1)], u0 = zr[n], up1 = zr[(n + 1) & (Nx - 1)],
                     up2 = zr[(n + 2) & (Nx - 1)];
        const double su = 81.0 * hy * (up1 - u0) - 3.0 * hy * (up2 - um1);
        if (r == 0) a[m].x = su;
        else a[m].y = has1 ? su : 0.0;
      }
    }
    __syncthreads();
  }
  // the v terms: rows y0-1 .. y0+3 of this thread's columns
  if (act) {
    const size_t N = (size_t)Nx * Ny;
    const size_t rm1 = N + (size_t)Nx * wrap(y0 - 1, Ny), r0 = N + (size_t)Nx * y0,
                 r1 = N + (size_t)Nx * wrap(y0 + 1, Ny), r2 = N + (size_t)Nx * wrap(y0 + 2, Ny),
                 r3 = N + (size_t)Nx * wrap(y0 + 3, Ny);
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int n = tid + m * Q;
      // branch-free (row y0 + 3 wraps; without a row y1 the value is
      // dropped), so that both rows' terms are formed as the loads arrive
      const double vm1 = W(rm1 + n), v0 = W(r0 + n), vp1 = W(r1 + n), vp2 = W(r2 + n), vp3 = W(r3 + n);
      a[m].x = D.oscale * (a[m].x + 81.0 * hx * (vp1 - v0) - 3.0 * hx * (vp2 - vm1));
      const double s1 = D.oscale * (a[m].y + 81.0 * hx * (vp2 - vp1) - 3.0 * hx * (vp3 - v0));
      a[m].y = has1 ? s1 : 0.0;
    }
  }
  fft_split_passes<E, false>(a, zr, Nx, P, tw);
  r2c_split_store<E>(a, zr, Nx, Ny, Nh, y0, has1, H);
}

// k_fftr_c2r_rows (transposed spectra) for Nx = E^P with the split exchange:
// the full row a + i b is staged through an Nx-double LDS array, real parts
// first, then imaginary parts, each half-spectrum value read from memory once
// (kept in registers between the two halves), then transformed in registers;
// ACC: the rows are added into out (their old values loaded before the
// transform).  Nx / E threads.
template <int E, bool ACC>
__global__ void __launch_bounds__(4096 / E) k_fftr_c2r_rows_s(int Nx, int Ny, int Nh, int P,
                                                           const double2* __restrict__ tw, const double2* __restrict__ H,
                                                           double* __restrict__ out,
                                                           const unsigned long long* __restrict__ fmax2) {
  extern __shared__ double zr[];
  const int y0 = 2 * xcd_row_pair(blockIdx.x, gridDim.x), y1 = y0 + 1;
  const bool has1 = y1 < Ny;
  const int tid = threadIdx.x, T = blockDim.x, Q = Nx / E;
  const bool act = tid < Q;
  double* o0 = out + (size_t)y0 * Nx;
  double* o1 = out + (size_t)(has1 ? y1 : y0) * Nx;
  if (fmax2_tiny(fmax2)) {
    // dp = 0: nothing to add, or zero rows
    if (!ACC)
      for (int n = tid; n < Nx; n += T) {
        o0[n] = 0.0;
        if (has1) o1[n] = 0.0;
      }
    return;
  }
  const double2* H0 = H + y0;
  const double2* H1 = H + (has1 ? y1 : y0);
  // half-spectrum values k = tid + s T, s < KS (k < Nh), loaded once
  constexpr int KS = E / 2 + 1;
  double2 ha[KS], hb[KS];
#pragma unroll
  for (int s = 0; s < KS; ++s) {
    const int k = tid + s * T;
    ha[s] = hb[s] = make_double2(0.0, 0.0);
    if (k < Nh) {
      ha[s] = H0[(size_t)k * Ny];
      hb[s] = has1 ? H1[(size_t)k * Ny] : make_double2(0.0, 0.0);
      if (k == 0 || 2 * k == Nx) { ha[s].y = 0.0; hb[s].y = 0.0; }
    }
  }
  double2 pre[E];
  if (ACC && act) {
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int n = tid + m * Q;
      pre[m] = make_double2(o0[n], has1 ? o1[n] : 0.0);
    }
  }
  double2 a[E];
  // z[k] = a + i b, z[Nx - k] = conj(a) + i conj(b): real parts, then imaginary
#pragma unroll
  for (int s = 0; s < KS; ++s) {
    const int k = tid + s * T;
    if (k < Nh) {
      zr[swz(k)] = ha[s].x - hb[s].y;
      if (k > 0 && 2 * k < Nx) zr[swz(Nx - k)] = ha[s].x + hb[s].y;
    }
  }
  __syncthreads();
  if (act) {
#pragma unroll
    for (int m = 0; m < E; ++m) a[m].x = zr[swz(tid + m * Q)];
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < KS; ++s) {
    const int k = tid + s * T;
    if (k < Nh) {
      zr[swz(k)] = ha[s].y + hb[s].x;
      if (k > 0 && 2 * k < Nx) zr[swz(Nx - k)] = hb[s].x - ha[s].y;
    }
  }
  __syncthreads();
  if (act) {
#pragma unroll
    for (int m = 0; m < E; ++m) a[m].y = zr[swz(tid + m * Q)];
  }
  fft_split_passes<E, true>(a, zr, Nx, P, tw);
  if (act) {
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int n = tid + m * Q;
      o0[n] = ACC ? pre[m].x + a[m].x : a[m].x;
      if (has1) o1[n] = ACC ? pre[m].y + a[m].y : a[m].y;
    }
  }
}

// one block per kx: forward FFT along y, division by -Ahat in registers,
// inverse FFT along y, in place on HT (Nh x Ny) (power-of-two Ny in
// [E, 4096], max(64, Ny / E) threads)
template <int E>
__global__ void __launch_bounds__(fft_block<E>(0)) k_fftr_cols_solve(int Nx, int Ny, int log2N, const double2* __restrict__ tw,
                                                         const double* __restrict__ ax, const double* __restrict__ ay,
                                                         double c, double2* __restrict__ HT,
                                                         unsigned long long* __restrict__ fmax2) {
  extern __shared__ double2 z[];
  const int kx = blockIdx.x;
  double2* row = HT + (size_t)kx * Ny;
  const double scale = 1.0 / ((double)Nx * (double)Ny);
  const double axk = ax[kx];
  double2 v[E];
  double m2 = 0.0;
  auto ldg = [&](int, int n) { return row[n]; };
  auto stv = [&](int s, int ky, double2 x) {
    m2 = fmax(m2, x.x * x.x + x.y * x.y);
    double Ah = c * (axk + ay[ky]);
    if (kx == 0 && ky == 0) Ah = 1.0;
    const double sc = -scale / Ah;
    v[s] = make_double2(x.x * sc, x.y * sc);
  };
  fft_reg<E, false, false>(z, Ny, log2N, tw, ldg, stv);
  fmax2_raise(fmax2, m2);
  __syncthreads();
  auto ldv = [&](int s, int) { return v[s]; };
  auto stg = [&](int, int n, double2 x) { row[n] = x; };
  fft_reg<E, true, false>(z, Ny, log2N, tw, ldv, stg);
}

// rows y0 = 2b and y1 = 2b+1 packed as one complex row; half spectra out
template <int C>
__global__ void __launch_bounds__(C == 1 ? 1024 : 256) k_fft_r2c_rows(int Nx, int Ny, int Nh, bool pow2, int log2N,
                                                     const double2* __restrict__ tw, const double* __restrict__ f,
                                                     double2* __restrict__ H, unsigned long long* __restrict__ fmax2) {
  extern __shared__ double2 sm[];
  double2* z = sm;
  double2* scratch = sm + Nx;
  fmax2_reset(fmax2);
  const int y0 = 2 * blockIdx.x, y1 = y0 + 1;
  const bool has1 = y1 < Ny;
  for (int n = threadIdx.x; n < Nx; n += blockDim.x)
    z[n] = make_double2(f[(size_t)y0 * Nx + n], has1 ? f[(size_t)y1 * Nx + n] : 0.0);
  __syncthreads();
  lds_transform<false, C>(z, scratch, Nx, pow2, log2N, tw);
  for (int k = threadIdx.x; k < Nh; k += blockDim.x) {
    const double2 a = z[k], b = conjd(z[(Nx - k) % Nx]);
    H[(size_t)y0 * Nh + k] = make_double2(0.5 * (a.x + b.x), 0.5 * (a.y + b.y));
    // (a - b) / (2i) = -i/2 (a - b)
    if (has1) H[(size_t)y1 * Nh + k] = make_double2(0.5 * (a.y - b.y), -0.5 * (a.x - b.x));
  }
}

template <int C>
__global__ void __launch_bounds__(C == 1 ? 1024 : 256) k_fft_c2r_rows(int Nx, int Ny, int Nh, bool pow2, int log2N,
                                                     const double2* __restrict__ tw, const double2* __restrict__ H,
                                                     double* __restrict__ out,
                                                     const unsigned long long* __restrict__ fmax2) {
  extern __shared__ double2 sm[];
  double2* z = sm;
  double2* scratch = sm + Nx;
  const int y0 = 2 * blockIdx.x, y1 = y0 + 1;
  const bool has1 = y1 < Ny;
  if (fmax2_tiny(fmax2)) {
    for (int n = threadIdx.x; n < Nx; n += blockDim.x) {
      out[(size_t)y0 * Nx + n] = 0.0;
      if (has1) out[(size_t)y1 * Nx + n] = 0.0;
    }
    return;
  }
  for (int k = threadIdx.x; k < Nx; k += blockDim.x) {
    double2 a, b;
    if (k < Nh) {
      a = H[(size_t)y0 * Nh + k];
      b = has1 ? H[(size_t)y1 * Nh + k] : make_double2(0.0, 0.0);
    } else {
      a = conjd(H[(size_t)y0 * Nh + (Nx - k)]);
      b = has1 ? conjd(H[(size_t)y1 * Nh + (Nx - k)]) : make_double2(0.0, 0.0);
    }
    if (k == 0 || (2 * k == Nx)) { a.y = 0.0; b.y = 0.0; }
    // Z = a + i b
    z[k] = make_double2(a.x - b.y, a.y + b.x);
  }
  __syncthreads();
  lds_transform<true, C>(z, scratch, Nx, pow2, log2N, tw);
  for (int n = threadIdx.x; n < Nx; n += blockDim.x) {
    out[(size_t)y0 * Nx + n] = z[n].x;
    if (has1) out[(size_t)y1 * Nx + n] = z[n].y;
  }
}

// complex transpose (rows x cols) -> (cols x rows), 32x32 tiles
__global__ void k_transpose_c(const double2* __restrict__ in, double2* __restrict__ out, int rows, int cols) {
  __shared__ double2 tile[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  for (int dy = threadIdx.y; dy < 32; dy += blockDim.y) {
    const int r = by + dy, c = bx + threadIdx.x;
    if (r < rows && c < cols) tile[dy][threadIdx.x] = in[(size_t)r * cols + c];
  }
  __syncthreads();
  for (int dy = threadIdx.y; dy < 32; dy += blockDim.y) {
    const int c = bx + dy, r = by + threadIdx.x;
    if (r < rows && c < cols) out[(size_t)c * rows + r] = tile[threadIdx.x][dy];
  }
}

// one block per kx: forward FFT along y, divide by -Ahat, inverse FFT along y
template <int C>
__global__ void __launch_bounds__(C == 1 ? 1024 : 256) k_fft_cols_solve(int Nx, int Ny, bool pow2, int log2N,
                                                       const double2* __restrict__ tw, const double* __restrict__ ax,
                                                       const double* __restrict__ ay, double c,
                                                       double2* __restrict__ HT,
                                                       unsigned long long* __restrict__ fmax2) {
  extern __shared__ double2 sm[];
  double2* z = sm;
  double2* scratch = sm + Ny;
  const int kx = blockIdx.x;
  double2* row = HT + (size_t)kx * Ny;
  for (int n = threadIdx.x; n < Ny; n += blockDim.x) z[n] = row[n];
  __syncthreads();
  lds_transform<false, C>(z, scratch, Ny, pow2, log2N, tw);
  double m2 = 0.0;
  for (int ky = threadIdx.x; ky < Ny; ky += blockDim.x) m2 = fmax(m2, z[ky].x * z[ky].x + z[ky].y * z[ky].y);
  fmax2_raise(fmax2, m2);
  const double scale = 1.0 / ((double)Nx * (double)Ny);
  for (int ky = threadIdx.x; ky < Ny; ky += blockDim.x) {
    double Ah = c * (ax[kx] + ay[ky]);
    if (kx == 0 && ky == 0) Ah = 1.0;
    const double s = -scale / Ah;
    z[ky] = make_double2(z[ky].x * s, z[ky].y * s);
  }
  __syncthreads();
  lds_transform<true, C>(z, scratch, Ny, pow2, log2N, tw);
  for (int n = threadIdx.x; n < Ny; n += blockDim.x) row[n] = z[n];
}

// ===========================================================================
// Laplacian and Jacobi-PCG (cg.F:158-280 stopping rule ||r||^2/||f||^2 <= tol^2)
// ===========================================================================
__global__ void __launch_bounds__(256) k_lap(DLap L, const double* __restrict__ p, double* __restrict__ q) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= L.nx || j >= L.ny) return;
  q[(size_t)i + (size_t)L.nx * j] = lap_at(L, p, i, j);
}

// CG state on the device: s[0]=rz, s[1]=nf2 (|f|^2), s[2]=alpha, s[3]=beta,
// s[4]=relres^2, s[5]=tol^2 ; flag[0]=done, flag[1]=iterations, flag[2]=failed
struct CGState {
  double* s;
  int* flag;
  double* part;   // per-block partial sums (2 per block)
  int nblk;
  int maxit;
  double* e;      // two-launch iteration: scalar slots and p.q partials (4 + nblk)
};

// 1-D grid over n points, 256 threads: q = A p on (i, j) = (idx % nx, idx / nx)
__global__ void __launch_bounds__(256) k_cg_apply(DLap L, CGState S, const double* __restrict__ p,
                                                 double* __restrict__ q) {
  __shared__ double red[4];
  if (S.flag[0]) return;
  const size_t n = (size_t)L.nx * L.ny;
  double acc = 0.0;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < n; idx += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(idx % L.nx), j = (int)(idx / L.nx);
    const double a = lap_at(L, p, i, j);
    q[idx] = a;
    acc += p[idx] * a;
  }
  const double s = block_sum(acc, red);
  if (threadIdx.x == 0) S.part[blockIdx.x] = s;
}

__global__ void k_cg_alpha(CGState S) {
  __shared__ double red[4];
  if (S.flag[0]) return;
  double acc = 0.0;
  for (int b = threadIdx.x; b < S.nblk; b += blockDim.x) acc += S.part[b];
  const double pq = block_sum(acc, red);
  if (threadIdx.x == 0) S.s[2] = S.s[0] / pq;
}

__global__ void __launch_bounds__(256) k_cg_update(DLap L, CGState S, const double* __restrict__ p,
                                                  const double* __restrict__ q, double* __restrict__ x,
                                                  double* __restrict__ r, double* __restrict__ z) {
  __shared__ double red[4];
  if (S.flag[0]) return;
  const size_t n = (size_t)L.nx * L.ny;
  const double alpha = S.s[2];
  double rr = 0.0, rz = 0.0;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < n; idx += (size_t)gridDim.x * blockDim.x) {
    x[idx] += alpha * p[idx];
    const double rn = r[idx] - alpha * q[idx];
    r[idx] = rn;
    const double zn = L.dinv[idx] * rn;
    z[idx] = zn;
    rr += rn * rn;
    rz += rn * zn;
  }
  const double a = block_sum(rr, red);
  const double b = block_sum(rz, red);
  if (threadIdx.x == 0) { S.part[2 * blockIdx.x] = a; S.part[2 * blockIdx.x + 1] = b; }
}

__global__ void k_cg_beta(CGState S) {
  __shared__ double red[4];
  if (S.flag[0]) return;
  double a = 0.0, b = 0.0;
  for (int k = threadIdx.x; k < S.nblk; k += blockDim.x) { a += S.part[2 * k]; b += S.part[2 * k + 1]; }
  const double rr = block_sum(a, red);
  const double rz = block_sum(b, red);
  if (threadIdx.x == 0) {
    const int it = S.flag[1] + 1;
    S.flag[1] = it;
    const double rel = rr / S.s[1];
    S.s[4] = rel;
    if (rel <= S.s[5]) {
      S.flag[0] = 1;
    } else if (it >= S.maxit) {
      S.flag[0] = 1;
      S.flag[2] = 1;
    }
    S.s[3] = rz / S.s[0];
    S.s[0] = rz;
  }
}

__global__ void __launch_bounds__(256) k_cg_dir(CGState S, size_t n, const double* __restrict__ z,
                                               double* __restrict__ p) {
  if (S.flag[0]) return;
  const double beta = S.s[3];
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < n; idx += (size_t)gridDim.x * blockDim.x)
    p[idx] = z[idx] + beta * p[idx];
}

// ---- the same iteration in two launches --------------------------------
// Each launch first reduces the partial sums the previous launch left, in
// every block (the same order in all of them, so every block forms the same
// scalars; no inter-block synchronisation inside a launch):
// (1) k_cg2_apply: from the partials of r.r and r.z, the stopping rule of the
// previous iteration and beta = rz / rz_prev (0 before the first iteration:
// p = z); then p = z + beta p, q = A p = A z + beta A p_old and the partials
// of p.q.  p is double-buffered (p_in -> p_out): A p_old reads p at the
// stencil neighbours, which other threads of the launch rewrite.
// (2) k_cg2_update: alpha = rz / p.q from those partials, then x += alpha p,
// r -= alpha q, z = dinv r and the partials of r.r and r.z.
// Scalars pass between the launches through two slots written by block 0
// of one launch and read by the next: e[0..1] (rz_prev, iterations done) for
// k_cg2_apply, e[2..3] (rz, iterations done) for k_cg2_update; e + 4 holds
// the p.q partials.  k_cg2_check applies the stopping rule at the end of a
// replayed chunk.
// the block sum handed to every thread of the block
__device__ __forceinline__ double block_sum_all(double v, double* red, double* out) {
  const double s = block_sum(v, red);
  if (threadIdx.x == 0) *out = s;
  __syncthreads();
  return *out;
}

// the stopping rule after an update (cg.F:247): true when the solve ends;
// block 0 records the outcome
__device__ __forceinline__ bool cg2_stop(const CGState& S, double rr, int it) {
  const double rel = rr / S.s[1];
  const bool conv = rel <= S.s[5], fail = !conv && it >= S.maxit;
  if ((conv || fail) && blockIdx.x == 0 && threadIdx.x == 0) {
    S.s[4] = rel;
    S.flag[1] = it;
    S.flag[2] = fail ? 1 : 0;
    S.flag[0] = 1;
  }
  return conv || fail;
}

// kron(Y, X) at (f, g) with KA entries per row of Y and KB of X known at
// compile time: the index and weight loads of all entries issue together
template <int KA, int KB>
__device__ __forceinline__ double kron_fixed_at(const DKron& K, const double* __restrict__ in, int f, int g) {
  const int ld = K.X.ncols;
  int cy[KA], cx[KB];
  double wy[KA], wx[KB];
#pragma unroll
  for (int a = 0; a < KA; ++a) { cy[a] = K.Y.col[(size_t)g * KA + a]; wy[a] = K.Y.val[(size_t)g * KA + a]; }
#pragma unroll
  for (int b = 0; b < KB; ++b) { cx[b] = K.X.col[(size_t)f * KB + b]; wx[b] = K.X.val[(size_t)f * KB + b]; }
  double acc = 0.0;
#pragma unroll
  for (int a = 0; a < KA; ++a) {
    const double* row = in + (size_t)cy[a] * ld;
    double s = 0.0;
#pragma unroll
    for (int b = 0; b < KB; ++b) s += wx[b] * row[cx[b]];
    acc += wy[a] * s;
  }
  return acc;
}

// A x at (i, j); TRI: the pressure Laplacian's factor widths
// (kron(diag, tridiagonal) + kron(tridiagonal, diag)) fixed at compile time
template <bool TRI>
__device__ __forceinline__ double lap_cg2(const DLap& L, const double* __restrict__ x, int i, int j) {
  if (TRI) return kron_fixed_at<1, 3>(L.L1, x, i, j) + kron_fixed_at<3, 1>(L.L2, x, i, j);
  return lap_at(L, x, i, j);
}

// ONE: the grid covers every cell once (one cell per thread), so the cell's
// loads and stencils are issued before the reduction of the partials and
// overlap its latency; otherwise a grid-stride loop after the reduction.
// A p = A z + beta A p_old (linearity: A z and A p_old need no beta)
template <bool ONE, bool TRI>
__global__ void __launch_bounds__(256) k_cg2_apply(DLap L, CGState S, const double* __restrict__ z,
                                                  const double* __restrict__ p_in, double* __restrict__ p_out,
                                                  double* __restrict__ q) {
  __shared__ double red[4];
  __shared__ double bc;
  if (S.flag[0]) return;
  const size_t n = (size_t)L.nx * L.ny;
  const size_t idx0 = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  double zc = 0.0, pc = 0.0, az = 0.0, ap = 0.0;
  if (ONE && idx0 < n) {
    const int i = (int)(idx0 % L.nx), j = (int)(idx0 / L.nx);
    zc = z[idx0];
    pc = p_in[idx0];
    az = lap_cg2<TRI>(L, z, i, j);
    ap = lap_cg2<TRI>(L, p_in, i, j);
  }
  const int it = (int)S.e[1];
  double ta = 0.0, tb = 0.0;
  for (int k = threadIdx.x; k < S.nblk; k += blockDim.x) { ta += S.part[2 * k]; tb += S.part[2 * k + 1]; }
  const double rr = block_sum_all(ta, red, &bc);
  const double rz = block_sum_all(tb, red, &bc);
  double beta = 0.0;
  if (it > 0) {
    if (cg2_stop(S, rr, it)) return;
    beta = rz / S.e[0];
  }
  double acc = 0.0;
  if (ONE) {
    if (idx0 < n) {
      const double pn = zc + beta * pc;
      p_out[idx0] = pn;
      const double a = az + beta * ap;
      q[idx0] = a;
      acc = pn * a;
    }
  } else {
    for (size_t idx = idx0; idx < n; idx += (size_t)gridDim.x * blockDim.x) {
      const int i = (int)(idx % L.nx), j = (int)(idx / L.nx);
      const double pn = z[idx] + beta * p_in[idx];
      p_out[idx] = pn;
      const double a = lap_cg2<TRI>(L, z, i, j) + beta * lap_cg2<TRI>(L, p_in, i, j);
      q[idx] = a;
      acc += pn * a;
    }
  }
  const double s = block_sum(acc, red);
  if (threadIdx.x == 0) {
    S.e[4 + blockIdx.x] = s;
    if (blockIdx.x == 0) {
      S.e[2] = rz;
      S.e[3] = (double)it;
    }
  }
}

template <bool ONE>
__global__ void __launch_bounds__(256) k_cg2_update(DLap L, CGState S, const double* __restrict__ p,
                                                   const double* __restrict__ q, double* __restrict__ x,
                                                   double* __restrict__ r, double* __restrict__ z) {
  __shared__ double red[4];
  __shared__ double bc;
  if (S.flag[0]) return;
  const size_t n = (size_t)L.nx * L.ny;
  const size_t idx0 = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  double pv = 0.0, qv = 0.0, xv = 0.0, rv = 0.0, dv = 0.0;
  if (ONE && idx0 < n) {
    pv = p[idx0];
    qv = q[idx0];
    xv = x[idx0];
    rv = r[idx0];
    dv = L.dinv[idx0];
  }
  double t = 0.0;
  for (int k = threadIdx.x; k < S.nblk; k += blockDim.x) t += S.e[4 + k];
  const double pq = block_sum_all(t, red, &bc);
  const double rz_cur = S.e[2];
  const double alpha = rz_cur / pq;
  double rr = 0.0, rz = 0.0;
  if (ONE) {
    if (idx0 < n) {
      x[idx0] = xv + alpha * pv;
      const double rn = rv - alpha * qv;
      r[idx0] = rn;
      const double zn = dv * rn;
      z[idx0] = zn;
      rr = rn * rn;
      rz = rn * zn;
    }
  } else {
    for (size_t idx = idx0; idx < n; idx += (size_t)gridDim.x * blockDim.x) {
      x[idx] += alpha * p[idx];
      const double rn = r[idx] - alpha * q[idx];
      r[idx] = rn;
      const double zn = L.dinv[idx] * rn;
      z[idx] = zn;
      rr += rn * rn;
      rz += rn * zn;
    }
  }
  const double a = block_sum(rr, red);
  const double b = block_sum(rz, red);
  if (threadIdx.x == 0) {
    S.part[2 * blockIdx.x] = a;
    S.part[2 * blockIdx.x + 1] = b;
    if (blockIdx.x == 0) {
      S.e[0] = rz_cur;
      S.e[1] = S.e[3] + 1.0;
    }
  }
}

// the stopping rule after the last update of a replayed chunk (one block)
__global__ void __launch_bounds__(256) k_cg2_check(CGState S) {
  __shared__ double red[4];
  __shared__ double bc;
  if (S.flag[0]) return;
  const int it = (int)S.e[1];
  if (it == 0) return;
  double ta = 0.0;
  for (int k = threadIdx.x; k < S.nblk; k += blockDim.x) ta += S.part[2 * k];
  const double rr = block_sum_all(ta, red, &bc);
  cg2_stop(S, rr, it);
}

// start from x0: r = f - A x0, z = dinv r, p = z; partials of f.f and r.z
__global__ void __launch_bounds__(256) k_cg_init(DLap L, CGState S, const double* __restrict__ f,
                                                const double* __restrict__ x, double* __restrict__ r,
                                                double* __restrict__ z, double* __restrict__ p, int x0_zero) {
  __shared__ double red[4];
  const size_t n = (size_t)L.nx * L.ny;
  double ff = 0.0, rz = 0.0;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < n; idx += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(idx % L.nx), j = (int)(idx / L.nx);
    const double rn = x0_zero ? f[idx] : f[idx] - lap_at(L, x, i, j);
    r[idx] = rn;
    const double zn = L.dinv[idx] * rn;
    z[idx] = zn;
    p[idx] = zn;
    ff += f[idx] * f[idx];
    rz += rn * zn;
  }
  const double a = block_sum(ff, red);
  const double b = block_sum(rz, red);
  if (threadIdx.x == 0) { S.part[2 * blockIdx.x] = a; S.part[2 * blockIdx.x + 1] = b; }
}

__global__ void k_cg_init_fin(CGState S, double tol2) {
  __shared__ double red[4];
  double a = 0.0, b = 0.0;
  for (int k = threadIdx.x; k < S.nblk; k += blockDim.x) { a += S.part[2 * k]; b += S.part[2 * k + 1]; }
  const double ff = block_sum(a, red);
  const double rz = block_sum(b, red);
  if (threadIdx.x == 0) {
    S.s[0] = rz;
    S.s[1] = ff;
    S.s[4] = 1.0;
    S.s[5] = tol2;
    S.flag[1] = 0;
    S.flag[2] = 0;
    if (S.e) {       // the two-launch iteration starts at iteration 0
      S.e[0] = rz;
      S.e[1] = 0.0;
    }
    // cg.F:222: an already tiny right-hand side returns x0 (absolute test)
    S.flag[0] = (ff <= tol2) ? 1 : 0;
  }
}

// ---- PCG with a preconditioner applied outside these kernels (the
// fast-diagonalization solve): same scalars and stopping rule, z supplied
// ---------------------------------------------------------------------------
// r = f (x = 0 set by the caller); partials of f.f; clears the done flag of
// the previous solve, which k_cgx_rz reads before k_cg_init_fin sets it anew
__global__ void __launch_bounds__(256) k_cgx_init(CGState S, size_t n, const double* __restrict__ f,
                                                 double* __restrict__ r) {
  __shared__ double red[4];
  if (blockIdx.x == 0 && threadIdx.x == 0) S.flag[0] = 0;
  double ff = 0.0;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < n; idx += (size_t)gridDim.x * blockDim.x) {
    const double fn = f[idx];
    r[idx] = fn;
    ff += fn * fn;
  }
  const double a = block_sum(ff, red);
  if (threadIdx.x == 0) S.part[2 * blockIdx.x] = a;
}

// partials of r.z; on the first call also p = z
__global__ void __launch_bounds__(256) k_cgx_rz(CGState S, size_t n, const double* __restrict__ r,
                                               const double* __restrict__ z, double* __restrict__ p, int first) {
  __shared__ double red[4];
  if (S.flag[0]) return;
  double rz = 0.0;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < n; idx += (size_t)gridDim.x * blockDim.x) {
    const double zn = z[idx];
    rz += r[idx] * zn;
    if (first) p[idx] = zn;
  }
  const double b = block_sum(rz, red);
  if (threadIdx.x == 0) S.part[2 * blockIdx.x + 1] = b;
}

// x += alpha p, r -= alpha q; partials of r.r
__global__ void __launch_bounds__(256) k_cgx_update(CGState S, size_t n, const double* __restrict__ p,
                                                   const double* __restrict__ q, double* __restrict__ x,
                                                   double* __restrict__ r) {
  __shared__ double red[4];
  if (S.flag[0]) return;
  const double alpha = S.s[2];
  double rr = 0.0;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < n; idx += (size_t)gridDim.x * blockDim.x) {
    x[idx] += alpha * p[idx];
    const double rn = r[idx] - alpha * q[idx];
    r[idx] = rn;
    rr += rn * rn;
  }
  const double a = block_sum(rr, red);
  if (threadIdx.x == 0) S.part[2 * blockIdx.x] = a;
}

// iteration count and the convergence test of cg.F:247 on ||r||^2
__global__ void k_cgx_test(CGState S) {
  __shared__ double red[4];
  if (S.flag[0]) return;
  double a = 0.0;
  for (int k = threadIdx.x; k < S.nblk; k += blockDim.x) a += S.part[2 * k];
  const double rr = block_sum(a, red);
  if (threadIdx.x == 0) {
    const int it = S.flag[1] + 1;
    S.flag[1] = it;
    const double rel = rr / S.s[1];
    S.s[4] = rel;
    if (rel <= S.s[5]) {
      S.flag[0] = 1;
    } else if (it >= S.maxit) {
      S.flag[0] = 1;
      S.flag[2] = 1;
    }
  }
}

// beta = (r.z)_new / (r.z)_old
__global__ void k_cgx_beta(CGState S) {
  __shared__ double red[4];
  if (S.flag[0]) return;
  double b = 0.0;
  for (int k = threadIdx.x; k < S.nblk; k += blockDim.x) b += S.part[2 * k + 1];
  const double rz = block_sum(b, red);
  if (threadIdx.x == 0) {
    S.s[3] = rz / S.s[0];
    S.s[0] = rz;
  }
}

// one workgroup per system: the whole PCG loop in one launch (small systems;
// vectors in global memory, reductions in LDS)
__global__ void __launch_bounds__(1024) k_cg_single(DLap L, const double* __restrict__ f, double* __restrict__ x,
                                                   double* __restrict__ r, double* __restrict__ z,
                                                   double* __restrict__ p, double* __restrict__ q, double tol2,
                                                   int maxit, int x0_zero, int* __restrict__ out_it,
                                                   double* __restrict__ out_rel) {
  __shared__ double red[16];
  __shared__ double bc[2];
  const int n = L.nx * L.ny;
  const int T = blockDim.x;
  double ff = 0.0, rz = 0.0;
  for (int idx = threadIdx.x; idx < n; idx += T) {
    const int i = idx % L.nx, j = idx / L.nx;
    const double rn = x0_zero ? f[idx] : f[idx] - lap_at(L, x, i, j);
    if (x0_zero) x[idx] = 0.0;
    r[idx] = rn;
    const double zn = L.dinv[idx] * rn;
    z[idx] = zn;
    p[idx] = zn;
    ff += f[idx] * f[idx];
    rz += rn * zn;
  }
  ff = block_sum(ff, red);
  rz = block_sum(rz, red);
  if (threadIdx.x == 0) { bc[0] = ff; bc[1] = rz; }
  __syncthreads();
  ff = bc[0];
  rz = bc[1];
  if (ff <= tol2) {
    if (threadIdx.x == 0) { *out_it = 0; *out_rel = 0.0; }
    return;
  }
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
    double rr = 0.0, rzn = 0.0;
    for (int idx = threadIdx.x; idx < n; idx += T) {
      x[idx] += alpha * p[idx];
      const double rn = r[idx] - alpha * q[idx];
      r[idx] = rn;
      const double zn = L.dinv[idx] * rn;
      z[idx] = zn;
      rr += rn * rn;
      rzn += rn * zn;
    }
    rr = block_sum(rr, red);
    rzn = block_sum(rzn, red);
    if (threadIdx.x == 0) { bc[0] = rr; bc[1] = rzn; }
    __syncthreads();
    ++it;
    rel = bc[0] / ff;
    if (rel <= tol2 || it >= maxit) break;
    const double beta = bc[1] / rz;
    rz = bc[1];
    for (int idx = threadIdx.x; idx < n; idx += T) p[idx] = z[idx] + beta * p[idx];
    __syncthreads();
  }
  if (threadIdx.x == 0) { *out_it = it; *out_rel = rel; }
}

// ===========================================================================
// elementwise and reductions
// ===========================================================================
__global__ void k_axpy(size_t n, double a, const double* __restrict__ x, double* __restrict__ y) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    y[i] += a * x[i];
}

// partial[b*4 + {0,1,2,3}] = max|x|, sum(w x), sum(w x^2), (unused) over a
// range with per-point weight w(i, j) = wx[i] * wy[j] (w = 1 if wx null)
__global__ void __launch_bounds__(256) k_reduce_stats(const double* __restrict__ x, int nx, int ny,
                                                     const double* __restrict__ wx, const double* __restrict__ wy,
                                                     double* __restrict__ part) {
  __shared__ double red[4];
  const size_t n = (size_t)nx * ny;
  double mx = 0.0, s1 = 0.0, s2 = 0.0;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < n; idx += (size_t)gridDim.x * blockDim.x) {
    const double xv = x[idx];
    mx = fmax(mx, fabs(xv));
    const double w = wx ? wx[idx % nx] * wy[idx / nx] : 1.0;
    s1 += w * xv;
    s2 += w * xv * xv;
  }
  const double a = block_max(mx, red);
  const double b = block_sum(s1, red);
  const double c = block_sum(s2, red);
  if (threadIdx.x == 0) { part[blockIdx.x * 4 + 0] = a; part[blockIdx.x * 4 + 1] = b; part[blockIdx.x * 4 + 2] = c; }
}

// ===========================================================================
// launchers
// ===========================================================================
static inline dim3 grid2(int nx, int ny, dim3 b) { return dim3((nx + b.x - 1) / b.x, (ny + b.y - 1) / b.y); }
// null or 16-byte aligned (the two-cells-per-thread kernels load and store
// the pairs at even offsets 16 bytes at a time)
static inline bool aligned16(const void* p) { return ((uintptr_t)p & 15) == 0; }

thread_local KernelTimer* g_ktimer = nullptr;
thread_local KernelTimer* g_fft_timer[4] = {nullptr, nullptr, nullptr, nullptr};

// launch one kernel of the FFT solve, with kernel timestamps into timer k
// while profiling
template <typename F, typename... Args>
static inline void fft_launch(int k, F kernel, dim3 grid, dim3 block, unsigned shmem, hipStream_t st, Args... args) {
  if (KernelTimer* t = g_fft_timer[k]) {
    hipExtLaunchKernelGGL(kernel, grid, block, shmem, st, t->launches == 0 ? t->start : nullptr, t->stop, 0u, args...);
    ++t->launches;
  } else {
    hipLaunchKernelGGL(kernel, grid, block, shmem, st, args...);
  }
}



// rows per strip of the row-streaming kernels: 64, halved (down to 8) while
// the grid would give fewer than 4096 strips (one wave each, 16 per CU)
static int strip_height(int Nx, int Ny) {
  const long sx = (Nx + kStripW - 1) / kStripW;
  int H = 64;
  while (H > 8 && sx * ((Ny + H - 1) / H) < 4096) H /= 2;
  return H;
}

static PerGrid per_grid(int order, int Nx, int Ny, double hx, double hy, double invRe) {
  return PerGrid{Nx, Ny, hx, hy, invRe, 1.0 / hx, 1.0 / hy, 1.0 / ((order == 2 ? 1.0 : 72.0) * hx * hy)};
}

void launch_rhs_generic(hipStream_t st, const DComp* dCu, const DComp* dCv, int nux, int nuy, int nvx, int nvy,
                        const double* u, const double* v, const double* p, int mode, double* out_u, double* out_v,
                        const StageCombine& scu, const StageCombine& scv, double* vt_u, double* vt_v) {
  dim3 b(64, 4);
  launch_timed(k_rhs_generic<0>, grid2(nux, nuy, b), b, 0, st, dCu, u, v, p, mode, out_u, scu, vt_u);
  launch_timed(k_rhs_generic<1>, grid2(nvx, nvy, b), b, 0, st, dCv, u, v, p, mode, out_v, scv, vt_v);
}

int rhs_split_rows() { return kRhsRows; }

void launch_rhs_split(hipStream_t st, const DComp* dCu, const DComp* dCv, const RTile* tiles, int ntiles,
                      const double* u, const double* v, const double* p, int mode, double* out_u, double* out_v,
                      const StageCombine& scu, const StageCombine& scv, double* vt_u, double* vt_v) {
  launch_timed(k_rhs_split, dim3(ntiles), dim3(256), 0, st, dCu, dCv, tiles, u, v, p, mode, out_u, out_v, scu, scv,
               vt_u, vt_v);
}

void launch_div_generic(hipStream_t st, const DComp* dCu, const DComp* dCv, int npx, int npy, const double* A,
                        const double* B, double ca, double cb, double ycoef, double oscale, int nu, double* f) {
  dim3 b(64, 4);
  launch_timed(k_div_generic, grid2(npx, npy, b), b, 0, st, dCu, dCv, npx, npy, A, B, ca, cb, ycoef, oscale,
                     nu, f);
}

void launch_sub_grad_generic(hipStream_t st, const DComp* dCu, const DComp* dCv, int nu, int nv, const double* k1,
                             const double* p, double* tV) {
  launch_timed(k_sub_grad_generic, dim3((nu + nv + 255) / 256), dim3(256), 0, st, dCu, dCv, nu, nv, k1, p, tV);
}

void launch_proj_generic(hipStream_t st, const DComp* dCu, const DComp* dCv, int nux, int nuy, int nvx, int nvy,
                         int nu, const double* Vn, const double* Vt, const double* dp, double dt, double c,
                         double* V) {
  dim3 b(64, 4);
  launch_timed(k_proj_generic<0>, grid2(nux, nuy, b), b, 0, st, dCu, Vn, Vt, dp, dt, c, V);
  launch_timed(k_proj_generic<1>, grid2(nvx, nvy, b), b, 0, st, dCv, Vn + nu, Vt ? Vt + nu : nullptr, dp, dt,
                     c, V + nu);
}

static bool k_all_aligned(const StageCombine& sc) {
  for (int q = 0; q < sc.n; ++q)
    if (!aligned16(sc.k[q])) return false;
  return true;
}

bool rhs_per_dual_ok(int order, int Nx) { return order == 4 ? Nx % kStripW2 == 0 : Nx % 2 == 0; }

void launch_rhs_per(hipStream_t st, int order, int Nx, int Ny, double hx, double hy, double invRe, const double* u,
                    const double* v, const double* p, int mode, double* out, const StageCombine& sc, double* vt,
                    double* out2) {
  const PerGrid G = per_grid(order, Nx, Ny, hx, hy, invRe);
  if (out2 && !(rhs_per_dual_ok(order, Nx) && aligned16(u) && aligned16(v) && aligned16(out) && aligned16(vt) &&
                aligned16(out2) && k_all_aligned(sc)))
    throw std::runtime_error("right-hand side with a second output: Nx even (a multiple of 128 at fourth order), "
                             "aligned arrays");
  const int H = strip_height(Nx, Ny);
  dim3 b(64, 4);
  bool pair = order == 2 && Nx % 2 == 0 && aligned16(u) && aligned16(v) && aligned16(out) && aligned16(vt);
  for (int q = 0; q < sc.n; ++q) pair = pair && aligned16(sc.k[q]);
  if (pair && out2)
    launch_timed(k_rhs_per2x<true, true>, grid2(Nx / 2, Ny, b), b, 0, st, G, u, v, p, mode, out, sc, vt, out2);
  else if (pair)   // two cells per thread: 9.8-10.0 against 11.7 us at 1024^2
    launch_timed(k_rhs_per2x<true, false>, grid2(Nx / 2, Ny, b), b, 0, st, G, u, v, p, mode, out, sc, vt, out2);
  else if (order == 2)
    launch_timed(k_rhs_per<2, true>, grid2(Nx, Ny, b), b, 0, st, G, u, v, p, mode, out, sc, vt);
  else if (Nx % kStripW2 == 0 && aligned16(u) && aligned16(v) && aligned16(out) && aligned16(vt) &&
           k_all_aligned(sc)) {
    // 64-row strips: at 4096^2 the 2048 strips are all resident (2 per SIMD)
    const int H2 = 64;
    if (out2)
      launch_timed(k_rhs_per4s2<true, true>, dim3(Nx / kStripW2, (Ny + H2 - 1) / H2), dim3(64), 0, st, G, u, v, p, mode,
                   out, sc, vt, H2, out2);
    else
      launch_timed(k_rhs_per4s2<true, false>, dim3(Nx / kStripW2, (Ny + H2 - 1) / H2), dim3(64), 0, st, G, u, v, p,
                   mode, out, sc, vt, H2, out2);
  } else
    launch_timed(k_rhs_per4s<true>, dim3((Nx + kStripW - 1) / kStripW, (Ny + H - 1) / H), dim3(kStripW), 0, st, G, u, v, p,
                 mode, out, sc, vt, H);
}

void launch_div_per(hipStream_t st, int order, int Nx, int Ny, double hx, double hy, const double* A, const double* B,
                    double ca, double cb, double oscale, double* f) {
  const PerGrid G = per_grid(order, Nx, Ny, hx, hy, 0.0);
  const int H = strip_height(Nx, Ny);
  dim3 b(64, 4);
  if (order == 2)
    launch_timed(k_div_per<2, true>, grid2(Nx, Ny, b), b, 0, st, G, A, B, ca, cb, oscale, f);
  else
    launch_timed(k_div_per4s<true>, dim3((Nx + kStripW - 1) / kStripW, (Ny + H - 1) / H), dim3(kStripW), 0, st, G, A, B,
                 ca, cb, oscale, H, f);
}

void launch_proj_per(hipStream_t st, int order, int Nx, int Ny, double hx, double hy, const double* Vn,
                     const double* Vt, const double* dp, double dt, double c, double* V) {
  const PerGrid G = per_grid(order, Nx, Ny, hx, hy, 0.0);
  dim3 b(64, 4);
  if (order == 2)
    launch_timed(k_proj_per<2, true>, grid2(Nx, Ny, b), b, 0, st, G, Vn, Vt, dp, dt, c, V);
  else if (Nx % 2 == 0 && aligned16(Vn) && aligned16(Vt) && aligned16(dp) && aligned16(V))   // 174 -> 161.5 us at 4096^2
    launch_timed(k_proj_per4x<false>, grid2(Nx / 2, Ny, b), b, 0, st, G, Vn, Vt, dp, dt, c, V);
  else   // banded blocks measured slower here (4096^2: 182 against 174 us)
    launch_timed(k_proj_per<4, false>, grid2(Nx, Ny, b), b, 0, st, G, Vn, Vt, dp, dt, c, V);
}

static int ilog2(int n) {
  int l = 0;
  while ((1 << l) < n) ++l;
  return l;
}

// threads per transform: one radix-4 butterfly per thread and pass for the
// power-of-two lengths (N / 4 threads, at least 64); 256 for the direct DFT
static int fft_threads(int N, bool pow2) { return pow2 ? std::max(64, std::min(1024, N / 4)) : 256; }
// register-butterfly transforms: power-of-two lengths 16 .. 4096.  Elements
// per thread E (the radix of the main passes) by length and kernel, from
// timings on MI355X: more elements per thread mean fewer passes through LDS
// but more registers and fewer waves; at 1024 (half-spectrum 513 x 1024) the
// wave count decides, and the column solve at 4096 is bound by its 64 KB LDS
// per workgroup, where 1024 light threads hide latency best.
//   rows (r2c, c2r):  E = 16 at 4096, 4 at 1024, 8 otherwise
//   column solve:     E = 4 at 1024 and 4096, 8 otherwise
static bool fftr_ok(int N, bool pow2) { return pow2 && N >= 16 && N <= 4096; }
static int fft_elems(int N, bool cols) {
  if (N == 1024) return 4;
  if (N == 4096) return cols ? 4 : 16;
  return 8;
}

// Length 4096: the transforms with the split exchange (radix 16, 3 passes,
// 32 KB of LDS per transform instead of 64 KB: three workgroups per CU instead
// of two; at 4096^2 column solve 110 -> 72 us, forward rows 97 -> 92 us,
// inverse rows 62 -> 62 us, with p += dp 96 -> 91 us).  The inverse rows stage
// the full row through LDS as before, each half-spectrum value read once: built
// in registers from the half spectra directly (every value needed by two
// threads, read twice) they ran at 79 us (158 us with p += dp).
constexpr int kSplitE = 16, kSplitP = 3;
static bool split_fft(int N) { return N == 4096; }

// TR: the row transforms store (load) the spectra transposed; D: the first
// transform computes its rows as the divergence of D->A, D->B (f unused);
// acc: the last transform adds dp into its output
template <int E, bool TR>
static void fftr_launch(hipStream_t st, const FFTPlan& P, const double* f, double* dp, int which, const DivSrc* D,
                        bool acc) {
  const int Nx = P.Nx, Ny = P.Ny, Nh = P.Nh;
  double2* H = (double2*)(TR ? P.HT : P.H);
  double2* HT = (double2*)P.HT;
  const dim3 gr((Ny + 1) / 2), bl(std::max(64, Nx / E));
  const size_t sh = (size_t)Nx * sizeof(double2);
  if (which == 0) {
    // the divergence rows are staged by a plain loop before the transform:
    // as many threads as the launch bound allows keep more of its loads in
    // flight (the transform's passes leave the extra threads idle)
    if (D && D->order == 4) {
      if (!TR || !split_fft(Nx)) throw std::runtime_error("fourth-order fused divergence: 4096-point rows only");
      if (D->B)
        fft_launch(0, k_fftr_r2c_rows_sdiv4<kSplitE, true>, gr, dim3(Nx / kSplitE), (unsigned)(Nx * sizeof(double)),
                   st, Nx, Ny, Nh, kSplitP, (const double2*)P.twx, *D, H, P.fmax2);
      else
        fft_launch(0, k_fftr_r2c_rows_sdiv4<kSplitE, false>, gr, dim3(Nx / kSplitE), (unsigned)(Nx * sizeof(double)),
                   st, Nx, Ny, Nh, kSplitP, (const double2*)P.twx, *D, H, P.fmax2);
    } else if (D)
      fft_launch(0, k_fftr_r2c_rows<E, TR, true>, gr, dim3(fft_block<E>(0)), sh, st, Nx, Ny, Nh, ilog2(Nx),
                 (const double2*)P.twx, (const double*)nullptr, *D, H, P.fmax2);
    else if (TR && split_fft(Nx))
      fft_launch(0, k_fftr_r2c_rows_s<kSplitE>, gr, dim3(Nx / kSplitE), (unsigned)(Nx * sizeof(double)), st, Nx, Ny,
                 Nh, kSplitP, (const double2*)P.twx, f, H, P.fmax2);
    else
      fft_launch(0, k_fftr_r2c_rows<E, TR, false>, gr, bl, sh, st, Nx, Ny, Nh, ilog2(Nx), (const double2*)P.twx, f,
                 DivSrc{nullptr, nullptr, 0.0, 0.0, 0.0, 0.0, 0.0, 0, 2}, H, P.fmax2);
  } else if (which == 1) {
    if (split_fft(Ny))
      fft_launch(1, k_fftr_cols_solve_s<kSplitE>, dim3(Nh), dim3(Ny / kSplitE), (unsigned)(Ny * sizeof(double)), st,
                 Nx, Ny, kSplitP, (const double2*)P.twy, (const double*)P.ax, (const double*)P.ay, P.c, HT, P.fmax2);
    else
      fft_launch(1, k_fftr_cols_solve<E>, dim3(Nh), dim3(std::max(64, Ny / E)), (unsigned)(Ny * sizeof(double2)), st,
                 Nx, Ny, ilog2(Ny), (const double2*)P.twy, (const double*)P.ax, (const double*)P.ay, P.c, HT, P.fmax2);
  } else {
    // the half spectra are staged by a plain loop before the transform (as
    // the divergence rows above): threads up to the launch bound
    const dim3 blc(fft_block<E>(0));
    const unsigned long long* fm = P.fmax2;
    if (TR && split_fft(Nx)) {
      const unsigned shs = (unsigned)(Nx * sizeof(double));
      if (acc)
        fft_launch(3, k_fftr_c2r_rows_s<kSplitE, true>, gr, dim3(Nx / kSplitE), shs, st, Nx, Ny, Nh, kSplitP,
                   (const double2*)P.twx, (const double2*)H, dp, fm);
      else
        fft_launch(2, k_fftr_c2r_rows_s<kSplitE, false>, gr, dim3(Nx / kSplitE), shs, st, Nx, Ny, Nh, kSplitP,
                   (const double2*)P.twx, (const double2*)H, dp, fm);
    } else if (acc)
      fft_launch(3, k_fftr_c2r_rows<E, TR, true>, gr, blc, sh, st, Nx, Ny, Nh, ilog2(Nx), (const double2*)P.twx,
                 (const double2*)H, dp, fm);
    else
      fft_launch(2, k_fftr_c2r_rows<E, TR, false>, gr, blc, sh, st, Nx, Ny, Nh, ilog2(Nx), (const double2*)P.twx,
                 (const double2*)H, dp, fm);
  }
}

template <bool TR>
static void fftr(hipStream_t st, const FFTPlan& P, const double* f, double* dp, int which,
                 const DivSrc* D = nullptr, bool acc = false) {
  const int E = which == 1 ? fft_elems(P.Ny, true) : fft_elems(P.Nx, false);
  if (E == 4) fftr_launch<4, TR>(st, P, f, dp, which, D, acc);
  else if (E == 8) fftr_launch<8, TR>(st, P, f, dp, which, D, acc);
  else fftr_launch<16, TR>(st, P, f, dp, which, D, acc);
}

bool fft_register_lengths(int Nx, int Ny) {
  auto pow2 = [](int n) { return n > 0 && (n & (n - 1)) == 0; };
  return fftr_ok(Nx, pow2(Nx)) && fftr_ok(Ny, pow2(Ny));
}

bool fft_fused_div4(int Nx, int Ny) { return fft_register_lengths(Nx, Ny) && split_fft(Nx); }

void launch_fft_poisson_fused(hipStream_t st, const FFTPlan& P, const double* f, double hx, double hy,
                              const double* A, const double* B, double ca, double cb, double oscale, double* out,
                              bool acc, int order) {
  if (!fft_register_lengths(P.Nx, P.Ny))
    throw std::runtime_error("fused FFT solve: needs power-of-two lengths 16 .. 4096");
  const DivSrc D{A, B, ca, cb, oscale, hx, hy, aligned16(A) && aligned16(B) ? 1 : 0, order};
  fftr<true>(st, P, f, out, 0, A ? &D : nullptr);
  fftr<true>(st, P, f, out, 1);
  fftr<true>(st, P, f, out, 2, nullptr, acc);
}

void launch_fft_poisson(hipStream_t st, const FFTPlan& P, const double* f, double* dp) {
  const int Nx = P.Nx, Ny = P.Ny, Nh = P.Nh;
  const int Tx = fft_threads(Nx, P.pow2x), Ty = fft_threads(Ny, P.pow2y);
  const bool fx = fftr_ok(Nx, P.pow2x), fy = fftr_ok(Ny, P.pow2y);
  size_t shx = (size_t)(P.pow2x ? Nx : 2 * Nx) * sizeof(double2);
  size_t shy = (size_t)(P.pow2y ? Ny : 2 * Ny) * sizeof(double2);
  double2* H = (double2*)P.H;
  double2* HT = (double2*)P.HT;
  if (fx && fy) {
    // register transforms in both directions: the row transforms store and
    // load the spectra transposed, no separate transposes
    fftr<true>(st, P, f, dp, 0);
    fftr<true>(st, P, f, dp, 1);
    fftr<true>(st, P, f, dp, 2);
    return;
  }
  if (fx)
    fftr<false>(st, P, f, dp, 0);
  else
    hipLaunchKernelGGL(k_fft_r2c_rows<1>, dim3((Ny + 1) / 2), dim3(Tx), shx, st, Nx, Ny, Nh, P.pow2x, ilog2(Nx),
                       (const double2*)P.twx, f, H, P.fmax2);
  dim3 tb(32, 8);
  hipLaunchKernelGGL(k_transpose_c, dim3((Nh + 31) / 32, (Ny + 31) / 32), tb, 0, st, H, HT, Ny, Nh);
  if (fy)
    fftr<false>(st, P, f, dp, 1);   // in place on HT
  else
    hipLaunchKernelGGL(k_fft_cols_solve<1>, dim3(Nh), dim3(Ty), shy, st, Nx, Ny, P.pow2y, ilog2(Ny),
                       (const double2*)P.twy, P.ax, P.ay, P.c, HT, P.fmax2);
  hipLaunchKernelGGL(k_transpose_c, dim3((Ny + 31) / 32, (Nh + 31) / 32), tb, 0, st, HT, H, Nh, Ny);
  if (fx)
    fftr<false>(st, P, f, dp, 2);
  else
    hipLaunchKernelGGL(k_fft_c2r_rows<1>, dim3((Ny + 1) / 2), dim3(Tx), shx, st, Nx, Ny, Nh, P.pow2x, ilog2(Nx),
                       (const double2*)P.twx, H, dp, (const unsigned long long*)P.fmax2);
}

// the early exit of pressure_poisson.m:32-38 for the matrix solvers: max |f|
// as kExitBlocks partial maxima, then dp = 0 when their maximum is below tol
constexpr int kExitBlocks = 64;
__global__ void __launch_bounds__(256) k_absmax_part(const double* __restrict__ f, size_t n,
                                                     double* __restrict__ part) {
  double m = 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    m = fmax(m, fabs(f[i]));
  __shared__ double w[4];
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
  if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = fmax(fmax(w[0], w[1]), fmax(w[2], w[3]));
}

__global__ void __launch_bounds__(256) k_zero_if_small(const double* __restrict__ part, double tol,
                                                      double* __restrict__ x, size_t n) {
  __shared__ int small;
  if (threadIdx.x < 64) {
    double m = part[threadIdx.x];
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
    if (threadIdx.x == 0) small = m < tol;
  }
  __syncthreads();
  if (!small) return;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    x[i] = 0.0;
}

void launch_absmax_part(hipStream_t st, const double* f, size_t n, double* part) {
  hipLaunchKernelGGL(k_absmax_part, dim3(kExitBlocks), dim3(256), 0, st, f, n, part);
}

void launch_zero_if_small(hipStream_t st, const double* part, double tol, double* x, size_t n) {
  const unsigned nb = (unsigned)std::min<size_t>(1024, (n + 255) / 256);
  hipLaunchKernelGGL(k_zero_if_small, dim3(std::max(1u, nb)), dim3(256), 0, st, part, tol, x, n);
}

void launch_lap(hipStream_t st, const DLap& L, const double* p, double* q) {
  dim3 b(64, 4);
  hipLaunchKernelGGL(k_lap, grid2(L.nx, L.ny, b), b, 0, st, L, p, q);
}

void launch_cg_single(hipStream_t st, const DLap& L, const double* f, double* x, double* r, double* z, double* p,
                      double* q, double tol2, int maxit, int x0_zero, int* out_it, double* out_rel) {
  hipLaunchKernelGGL(k_cg_single, dim3(1), dim3(1024), 0, st, L, f, x, r, z, p, q, tol2, maxit, x0_zero, out_it,
                     out_rel);
}

void launch_cg_init(hipStream_t st, const DLap& L, const CGStateH& S, const double* f, const double* x, double* r,
                    double* z, double* p, int x0_zero, double tol2) {
  CGState D{S.s, S.flag, S.part, S.nblk, S.maxit, S.e};
  hipLaunchKernelGGL(k_cg_init, dim3(S.nblk), dim3(256), 0, st, L, D, f, x, r, z, p, x0_zero);
  hipLaunchKernelGGL(k_cg_init_fin, dim3(1), dim3(256), 0, st, D, tol2);
}

void launch_cg_iter(hipStream_t st, const DLap& L, const CGStateH& S, double* x, double* r, double* z, double* p,
                    double* q) {
  CGState D{S.s, S.flag, S.part, S.nblk, S.maxit, S.e};
  const size_t n = (size_t)L.nx * L.ny;
  hipLaunchKernelGGL(k_cg_apply, dim3(S.nblk), dim3(256), 0, st, L, D, p, q);
  hipLaunchKernelGGL(k_cg_alpha, dim3(1), dim3(256), 0, st, D);
  hipLaunchKernelGGL(k_cg_update, dim3(S.nblk), dim3(256), 0, st, L, D, p, q, x, r, z);
  hipLaunchKernelGGL(k_cg_beta, dim3(1), dim3(256), 0, st, D);
  hipLaunchKernelGGL(k_cg_dir, dim3(S.nblk), dim3(256), 0, st, D, n, z, p);
}

void launch_cg_iter2(hipStream_t st, const DLap& L, const CGStateH& S, double* x, double* r, double* z,
                     const double* p_in, double* p_out, double* q) {
  CGState D{S.s, S.flag, S.part, S.nblk, S.maxit, S.e};
  const size_t n = (size_t)L.nx * L.ny;
  const bool one = (size_t)S.nblk * 256 >= n;
  // the pressure Laplacian: kron(diag, tridiagonal) + kron(tridiagonal, diag)
  const bool tri = L.L1.Y.K == 1 && L.L1.X.K == 3 && L.L2.Y.K == 3 && L.L2.X.K == 1;
  const dim3 g(S.nblk), t(256);
  const double* zc = z;
  if (one && tri) hipLaunchKernelGGL((k_cg2_apply<true, true>), g, t, 0, st, L, D, zc, p_in, p_out, q);
  else if (one) hipLaunchKernelGGL((k_cg2_apply<true, false>), g, t, 0, st, L, D, zc, p_in, p_out, q);
  else if (tri) hipLaunchKernelGGL((k_cg2_apply<false, true>), g, t, 0, st, L, D, zc, p_in, p_out, q);
  else hipLaunchKernelGGL((k_cg2_apply<false, false>), g, t, 0, st, L, D, zc, p_in, p_out, q);
  const double *pc = p_out, *qc = q;
  if (one) hipLaunchKernelGGL(k_cg2_update<true>, g, t, 0, st, L, D, pc, qc, x, r, z);
  else hipLaunchKernelGGL(k_cg2_update<false>, g, t, 0, st, L, D, pc, qc, x, r, z);
}

void launch_cg2_check(hipStream_t st, const CGStateH& S) {
  CGState D{S.s, S.flag, S.part, S.nblk, S.maxit, S.e};
  hipLaunchKernelGGL(k_cg2_check, dim3(1), dim3(256), 0, st, D);
}

void launch_cgx(hipStream_t st, const DLap& L, const CGStateH& S, int phase, const double* a, double* b, double* c,
                double* d, double tol2) {
  CGState D{S.s, S.flag, S.part, S.nblk, S.maxit, S.e};
  const size_t n = (size_t)L.nx * L.ny;
  const dim3 g(S.nblk), t(256);
  switch (phase) {
    case 0:   // r (b) = f (a), f.f
      hipLaunchKernelGGL(k_cgx_init, g, t, 0, st, D, n, a, b);
      break;
    case 1:   // first r.z (r a, z b), p (c) = z, then the initial scalars
      hipLaunchKernelGGL(k_cgx_rz, g, t, 0, st, D, n, a, b, c, 1);
      hipLaunchKernelGGL(k_cg_init_fin, dim3(1), t, 0, st, D, tol2);
      break;
    case 2:   // q (b) = A p (a), alpha, x (c) and r (d) updated, convergence test
      hipLaunchKernelGGL(k_cg_apply, g, t, 0, st, L, D, a, b);
      hipLaunchKernelGGL(k_cg_alpha, dim3(1), t, 0, st, D);
      hipLaunchKernelGGL(k_cgx_update, g, t, 0, st, D, n, a, b, c, d);
      hipLaunchKernelGGL(k_cgx_test, dim3(1), t, 0, st, D);
      break;
    default:  // r.z (r a, z b), beta, p (c) = z + beta p
      hipLaunchKernelGGL(k_cgx_rz, g, t, 0, st, D, n, a, b, c, 0);
      hipLaunchKernelGGL(k_cgx_beta, dim3(1), t, 0, st, D);
      hipLaunchKernelGGL(k_cg_dir, g, t, 0, st, D, n, b, c);
      break;
  }
}

// counts a solve whose convergence test has not passed (flag[0] = 0) or
// that stopped at maxit (flag[2] = 1) into missed[0] (the device-iteration
// mode of the fast-diagonalization PCG)
__global__ void k_cgx_missed(const int* __restrict__ flag, int* __restrict__ missed) {
  if (threadIdx.x == 0 && (!flag[0] || flag[2])) missed[0] += 1;
}
void launch_cgx_missed(hipStream_t st, const CGStateH& S, int* missed) {
  hipLaunchKernelGGL(k_cgx_missed, dim3(1), dim3(64), 0, st, S.flag, missed);
}

void launch_axpy(hipStream_t st, size_t n, double a, const double* x, double* y) {
  int nb = (int)std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_axpy, dim3(nb), dim3(256), 0, st, n, a, x, y);
}

void launch_reduce_stats(hipStream_t st, const double* x, int nx, int ny, const double* wx, const double* wy,
                         double* part, int nblk) {
  hipLaunchKernelGGL(k_reduce_stats, dim3(nblk), dim3(256), 0, st, x, nx, ny, wx, wy, part);
}

}  // namespace ins2d
