// Closed-form interior stencils of the generic path on uniform grids, and the
// host-side classification of the points where they apply.
//
// On a uniform grid every interior row of the 1-D operator factors is the
// same two-point average, interpolation or difference, so away from the
// boundaries the momentum right-hand side of the generic kron-factored
// operators equals the closed form of the periodic kernels (the interior
// formulas of operator_averaging.m, operator_interpolation.m and
// operator_convection_diffusion.m), evaluated here with explicit neighbour
// indices instead of wrap-around.  Which points are interior is decided on
// the host numerically: a point is marked where the closed form (its indices
// in range) reproduces the generic operators on random fields, so a boundary
// closure that changes a row, or a layout the index offsets do not describe,
// leaves the point on the generic path.
#pragma once

#include <algorithm>
#include <cmath>
#include <random>
#include <vector>

#include "device.hpp"
#include "ops1d.hpp"

namespace ins2d {

// Grid-index accessors: U(a, b) is u at (x_a, yp_b), V(a, b) is v at
// (xp_a, y_b); g holds the spacings (constants on uniform grids, else the
// arrays: hx_I = x_{I+1} - x_I, gx_I = xp_I - xp_{I-1}, the same in y) and 1/Re.
// The convective cross terms interpolate with the neighbouring cell widths
// (operator_interpolation.m mat_hx2/mat_hy2), the diffusive fluxes divide by
// the distance between the two values they difference (operator_convection_
// diffusion.m Su_ux = 1/hxd, Su_uy = 1/gyd, ...) and are weighted by the face
// length (Dux = diag(hy), Duy = diag(gx), ...).
struct Spacing {
  const UniForm& g;
  __host__ __device__ double hx(int I) const { return g.ahx ? g.ahx[I] : g.hx; }
  __host__ __device__ double hy(int J) const { return g.ahy ? g.ahy[J] : g.hy; }
  __host__ __device__ double gx(int I) const { return g.agx ? g.agx[I] : g.hx; }
  __host__ __device__ double gy(int J) const { return g.agy ? g.agy[J] : g.hy; }
  __host__ __device__ double rhx(int I) const { return g.arhx ? g.arhx[I] : g.rhx; }
  __host__ __device__ double rhy(int J) const { return g.arhy ? g.arhy[J] : g.rhy; }
  __host__ __device__ double rgx(int I) const { return g.argx ? g.argx[I] : g.rhx; }
  __host__ __device__ double rgy(int J) const { return g.argy ? g.argy[J] : g.rhy; }
};

// u-momentum at (x_I, yp_J): -(du2dx + duvdy) + d2u / Re
template <class UF, class VF>
__host__ __device__ __forceinline__ double uni2_u(const UniForm& g, UF U, VF V, int I, int J) {
  const Spacing s{g};
  const double hyJ = s.hy(J), hxw = s.hx(I - 1), hxe = s.hx(I);
  const double u0 = U(I, J), ue = U(I + 1, J), uw = U(I - 1, J), un = U(I, J + 1), us = U(I, J - 1);
  const double v0 = V(I, J), vw = V(I - 1, J), vn = V(I, J + 1), v_wn = V(I - 1, J + 1);
  const double Ue = 0.5 * (u0 + ue), Uw = 0.5 * (uw + u0);
  const double du2dx = (hyJ * Ue) * Ue - (hyJ * Uw) * Uw;
  const double Vn = 0.5 * (hxw * v_wn + hxe * vn), Vs = 0.5 * (hxw * vw + hxe * v0);
  const double duvdy = Vn * (0.5 * (u0 + un)) - Vs * (0.5 * (us + u0));
  const double d2u = hyJ * ((ue - u0) * s.rhx(I) - (u0 - uw) * s.rhx(I - 1)) +
                     s.gx(I) * ((un - u0) * s.rgy(J + 1) - (u0 - us) * s.rgy(J));
  return -(du2dx + duvdy) + g.invRe * d2u;
}
// v-momentum at (xp_I, y_J)
template <class UF, class VF>
__host__ __device__ __forceinline__ double uni2_v(const UniForm& g, UF U, VF V, int I, int J) {
  const Spacing s{g};
  const double hxI = s.hx(I), hys = s.hy(J - 1), hyn = s.hy(J);
  const double v0 = V(I, J), ve = V(I + 1, J), vw = V(I - 1, J), vn = V(I, J + 1), vs = V(I, J - 1);
  const double u0 = U(I, J), ue = U(I + 1, J), us = U(I, J - 1), u_se = U(I + 1, J - 1);
  const double Ue2 = 0.5 * (hys * u_se + hyn * ue), Uw2 = 0.5 * (hys * us + hyn * u0);
  const double duvdx = Ue2 * (0.5 * (v0 + ve)) - Uw2 * (0.5 * (vw + v0));
  const double Vn2 = 0.5 * (v0 + vn), Vs2 = 0.5 * (vs + v0);
  const double dv2dy = (hxI * Vn2) * Vn2 - (hxI * Vs2) * Vs2;
  const double d2v = s.gy(J) * ((ve - v0) * s.rgx(I + 1) - (v0 - vw) * s.rgx(I)) +
                     hxI * ((vn - v0) * s.rhy(J) - (v0 - vs) * s.rhy(J - 1));
  return -(duvdx + dv2dy) + g.invRe * d2v;
}

// host copies of the spacing arrays of a stretched grid (UniForm points at
// them while the host classifies; the caller uploads them)
struct UniArrays {
  std::vector<double> hx, hy, gx, gy, rhx, rhy, rgx, rgy;
};

// ---------------------------------------------------------------- host side
namespace uhost {

inline double kron_at(const HKron& K, const std::vector<double>& in, int f, int g) {
  double acc = 0.0;
  for (int a = 0; a < K.Y.K; ++a) {
    const double wy = K.Y.val[(size_t)g * K.Y.K + a];
    const int cy = K.Y.col[(size_t)g * K.Y.K + a];
    double s = 0.0;
    for (int b = 0; b < K.X.K; ++b)
      s += K.X.val[(size_t)f * K.X.K + b] * in[(size_t)cy * K.X.ncols + K.X.col[(size_t)f * K.X.K + b]];
    acc += wy * s;
  }
  return acc;
}

template <class FACE>
double apply_d(const HKron& D, int i, int j, FACE face) {
  double acc = 0.0;
  for (int a = 0; a < D.Y.K; ++a) {
    const double wy = D.Y.val[(size_t)j * D.Y.K + a];
    if (wy == 0.0) continue;
    const int g = D.Y.col[(size_t)j * D.Y.K + a];
    double s = 0.0;
    for (int b = 0; b < D.X.K; ++b) {
      const double wx = D.X.val[(size_t)i * D.X.K + b];
      if (wx != 0.0) s += wx * face(D.X.col[(size_t)i * D.X.K + b], g);
    }
    acc += wy * s;
  }
  return acc;
}

// -sum conv + sum diff of the generic operators at (i, j), boundary terms left
// out (the interior points have none)
inline double generic_terms(const HComp& C, const std::vector<double>& u, const std::vector<double>& v,
                            const std::vector<double>& own, int i, int j) {
  double acc = 0.0;
  for (const HConv& T : C.conv) {
    const std::vector<double>& cf = T.icomp == 0 ? u : v;
    acc -= T.scale * apply_d(T.D, i, j, [&](int f, int g) { return kron_at(T.I, cf, f, g) * kron_at(T.A, own, f, g); });
  }
  for (const HDiff& T : C.diff)
    acc += apply_d(T.D[0], i, j, [&](int f, int g) { return T.sscale[0] * kron_at(T.S[0], own, f, g); });
  return acc;
}

}  // namespace uhost

// Per-point maps (1: closed form applies) of the u and v components, the
// constants of the closed form and, for a stretched grid, its spacing arrays
// in A (g points at A's host copies on return); empty maps when the
// operators are not the second-order two-term ones.
inline void uniform_maps(const Plan& P, UniForm& g, UniArrays& A, std::vector<unsigned char>& fu,
                         std::vector<unsigned char>& fv) {
  fu.clear();
  fv.clear();
  const int Nx = P.Npx, Ny = P.Npy;
  double hxmin = 1e300, hxmax = 0, hymin = 1e300, hymax = 0;
  for (double h : P.hx) { hxmin = std::min(hxmin, h); hxmax = std::max(hxmax, h); }
  for (double h : P.hy) { hymin = std::min(hymin, h); hymax = std::max(hymax, h); }
  if (P.cu.conv.size() != 2 || P.cv.conv.size() != 2 || P.cu.diff.size() != 2 || P.cv.diff.size() != 2) return;
  g.hx = (P.cd.x.back() - P.cd.x.front()) / Nx;
  g.hy = (P.cd.y.back() - P.cd.y.front()) / Ny;
  g.rhx = 1.0 / g.hx;
  g.rhy = 1.0 / g.hy;
  g.ahx = g.ahy = g.agx = g.agy = g.arhx = g.arhy = g.argx = g.argy = nullptr;
  // grid lines built as x0 + i h differ from uniform in the last digits:
  // those grids take the constants, others the arrays
  if (hxmax - hxmin > 1e-12 * hxmax || hymax - hymin > 1e-12 * hymax) {
    A.hx = P.hx;
    A.hy = P.hy;
    A.gx.assign(Nx + 1, 0.0);
    A.gy.assign(Ny + 1, 0.0);
    for (int i = 1; i < Nx; ++i) A.gx[i] = 0.5 * (P.hx[i - 1] + P.hx[i]);
    for (int j = 1; j < Ny; ++j) A.gy[j] = 0.5 * (P.hy[j - 1] + P.hy[j]);
    A.gx[0] = A.gx[Nx] = A.gy[0] = A.gy[Ny] = 1.0;   // boundary lines: never read by interior points
    auto recip = [](const std::vector<double>& a) {
      std::vector<double> r(a.size());
      for (size_t k = 0; k < a.size(); ++k) r[k] = 1.0 / a[k];
      return r;
    };
    A.rhx = recip(A.hx);
    A.rhy = recip(A.hy);
    A.rgx = recip(A.gx);
    A.rgy = recip(A.gy);
    g.ahx = A.hx.data(); g.ahy = A.hy.data(); g.agx = A.gx.data(); g.agy = A.gy.data();
    g.arhx = A.rhx.data(); g.arhy = A.rhy.data(); g.argx = A.rgx.data(); g.argy = A.rgy.data();
  }
  g.invRe = P.cd.Re >= 1e100 ? 0.0 : 1.0 / P.cd.Re;
  // offsets: u array column 0 sits on grid line ou, v array row 0 on line ov
  auto nearest = [](const std::vector<double>& lines, double z) {
    int best = 0;
    for (int k = 1; k < (int)lines.size(); ++k)
      if (std::fabs(lines[k] - z) < std::fabs(lines[best] - z)) best = k;
    return best;
  };
  std::vector<double> xl(P.cd.x), yl(P.cd.y);
  g.ou = P.xin.empty() ? 0 : nearest(xl, P.xin[0]);
  g.ov = P.yin.empty() ? 0 : nearest(yl, P.yin[0]);
  g.nux = P.Nux_in;
  g.nuy = P.Nuy_in;
  g.nvx = P.Nvx_in;
  g.nvy = P.Nvy_in;
  // two random fields; a point is interior where the closed form matches the
  // generic operators to round-off on both
  std::mt19937_64 rng(12345);
  std::uniform_real_distribution<double> dist(-1.0, 1.0);
  std::vector<double> u[2], v[2];
  for (int r = 0; r < 2; ++r) {
    u[r].resize(P.Nu);
    v[r].resize(P.Nv);
    for (auto& x : u[r]) x = dist(rng);
    for (auto& x : v[r]) x = dist(rng);
  }
  fu.assign(P.Nu, 0);
  fv.assign(P.Nv, 0);
  // the size of the terms of one point for O(1) fields (round-off is 1e-16 of it)
  const double hmax = std::max(hxmax, hymax), hmin = std::min(hxmin, hymin);
  const double scale = 4.0 * 2.0 * hmax + g.invRe * 8.0 * hmax / hmin;
  for (int r = 0; r < 2; ++r) {
    const auto& uu = u[r];
    const auto& vv = v[r];
    bool oob = false;
    auto U = [&](int a, int b) {
      const int ia = a - g.ou;
      if (ia < 0 || ia >= g.nux || b < 0 || b >= g.nuy) { oob = true; return 0.0; }
      return uu[(size_t)ia + (size_t)g.nux * b];
    };
    auto V = [&](int a, int b) {
      const int jb = b - g.ov;
      if (a < 0 || a >= g.nvx || jb < 0 || jb >= g.nvy) { oob = true; return 0.0; }
      return vv[(size_t)a + (size_t)g.nvx * jb];
    };
    for (int j = 0; j < g.nuy; ++j)
      for (int i = 0; i < g.nux; ++i) {
        unsigned char& m = fu[(size_t)i + (size_t)g.nux * j];
        if (r == 1 && !m) continue;
        // spacings read: hx(I - 1), hx(I), gx(I), hy(J), gy(J), gy(J + 1)
        const int I = i + g.ou;
        if (I < 1 || I > Nx - 1 || j > Ny - 1) { m = 0; continue; }
        oob = false;
        const double c = uni2_u(g, U, V, i + g.ou, j);
        const double ref = uhost::generic_terms(P.cu, uu, vv, uu, i, j);
        m = (!oob && std::fabs(c - ref) <= 1e-12 * scale) ? 1 : 0;
      }
    for (int j = 0; j < g.nvy; ++j)
      for (int i = 0; i < g.nvx; ++i) {
        unsigned char& m = fv[(size_t)i + (size_t)g.nvx * j];
        if (r == 1 && !m) continue;
        // spacings read: hx(I), gx(I), gx(I + 1), hy(J - 1), hy(J), gy(J)
        const int J = j + g.ov;
        if (i > Nx - 1 || J < 1 || J > Ny - 1) { m = 0; continue; }
        oob = false;
        const double c = uni2_v(g, U, V, i, j + g.ov);
        const double ref = uhost::generic_terms(P.cv, uu, vv, vv, i, j);
        m = (!oob && std::fabs(c - ref) <= 1e-12 * scale) ? 1 : 0;
      }
  }
}

// Split tiles of one component (nx x ny points, closed-form map f): tiles
// over the rectangle where the closed form applies (tx wide, 256 / tx rows
// per pass of a 256-thread block)
// and tiles over the ring around it (up to 64 wide); false (nothing
// appended) when the marked points do not fill a rectangle.
inline bool split_tiles(int inst, int comp, int nx, int ny, const std::vector<unsigned char>& f, int tx,
                        std::vector<RTile>& out, int rows_per_thread = 1) {
  if (f.size() != (size_t)nx * ny) return false;
  int i0 = nx, i1 = 0, j0 = ny, j1 = 0;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i)
      if (f[(size_t)i + (size_t)nx * j]) {
        i0 = std::min(i0, i); i1 = std::max(i1, i + 1);
        j0 = std::min(j0, j); j1 = std::max(j1, j + 1);
      }
  if (i0 >= i1 || j0 >= j1) return false;
  for (int j = j0; j < j1; ++j)
    for (int i = i0; i < i1; ++i)
      if (!f[(size_t)i + (size_t)nx * j]) return false;
  auto cover = [&](int a0, int a1, int b0, int b1, int kind) {
    if (a0 >= a1 || b0 >= b1) return;
    // interior tiles tx x (256 / tx) (more rows when the rectangle is
    // narrower than tx), ring tiles up to 64 wide; w h <= 256 either way,
    // times rows_per_thread rows (each thread then takes rows q / w,
    // q / w + 256 / w, ...)
    const int w = std::min(a1 - a0, kind ? tx : 64);
    const int h = std::max(1, std::min(b1 - b0, rows_per_thread * (256 / w)));
    for (int b = b0; b < b1; b += h)
      for (int a = a0; a < a1; a += w)
        out.push_back(RTile{inst, comp | kind, a, b, std::min(w, a1 - a), std::min(h, b1 - b)});
  };
  cover(i0, i1, j0, j1, 2);   // closed-form tiles first (callers sort a list by this bit)
  cover(0, i0, 0, ny, 0);
  cover(i1, nx, 0, ny, 0);
  cover(i0, i1, 0, j0, 0);
  cover(i0, i1, j1, ny, 0);
  return true;
}

// closed-form tiles first; returns their number
inline int order_split_tiles(std::vector<RTile>& t) {
  std::stable_partition(t.begin(), t.end(), [](const RTile& r) { return (r.comp & 2) != 0; });
  int n = 0;
  for (const RTile& r : t) n += (r.comp & 2) ? 1 : 0;
  return n;
}

}  // namespace ins2d
