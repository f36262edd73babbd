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

#include <cmath>
#include <random>
#include <vector>

#include "device.hpp"
#include "ops1d.hpp"

namespace ins2d {

// Grid-index accessors: U(a, b) is u at (x_a, yp_b), V(a, b) is v at
// (xp_a, y_b); g holds hx, hy, 1/hx, 1/hy, 1/Re.
// u-momentum at (x_I, yp_J): -(du2dx + duvdy) + d2u / Re
template <class UF, class VF>
__host__ __device__ __forceinline__ double uni2_u(const UniForm& g, UF U, VF V, int I, int J) {
  const double hx = g.hx, hy = g.hy;
  const double u0 = U(I, J), ue = U(I + 1, J), uw = U(I - 1, J), un = U(I, J + 1), us = U(I, J - 1);
  const double v0 = V(I, J), vw = V(I - 1, J), vn = V(I, J + 1), v_wn = V(I - 1, J + 1);
  const double Ue = 0.5 * (u0 + ue), Uw = 0.5 * (uw + u0);
  const double du2dx = (hy * Ue) * Ue - (hy * Uw) * Uw;
  const double Vn = 0.5 * hx * (v_wn + vn), Vs = 0.5 * hx * (vw + v0);
  const double duvdy = Vn * (0.5 * (u0 + un)) - Vs * (0.5 * (us + u0));
  const double d2u = hy * ((ue - u0) * g.rhx - (u0 - uw) * g.rhx) + hx * ((un - u0) * g.rhy - (u0 - us) * g.rhy);
  return -(du2dx + duvdy) + g.invRe * d2u;
}
// v-momentum at (xp_I, y_J)
template <class UF, class VF>
__host__ __device__ __forceinline__ double uni2_v(const UniForm& g, UF U, VF V, int I, int J) {
  const double hx = g.hx, hy = g.hy;
  const double v0 = V(I, J), ve = V(I + 1, J), vw = V(I - 1, J), vn = V(I, J + 1), vs = V(I, J - 1);
  const double u0 = U(I, J), ue = U(I + 1, J), us = U(I, J - 1), u_se = U(I + 1, J - 1);
  const double Ue2 = 0.5 * hy * (u_se + ue), Uw2 = 0.5 * hy * (us + u0);
  const double duvdx = Ue2 * (0.5 * (v0 + ve)) - Uw2 * (0.5 * (vw + v0));
  const double Vn2 = 0.5 * (v0 + vn), Vs2 = 0.5 * (vs + v0);
  const double dv2dy = (hx * Vn2) * Vn2 - (hx * Vs2) * Vs2;
  const double d2v = hy * ((ve - v0) * g.rhx - (v0 - vw) * g.rhx) + hx * ((vn - v0) * g.rhy - (v0 - vs) * g.rhy);
  return -(duvdx + dv2dy) + g.invRe * d2v;
}

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

// Per-point maps (1: closed form applies) of the u and v components, and the
// constants of the closed form; empty maps when the grid is not uniform.
inline void uniform_maps(const Plan& P, UniForm& g, std::vector<unsigned char>& fu, std::vector<unsigned char>& fv) {
  fu.clear();
  fv.clear();
  const int Nx = P.Npx, Ny = P.Npy;
  double hxmin = 1e300, hxmax = 0, hymin = 1e300, hymax = 0;
  for (double h : P.hx) { hxmin = std::min(hxmin, h); hxmax = std::max(hxmax, h); }
  for (double h : P.hy) { hymin = std::min(hymin, h); hymax = std::max(hymax, h); }
  // grid lines built as x0 + i h differ from uniform in the last digits
  if (hxmax - hxmin > 1e-12 * hxmax || hymax - hymin > 1e-12 * hymax) return;
  if (P.cu.conv.size() != 2 || P.cv.conv.size() != 2 || P.cu.diff.size() != 2 || P.cv.diff.size() != 2) return;
  g.hx = (P.cd.x.back() - P.cd.x.front()) / Nx;
  g.hy = (P.cd.y.back() - P.cd.y.front()) / Ny;
  g.rhx = 1.0 / g.hx;
  g.rhy = 1.0 / g.hy;
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
        oob = false;
        const double c = uni2_u(g, U, V, i + g.ou, j);
        const double ref = uhost::generic_terms(P.cu, uu, vv, uu, i, j);
        const double scale = 4.0 * (g.hy + g.hx) + g.invRe * 4.0 * (g.hy * g.rhx + g.hx * g.rhy);
        m = (!oob && std::fabs(c - ref) <= 1e-12 * scale) ? 1 : 0;
      }
    for (int j = 0; j < g.nvy; ++j)
      for (int i = 0; i < g.nvx; ++i) {
        unsigned char& m = fv[(size_t)i + (size_t)g.nvx * j];
        if (r == 1 && !m) continue;
        oob = false;
        const double c = uni2_v(g, U, V, i, j + g.ov);
        const double ref = uhost::generic_terms(P.cv, uu, vv, vv, i, j);
        const double scale = 4.0 * (g.hy + g.hx) + g.invRe * 4.0 * (g.hy * g.rhx + g.hx * g.rhy);
        m = (!oob && std::fabs(c - ref) <= 1e-12 * scale) ? 1 : 0;
      }
  }
}

}  // namespace ins2d
