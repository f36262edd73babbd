// Fourth-order operators on grids with walls: the alfa (1h) - (3h)
// combinations of operator_mesh.m, operator_averaging.m,
// operator_interpolation.m, operator_divergence.m and
// operator_convection_diffusion.m (order4 = 1), each direction either
// Dirichlet at both ends or periodic, any tensor-product grid.
//
// The wide closures (three ghost values per side) and the products of
// closures make these operators products of several Kronecker factors with
// non-separable volume weights in between, so they are assembled here as
// 2-D sparse matrices and applied on the device in CSR form.
#pragma once

#include <vector>

#include "ops1d.hpp"

namespace ins2d {

struct W4Ops {
  int Nx = 0, Ny = 0, nux = 0, nuy = 0, nvx = 0, nvy = 0;
  int Nu = 0, Nv = 0, NV = 0, Np = 0;
  int Nf = 0;                            // face values of the eight convective products
  std::vector<double> xin, yin;          // u points in x, v points in y
  std::vector<double> hx, hy, gx, gy;    // second-order cell widths and line spacings
  std::vector<double> Om;                // NV volumes alfa Om1 - Om3
  // convection: w = (If V + yI) .* (Af V + yA) for the products
  // [ux, uy, vx, vy] of the 1h stencils then the 3h stencils, and
  // conv = Cw w (alfa C1 on the 1h products, -C3 on the 3h ones)
  Sp Af, If;   // Nf x NV
  Sp Cw;       // NV x Nf
  Sp Diff;     // NV x NV, diag(Diffu, Diffv)
  Sp G;        // NV x Np, -M^T
  Sp M;        // Np x NV
  Sp A;        // Np x Np, M Om^-1 G
  bool singular = true;   // A has the constant null vector (no pressure boundary)
  // boundary vectors, linear in the boundary samples:
  // y = sum_src T[src] * sample[src] (empty matrices for unused sources)
  std::vector<Sp> TyA, TyI, TyD, TyM;
};

// throws std::runtime_error for boundary combinations the fourth-order
// closures do not define (pressure or symmetry boundaries)
void build_w4(const CaseDesc& cd, W4Ops& W);
// the boundary vectors yA, yI (Nf), yD (NV) and yM (Np) at the given samples
void w4_bc_vectors(const W4Ops& W, const BCSamples& S, std::vector<double>& yA, std::vector<double>& yI,
                   std::vector<double>& yD, std::vector<double>& yM);
// sample lengths of each source (Nx + 1, xin, yp, ...)
int w4_sample_len(const W4Ops& W, int src);

}  // namespace ins2d
