// Host-side construction of the 1-D factors of the INS2D staggered-grid
// operators.
//
// Every 2-D operator the reference assembles with kron() in
// spatial/operators/*.m is a Kronecker product kron(Y, X) of two 1-D sparse
// operators (or a product of such products, which is again one), and every
// boundary-condition vector it builds in set_bc_vectors.m is a sum of at most
// four outer products of 1-D vectors.  This file builds those 1-D pieces; the
// device kernels apply them matrix-free, so no 2-D matrix is ever stored.
#pragma once

#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace ins2d {

// boundary-condition kinds (boundary_conditions.m:17-40)
enum BCKind : int { BC_DIR = 0, BC_PER = 1, BC_PRES = 2, BC_SYM = 3 };

// ---------------------------------------------------------------------------
// small sparse matrix (rows of (col, val) pairs), enough for 1-D operators
// ---------------------------------------------------------------------------
struct Sp {
  int m = 0, n = 0;
  std::vector<std::map<int, double>> r;
  Sp() = default;
  Sp(int m_, int n_) : m(m_), n(n_), r(m_) {}
  void add(int i, int j, double v) {
    if (i < 0 || i >= m || j < 0 || j >= n) throw std::runtime_error("Sp::add out of range");
    r[i][j] += v;
  }
  void set(int i, int j, double v) { r[i][j] = v; }
  double get(int i, int j) const {
    auto it = r[i].find(j);
    return it == r[i].end() ? 0.0 : it->second;
  }
  void scale_row(int i, double s) {
    for (auto& kv : r[i]) kv.second *= s;
  }
  int max_row_nnz() const;
  void prune(double tol = 0.0);
};

Sp eye(int n);
Sp diag(const std::vector<double>& d);
Sp mul(const Sp& A, const Sp& B);
Sp add(const Sp& A, double a, const Sp& B, double b);
Sp scale(const Sp& A, double s);
Sp transpose(const Sp& A);
std::vector<double> matvec(const Sp& A, const std::vector<double>& x);
// MATLAB spdiags(B, d, m, n): columns of B given as vectors (indexing by
// column for m >= n, by row for m < n, as MATLAB documents)
Sp spdiags(const std::vector<std::vector<double>>& cols, const std::vector<int>& offs, int m, int n);
Sp spdiags_const(const std::vector<double>& vals, const std::vector<int>& offs, int m, int n);

// boundary closure: u_ext = B1D u_in + Btemp ybc (BC_general.m:1-7)
struct Closure {
  Sp B1D, Btemp;
  std::vector<double> ybc1, ybc2;
};
// the closure of the boundary rows Bbc u_ext = ybc (Bb: Nt x Nb places the Nb
// boundary unknowns in u_ext, interior unknown r sits at row r - bin_off)
Closure make_closure(int Nt, int Nin, int Nb, int bin_off, const Sp& Bb, const Sp& Bbc,
                     const std::vector<double>& y1, const std::vector<double>& y2);
Closure bc_general(int Nt, int Nin, int Nb, int bc1, int bc2, double h1, double h2);
Closure bc_general_stag(int Nt, int Nin, int Nb, int bc1, int bc2, double h1, double h2, bool diff_variant);

// ---------------------------------------------------------------------------
// host description of the kron-factored operators
// ---------------------------------------------------------------------------
struct HOp1 {            // 1-D operator in ELL form
  int nrows = 0, ncols = 0, K = 0;   // K: stored entries per row (row stride)
  int kmax = 0;                      // most nonzero entries of any row
  std::vector<int> col;
  std::vector<double> val;
};
HOp1 to_ell(const Sp& A, int min_k = 1);

struct HKron {           // kron(Y, X)
  HOp1 X, Y;
};

// y(f,g) = sum_t ax[t][f] * ay[t][g]; the boundary samples enter through the
// "source" index: ax or ay is a fixed structural vector times a sampled
// boundary vector (set_bc_vectors.m:53-79)
enum Src : int {
  S_NONE = -1, S_uLo = 0, S_uUp, S_uLo_i, S_uUp_i, S_uLe_i, S_uRi_i,
  S_vLe, S_vRi, S_vLo_i, S_vUp_i, S_vLe_i, S_vRi_i, S_NSRC
};
// one factor of an outer-product term: either a structural vector, or a
// sampled boundary vector (optionally weighted pointwise, then mapped by a
// 1-D operator)
struct HFac {
  std::vector<double> s;   // structural vector (used when src == S_NONE)
  int src = S_NONE;
  std::vector<double> w;   // pointwise weight on the sample (optional)
  Sp map;                  // 1-D operator applied after weighting (optional)
  bool has_map = false;
};
struct HYTerm {
  HFac x, y;
};
struct HYVec {
  std::vector<HYTerm> t;
};

struct HConv {           // D * [(I c + yI) .* (A u + yA)] * scale
  HKron D, I, A;
  HYVec yI, yA;
  int icomp = 0;          // component read by I (0 = u, 1 = v)
  double scale = 1.0;
};
struct HDiff {           // sum_d dscale_d D_d [ W .* (sum_s sscale_s S_s u + yS) ]
  HKron D[2], S[2];
  double dscale[2] = {1, 1}, sscale[2] = {1, 1};
  int nD = 1, nS = 1;
  HYVec yS;
  bool has_w = false;     // W(f,g) = 1 / (wx1[f] wy1[g] - wx3[f] wy3[g])
  std::vector<double> wx1, wy1, wx3, wy3;
};
struct HComp {           // one momentum component
  int nx = 0, ny = 0;     // interior points (Nux_in x Nuy_in or Nvx_in x Nvy_in)
  std::vector<HConv> conv;
  std::vector<HDiff> diff;
  // volume: om(i,j) = ox[i] * oy[j]
  std::vector<double> ox, oy;
  HYVec yp;               // y_px / y_py
  HKron G;                // pressure gradient of this component (G = -M^T)
  HKron M;                // divergence contribution (this component -> p)
  HYVec yM;               // divergence boundary vector from this component
};

struct CaseDesc {
  int Nx = 0, Ny = 0;
  std::vector<double> x, y;     // grid lines (Nx+1, Ny+1)
  int u_left, u_right, u_low, u_up, v_left, v_right, v_low, v_up;
  double Re = 1.0;
  int order4 = 0;
  double alfa = 81.0, beta = 9.0 / 8.0;
};

struct Plan {
  CaseDesc cd;
  // grid sizes (operator_mesh.m:54-110)
  int Npx, Npy, Np, Nux_in, Nuy_in, Nvx_in, Nvy_in, Nu, Nv, NV;
  std::vector<double> hx, hy, gx, gy, xp, yp, xin, yin;
  std::vector<double> hxi, hyi, hxd, hyd, gxi, gyi, gxd, gyd;
  bool periodic = false, uniform = false;
  HComp cu, cv;
  // pressure Laplacian (second order): A = sum of two kron products
  // A = kron(diag(-hy), Lx) + kron(Ly, diag(-hx)), plus its diagonal
  HKron L1, L2;
  std::vector<double> Adiag;
  bool pres_bc = false;   // any 'pres' boundary (A nonsingular)
};

void build_plan(const CaseDesc& cd, Plan& P);

// factors of the fast-diagonalization direct Poisson solve (fdm_host.cpp):
// Lx Vx = diag(hx) Vx diag(lx), Vx^T diag(hx) Vx = I, the same in y; (nx0,
// ny0) is the dropped null mode of a singular A (-1 when A is nonsingular)
struct FDMHost {
  int nx = 0, ny = 0, nx0 = -1, ny0 = -1;
  std::vector<double> Vx, VxT, Vy, VyT, lx, ly;
};
void build_fdm_host(const Plan& P, FDMHost& F);
// the same factors for the Galerkin coarse operator P^T A P with aggregates
// of agg x agg cells (the last aggregate of a row or column may be smaller);
// VxT/VyT are left empty
void build_coarse_host(const Plan& P, int agg, FDMHost& C);

// IC(0) factor of the pressure matrix in the reference's form
// (make_cholinc.m:15-36, cg.F:394-430): M = (D + L) D^-1 (D + L^T), L the
// strict lower triangle of A, D_i = a_ii - sum_{c < i} a_ic^2 / D_c.  Lower
// entries of cell (i, j): west (cw), south (cs), and for periodic directions
// the wrap of the last column to the first (cwx[j]) and of the last row to the
// first (cwy[i]); empty vectors when absent.
struct IC0Host {
  std::vector<double> D, cw, cs, cwx, cwy;
};
void build_ic0_host(const Plan& P, IC0Host& H);

// sampled boundary values (set_bc_vectors.m:53-79), host copies
struct BCSamples {
  std::vector<double> s[S_NSRC];
};
// evaluate a YVec against samples: returns nt terms (ax, ay)
void eval_yvec(const HYVec& Y, const BCSamples& S, std::vector<std::vector<double>>& ax,
               std::vector<std::vector<double>>& ay);
int sample_len(const Plan& P, int src);

}  // namespace ins2d
