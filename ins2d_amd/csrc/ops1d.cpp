// Host-side construction of the 1-D factors of the INS2D operators; see
// ops1d.hpp.  Each builder cites the reference lines whose kron() assembly it
// factors.
#include "ops1d.hpp"

#include <algorithm>
#include <cmath>

namespace ins2d {

int Sp::max_row_nnz() const {
  size_t k = 0;
  for (auto& row : r) k = std::max(k, row.size());
  return (int)k;
}

void Sp::prune(double tol) {
  for (auto& row : r)
    for (auto it = row.begin(); it != row.end();)
      if (std::fabs(it->second) <= tol) it = row.erase(it); else ++it;
}

Sp eye(int n) {
  Sp A(n, n);
  for (int i = 0; i < n; ++i) A.set(i, i, 1.0);
  return A;
}

Sp diag(const std::vector<double>& d) {
  Sp A((int)d.size(), (int)d.size());
  for (int i = 0; i < (int)d.size(); ++i) A.set(i, i, d[i]);
  return A;
}

Sp mul(const Sp& A, const Sp& B) {
  if (A.n != B.m) throw std::runtime_error("mul: shape mismatch");
  Sp C(A.m, B.n);
  for (int i = 0; i < A.m; ++i)
    for (auto& ka : A.r[i])
      for (auto& kb : B.r[ka.first]) C.r[i][kb.first] += ka.second * kb.second;
  C.prune();
  return C;
}

Sp add(const Sp& A, double a, const Sp& B, double b) {
  if (A.m != B.m || A.n != B.n) throw std::runtime_error("add: shape mismatch");
  Sp C(A.m, A.n);
  for (int i = 0; i < A.m; ++i) {
    for (auto& kv : A.r[i]) C.r[i][kv.first] += a * kv.second;
    for (auto& kv : B.r[i]) C.r[i][kv.first] += b * kv.second;
  }
  C.prune();
  return C;
}

Sp scale(const Sp& A, double s) {
  Sp C = A;
  for (auto& row : C.r)
    for (auto& kv : row) kv.second *= s;
  return C;
}

Sp transpose(const Sp& A) {
  Sp C(A.n, A.m);
  for (int i = 0; i < A.m; ++i)
    for (auto& kv : A.r[i]) C.r[kv.first][i] = kv.second;
  return C;
}

std::vector<double> matvec(const Sp& A, const std::vector<double>& x) {
  if ((int)x.size() != A.n) throw std::runtime_error("matvec: shape mismatch");
  std::vector<double> y(A.m, 0.0);
  for (int i = 0; i < A.m; ++i)
    for (auto& kv : A.r[i]) y[i] += kv.second * x[kv.first];
  return y;
}

Sp spdiags(const std::vector<std::vector<double>>& cols, const std::vector<int>& offs, int m, int n) {
  Sp A(m, n);
  for (size_t k = 0; k < cols.size(); ++k) {
    int d = offs[k];
    const auto& B = cols[k];
    if (m >= n) {
      for (int j = 0; j < n; ++j) {
        int i = j - d;
        if (i >= 0 && i < m && j < (int)B.size()) A.r[i][j] = B[j];
      }
    } else {
      for (int i = 0; i < m; ++i) {
        int j = i + d;
        if (j >= 0 && j < n && i < (int)B.size()) A.r[i][j] = B[i];
      }
    }
  }
  A.prune();
  return A;
}

Sp spdiags_const(const std::vector<double>& vals, const std::vector<int>& offs, int m, int n) {
  std::vector<std::vector<double>> cols;
  int L = std::max(m, n) + 8;
  for (double v : vals) cols.push_back(std::vector<double>(L, v));
  return spdiags(cols, offs, m, n);
}

// u_ext = (Bin - Btemp Bbc Bin) u_in + Btemp ybc, Btemp = Bb (Bbc Bb)^-1
// (spatial/boundaryconditions/proposed/BC_general.m:1-7,124-125)
Closure make_closure(int Nt, int Nin, int Nb, int bin_off, const Sp& Bb, const Sp& Bbc,
                     const std::vector<double>& y1, const std::vector<double>& y2) {
  Sp Bin = spdiags_const({1.0}, {bin_off}, Nt, Nin);
  // S = Bbc Bb (Nb x Nb), dense inverse by Gauss-Jordan
  Sp S = mul(Bbc, Bb);
  std::vector<double> a(Nb * Nb, 0.0), inv(Nb * Nb, 0.0);
  for (int i = 0; i < Nb; ++i) {
    for (int j = 0; j < Nb; ++j) a[i * Nb + j] = S.get(i, j);
    inv[i * Nb + i] = 1.0;
  }
  for (int c = 0; c < Nb; ++c) {
    int piv = c;
    for (int i = c + 1; i < Nb; ++i)
      if (std::fabs(a[i * Nb + c]) > std::fabs(a[piv * Nb + c])) piv = i;
    if (std::fabs(a[piv * Nb + c]) < 1e-300) throw std::runtime_error("singular boundary closure");
    for (int j = 0; j < Nb; ++j) {
      std::swap(a[c * Nb + j], a[piv * Nb + j]);
      std::swap(inv[c * Nb + j], inv[piv * Nb + j]);
    }
    double p = a[c * Nb + c];
    for (int j = 0; j < Nb; ++j) { a[c * Nb + j] /= p; inv[c * Nb + j] /= p; }
    for (int i = 0; i < Nb; ++i) {
      if (i == c) continue;
      double f = a[i * Nb + c];
      if (f == 0.0) continue;
      for (int j = 0; j < Nb; ++j) { a[i * Nb + j] -= f * a[c * Nb + j]; inv[i * Nb + j] -= f * inv[c * Nb + j]; }
    }
  }
  Sp Sinv(Nb, Nb);
  for (int i = 0; i < Nb; ++i)
    for (int j = 0; j < Nb; ++j)
      if (inv[i * Nb + j] != 0.0) Sinv.set(i, j, inv[i * Nb + j]);
  Closure C;
  C.Btemp = mul(Bb, Sinv);
  C.B1D = add(Bin, 1.0, mul(C.Btemp, mul(Bbc, Bin)), -1.0);
  C.ybc1 = y1;
  C.ybc2 = y2;
  return C;
}

static Closure identity_closure(int Nt) {
  Closure C;
  C.B1D = eye(Nt);
  C.Btemp = Sp(Nt, 2);
  C.ybc1 = {0, 0};
  C.ybc2 = {0, 0};
  return C;
}

// ghosts on grid lines (BC_general.m:1-130)
Closure bc_general(int Nt, int Nin, int Nb, int bc1, int bc2, double h1, double h2) {
  if (Nt != Nin + Nb) throw std::runtime_error("bc_general: Nt != Nin + Nb");
  if (Nb == 0) return identity_closure(Nt);
  Sp Bbc(Nb, Nt), Bb(Nt, Nb);
  std::vector<double> y1(Nb, 0.0), y2(Nb, 0.0);
  int off = -1;
  if (Nb == 2) {
    Bb.set(0, 0, 1.0);
    Bb.set(Nt - 1, 1, 1.0);
    switch (bc1) {
      case BC_DIR: Bbc.set(0, 0, 1.0); y1[0] = 1.0; break;
      case BC_PRES: Bbc.set(0, 0, -1.0); Bbc.set(0, 1, 1.0); y1[0] = 2 * h1; break;
      case BC_PER:
        Bbc.set(0, 0, -1.0); Bbc.set(0, Nt - 2, 1.0); Bbc.set(1, 1, -1.0); Bbc.set(1, Nt - 1, 1.0); break;
      default: throw std::runtime_error("bc_general: boundary type not implemented");
    }
    switch (bc2) {
      case BC_DIR: Bbc.set(1, Nt - 1, 1.0); y2[1] = 1.0; break;
      case BC_PRES: Bbc.set(1, Nt - 2, -1.0); Bbc.set(1, Nt - 1, 1.0); y2[1] = 2 * h2; break;
      case BC_PER:
        Bbc.set(0, 0, -1.0); Bbc.set(0, Nt - 2, 1.0); Bbc.set(1, 1, -1.0); Bbc.set(1, Nt - 1, 1.0); break;
      default: throw std::runtime_error("bc_general: boundary type not implemented");
    }
  } else if (Nb == 1) {
    switch (bc1) {
      case BC_DIR: Bbc.set(0, 0, 1.0); y1[0] = 1.0; Bb.set(0, 0, 1.0); break;
      case BC_PRES: off = 0; break;
      case BC_PER: off = 0; Bbc.set(0, 0, -1.0); Bbc.set(0, Nt - 1, 1.0); Bb.set(Nt - 1, 0, 1.0); break;
      default: throw std::runtime_error("bc_general: boundary type not implemented");
    }
    switch (bc2) {
      case BC_DIR: Bbc.set(0, Nt - 1, 1.0); y2[0] = 1.0; Bb.set(Nt - 1, 0, 1.0); break;
      case BC_PRES: break;
      case BC_PER: off = 0; Bbc.set(0, 0, -1.0); Bbc.set(0, Nt - 1, 1.0); Bb.set(Nt - 1, 0, 1.0); break;
      default: throw std::runtime_error("bc_general: boundary type not implemented");
    }
  } else {
    throw std::runtime_error("bc_general: unsupported number of boundary points");
  }
  return make_closure(Nt, Nin, Nb, off, Bb, Bbc, y1, y2);
}

// ghosts on staggered lines (BC_general_stag.m:1-138; BC_diff_stag.m:1-160
// has the same Nb = 2 closures)
Closure bc_general_stag(int Nt, int Nin, int Nb, int bc1, int bc2, double h1, double h2, bool diff_variant) {
  if (Nt != Nin + Nb) throw std::runtime_error("bc_general_stag: Nt != Nin + Nb");
  if (Nb == 0) return identity_closure(Nt);
  Sp Bbc(Nb, Nt), Bb(Nt, Nb);
  std::vector<double> y1(Nb, 0.0), y2(Nb, 0.0);
  int off = -1;
  if (Nb == 2) {
    Bb.set(0, 0, 1.0);
    Bb.set(Nt - 1, 1, 1.0);
    switch (bc1) {
      case BC_DIR: Bbc.set(0, 0, 0.5); Bbc.set(0, 1, 0.5); y1[0] = 1.0; break;
      case BC_SYM: Bbc.set(0, 0, -1.0); Bbc.set(0, 1, 1.0); y1[0] = h1; break;
      case BC_PER:
        Bbc.set(0, 0, -1.0); Bbc.set(0, Nt - 2, 1.0); Bbc.set(1, 1, -1.0); Bbc.set(1, Nt - 1, 1.0); break;
      default: throw std::runtime_error("bc_general_stag: boundary type not implemented");
    }
    switch (bc2) {
      case BC_DIR: Bbc.set(1, Nt - 2, 0.5); Bbc.set(1, Nt - 1, 0.5); y2[1] = 1.0; break;
      case BC_SYM: Bbc.set(1, Nt - 2, -1.0); Bbc.set(1, Nt - 1, 1.0); y2[1] = h2; break;
      case BC_PER:
        Bbc.set(0, 0, -1.0); Bbc.set(0, Nt - 2, 1.0); Bbc.set(1, 1, -1.0); Bbc.set(1, Nt - 1, 1.0); break;
      default: throw std::runtime_error("bc_general_stag: boundary type not implemented");
    }
  } else if (Nb == 1 && !diff_variant) {
    switch (bc1) {
      case BC_DIR: Bbc.set(0, 0, 0.5); Bbc.set(0, 1, 0.5); y1[0] = 1.0; Bb.set(0, 0, 1.0); break;
      case BC_SYM: Bbc.set(0, 0, -1.0); Bbc.set(0, 1, 1.0); y1[0] = h1; break;
      case BC_PER: Bbc.set(0, 0, -1.0); Bbc.set(0, Nt - 1, 1.0); Bb.set(0, 0, 1.0); break;
      default: throw std::runtime_error("bc_general_stag: boundary type not implemented");
    }
    switch (bc2) {
      case BC_DIR: Bbc.set(0, Nt - 2, 0.5); Bbc.set(0, Nt - 1, 0.5); y2[0] = 1.0; Bb.set(Nt - 1, 0, 1.0); break;
      case BC_SYM: Bbc.set(0, Nt - 2, -1.0); Bbc.set(0, Nt - 1, 1.0); y2[0] = h2; break;
      case BC_PER: Bbc.set(0, 0, -1.0); Bbc.set(0, Nt - 1, 1.0); Bb.set(0, 0, 1.0); break;
      default: throw std::runtime_error("bc_general_stag: boundary type not implemented");
    }
  } else {
    throw std::runtime_error("bc_general_stag: unsupported number of boundary points");
  }
  return make_closure(Nt, Nin, Nb, off, Bb, Bbc, y1, y2);
}

HOp1 to_ell(const Sp& A, int min_k) {
  HOp1 o;
  o.nrows = A.m;
  o.ncols = A.n;
  o.kmax = std::max(1, A.max_row_nnz());
  o.K = std::max(min_k, o.kmax);
  o.col.assign((size_t)o.nrows * o.K, 0);
  o.val.assign((size_t)o.nrows * o.K, 0.0);
  for (int i = 0; i < A.m; ++i) {
    int k = 0;
    for (auto& kv : A.r[i]) {
      o.col[(size_t)i * o.K + k] = kv.first;
      o.val[(size_t)i * o.K + k] = kv.second;
      ++k;
    }
    // padded entries point at the row's first column (or 0) with weight 0
    for (; k < o.K; ++k) o.col[(size_t)i * o.K + k] = (A.r[i].empty() ? 0 : A.r[i].begin()->first);
  }
  return o;
}

// min_k 2 for the operators of the stencil kernels, which read exactly two
// (zero-padded) entries of every factor
static HKron kron_of(const Sp& Y, const Sp& X, int min_k = 2) {
  HKron K;
  K.X = to_ell(X, min_k);
  K.Y = to_ell(Y, min_k);
  return K;
}

// the momentum, divergence and gradient factors of a component must have
// exactly two entries per row (the form the generic stencil kernels read)
static void require_two_per_row(const HComp& C, const char* name) {
  auto chk = [&](const HKron& k) {
    if (k.X.K != 2 || k.Y.K != 2)
      throw std::runtime_error(std::string("operator of ") + name + " with more than two entries per row");
  };
  for (auto& t : C.conv) {
    chk(t.D);
    chk(t.I);
    chk(t.A);
  }
  for (auto& t : C.diff) {
    chk(t.D[0]);
    chk(t.S[0]);
  }
  chk(C.G);
  chk(C.M);
}

static std::vector<double> vmul(const std::vector<double>& a, const std::vector<double>& b) {
  std::vector<double> c(a.size());
  for (size_t i = 0; i < a.size(); ++i) c[i] = a[i] * b[i];
  return c;
}

static HFac fac_struct(std::vector<double> s) {
  HFac f;
  f.s = std::move(s);
  return f;
}
static HFac fac_sample(int src, const std::vector<double>& w = {}) {
  HFac f;
  f.src = src;
  f.w = w;
  return f;
}
static HFac fac_mapped(int src, const Sp& map) {
  HFac f;
  f.src = src;
  f.map = map;
  f.has_map = true;
  return f;
}
static bool all_zero(const std::vector<double>& v) {
  for (double x : v)
    if (x != 0.0) return false;
  return true;
}
// term with an x-structural factor (skipped when that factor is zero)
static void push_xs(HYVec& Y, std::vector<double> sx, HFac fy) {
  if (all_zero(sx)) return;
  HYTerm t;
  t.x = fac_struct(std::move(sx));
  t.y = std::move(fy);
  Y.t.push_back(std::move(t));
}
static void push_ys(HYVec& Y, HFac fx, std::vector<double> sy) {
  if (all_zero(sy)) return;
  HYTerm t;
  t.x = std::move(fx);
  t.y = fac_struct(std::move(sy));
  Y.t.push_back(std::move(t));
}

int sample_len(const Plan& P, int src) {
  switch (src) {
    case S_uLo: case S_uUp: return P.cd.Nx + 1;
    case S_uLo_i: case S_uUp_i: return (int)P.xin.size();
    case S_uLe_i: case S_uRi_i: return P.cd.Ny;
    case S_vLe: case S_vRi: return P.cd.Ny + 1;
    case S_vLo_i: case S_vUp_i: return P.cd.Nx;
    case S_vLe_i: case S_vRi_i: return (int)P.yin.size();
    default: return 0;
  }
}

static std::vector<double> eval_fac(const HFac& f, const BCSamples& S) {
  if (f.src == S_NONE) return f.s;
  std::vector<double> v = S.s[f.src];
  if (!f.w.empty()) v = vmul(v, f.w);
  if (f.has_map) v = matvec(f.map, v);
  return v;
}

void eval_yvec(const HYVec& Y, const BCSamples& S, std::vector<std::vector<double>>& ax,
               std::vector<std::vector<double>>& ay) {
  ax.clear();
  ay.clear();
  for (auto& t : Y.t) {
    ax.push_back(eval_fac(t.x, S));
    ay.push_back(eval_fac(t.y, S));
  }
}

static bool is(int a, int b) { return a == b; }

void build_plan(const CaseDesc& cd, Plan& P) {
  if (cd.order4) throw std::runtime_error("fourth-order operators use the periodic uniform kernels");
  P.cd = cd;
  const int Nx = cd.Nx, Ny = cd.Ny;
  if ((int)cd.x.size() != Nx + 1 || (int)cd.y.size() != Ny + 1) throw std::runtime_error("grid size mismatch");
  // ---- mesh_generation.m:17-39
  P.hx.resize(Nx); P.hy.resize(Ny); P.xp.resize(Nx); P.yp.resize(Ny);
  for (int i = 0; i < Nx; ++i) { P.hx[i] = cd.x[i + 1] - cd.x[i]; P.xp[i] = (cd.x[i] + cd.x[i + 1]) / 2; }
  for (int j = 0; j < Ny; ++j) { P.hy[j] = cd.y[j + 1] - cd.y[j]; P.yp[j] = (cd.y[j] + cd.y[j + 1]) / 2; }
  P.gx.assign(Nx + 1, 0.0); P.gy.assign(Ny + 1, 0.0);
  P.gx[0] = P.hx[0] / 2; P.gx[Nx] = P.hx[Nx - 1] / 2;
  for (int i = 1; i < Nx; ++i) P.gx[i] = (P.hx[i - 1] + P.hx[i]) / 2;
  P.gy[0] = P.hy[0] / 2; P.gy[Ny] = P.hy[Ny - 1] / 2;
  for (int j = 1; j < Ny; ++j) P.gy[j] = (P.hy[j - 1] + P.hy[j]) / 2;
  const auto &hx = P.hx, &hy = P.hy, &gx = P.gx, &gy = P.gy;

  // ---- operator_mesh.m:54-110
  P.Npx = Nx; P.Npy = Ny; P.Np = Nx * Ny;
  int Nux_in = Nx + 1;
  if (is(cd.u_left, BC_DIR) || is(cd.u_left, BC_SYM)) Nux_in--;
  if (is(cd.u_right, BC_DIR) || is(cd.u_right, BC_SYM)) Nux_in--;
  if (is(cd.u_left, BC_PER) && is(cd.u_right, BC_PER)) Nux_in--;
  int Nvy_in = Ny + 1;
  if (is(cd.v_low, BC_DIR) || is(cd.v_low, BC_SYM)) Nvy_in--;
  if (is(cd.v_up, BC_DIR) || is(cd.v_up, BC_SYM)) Nvy_in--;
  if (is(cd.v_low, BC_PER) && is(cd.v_up, BC_PER)) Nvy_in--;
  const int Nuy_in = Ny, Nvx_in = Nx;
  const int Nux_t = Nux_in + 2, Nuy_t = Nuy_in + 2, Nvx_t = Nvx_in + 2, Nvy_t = Nvy_in + 2;
  P.Nux_in = Nux_in; P.Nuy_in = Nuy_in; P.Nvx_in = Nvx_in; P.Nvy_in = Nvy_in;
  P.Nu = Nux_in * Nuy_in; P.Nv = Nvx_in * Nvy_in; P.NV = P.Nu + P.Nv;

  // x-direction metrics (operator_mesh.m:176-257)
  std::vector<double> gxd = gx;
  gxd[0] = hx[0]; gxd[Nx] = hx[Nx - 1];
  std::vector<double> hxd, gxi;
  int dposx;
  const int L = cd.u_left, R = cd.u_right;
  if (L == BC_DIR && R == BC_DIR) {
    P.xin.assign(cd.x.begin() + 1, cd.x.end() - 1); hxd = hx; gxi.assign(gx.begin() + 1, gx.end() - 1); dposx = 1;
  } else if (L == BC_DIR && R == BC_PRES) {
    P.xin.assign(cd.x.begin() + 1, cd.x.end()); hxd = hx; hxd.push_back(hx[Nx - 1]);
    gxi.assign(gx.begin() + 1, gx.end()); dposx = 1;
  } else if (L == BC_PRES && R == BC_DIR) {
    P.xin.assign(cd.x.begin(), cd.x.end() - 1); hxd = hx; hxd.insert(hxd.begin(), hx[0]);
    gxi.assign(gx.begin(), gx.end() - 1); dposx = 0;
  } else if (L == BC_PRES && R == BC_PRES) {
    P.xin = cd.x; hxd = hx; hxd.insert(hxd.begin(), hx[0]); hxd.push_back(hx[Nx - 1]); gxi = gx; dposx = 0;
  } else if (L == BC_PER && R == BC_PER) {
    P.xin.assign(cd.x.begin(), cd.x.end() - 1); hxd = hx; hxd.insert(hxd.begin(), hx[Nx - 1]);
    gxi.assign(gx.begin() + 1, gx.end() - 1); gxi.insert(gxi.begin(), gx[0] + gx[Nx]);
    gxd[0] = (hx[0] + hx[Nx - 1]) / 2; gxd[Nx] = (hx[0] + hx[Nx - 1]) / 2; dposx = 0;
  } else {
    throw std::runtime_error("unsupported u left/right boundary combination");
  }
  Sp Bvux = spdiags_const({1.0}, {dposx}, Nux_in, Nvx_t - 1);

  std::vector<double> gyd = gy;
  gyd[0] = hy[0]; gyd[Ny] = hy[Ny - 1];
  std::vector<double> hyd, gyi;
  int dposy;
  const int Lo = cd.v_low, Up = cd.v_up;
  if (Lo == BC_DIR && Up == BC_DIR) {
    P.yin.assign(cd.y.begin() + 1, cd.y.end() - 1); hyd = hy; gyi.assign(gy.begin() + 1, gy.end() - 1); dposy = 1;
  } else if (Lo == BC_DIR && Up == BC_PRES) {
    P.yin.assign(cd.y.begin() + 1, cd.y.end()); hyd = hy; hyd.push_back(hy[Ny - 1]);
    gyi.assign(gy.begin() + 1, gy.end()); dposy = 1;
  } else if (Lo == BC_PRES && Up == BC_DIR) {
    P.yin.assign(cd.y.begin(), cd.y.end() - 1); hyd = hy; hyd.insert(hyd.begin(), hy[0]);
    gyi.assign(gy.begin(), gy.end() - 1); dposy = 0;
  } else if (Lo == BC_PRES && Up == BC_PRES) {
    P.yin = cd.y; hyd = hy; hyd.insert(hyd.begin(), hy[0]); hyd.push_back(hy[Ny - 1]); gyi = gy; dposy = 0;
  } else if (Lo == BC_PER && Up == BC_PER) {
    P.yin.assign(cd.y.begin(), cd.y.end() - 1); hyd = hy; hyd.insert(hyd.begin(), hy[Ny - 1]);
    gyi.assign(gy.begin() + 1, gy.end() - 1); gyi.insert(gyi.begin(), gy[0] + gy[Ny]);
    gyd[0] = (hy[0] + hy[Ny - 1]) / 2; gyd[Ny] = (hy[0] + hy[Ny - 1]) / 2; dposy = 0;
  } else {
    throw std::runtime_error("unsupported v low/up boundary combination");
  }
  Sp Buvy = spdiags_const({1.0}, {dposy}, Nvy_in, Nuy_t - 1);
  std::vector<double> hxi = hx, hyi = hy;
  P.hxi = hxi; P.hyi = hyi; P.hxd = hxd; P.hyd = hyd; P.gxi = gxi; P.gyi = gyi; P.gxd = gxd; P.gyd = gyd;
  P.periodic = (L == BC_PER && R == BC_PER && Lo == BC_PER && Up == BC_PER && cd.u_low == BC_PER &&
                cd.u_up == BC_PER && cd.v_left == BC_PER && cd.v_right == BC_PER);
  double hmax = 0, hmin = 1e300;
  for (double h : hx) { hmax = std::max(hmax, h); hmin = std::min(hmin, h); }
  bool ux = (hmax - hmin) <= 1e-14;
  hmax = 0; hmin = 1e300;
  for (double h : hy) { hmax = std::max(hmax, h); hmin = std::min(hmin, h); }
  P.uniform = ux && (hmax - hmin) <= 1e-14;
  P.pres_bc = (L == BC_PRES || R == BC_PRES || Lo == BC_PRES || Up == BC_PRES);

  const double w = 0.5;
  auto A1 = [&](int n) { return spdiags_const({w, w}, {0, 1}, n - 1, n); };
  auto D1 = [&](int nr, int nc) { return spdiags_const({-1.0, 1.0}, {0, 1}, nr, nc); };
  auto S1 = [&](const std::vector<double>& h, int nr, int nc) {
    std::vector<double> inv(h.size());
    for (size_t i = 0; i < h.size(); ++i) inv[i] = 1.0 / h[i];
    std::vector<double> ninv(inv);
    for (auto& v : ninv) v = -v;
    return spdiags({ninv, inv}, {0, 1}, nr, nc);
  };
  std::vector<double> hx2(Nx + 2), hy2(Ny + 2);
  bool perx = (L == BC_PER && R == BC_PER), pery = (Lo == BC_PER && Up == BC_PER);
  hx2[0] = perx ? hx[Nx - 1] : hx[0]; hx2[Nx + 1] = perx ? hx[0] : hx[Nx - 1];
  hy2[0] = pery ? hy[Ny - 1] : hy[0]; hy2[Ny + 1] = pery ? hy[0] : hy[Ny - 1];
  for (int i = 0; i < Nx; ++i) hx2[i + 1] = hx[i];
  for (int j = 0; j < Ny; ++j) hy2[j + 1] = hy[j];

  // closures shared by several operators
  Closure cu_x = bc_general(Nux_t, Nux_in, 2, cd.u_left, cd.u_right, hx[0], hx[Nx - 1]);
  Closure cu_y = bc_general_stag(Nuy_t, Nuy_in, 2, cd.u_low, cd.u_up, hy[0], hy[Ny - 1], false);
  Closure cu_yd = bc_general_stag(Nuy_t, Nuy_in, 2, cd.u_low, cd.u_up, hy[0], hy[Ny - 1], true);
  Closure cv_x = bc_general_stag(Nvx_t, Nvx_in, 2, cd.v_left, cd.v_right, hx[0], hx[Nx - 1], false);
  Closure cv_xd = bc_general_stag(Nvx_t, Nvx_in, 2, cd.v_left, cd.v_right, hx[0], hx[Nx - 1], true);
  Closure cv_y = bc_general(Nvy_t, Nvy_in, 2, cd.v_low, cd.v_up, hy[0], hy[Ny - 1]);

  // ================= u component ==========================================
  HComp& U = P.cu;
  U.nx = Nux_in; U.ny = Nuy_in;
  {
    // du2dx = Cux ((Iu_ux u + yIu_ux) .* (Au_ux u + yAu_ux))   (convection.m:198-200)
    HConv c;
    Sp Ax = mul(A1(Nux_t), cu_x.B1D);                      // operator_averaging.m:33-42
    Sp Ab = mul(A1(Nux_t), cu_x.Btemp);
    c.D = kron_of(eye(Nuy_in), D1(Nux_t - 2, Nux_t - 1));   // operator_convection_diffusion.m:84-86
    c.A = kron_of(eye(Nuy_in), Ax);
    push_xs(c.yA, matvec(Ab, cu_x.ybc1), fac_sample(S_uLe_i));
    push_xs(c.yA, matvec(Ab, cu_x.ybc2), fac_sample(S_uRi_i));
    c.I = kron_of(diag(hyi), Ax);                          // operator_interpolation.m:67-76
    push_xs(c.yI, matvec(Ab, cu_x.ybc1), fac_sample(S_uLe_i, hyi));
    push_xs(c.yI, matvec(Ab, cu_x.ybc2), fac_sample(S_uRi_i, hyi));
    c.icomp = 0;
    U.conv.push_back(c);
  }
  {
    // duvdy = Cuy ((Iv_uy v + yIv_uy) .* (Au_uy u + yAu_uy))   (convection.m:202-204)
    HConv c;
    Sp Ay = mul(A1(Nuy_t), cu_y.B1D);                      // operator_averaging.m:46-55
    Sp Ab = mul(A1(Nuy_t), cu_y.Btemp);
    c.D = kron_of(D1(Nuy_t - 2, Nuy_t - 1), eye(Nux_in));   // operator_convection_diffusion.m:93-95
    c.A = kron_of(Ay, eye(Nux_in));
    push_ys(c.yA, fac_sample(S_uLo_i), matvec(Ab, cu_y.ybc1));
    push_ys(c.yA, fac_sample(S_uUp_i), matvec(Ab, cu_y.ybc2));
    // Iv_uy (operator_interpolation.m:79-113), incl. the 'pres' half-cell fix
    Sp I1 = A1(Nvx_t);
    if (cd.u_right == BC_PRES) I1.scale_row(I1.m - 1, 0.5);
    if (cd.u_left == BC_PRES) I1.scale_row(0, 0.5);
    Sp Xr = mul(mul(Bvux, I1), diag(hx2));
    Closure lr = cv_x;
    int Nb = Nuy_in + 1 - Nvy_in;
    Closure lu = bc_general(Nuy_in + 1, Nvy_in, Nb, cd.v_low, cd.v_up, hy[0], hy[Ny - 1]);
    Sp X = mul(Xr, lr.B1D);
    c.I = kron_of(lu.B1D, X);
    Sp XrB = mul(Xr, lr.Btemp);
    push_xs(c.yI, matvec(XrB, lr.ybc1), fac_sample(S_vLe));   // set_bc_vectors.m:445-446
    push_xs(c.yI, matvec(XrB, lr.ybc2), fac_sample(S_vRi));
    push_ys(c.yI, fac_mapped(S_vLo_i, X), matvec(lu.Btemp, lu.ybc1));   // :450-451
    push_ys(c.yI, fac_mapped(S_vUp_i, X), matvec(lu.Btemp, lu.ybc2));
    c.icomp = 1;
    U.conv.push_back(c);
  }
  {
    // d2u = Dux (Su_ux u + ySu_ux)/Re + Duy (Su_uy u + ySu_uy)/Re (operator_convection_diffusion.m:460)
    HDiff d;
    d.D[0] = kron_of(diag(hyi), D1(Nux_t - 2, Nux_t - 1));
    Sp S = S1(hxd, Nux_t - 1, Nux_t);                      // :133-142
    d.S[0] = kron_of(eye(Ny), mul(S, cu_x.B1D));
    d.sscale[0] = 1.0 / cd.Re;
    Sp Sb = scale(mul(S, cu_x.Btemp), 1.0 / cd.Re);
    push_xs(d.yS, matvec(Sb, cu_x.ybc1), fac_sample(S_uLe_i));  // set_bc_vectors.m:268-269
    push_xs(d.yS, matvec(Sb, cu_x.ybc2), fac_sample(S_uRi_i));
    U.diff.push_back(d);
  }
  {
    HDiff d;
    d.D[0] = kron_of(D1(Nuy_t - 2, Nuy_t - 1), diag(gxi));
    Sp S = S1(gyd, Nuy_t - 1, Nuy_t);                      // :147-158 (BC_diff_stag)
    d.S[0] = kron_of(mul(S, cu_yd.B1D), eye(Nux_in));
    d.sscale[0] = 1.0 / cd.Re;
    Sp Sb = scale(mul(S, cu_yd.Btemp), 1.0 / cd.Re);
    push_ys(d.yS, fac_sample(S_uLo_i), matvec(Sb, cu_yd.ybc1));
    push_ys(d.yS, fac_sample(S_uUp_i), matvec(Sb, cu_yd.ybc2));
    U.diff.push_back(d);
  }
  U.ox = gxi; U.oy = hyi;                                 // Omu = kron(hyi, gxi)

  // ================= v component ==========================================
  HComp& Vc = P.cv;
  Vc.nx = Nvx_in; Vc.ny = Nvy_in;
  {
    // duvdx = Cvx ((Iu_vx u + yIu_vx) .* (Av_vx v + yAv_vx))   (convection.m:206-208)
    HConv c;
    Sp Ax = mul(A1(Nvx_t), cv_x.B1D);                      // operator_averaging.m:61-70
    Sp Ab = mul(A1(Nvx_t), cv_x.Btemp);
    c.D = kron_of(eye(Nvy_in), D1(Nvx_t - 2, Nvx_t - 1));   // operator_convection_diffusion.m:108-110
    c.A = kron_of(eye(Nvy_in), Ax);
    push_xs(c.yA, matvec(Ab, cv_x.ybc1), fac_sample(S_vLe_i));
    push_xs(c.yA, matvec(Ab, cv_x.ybc2), fac_sample(S_vRi_i));
    // Iu_vx (operator_interpolation.m:119-150)
    Sp I1 = A1(Nuy_t);
    if (cd.v_up == BC_PRES) I1.scale_row(I1.m - 1, 0.5);
    if (cd.v_low == BC_PRES) I1.scale_row(0, 0.5);
    Sp Yr = mul(mul(Buvy, I1), diag(hy2));
    Closure lu = cu_y;
    int Nb = Nvx_in + 1 - Nux_in;
    Closure lr = bc_general(Nvx_in + 1, Nux_in, Nb, cd.u_left, cd.u_right, hx[0], hx[Nx - 1]);
    Sp Y = mul(Yr, lu.B1D);
    c.I = kron_of(Y, lr.B1D);
    Sp YrB = mul(Yr, lu.Btemp);
    push_ys(c.yI, fac_sample(S_uLo), matvec(YrB, lu.ybc1));     // set_bc_vectors.m:487-488
    push_ys(c.yI, fac_sample(S_uUp), matvec(YrB, lu.ybc2));
    push_xs(c.yI, matvec(lr.Btemp, lr.ybc1), fac_mapped(S_uLe_i, Y));   // :492-493
    push_xs(c.yI, matvec(lr.Btemp, lr.ybc2), fac_mapped(S_uRi_i, Y));
    c.icomp = 0;
    Vc.conv.push_back(c);
  }
  {
    // dv2dy = Cvy ((Iv_vy v + yIv_vy) .* (Av_vy v + yAv_vy))   (convection.m:210-212)
    HConv c;
    Sp Ay = mul(A1(Nvy_t), cv_y.B1D);                      // operator_averaging.m:74-83
    Sp Ab = mul(A1(Nvy_t), cv_y.Btemp);
    c.D = kron_of(D1(Nvy_t - 2, Nvy_t - 1), eye(Nvx_in));   // operator_convection_diffusion.m:117-119
    c.A = kron_of(Ay, eye(Nvx_in));
    push_ys(c.yA, fac_sample(S_vLo_i), matvec(Ab, cv_y.ybc1));
    push_ys(c.yA, fac_sample(S_vUp_i), matvec(Ab, cv_y.ybc2));
    c.I = kron_of(Ay, diag(hxi));                          // operator_interpolation.m:153-162
    push_ys(c.yI, fac_sample(S_vLo_i, hxi), matvec(Ab, cv_y.ybc1));
    push_ys(c.yI, fac_sample(S_vUp_i, hxi), matvec(Ab, cv_y.ybc2));
    c.icomp = 1;
    Vc.conv.push_back(c);
  }
  {
    HDiff d;
    d.D[0] = kron_of(diag(gyi), D1(Nvx_t - 2, Nvx_t - 1));
    Sp S = S1(gxd, Nvx_t - 1, Nvx_t);                      // :225-236 (BC_diff_stag)
    d.S[0] = kron_of(eye(Nvy_in), mul(S, cv_xd.B1D));
    d.sscale[0] = 1.0 / cd.Re;
    Sp Sb = scale(mul(S, cv_xd.Btemp), 1.0 / cd.Re);
    push_xs(d.yS, matvec(Sb, cv_xd.ybc1), fac_sample(S_vLe_i));   // set_bc_vectors.m:328-329
    push_xs(d.yS, matvec(Sb, cv_xd.ybc2), fac_sample(S_vRi_i));
    Vc.diff.push_back(d);
  }
  {
    HDiff d;
    d.D[0] = kron_of(D1(Nvy_t - 2, Nvy_t - 1), diag(hxi));
    Sp S = S1(hyd, Nvy_t - 1, Nvy_t);                      // :241-250
    d.S[0] = kron_of(mul(S, cv_y.B1D), eye(Nx));
    d.sscale[0] = 1.0 / cd.Re;
    Sp Sb = scale(mul(S, cv_y.Btemp), 1.0 / cd.Re);
    push_ys(d.yS, fac_sample(S_vLo_i), matvec(Sb, cv_y.ybc1));
    push_ys(d.yS, fac_sample(S_vUp_i), matvec(Sb, cv_y.ybc2));
    Vc.diff.push_back(d);
  }
  Vc.ox = hxi; Vc.oy = gyi;                               // Omv = kron(gyi, hxi)

  // ================= divergence / gradient (operator_divergence.m:54-167) ==
  auto dpos = [](int lo, int hi) {
    int pos = 0;
    if (hi == BC_PRES && lo == BC_PRES) pos = 1;
    if (hi != BC_PRES && lo == BC_PRES) pos = 1;
    if (hi == BC_PRES && lo != BC_PRES) pos = 0;
    if (hi == BC_PER && lo == BC_PER) pos = 1;
    return pos;
  };
  Sp BMx = spdiags_const({1.0}, {dpos(L, R)}, P.Npx, Nux_t - 1);
  Sp M1x = mul(BMx, D1(Nux_t - 1, Nux_t));
  Sp Mx1 = mul(M1x, cu_x.B1D);                              // Npx x Nux_in
  Sp BMy = spdiags_const({1.0}, {dpos(Lo, Up)}, P.Npy, Nvy_t - 1);
  Sp M1y = mul(BMy, D1(Nvy_t - 1, Nvy_t));
  Sp My1 = mul(M1y, cv_y.B1D);                              // Npy x Nvy_in
  U.M = kron_of(diag(hy), Mx1);
  push_xs(U.yM, matvec(mul(M1x, cu_x.Btemp), cu_x.ybc1), fac_sample(S_uLe_i, hy));   // set_bc_vectors.m:100-101
  push_xs(U.yM, matvec(mul(M1x, cu_x.Btemp), cu_x.ybc2), fac_sample(S_uRi_i, hy));
  Vc.M = kron_of(My1, diag(hx));
  push_ys(Vc.yM, fac_sample(S_vLo_i, hx), matvec(mul(M1y, cv_y.Btemp), cv_y.ybc1));   // :109-110
  push_ys(Vc.yM, fac_sample(S_vUp_i, hx), matvec(mul(M1y, cv_y.Btemp), cv_y.ybc2));
  U.G = kron_of(diag(hy), scale(transpose(Mx1), -1.0));
  Vc.G = kron_of(scale(transpose(My1), -1.0), diag(hx));

  // pressure boundary vectors y_px, y_py (set_bc_vectors.m:156-176); p_inf = 0
  // (boundary_conditions.m:107) makes them vanish, so they are not stored.

  // ---- Laplacian A = M Om^-1 G = -(kron(diag(hy), Lx) + kron(Ly, diag(hx)))
  Sp Lx = mul(mul(Mx1, diag(std::vector<double>(gxi.size()))), transpose(Mx1));
  {
    std::vector<double> ig(gxi.size());
    for (size_t i = 0; i < gxi.size(); ++i) ig[i] = 1.0 / gxi[i];
    Lx = mul(mul(Mx1, diag(ig)), transpose(Mx1));
  }
  Sp Ly;
  {
    std::vector<double> ig(gyi.size());
    for (size_t i = 0; i < gyi.size(); ++i) ig[i] = 1.0 / gyi[i];
    Ly = mul(mul(My1, diag(ig)), transpose(My1));
  }
  std::vector<double> mhy(Ny), mhx(Nx);
  for (int j = 0; j < Ny; ++j) mhy[j] = -hy[j];
  for (int i = 0; i < Nx; ++i) mhx[i] = -hx[i];
  P.L1 = kron_of(diag(mhy), Lx, 1);
  P.L2 = kron_of(Ly, diag(mhx), 1);
  P.Adiag.resize(P.Np);
  for (int j = 0; j < Ny; ++j)
    for (int i = 0; i < Nx; ++i) P.Adiag[i + (size_t)Nx * j] = -(hy[j] * Lx.get(i, i) + hx[i] * Ly.get(j, j));
  require_two_per_row(P.cu, "u");
  require_two_per_row(P.cv, "v");
}

}  // namespace ins2d
