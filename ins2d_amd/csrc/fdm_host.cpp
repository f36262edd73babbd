// Host set-up of the fast-diagonalization direct pressure solve (fdm.hip)
// Direct pressure solve on tensor-product grids by fast diagonalization
// (Lynch, Rice & Thomas, Numer. Math. 6, 1964).
//
// The second-order pressure matrix of a rectangular INS2D grid is a sum of
// two Kronecker products (operator_divergence.m:54-167 builds A = M Om^-1 G
// from 1-D factors):
//     A = -(Hy (x) Lx + Ly (x) Hx),  Hx = diag(hx), Hy = diag(hy),
//     Lx = Mx diag(1/gx) Mx^T,  Ly = My diag(1/gy) My^T   (symmetric PSD).
// With the generalized eigen-decompositions Lx Vx = Hx Vx Lamx, Vx^T Hx Vx = I
// (and the same in y),
//     A^-1 f = -(Vy (x) Vx) diag(1 / (lamx_i + lamy_j)) (Vy (x) Vx)^T f,
// i.e. with F the Npx x Npy array of f (x fastest):
//     P = Vx [ -(Vx^T F Vy) ./ (lamx_i + lamy_j) ] Vy^T,
// four dense FP64 GEMMs on the GPU.  When A is singular (no 'pres' boundary)
// the mode with lamx = lamy = 0 (the constant) is dropped, which returns the
// solution with zero area-weighted mean; the reference's direct solve of
// such a system fixes an arbitrary constant (pressure_poisson.m:12 with
// main.m:34 silencing the singular-matrix warning).
//
// The 1-D eigenproblems are solved once on the host: Householder reduction to
// tridiagonal form (skipped when the operator is already tridiagonal, i.e. not
// periodic) and the implicit QL iteration with Wilkinson shifts, accumulating
// the eigenvectors.
#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>
#include <vector>

#include "ops1d.hpp"

namespace ins2d {

namespace {

// symmetric eigen-decomposition of the n x n column-major matrix Z (in: the
// matrix, out: orthonormal eigenvectors as columns), eigenvalues in w
void sym_eig(int n, std::vector<double>& Z, std::vector<double>& w) {
  std::vector<double> d(n), e(n, 0.0);
  auto at = [&](int i, int j) -> double& { return Z[(size_t)i + (size_t)n * j]; };
  bool tridiag = true;
  for (int j = 0; j < n && tridiag; ++j)
    for (int i = 0; i < n; ++i)
      if (std::abs(i - j) > 1 && at(i, j) != 0.0) { tridiag = false; break; }
  if (tridiag) {
    for (int i = 0; i < n; ++i) d[i] = at(i, i);
    for (int i = 0; i + 1 < n; ++i) e[i] = at(i + 1, i);
    std::fill(Z.begin(), Z.end(), 0.0);
    for (int i = 0; i < n; ++i) at(i, i) = 1.0;
  } else {
    // Householder: A <- H_k A H_k zeroes column k below the subdiagonal;
    // Q = H_0 H_1 ... accumulated in Q
    std::vector<double> A = Z, Q((size_t)n * n, 0.0), v(n), p(n), q(n);
    auto a = [&](int i, int j) -> double& { return A[(size_t)i + (size_t)n * j]; };
    for (int i = 0; i < n; ++i) Q[(size_t)i * (n + 1)] = 1.0;
    for (int k = 0; k + 2 < n; ++k) {
      double sig = 0.0;
      for (int i = k + 1; i < n; ++i) sig += a(i, k) * a(i, k);
      const double nrm = std::sqrt(sig);
      if (nrm == 0.0) continue;
      const double x0 = a(k + 1, k);
      const double alpha = x0 > 0 ? -nrm : nrm;
      for (int i = 0; i < n; ++i) v[i] = 0.0;
      v[k + 1] = x0 - alpha;
      for (int i = k + 2; i < n; ++i) v[i] = a(i, k);
      double vv = 0.0;
      for (int i = k + 1; i < n; ++i) vv += v[i] * v[i];
      if (vv == 0.0) continue;
      const double beta = 2.0 / vv;
      // p = beta A v (trailing block), K = beta/2 v.p, q = p - K v
      for (int i = k + 1; i < n; ++i) p[i] = 0.0;
      for (int j = k + 1; j < n; ++j) {
        const double vj = v[j];
        if (vj == 0.0) continue;
        for (int i = k + 1; i < n; ++i) p[i] += a(i, j) * vj;
      }
      double vp = 0.0;
      for (int i = k + 1; i < n; ++i) { p[i] *= beta; vp += v[i] * p[i]; }
      const double K = 0.5 * beta * vp;
      for (int i = k + 1; i < n; ++i) q[i] = p[i] - K * v[i];
      for (int j = k + 1; j < n; ++j)
        for (int i = k + 1; i < n; ++i) a(i, j) -= v[i] * q[j] + q[i] * v[j];
      a(k + 1, k) = alpha;
      a(k, k + 1) = alpha;
      for (int i = k + 2; i < n; ++i) { a(i, k) = 0.0; a(k, i) = 0.0; }
      // Q <- Q H: Q[:, k+1:] -= beta (Q v) v^T
      for (int r = 0; r < n; ++r) {
        double s = 0.0;
        for (int i = k + 1; i < n; ++i) s += Q[(size_t)r + (size_t)n * i] * v[i];
        s *= beta;
        if (s == 0.0) continue;
        for (int i = k + 1; i < n; ++i) Q[(size_t)r + (size_t)n * i] -= s * v[i];
      }
    }
    for (int i = 0; i < n; ++i) d[i] = a(i, i);
    for (int i = 0; i + 1 < n; ++i) e[i] = a(i + 1, i);
    Z = Q;
  }
  // implicit QL with Wilkinson shifts on (d, e); rotations applied to the
  // columns of Z (contiguous in column-major storage)
  const double eps = std::numeric_limits<double>::epsilon();
  for (int l = 0; l < n; ++l) {
    int iter = 0;
    while (true) {
      int m = l;
      for (; m < n - 1; ++m) {
        const double dd = std::abs(d[m]) + std::abs(d[m + 1]);
        if (std::abs(e[m]) <= eps * dd) break;
      }
      if (m == l) break;
      if (++iter > 60) throw std::runtime_error("eigenvalue iteration did not converge");
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = std::hypot(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + (g >= 0 ? r : -r));
      double s = 1.0, c = 1.0, pp = 0.0;
      int i = m - 1;
      bool underflow = false;
      for (; i >= l; --i) {
        const double f = s * e[i], b = c * e[i];
        r = std::hypot(f, g);
        e[i + 1] = r;
        if (r == 0.0) {
          d[i + 1] -= pp;
          e[m] = 0.0;
          underflow = true;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - pp;
        r = (d[i] - g) * s + 2.0 * c * b;
        pp = s * r;
        d[i + 1] = g + pp;
        g = c * r - b;
        double* zi = &Z[(size_t)n * i];
        double* zi1 = &Z[(size_t)n * (i + 1)];
        for (int k = 0; k < n; ++k) {
          const double t = zi1[k];
          zi1[k] = s * zi[k] + c * t;
          zi[k] = c * zi[k] - s * t;
        }
      }
      if (underflow) continue;
      d[l] -= pp;
      e[l] = g;
      e[m] = 0.0;
    }
  }
  w = d;
}

std::vector<double> dense_of(const HOp1& o) {
  std::vector<double> M((size_t)o.nrows * o.ncols, 0.0);
  for (int i = 0; i < o.nrows; ++i)
    for (int k = 0; k < o.K; ++k) M[(size_t)i + (size_t)o.nrows * o.col[(size_t)i * o.K + k]] += o.val[(size_t)i * o.K + k];
  return M;
}

// generalized problem L v = lam diag(h) v, V^T diag(h) V = I (L dense n x n)
void gen_eig_dense(int n, const std::vector<double>& Ld, const std::vector<double>& h, std::vector<double>& V,
                   std::vector<double>& lam) {
  if ((int)h.size() != n || Ld.size() != (size_t)n * n) throw std::runtime_error("fast diagonalization: bad 1-D operator");
  std::vector<double> S = Ld;
  double amax = 0.0, asym = 0.0;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      amax = std::max(amax, std::abs(S[(size_t)i + (size_t)n * j]));
      asym = std::max(asym, std::abs(S[(size_t)i + (size_t)n * j] - S[(size_t)j + (size_t)n * i]));
    }
  if (asym > 1e-12 * amax) throw std::runtime_error("fast diagonalization needs a symmetric 1-D Laplacian");
  std::vector<double> rs(n);
  for (int i = 0; i < n; ++i) rs[i] = 1.0 / std::sqrt(h[i]);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) S[(size_t)i + (size_t)n * j] *= rs[i] * rs[j];
  sym_eig(n, S, lam);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) S[(size_t)i + (size_t)n * j] *= rs[i];
  V.swap(S);
  // Neumann or periodic in this direction: L 1 = 0 exactly, but QL returns the
  // zero eigenvalue only to within eps * max|lam|, which next to a small
  // eigenvalue of the other direction would perturb the solve; use the exact
  // pair (0, constant / sqrt(sum h)) instead
  double rmax = 0.0;
  for (int i = 0; i < n; ++i) {
    double r = 0.0;
    for (int j = 0; j < n; ++j) r += Ld[(size_t)i + (size_t)n * j];
    rmax = std::max(rmax, std::abs(r));
  }
  if (rmax <= 64.0 * std::numeric_limits<double>::epsilon() * amax) {
    int k = 0;
    for (int i = 1; i < n; ++i)
      if (std::abs(lam[i]) < std::abs(lam[k])) k = i;
    double hs = 0.0;
    for (int i = 0; i < n; ++i) hs += h[i];
    lam[k] = 0.0;
    for (int i = 0; i < n; ++i) V[(size_t)i + (size_t)n * k] = 1.0 / std::sqrt(hs);
  }
}

void gen_eig(const HOp1& L, const std::vector<double>& h, std::vector<double>& V, std::vector<double>& lam) {
  if (L.ncols != L.nrows) throw std::runtime_error("fast diagonalization: bad 1-D operator");
  gen_eig_dense(L.nrows, dense_of(L), h, V, lam);
}

// Galerkin coarsening of a 1-D factor over aggregates of b consecutive
// points: Lc = P^T L P, hc = P^T h (P the 0/1 aggregation matrix)
void coarsen(const HOp1& L, const std::vector<double>& h, int b, std::vector<double>& Lc, std::vector<double>& hc,
             int& nc) {
  const int n = L.nrows;
  nc = (n + b - 1) / b;
  Lc.assign((size_t)nc * nc, 0.0);
  hc.assign(nc, 0.0);
  for (int i = 0; i < n; ++i) {
    hc[i / b] += h[i];
    for (int k = 0; k < L.K; ++k) {
      const int j = L.col[(size_t)i * L.K + k];
      Lc[(size_t)(i / b) + (size_t)nc * (j / b)] += L.val[(size_t)i * L.K + k];
    }
  }
}

}  // namespace

void build_coarse_host(const Plan& P, int agg, FDMHost& C) {
  std::vector<double> Lcx, Lcy, hcx, hcy;
  coarsen(P.L1.X, P.hx, agg, Lcx, hcx, C.nx);
  coarsen(P.L2.Y, P.hy, agg, Lcy, hcy, C.ny);
  gen_eig_dense(C.nx, Lcx, hcx, C.Vx, C.lx);
  gen_eig_dense(C.ny, Lcy, hcy, C.Vy, C.ly);
  C.nx0 = C.ny0 = -1;
  if (!P.pres_bc) {
    C.nx0 = (int)(std::min_element(C.lx.begin(), C.lx.end()) - C.lx.begin());
    C.ny0 = (int)(std::min_element(C.ly.begin(), C.ly.end()) - C.ly.begin());
  }
}

void build_fdm_host(const Plan& P, FDMHost& F) {
  F.nx = P.Npx;
  F.ny = P.Npy;
  // Lx is the x factor of kron(diag(-hy), Lx); Ly the y factor of kron(Ly, diag(-hx))
  gen_eig(P.L1.X, P.hx, F.Vx, F.lx);
  gen_eig(P.L2.Y, P.hy, F.Vy, F.ly);
  F.nx0 = F.ny0 = -1;
  if (!P.pres_bc) {
    // singular A: the constant mode has the smallest eigenvalue in each direction
    F.nx0 = (int)(std::min_element(F.lx.begin(), F.lx.end()) - F.lx.begin());
    F.ny0 = (int)(std::min_element(F.ly.begin(), F.ly.end()) - F.ly.begin());
  }
  auto transpose = [](const std::vector<double>& V, int n) {
    std::vector<double> T(V.size());
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) T[(size_t)j + (size_t)n * i] = V[(size_t)i + (size_t)n * j];
    return T;
  };
  F.VxT = transpose(F.Vx, F.nx);
  F.VyT = transpose(F.Vy, F.ny);
}

void build_ic0_host(const Plan& P, IC0Host& H) {
  const int nx = P.Npx, ny = P.Npy;
  const std::vector<double> Lx = dense_of(P.L1.X), Ly = dense_of(P.L2.Y);
  auto lx = [&](int a, int b) { return Lx[(size_t)a + (size_t)nx * b]; };
  auto ly = [&](int a, int b) { return Ly[(size_t)a + (size_t)ny * b]; };
  if (nx < 3 || ny < 3) throw std::runtime_error("IC(0) needs at least 3 cells per direction");
  for (int a = 0; a < nx; ++a)
    for (int b = 0; b < nx; ++b) {
      const int d = std::abs(a - b);
      if (d > 1 && d != nx - 1 && lx(a, b) != 0.0) throw std::runtime_error("IC(0): unexpected pressure stencil");
    }
  for (int a = 0; a < ny; ++a)
    for (int b = 0; b < ny; ++b) {
      const int d = std::abs(a - b);
      if (d > 1 && d != ny - 1 && ly(a, b) != 0.0) throw std::runtime_error("IC(0): unexpected pressure stencil");
    }
  const size_t n = (size_t)nx * ny;
  H.D.assign(n, 0.0);
  H.cw.assign(n, 0.0);
  H.cs.assign(n, 0.0);
  const bool wx = lx(nx - 1, 0) != 0.0, wy = ly(ny - 1, 0) != 0.0;
  H.cwx.assign(wx ? ny : 0, 0.0);
  H.cwy.assign(wy ? nx : 0, 0.0);
  // A((i,j),(i',j')) = -hy_j Lx(i,i') delta_jj' - hx_i Ly(j,j') delta_ii'
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const size_t r = (size_t)i + (size_t)nx * j;
      double d = -(P.hy[j] * lx(i, i) + P.hx[i] * ly(j, j));
      if (i > 0) {
        const double a = -P.hy[j] * lx(i, i - 1);
        H.cw[r] = a;
        d -= a * a / H.D[r - 1];
      }
      if (j > 0) {
        const double a = -P.hx[i] * ly(j, j - 1);
        H.cs[r] = a;
        d -= a * a / H.D[r - nx];
      }
      if (wx && i == nx - 1) {
        const double a = -P.hy[j] * lx(nx - 1, 0);
        H.cwx[j] = a;
        d -= a * a / H.D[(size_t)nx * j];
      }
      if (wy && j == ny - 1) {
        const double a = -P.hx[i] * ly(ny - 1, 0);
        H.cwy[i] = a;
        d -= a * a / H.D[i];
      }
      H.D[r] = d;
    }
}

}  // namespace ins2d
