// Host assembly of the fourth-order operators on grids with walls; see
// ops4.hpp.  Each block cites the reference lines whose assembly it
// restates.
#include "ops4.hpp"

#include <algorithm>
#include <cmath>
#include <initializer_list>
#include <string>

namespace ins2d {

namespace {

// kron(A, B): (A.m B.m) x (A.n B.n), entry (i B.m + k, j B.n + l) = A(i,j) B(k,l)
Sp kron(const Sp& A, const Sp& B) {
  Sp C(A.m * B.m, A.n * B.n);
  for (int i = 0; i < A.m; ++i)
    for (auto& a : A.r[i])
      for (int k = 0; k < B.m; ++k) {
        auto& row = C.r[(size_t)i * B.m + k];
        for (auto& b : B.r[k]) row[a.first * B.n + b.first] += a.second * b.second;
      }
  C.prune();
  return C;
}

// dst(roff + i, coff + j) += s * src(i, j)
void place(Sp& dst, const Sp& src, int roff, int coff, double s = 1.0) {
  if (roff + src.m > dst.m || coff + src.n > dst.n) throw std::runtime_error("place: block out of range");
  for (int i = 0; i < src.m; ++i)
    for (auto& kv : src.r[i]) dst.r[roff + i][coff + kv.first] += s * kv.second;
}

std::vector<double> recip(const std::vector<double>& v) {
  std::vector<double> r(v.size());
  for (size_t i = 0; i < v.size(); ++i) r[i] = 1.0 / v[i];
  return r;
}

// kron(a, b) of two vectors (a outer, b inner)
std::vector<double> vkron(const std::vector<double>& a, const std::vector<double>& b) {
  std::vector<double> c(a.size() * b.size());
  for (size_t i = 0; i < a.size(); ++i)
    for (size_t k = 0; k < b.size(); ++k) c[i * b.size() + k] = a[i] * b[k];
  return c;
}

std::vector<double> cat(std::initializer_list<std::vector<double>> parts) {
  std::vector<double> v;
  for (auto& p : parts) v.insert(v.end(), p.begin(), p.end());
  return v;
}

std::vector<double> slice(const std::vector<double>& v, int a, int b) {   // v[a..b)
  return std::vector<double>(v.begin() + a, v.begin() + b);
}

// closures with k boundary values at each end (Nb = 2k): periodic ends copy
// the k values from the other side, Dirichlet ends take the rows that lo / hi
// write (BC_int2.m, BC_div2.m, BC_av3.m, BC_int3.m, BC_diff3.m,
// BC_av_stag3.m, BC_diff_stag3.m, BC_int_mixed_stag2.m, BC_int_mixed_stag3.m)
template <class Lo, class Hi>
Closure wide(int Nt, int Nin, int Nb, int bc1, int bc2, int k, Lo lo, Hi hi, const char* name) {
  if (Nt != Nin + Nb) throw std::runtime_error(std::string(name) + ": Nt != Nin + Nb");
  if (Nb != 2 * k) throw std::runtime_error(std::string(name) + ": unsupported number of boundary points");
  Sp Bb(Nt, Nb), Bbc(Nb, Nt);
  std::vector<double> y1(Nb, 0.0), y2(Nb, 0.0);
  for (int r = 0; r < k; ++r) {
    Bb.set(r, r, 1.0);
    Bb.set(Nt - k + r, Nb - k + r, 1.0);
  }
  if (bc1 == BC_PER || bc2 == BC_PER) {
    if (bc1 != BC_PER || bc2 != BC_PER) throw std::runtime_error(std::string(name) + ": periodic at one end only");
    for (int r = 0; r < k; ++r) {
      Bbc.set(r, r, -1.0);
      Bbc.set(r, Nt - 2 * k + r, 1.0);
      Bbc.set(k + r, k + r, -1.0);
      Bbc.set(k + r, Nt - k + r, 1.0);
    }
  } else {
    if (bc1 != BC_DIR || bc2 != BC_DIR)
      throw std::runtime_error(std::string(name) + ": fourth order is defined for Dirichlet or periodic ends");
    lo(Bbc, y1);
    hi(Bbc, y2);
  }
  return make_closure(Nt, Nin, Nb, -k, Bb, Bbc, y1, y2);
}

// BC_int2.m / BC_div2.m: (u1 + u3)/2 = ub, u2 = ub
Closure bc_int2(int Nt, int Nin, int Nb, int bc1, int bc2) {
  const int n = Nt, b = Nb;
  return wide(
      Nt, Nin, Nb, bc1, bc2, 2,
      [&](Sp& E, std::vector<double>& y) { E.set(0, 0, .5); E.set(0, 2, .5); E.set(1, 1, 1.); y[0] = y[1] = 1.; },
      [&](Sp& E, std::vector<double>& y) {
        E.set(b - 1, n - 1, .5); E.set(b - 1, n - 3, .5); E.set(b - 2, n - 2, 1.); y[b - 2] = y[b - 1] = 1.;
      },
      "BC_int2");
}

// BC_av3.m / BC_int3.m / BC_diff3.m: (u1 + u5)/2 = (u2 + u4)/2 = u3 = ub
Closure bc_sym3(int Nt, int Nin, int Nb, int bc1, int bc2) {
  const int n = Nt, b = Nb;
  return wide(
      Nt, Nin, Nb, bc1, bc2, 3,
      [&](Sp& E, std::vector<double>& y) {
        E.set(0, 0, .5); E.set(0, 4, .5); E.set(1, 1, .5); E.set(1, 3, .5); E.set(2, 2, 1.);
        y[0] = y[1] = y[2] = 1.;
      },
      [&](Sp& E, std::vector<double>& y) {
        E.set(b - 1, n - 1, .5); E.set(b - 1, n - 5, .5); E.set(b - 2, n - 2, .5); E.set(b - 2, n - 4, .5);
        E.set(b - 3, n - 3, 1.);
        y[b - 3] = y[b - 2] = y[b - 1] = 1.;
      },
      "BC_av3");
}

// BC_av_stag3.m: u6 - u1 = 0, (u2 + u5)/2 = ub, u4 - u3 = 0
Closure bc_av_stag3(int Nt, int Nin, int Nb, int bc1, int bc2) {
  const int n = Nt, b = Nb;
  return wide(
      Nt, Nin, Nb, bc1, bc2, 3,
      [&](Sp& E, std::vector<double>& y) {
        E.set(0, 0, -1.); E.set(0, 5, 1.); E.set(1, 1, .5); E.set(1, 4, .5); E.set(2, 2, -1.); E.set(2, 3, 1.);
        y[1] = 1.;
      },
      [&](Sp& E, std::vector<double>& y) {
        E.set(b - 1, n - 1, -1.); E.set(b - 1, n - 6, 1.); E.set(b - 2, n - 2, .5); E.set(b - 2, n - 5, .5);
        E.set(b - 3, n - 3, -1.); E.set(b - 3, n - 4, 1.);
        y[b - 2] = 1.;
      },
      "BC_av_stag3");
}

// BC_diff_stag3.m: (u1 + u6)/2 = (u2 + u5)/2 = (u3 + u4)/2 = ub
Closure bc_diff_stag3(int Nt, int Nin, int Nb, int bc1, int bc2) {
  const int n = Nt, b = Nb;
  return wide(
      Nt, Nin, Nb, bc1, bc2, 3,
      [&](Sp& E, std::vector<double>& y) {
        E.set(0, 0, .5); E.set(0, 5, .5); E.set(1, 1, .5); E.set(1, 4, .5); E.set(2, 2, .5); E.set(2, 3, .5);
        y[0] = y[1] = y[2] = 1.;
      },
      [&](Sp& E, std::vector<double>& y) {
        E.set(b - 1, n - 1, .5); E.set(b - 1, n - 6, .5); E.set(b - 2, n - 2, .5); E.set(b - 2, n - 5, .5);
        E.set(b - 3, n - 3, .5); E.set(b - 3, n - 4, .5);
        y[b - 3] = y[b - 2] = y[b - 1] = 1.;
      },
      "BC_diff_stag3");
}

// BC_int_mixed_stag2.m / BC_int_mixed_stag3.m: u4 - u1 = 0, u3 - u2 = 0
// (no boundary values enter)
Closure bc_mixed_stag(int Nt, int Nin, int Nb, int bc1, int bc2) {
  const int n = Nt, b = Nb;
  return wide(
      Nt, Nin, Nb, bc1, bc2, 2,
      [&](Sp& E, std::vector<double>&) { E.set(0, 0, -1.); E.set(0, 3, 1.); E.set(1, 1, -1.); E.set(1, 2, 1.); },
      [&](Sp& E, std::vector<double>&) {
        E.set(b - 1, n - 1, -1.); E.set(b - 1, n - 4, 1.); E.set(b - 2, n - 2, -1.); E.set(b - 2, n - 3, 1.);
      },
      "BC_int_mixed_stag");
}

// BC_int_mixed2.m: Dirichlet u3 - u1 = 0, u2 = ub at each end (Nb = 4), or
// periodic with three wrapped values (Nb = 3)
Closure bc_int_mixed2(int Nt, int Nin, int Nb, int bc1, int bc2) {
  if (Nt != Nin + Nb) throw std::runtime_error("BC_int_mixed2: Nt != Nin + Nb");
  std::vector<double> y1(Nb, 0.0), y2(Nb, 0.0);
  Sp Bb(Nt, Nb), Bbc(Nb, Nt);
  if (bc1 == BC_DIR && bc2 == BC_DIR) {
    if (Nb != 4) throw std::runtime_error("BC_int_mixed2: unsupported number of boundary points");
    Bb.set(0, 0, 1.); Bb.set(1, 1, 1.); Bb.set(Nt - 2, Nb - 2, 1.); Bb.set(Nt - 1, Nb - 1, 1.);
    Bbc.set(0, 0, -1.); Bbc.set(0, 2, 1.); Bbc.set(1, 1, 1.);
    Bbc.set(Nb - 1, Nt - 1, -1.); Bbc.set(Nb - 1, Nt - 3, 1.); Bbc.set(Nb - 2, Nt - 2, 1.);
    y1[1] = 1.;
    y2[Nb - 2] = 1.;
    return make_closure(Nt, Nin, Nb, -2, Bb, Bbc, y1, y2);
  }
  if (bc1 == BC_PER && bc2 == BC_PER) {
    if (Nb != 3) throw std::runtime_error("BC_int_mixed2: unsupported number of boundary points");
    Bb.set(0, 0, 1.); Bb.set(Nt - 2, Nb - 2, 1.); Bb.set(Nt - 1, Nb - 1, 1.);
    Bbc.set(0, 0, -1.); Bbc.set(0, Nt - 3, 1.);
    Bbc.set(Nb - 2, 1, -1.); Bbc.set(Nb - 2, Nt - 2, 1.);
    Bbc.set(Nb - 1, 2, -1.); Bbc.set(Nb - 1, Nt - 1, 1.);
    return make_closure(Nt, Nin, Nb, -1, Bb, Bbc, y1, y2);
  }
  throw std::runtime_error("BC_int_mixed2: fourth order is defined for Dirichlet or periodic ends");
}

// the matrix K with K s = kron(s, ybc) (first) or kron(ybc, s) (second)
// for a sample s of length len
Sp kmat(const std::vector<double>& ybc, int len, bool first) {
  const int nb = (int)ybc.size();
  Sp K(len * nb, len);
  for (int i = 0; i < len; ++i)
    for (int k = 0; k < nb; ++k)
      if (ybc[k] != 0.0) K.set(first ? i * nb + k : k * len + i, i, ybc[k]);
  return K;
}

// ghost extension of a boundary sample by one value at each end
// (set_bc_vectors.m:455-474 for vLe/vRi, :497-516 for uLo/uUp): Dirichlet
// ends extrapolate linearly, periodic ends append 0.  The reference's
// upper end of the second sample (vRi_ext, uUp_ext) extrapolates from its
// first two values, not its last two; restated as written.
Sp ext_map(int lo_bc, int hi_bc, int n, bool second) {
  Sp E(n + 2, n);
  if (lo_bc == BC_DIR) { E.set(0, 0, 2.0); E.set(0, 1, -1.0); }
  else if (lo_bc != BC_PER) E.set(0, 1, 1.0);
  for (int i = 0; i < n; ++i) E.set(i + 1, i, 1.0);
  if (hi_bc == BC_DIR) {
    if (second) { E.add(n + 1, 0, 2.0); E.add(n + 1, 1, -1.0); }
    else { E.add(n + 1, n - 1, 2.0); E.add(n + 1, n - 2, -1.0); }
  } else if (hi_bc != BC_PER) {
    E.set(n + 1, n - 2, 1.0);
  }
  E.prune();
  return E;
}

struct YBuild {
  std::vector<Sp>* T;
  int tlen;
  const W4Ops* W;
  // T[lo] += s Beff K(ybc1) E, T[hi] += s Beff K(ybc2) E at row offset roff
  void pair(int roff, const Sp& Beff, const Closure& c, int lo, int hi, bool first, int len, const Sp* Elo = nullptr,
            const Sp* Ehi = nullptr, double s = 1.0) {
    add(roff, Beff, c.ybc1, lo, first, len, Elo, s);
    add(roff, Beff, c.ybc2, hi, first, len, Ehi, s);
  }
  void add(int roff, const Sp& Beff, const std::vector<double>& ybc, int src, bool first, int len, const Sp* E,
           double s) {
    Sp t = mul(Beff, kmat(ybc, len, first));
    if (E) t = mul(t, *E);
    Sp& dst = (*T)[src];
    if (dst.m == 0) dst = Sp(tlen, w4_sample_len(*W, src));
    if (t.n != dst.n) throw std::runtime_error("boundary term: sample length mismatch");
    place(dst, t, roff, 0, s);
  }
};

}  // namespace

int w4_sample_len(const W4Ops& W, int src) {
  switch (src) {
    case S_uLo: case S_uUp: return W.Nx + 1;
    case S_uLo_i: case S_uUp_i: return W.nux;
    case S_uLe_i: case S_uRi_i: return W.Ny;
    case S_vLe: case S_vRi: return W.Ny + 1;
    case S_vLo_i: case S_vUp_i: return W.Nx;
    default: return W.nvy;
  }
}

void build_w4(const CaseDesc& cd, W4Ops& W) {
  const int Nx = cd.Nx, Ny = cd.Ny;
  const double alfa = cd.alfa, beta = cd.beta;
  if ((int)cd.x.size() != Nx + 1 || (int)cd.y.size() != Ny + 1) throw std::runtime_error("grid size mismatch");
  if (Nx < 4 || Ny < 4) throw std::runtime_error("fourth order needs at least 4 cells per direction");
  auto pair_ok = [](int a, int b) { return (a == BC_DIR && b == BC_DIR) || (a == BC_PER && b == BC_PER); };
  if (!pair_ok(cd.u_left, cd.u_right) || !pair_ok(cd.v_left, cd.v_right) || !pair_ok(cd.u_low, cd.u_up) ||
      !pair_ok(cd.v_low, cd.v_up))
    throw std::runtime_error("fourth order is defined for Dirichlet or periodic boundaries (both ends alike)");
  const bool perx = cd.u_left == BC_PER, pery = cd.v_low == BC_PER;
  if ((cd.v_left == BC_PER) != perx || (cd.u_low == BC_PER) != pery)
    throw std::runtime_error("u and v must share the periodic directions");
  W.Nx = Nx;
  W.Ny = Ny;

  // ---- mesh_generation.m:17-39
  std::vector<double> hx(Nx), hy(Ny), gx(Nx + 1), gy(Ny + 1);
  for (int i = 0; i < Nx; ++i) hx[i] = cd.x[i + 1] - cd.x[i];
  for (int j = 0; j < Ny; ++j) hy[j] = cd.y[j + 1] - cd.y[j];
  gx[0] = hx[0] / 2; gx[Nx] = hx[Nx - 1] / 2;
  for (int i = 1; i < Nx; ++i) gx[i] = (hx[i - 1] + hx[i]) / 2;
  gy[0] = hy[0] / 2; gy[Ny] = hy[Ny - 1] / 2;
  for (int j = 1; j < Ny; ++j) gy[j] = (hy[j - 1] + hy[j]) / 2;
  W.hx = hx; W.hy = hy; W.gx = gx; W.gy = gy;

  // ---- operator_mesh.m:54-110 (sizes)
  const int nux = perx ? Nx : Nx - 1, nuy = Ny, nvx = Nx, nvy = pery ? Ny : Ny - 1;
  const int Nux_t = nux + 2, Nuy_t = nuy + 2, Nvx_t = nvx + 2, Nvy_t = nvy + 2;
  W.nux = nux; W.nuy = nuy; W.nvx = nvx; W.nvy = nvy;
  W.Nu = nux * nuy; W.Nv = nvx * nvy; W.NV = W.Nu + W.Nv; W.Np = Nx * Ny;
  if (perx) W.xin.assign(cd.x.begin(), cd.x.end() - 1); else W.xin.assign(cd.x.begin() + 1, cd.x.end() - 1);
  if (pery) W.yin.assign(cd.y.begin(), cd.y.end() - 1); else W.yin.assign(cd.y.begin() + 1, cd.y.end() - 1);

  // ---- operator_mesh.m:112-171: widths over three cells (h3) and line
  // spacings over three (g3)
  auto three = [](const std::vector<double>& h, bool per) {
    const int N = (int)h.size();
    std::vector<double> h3(N);
    for (int i = 1; i < N - 1; ++i) h3[i] = h[i - 1] + h[i] + h[i + 1];
    if (per) { h3[0] = h[N - 1] + h[0] + h[1]; h3[N - 1] = h[N - 2] + h[N - 1] + h[0]; }
    else { h3[0] = 2 * h[0] + h[1]; h3[N - 1] = h[N - 2] + 2 * h[N - 1]; }
    return h3;
  };
  auto gthree = [](const std::vector<double>& g, bool per) {
    const int N = (int)g.size() - 1;
    std::vector<double> g3(N + 1);
    for (int k = 2; k <= N - 2; ++k) g3[k] = g[k - 1] + g[k] + g[k + 1];
    if (per) {
      g3[0] = g[N - 1] + g[N] + g[0] + g[1];
      g3[1] = g[N] + g[0] + g[1] + g[2];
      g3[N - 1] = g[N - 2] + g[N - 1] + g[N] + g[0];
      g3[N] = g[N - 1] + g[N] + g[0] + g[1];
    } else {
      g3[0] = 2 * g[0] + 2 * g[1];
      g3[1] = 2 * g[0] + g[1] + g[2];
      g3[N - 1] = 2 * g[N] + g[N - 1] + g[N - 2];
      g3[N] = 2 * g[N] + 2 * g[N - 1];
    }
    return g3;
  };
  const std::vector<double> hx3 = three(hx, perx), hy3 = three(hy, pery);
  const std::vector<double> gx3 = gthree(gx, perx), gy3 = gthree(gy, pery);

  // ---- operator_mesh.m:176-330: metrics of the staggered volumes
  struct Dir {
    std::vector<double> hd13, hd3, gd13, gd3, gi, gi3;
    int dpos;
  };
  auto metrics = [](const std::vector<double>& h, const std::vector<double>& g, const std::vector<double>& h3,
                    const std::vector<double>& g3, bool per) {
    const int N = (int)h.size();
    Dir d;
    if (per) {
      d.gi = cat({{g[0] + g[N]}, slice(g, 1, N)});
      d.hd3 = cat({{h3[N - 2], h3[N - 1]}, h3, {h3[0]}});
      d.hd13 = cat({{h[N - 2], h[N - 1]}, h, {h[0]}});
      d.gd3 = cat({{g3[N - 1]}, g3, {g3[1]}});
      d.gd13 = cat({{g[N - 1], g[0] + g[N]}, slice(g, 1, N), {g[N] + g[0], g[1]}});
      d.gi3 = slice(g3, 0, N);
      d.dpos = 0;
    } else {
      d.gi = slice(g, 1, N);
      d.hd3 = cat({{h3[0]}, h3, {h3[N - 1]}});
      d.hd13 = cat({{h[0]}, h, {h[N - 1]}});
      d.gd3 = cat({{2 * g[0] + g[1] + g[2]}, g3, {2 * g[N] + g[N - 1] + g[N - 2]}});
      d.gd13 = cat({{g[1], 2 * g[0]}, slice(g, 1, N), {2 * g[N], g[N - 1]}});
      d.gi3 = slice(g3, 1, N);
      d.dpos = 1;
    }
    return d;
  };
  const Dir X = metrics(hx, gx, hx3, gx3, perx), Y = metrics(hy, gy, hy3, gy3, pery);
  const std::vector<double>&hxi = hx, &hyi = hy;
  const Sp Bvux = spdiags_const({1.0}, {X.dpos}, nux, Nvx_t - 1);
  const Sp Buvy = spdiags_const({1.0}, {Y.dpos}, nvy, Nuy_t - 1);

  // volumes (operator_mesh.m:332-400): Om = alfa Om1 - Om3
  auto comb = [&](const std::vector<double>& a1, const std::vector<double>& a3) {
    std::vector<double> o(a1.size());
    for (size_t k = 0; k < o.size(); ++k) o[k] = alfa * a1[k] - a3[k];
    return o;
  };
  const std::vector<double> Omu = comb(vkron(hyi, X.gi), vkron(hy3, X.gi3));
  const std::vector<double> Omv = comb(vkron(Y.gi, hxi), vkron(Y.gi3, hx3));
  const std::vector<double> Omux1 = vkron(hyi, X.hd13), Omux3 = vkron(hy3, X.hd3);
  const std::vector<double> Omuy1 = vkron(Y.gd13, X.gi), Omuy3 = vkron(Y.gd3, X.gi3);
  const std::vector<double> Omvx1 = vkron(Y.gi, X.gd13), Omvx3 = vkron(Y.gi3, X.gd3);
  const std::vector<double> Omvy1 = vkron(Y.hd13, hxi), Omvy3 = vkron(Y.hd3, hx3);
  const std::vector<double> Omux = comb(Omux1, Omux3), Omuy = comb(Omuy1, Omuy3);
  const std::vector<double> Omvx = comb(Omvx1, Omvx3), Omvy = comb(Omvy1, Omvy3);
  W.Om = cat({Omu, Omv});

  const Sp Iux = eye(nux), Iuy = eye(nuy), Ivx = eye(nvx), Ivy = eye(nvy);
  const int NV = W.NV, Nu = W.Nu, Np = W.Np;

  // face blocks: 1h products ux, uy, vx, vy, then the 3h ones
  const int nf[8] = {(Nux_t - 1) * nuy, (Nuy_t - 1) * nux, (Nvx_t - 1) * nvy, (Nvy_t - 1) * nvx,
                     (nux + 3) * nuy,   (nuy + 3) * nux,   (nvx + 3) * nvy,   (nvy + 3) * nvx};
  int fo[9] = {0};
  for (int k = 0; k < 8; ++k) fo[k + 1] = fo[k] + nf[k];
  W.Nf = fo[8];
  const int Nf = W.Nf;
  W.Af = Sp(Nf, NV);
  W.If = Sp(Nf, NV);
  W.Cw = Sp(NV, Nf);
  W.TyA.assign(S_NSRC, Sp());
  W.TyI.assign(S_NSRC, Sp());
  W.TyD.assign(S_NSRC, Sp());
  W.TyM.assign(S_NSRC, Sp());
  YBuild yA{&W.TyA, Nf, &W}, yI{&W.TyI, Nf, &W}, yD{&W.TyD, NV, &W}, yM{&W.TyM, Np, &W};
  const int Lx = Nx + 1, Ly = Ny + 1;   // lengths of uLo / vLe

  // ---- averaging (operator_averaging.m:20-158)
  {
    const double w = 0.5;
    auto A1 = [&](int n) { return spdiags_const({w, w}, {0, 1}, n - 1, n); };
    auto A3 = [&](int nin, int nt) { return spdiags_const({w, w}, {0, 3}, nin + 3, nt + 4); };
    Closure b = bc_general(Nux_t, nux, 2, cd.u_left, cd.u_right, hx[0], hx[Nx - 1]);
    place(W.Af, kron(Iuy, mul(A1(Nux_t), b.B1D)), fo[0], 0);
    yA.pair(fo[0], kron(Iuy, mul(A1(Nux_t), b.Btemp)), b, S_uLe_i, S_uRi_i, true, Ny);
    b = bc_general_stag(Nuy_t, nuy, 2, cd.u_low, cd.u_up, hy[0], hy[Ny - 1], false);
    place(W.Af, kron(mul(A1(Nuy_t), b.B1D), Iux), fo[1], 0);
    yA.pair(fo[1], kron(mul(A1(Nuy_t), b.Btemp), Iux), b, S_uLo_i, S_uUp_i, false, nux);
    b = bc_general_stag(Nvx_t, nvx, 2, cd.v_left, cd.v_right, hx[0], hx[Nx - 1], false);
    place(W.Af, kron(Ivy, mul(A1(Nvx_t), b.B1D)), fo[2], Nu);
    yA.pair(fo[2], kron(Ivy, mul(A1(Nvx_t), b.Btemp)), b, S_vLe_i, S_vRi_i, true, nvy);
    b = bc_general(Nvy_t, nvy, 2, cd.v_low, cd.v_up, hy[0], hy[Ny - 1]);
    place(W.Af, kron(mul(A1(Nvy_t), b.B1D), Ivx), fo[3], Nu);
    yA.pair(fo[3], kron(mul(A1(Nvy_t), b.Btemp), Ivx), b, S_vLo_i, S_vUp_i, false, Nx);

    b = bc_sym3(Nux_t + 4, nux, Nux_t + 4 - nux, cd.u_left, cd.u_right);
    place(W.Af, kron(Iuy, mul(A3(nux, Nux_t), b.B1D)), fo[4], 0);
    yA.pair(fo[4], kron(Iuy, mul(A3(nux, Nux_t), b.Btemp)), b, S_uLe_i, S_uRi_i, true, Ny);
    b = bc_av_stag3(Nuy_t + 4, nuy, Nuy_t + 4 - nuy, cd.u_low, cd.u_up);
    place(W.Af, kron(mul(A3(nuy, Nuy_t), b.B1D), Iux), fo[5], 0);
    yA.pair(fo[5], kron(mul(A3(nuy, Nuy_t), b.Btemp), Iux), b, S_uLo_i, S_uUp_i, false, nux);
    b = bc_av_stag3(Nvx_t + 4, nvx, Nvx_t + 4 - nvx, cd.v_left, cd.v_right);
    place(W.Af, kron(Ivy, mul(A3(nvx, Nvx_t), b.B1D)), fo[6], Nu);
    yA.pair(fo[6], kron(Ivy, mul(A3(nvx, Nvx_t), b.Btemp)), b, S_vLe_i, S_vRi_i, true, nvy);
    b = bc_sym3(Nvy_t + 4, nvy, Nvy_t + 4 - nvy, cd.v_low, cd.v_up);
    place(W.Af, kron(mul(A3(nvy, Nvy_t), b.B1D), Ivx), fo[7], Nu);
    yA.pair(fo[7], kron(mul(A3(nvy, Nvy_t), b.Btemp), Ivx), b, S_vLo_i, S_vUp_i, false, Nx);
  }

  // ---- interpolation (operator_interpolation.m:166-337)
  {
    const double w1 = 0.5 * beta, w2 = 0.5 * (1.0 - beta);
    auto I4 = [&](int m, int n) { return spdiags_const({w2, w1, w1, w2}, {0, 1, 2, 3}, m, n); };
    const Sp mat_hx = diag(hxi), mat_hy = diag(hyi), mat_hx3 = diag(hx3), mat_hy3 = diag(hy3);
    auto wide2 = [](const std::vector<double>& h, bool per) {
      const int N = (int)h.size();
      return per ? cat({{h[N - 2], h[N - 1]}, h, {h[0], h[1]}}) : cat({{h[1], h[0]}, h, {h[N - 1], h[N - 2]}});
    };
    auto wide4 = [](const std::vector<double>& h, const std::vector<double>& h3, bool per) {
      const int N = (int)h.size();
      return per ? cat({{h3[N - 2], h3[N - 1]}, h3, {h3[0], h3[1]}})
                 : cat({{h[0] + h[1] + h[2], 2 * h[0] + h[1]}, h3, {2 * h[N - 1] + h[N - 2], h[N - 1] + h[N - 2] + h[N - 3]}});
    };
    const Sp mat_hx2 = diag(wide2(hx, perx)), mat_hy2 = diag(wide2(hy, pery));
    const Sp mat_hx4 = diag(wide4(hx, hx3, perx)), mat_hy4 = diag(wide4(hy, hy3, pery));

    // Iu_ux, Iu_ux3
    Closure b = bc_int2(Nux_t + 2, nux, Nux_t + 2 - nux, cd.u_left, cd.u_right);
    Sp I = I4(Nux_t - 1, Nux_t + 2);
    place(W.If, kron(mat_hy, mul(I, b.B1D)), fo[0], 0);
    yI.pair(fo[0], kron(mat_hy, mul(I, b.Btemp)), b, S_uLe_i, S_uRi_i, true, Ny);
    b = bc_sym3(Nux_t + 4, nux, Nux_t + 4 - nux, cd.u_left, cd.u_right);
    I = I4(nux + 3, Nux_t + 4);
    place(W.If, kron(mat_hy3, mul(I, b.B1D)), fo[4], 0);
    yI.pair(fo[4], kron(mat_hy3, mul(I, b.Btemp)), b, S_uLe_i, S_uRi_i, true, Ny);

    // Iv_uy
    {
      I = mul(mul(Bvux, I4(Nvx_t - 1, Nvx_t + 2)), mat_hx2);
      const Sp I2D = kron(eye(Nuy_t - 1), I);
      Closure lu = bc_general(nuy + 1, nvy, nuy + 1 - nvy, cd.v_low, cd.v_up, hy[0], hy[Ny - 1]);
      const Sp luB2D = kron(lu.B1D, Ivx), luBbc = kron(lu.Btemp, Ivx);
      Closure lr = bc_mixed_stag(Nvx_t + 2, nvx, Nvx_t + 2 - nvx, cd.v_left, cd.v_right);
      const Sp lrB2D = mul(I2D, kron(eye(Nuy_t - 1), lr.B1D)), lrBbc = mul(I2D, kron(eye(Nuy_t - 1), lr.Btemp));
      place(W.If, mul(lrB2D, luB2D), fo[1], Nu);
      yI.pair(fo[1], lrBbc, lr, S_vLe, S_vRi, true, Ly);
      yI.pair(fo[1], mul(lrB2D, luBbc), lu, S_vLo_i, S_vUp_i, false, Nx);
    }
    // Iv_uy3 (vLe / vRi extended by one ghost value at each end)
    {
      I = mul(mul(Bvux, I4(Nvx_t - 1, Nvx_t + 2)), mat_hx4);
      const Sp I2D = kron(eye(Nuy_t + 1), I);
      Closure lu = bc_int_mixed2(nuy + 3, nvy, nuy + 3 - nvy, cd.v_low, cd.v_up);
      const Sp luB2D = kron(lu.B1D, Ivx), luBbc = kron(lu.Btemp, Ivx);
      Closure lr = bc_mixed_stag(Nvx_t + 2, nvx, Nvx_t + 2 - nvx, cd.v_left, cd.v_right);
      const Sp lrB2D = mul(I2D, kron(eye(Nuy_t + 1), lr.B1D)), lrBbc = mul(I2D, kron(eye(Nuy_t + 1), lr.Btemp));
      place(W.If, mul(lrB2D, luB2D), fo[5], Nu);
      const Sp Ele = ext_map(cd.v_low, cd.v_up, Ly, false), Eri = ext_map(cd.v_low, cd.v_up, Ly, true);
      yI.pair(fo[5], lrBbc, lr, S_vLe, S_vRi, true, Ly + 2, &Ele, &Eri);
      yI.pair(fo[5], mul(lrB2D, luBbc), lu, S_vLo_i, S_vUp_i, false, Nx);
    }
    // Iu_vx
    {
      I = mul(mul(Buvy, I4(Nuy_t - 1, Nuy_t + 2)), mat_hy2);
      const Sp I2D = kron(I, eye(Nvx_t - 1));
      Closure lu = bc_mixed_stag(Nuy_t + 2, nuy, Nuy_t + 2 - nuy, cd.u_low, cd.u_up);
      const Sp luB2D = mul(I2D, kron(lu.B1D, eye(Nvx_t - 1))), luBbc = mul(I2D, kron(lu.Btemp, eye(Nvx_t - 1)));
      Closure lr = bc_general(nvx + 1, nux, nvx + 1 - nux, cd.u_left, cd.u_right, hx[0], hx[Nx - 1]);
      const Sp lrB2D = kron(Iuy, lr.B1D), lrBbc = kron(Iuy, lr.Btemp);
      place(W.If, mul(luB2D, lrB2D), fo[2], 0);
      yI.pair(fo[2], luBbc, lu, S_uLo, S_uUp, false, Lx);
      yI.pair(fo[2], mul(luB2D, lrBbc), lr, S_uLe_i, S_uRi_i, true, Ny);
    }
    // Iu_vx3 (uLo / uUp extended)
    {
      I = mul(mul(Buvy, I4(Nuy_t - 1, Nuy_t + 2)), mat_hy4);
      const Sp I2D = kron(I, eye(Nvx_t + 1));
      Closure lu = bc_mixed_stag(Nuy_t + 2, nuy, Nuy_t + 2 - nuy, cd.u_low, cd.u_up);
      const Sp luB2D = mul(I2D, kron(lu.B1D, eye(Nvx_t + 1))), luBbc = mul(I2D, kron(lu.Btemp, eye(Nvx_t + 1)));
      Closure lr = bc_int_mixed2(nvx + 3, nux, nvx + 3 - nux, cd.u_left, cd.u_right);
      const Sp lrB2D = kron(Iuy, lr.B1D), lrBbc = kron(Iuy, lr.Btemp);
      place(W.If, mul(luB2D, lrB2D), fo[6], 0);
      const Sp Elo = ext_map(cd.u_left, cd.u_right, Lx, false), Eup = ext_map(cd.u_left, cd.u_right, Lx, true);
      yI.pair(fo[6], luBbc, lu, S_uLo, S_uUp, false, Lx + 2, &Elo, &Eup);
      yI.pair(fo[6], mul(luB2D, lrBbc), lr, S_uLe_i, S_uRi_i, true, Ny);
    }
    // Iv_vy, Iv_vy3
    b = bc_int2(Nvy_t + 2, nvy, Nvy_t + 2 - nvy, cd.v_low, cd.v_up);
    I = I4(Nvy_t - 1, Nvy_t + 2);
    place(W.If, kron(mul(I, b.B1D), mat_hx), fo[3], Nu);
    yI.pair(fo[3], kron(mul(I, b.Btemp), mat_hx), b, S_vLo_i, S_vUp_i, false, Nx);
    b = bc_sym3(Nvy_t + 4, nvy, Nvy_t + 4 - nvy, cd.v_low, cd.v_up);
    I = I4(nvy + 3, Nvy_t + 4);
    place(W.If, kron(mul(I, b.B1D), mat_hx3), fo[7], Nu);
    yI.pair(fo[7], kron(mul(I, b.Btemp), mat_hx3), b, S_vLo_i, S_vUp_i, false, Nx);
  }

  // ---- divergence and gradient (operator_divergence.m:20-120)
  {
    auto dpos = [](int lo, int hi) { return (lo == BC_PER && hi == BC_PER) ? 1 : 0; };   // no pressure ends
    const Sp BMx = spdiags_const({1.0}, {dpos(cd.u_left, cd.u_right)}, Nx, Nux_t - 1);
    const Sp BMy = spdiags_const({1.0}, {dpos(cd.v_low, cd.v_up)}, Ny, Nvy_t - 1);
    const Sp M1x = mul(BMx, spdiags_const({-1.0, 1.0}, {0, 1}, Nux_t - 1, Nux_t));
    const Sp M3x = mul(BMx, spdiags_const({-1.0, 1.0}, {0, 3}, Nux_t - 1, Nux_t + 2));
    const Sp M1y = mul(BMy, spdiags_const({-1.0, 1.0}, {0, 1}, Nvy_t - 1, Nvy_t));
    const Sp M3y = mul(BMy, spdiags_const({-1.0, 1.0}, {0, 3}, Nvy_t - 1, Nvy_t + 2));
    Closure b = bc_general(Nux_t, nux, 2, cd.u_left, cd.u_right, hx[0], hx[Nx - 1]);
    Closure b3 = bc_int2(Nux_t + 2, nux, Nux_t + 2 - nux, cd.u_left, cd.u_right);
    const Sp Mx = add(kron(diag(hyi), mul(M1x, b.B1D)), alfa, kron(diag(hy3), mul(M3x, b3.B1D)), -1.0);
    yM.pair(0, kron(diag(hyi), mul(M1x, b.Btemp)), b, S_uLe_i, S_uRi_i, true, Ny, nullptr, nullptr, alfa);
    yM.pair(0, kron(diag(hy3), mul(M3x, b3.Btemp)), b3, S_uLe_i, S_uRi_i, true, Ny, nullptr, nullptr, -1.0);
    b = bc_general(Nvy_t, nvy, 2, cd.v_low, cd.v_up, hy[0], hy[Ny - 1]);
    b3 = bc_int2(Nvy_t + 2, nvy, Nvy_t + 2 - nvy, cd.v_low, cd.v_up);
    const Sp My = add(kron(mul(M1y, b.B1D), diag(hxi)), alfa, kron(mul(M3y, b3.B1D), diag(hx3)), -1.0);
    yM.pair(0, kron(mul(M1y, b.Btemp), diag(hxi)), b, S_vLo_i, S_vUp_i, false, Nx, nullptr, nullptr, alfa);
    yM.pair(0, kron(mul(M3y, b3.Btemp), diag(hx3)), b3, S_vLo_i, S_vUp_i, false, Nx, nullptr, nullptr, -1.0);
    W.M = Sp(Np, NV);
    place(W.M, Mx, 0, 0);
    place(W.M, My, 0, Nu);
    W.G = scale(transpose(W.M), -1.0);
    W.A = mul(W.M, mul(diag(recip(W.Om)), W.G));
    double rs = 0.0, dg = 0.0;
    for (int i = 0; i < Np; ++i) {
      double s = 0.0;
      for (auto& kv : W.A.r[i]) s += kv.second;
      rs = std::max(rs, std::fabs(s));
      dg = std::max(dg, std::fabs(W.A.get(i, i)));
    }
    W.singular = rs <= 1e-10 * dg;
  }

  // ---- convection and diffusion (operator_convection_diffusion.m:257-449)
  {
    auto D1 = [](int nr, int nc, int o1, int o2) { return spdiags_const({-1.0, 1.0}, {o1, o2}, nr, nc); };
    // the (0, 3) difference with the one-sided Dirichlet rows of :300-315
    auto D13w = [&](int nr, int nc, bool lo_dir, bool hi_dir) {
      Sp D = D1(nr, nc, 0, 3);
      if (lo_dir) { D.set(0, 0, 1.0); D.set(0, 1, -2.0); }
      if (hi_dir) { D.set(nr - 1, nc - 2, 2.0); D.set(nr - 1, nc - 1, -1.0); }
      D.prune();
      return D;
    };
    // convection: conv = alfa C1 w1 - C3 w3 (convection.m:30-44)
    place(W.Cw, kron(Iuy, D1(Nux_t - 2, Nux_t - 1, 0, 1)), 0, fo[0], alfa);
    place(W.Cw, kron(D1(Nuy_t - 2, Nuy_t - 1, 0, 1), Iux), 0, fo[1], alfa);
    place(W.Cw, kron(Ivy, D1(Nvx_t - 2, Nvx_t - 1, 0, 1)), Nu, fo[2], alfa);
    place(W.Cw, kron(D1(Nvy_t - 2, Nvy_t - 1, 0, 1), Ivx), Nu, fo[3], alfa);
    const Sp D13ux = D1(Nux_t - 2, Nux_t + 1, 0, 3);
    const Sp D13uy = D13w(Nuy_t - 2, Nuy_t + 1, cd.u_low == BC_DIR, cd.u_up == BC_DIR);
    const Sp D13vx = D13w(Nvx_t - 2, Nvx_t + 1, cd.v_left == BC_DIR, cd.v_right == BC_DIR);
    const Sp D13vy = D1(Nvy_t - 2, Nvy_t + 1, 0, 3);
    place(W.Cw, kron(Iuy, D13ux), 0, fo[4], -1.0);
    place(W.Cw, kron(D13uy, Iux), 0, fo[5], -1.0);
    place(W.Cw, kron(Ivy, D13vx), Nu, fo[6], -1.0);
    place(W.Cw, kron(D13vy, Ivx), Nu, fo[7], -1.0);
    W.Af.prune(); W.If.prune(); W.Cw.prune();

    // diffusion: Diffu = 1/Re [Diffux_div (alfa Su_ux - Su_ux3) + Diffuy_div (alfa Su_uy - Su_uy3)]
    const double iRe = 1.0 / cd.Re;
    auto S = [](const std::vector<double>& h, int o1, int o2, int nr, int nc) {
      std::vector<double> inv = recip(h), ninv(inv);
      for (auto& v : ninv) v = -v;
      return spdiags({ninv, inv}, {o1, o2}, nr, nc);
    };
    auto divop = [&](const Sp& D, const Sp& D3, const std::vector<double>& om) {
      return mul(add(D, alfa, D3, -1.0), diag(recip(om)));
    };
    const Sp Dux = kron(diag(hyi), D1(Nux_t - 2, Nux_t + 1, 1, 2)), Dux3 = kron(diag(hy3), D13ux);
    const Sp Duy = kron(D1(Nuy_t - 2, Nuy_t + 1, 1, 2), diag(X.gi)), Duy3 = kron(D13uy, diag(X.gi3));
    const Sp Dvx = kron(diag(Y.gi), D1(Nvx_t - 2, Nvx_t + 1, 1, 2)), Dvx3 = kron(diag(Y.gi3), D13vx);
    const Sp Dvy = kron(D1(Nvy_t - 2, Nvy_t + 1, 1, 2), diag(hxi)), Dvy3 = kron(D13vy, diag(hx3));
    const Sp Dxu = scale(divop(Dux, Dux3, Omux), iRe), Dyu = scale(divop(Duy, Duy3, Omuy), iRe);
    const Sp Dxv = scale(divop(Dvx, Dvx3, Omvx), iRe), Dyv = scale(divop(Dvy, Dvy3, Omvy), iRe);

    // one direction of one component: alfa Su - Su3 and its boundary terms
    // (x-direction closures take samples kron(sample, ybc), y-direction
    // closures kron(ybc, sample))
    auto term = [&](const Sp& Dd, int roff, int coff, const Closure& b, const Sp& S1, const Sp& S3,
                    const std::vector<double>& om1, const std::vector<double>& om3, bool xdir, int nother, int lo,
                    int hi, int len) {
      const Sp O = xdir ? eye(nother) : Sp();
      auto k2 = [&](const Sp& A) { return xdir ? kron(O, A) : kron(A, eye(nother)); };
      const Sp Su = mul(diag(om1), k2(mul(S1, b.B1D)));
      const Sp Su3 = mul(diag(om3), k2(mul(S3, b.B1D)));
      place(W.Diff, mul(Dd, add(Su, alfa, Su3, -1.0)), roff, coff);
      yD.pair(roff, mul(Dd, mul(diag(om1), k2(mul(S1, b.Btemp)))), b, lo, hi, xdir, len, nullptr, nullptr, alfa);
      yD.pair(roff, mul(Dd, mul(diag(om3), k2(mul(S3, b.Btemp)))), b, lo, hi, xdir, len, nullptr, nullptr, -1.0);
    };
    W.Diff = Sp(NV, NV);
    Closure b = bc_sym3(Nux_t + 4, nux, Nux_t + 4 - nux, cd.u_left, cd.u_right);
    term(Dxu, 0, 0, b, S(X.hd13, 1, 2, nux + 3, Nux_t + 4), S(X.hd3, 0, 3, nux + 3, Nux_t + 4), Omux1, Omux3, true,
         nuy, S_uLe_i, S_uRi_i, Ny);
    b = bc_diff_stag3(Nuy_t + 4, nuy, Nuy_t + 4 - nuy, cd.u_low, cd.u_up);
    term(Dyu, 0, 0, b, S(Y.gd13, 1, 2, nuy + 3, Nuy_t + 4), S(Y.gd3, 0, 3, nuy + 3, Nuy_t + 4), Omuy1, Omuy3, false,
         nux, S_uLo_i, S_uUp_i, nux);
    b = bc_diff_stag3(Nvx_t + 4, nvx, Nvx_t + 4 - nvx, cd.v_left, cd.v_right);
    term(Dxv, Nu, Nu, b, S(X.gd13, 1, 2, nvx + 3, Nvx_t + 4), S(X.gd3, 0, 3, nvx + 3, Nvx_t + 4), Omvx1, Omvx3,
         true, nvy, S_vLe_i, S_vRi_i, nvy);
    b = bc_sym3(Nvy_t + 4, nvy, Nvy_t + 4 - nvy, cd.v_low, cd.v_up);
    term(Dyv, Nu, Nu, b, S(Y.hd13, 1, 2, nvy + 3, Nvy_t + 4), S(Y.hd3, 0, 3, nvy + 3, Nvy_t + 4), Omvy1, Omvy3,
         false, nvx, S_vLo_i, S_vUp_i, Nx);
    W.Diff.prune();
  }
}

void w4_bc_vectors(const W4Ops& W, const BCSamples& S, std::vector<double>& yA, std::vector<double>& yI,
                   std::vector<double>& yD, std::vector<double>& yM) {
  auto eval = [&](const std::vector<Sp>& T, int n, std::vector<double>& y) {
    y.assign(n, 0.0);
    for (int s = 0; s < S_NSRC; ++s) {
      if (T[s].m == 0) continue;
      if ((int)S.s[s].size() != T[s].n) throw std::runtime_error("boundary sample length mismatch");
      const std::vector<double> t = matvec(T[s], S.s[s]);
      for (int i = 0; i < n; ++i) y[i] += t[i];
    }
  };
  eval(W.TyA, W.Nf, yA);
  eval(W.TyI, W.Nf, yI);
  eval(W.TyD, W.NV, yD);
  eval(W.TyM, W.Np, yM);
}

}  // namespace ins2d
