// Test helper: builds the host operator plan of libins2d_hip (ops1d.cpp) for
// a few grids and boundary sets and prints, per case, how many points the
// closed-form classification (uniform.hpp) marks for the u and v momentum,
// with the grid-line offsets it found, and the boundary masks' count of
// points that read a boundary term (bcmask.hpp).  Host code only (built with
// hipcc for the HIP headers; no GPU call).
#include <cstdio>
#include <vector>

#include "../../ins2d_amd/csrc/bcmask.hpp"
#include "../../ins2d_amd/csrc/uniform.hpp"

using namespace ins2d;

// stretch: geometric growth factor of the cells in y; stretch_x the same in x
static void run(const char* name, int Nx, int Ny, double Lx, const int bu[4], const int bv[4], double stretch,
                double stretch_x = 1.0) {
  CaseDesc cd;
  cd.Nx = Nx;
  cd.Ny = Ny;
  cd.x.resize(Nx + 1);
  cd.y.resize(Ny + 1);
  {
    std::vector<double> dx(Nx);
    double h = 1.0, tot = 0.0;
    for (int i = 0; i < Nx; ++i) { dx[i] = h; tot += h; h *= stretch_x; }
    cd.x[0] = 0.0;
    for (int i = 0; i < Nx; ++i) cd.x[i + 1] = stretch_x == 1.0 ? (i + 1) * (Lx / Nx) : cd.x[i] + Lx * dx[i] / tot;
  }
  std::vector<double> dy(Ny);
  double h = 1.0, tot = 0.0;
  for (int j = 0; j < Ny; ++j) { dy[j] = h; tot += h; h *= stretch; }
  cd.y[0] = 0.0;
  for (int j = 0; j < Ny; ++j) cd.y[j + 1] = cd.y[j] + dy[j] / tot;
  cd.u_left = bu[0]; cd.u_right = bu[1]; cd.u_low = bu[2]; cd.u_up = bu[3];
  cd.v_left = bv[0]; cd.v_right = bv[1]; cd.v_low = bv[2]; cd.v_up = bv[3];
  cd.Re = 500.0;
  cd.order4 = 0;
  Plan P;
  build_plan(cd, P);
  UniForm g{};
  UniArrays A;
  std::vector<unsigned char> fu, fv;
  uniform_maps(P, g, A, fu, fv);
  size_t su = 0, sv = 0;
  for (auto m : fu) su += m;
  for (auto m : fv) sv += m;
  printf("%s %d %d %d %d %zu %zu %d %d\n", name, g.ou, g.ov, P.Nu, P.Nv, su, sv, P.Nux_in, P.Nvy_in);
}

int main() {
  // boundary kinds: 0 dir, 1 per, 2 pres, 3 sym
  const int dir[4] = {0, 0, 0, 0}, chu[4] = {0, 2, 0, 0}, chv[4] = {0, 3, 0, 0}, per[4] = {1, 1, 0, 0};
  run("cavity", 64, 48, 1.0, dir, dir, 1.0);
  run("channel", 96, 24, 4.0, chu, chv, 1.0);
  run("periodic_x", 64, 32, 2.0, per, per, 1.0);
  run("stretched", 48, 40, 1.0, dir, dir, 1.05);
  run("stretched_xy", 40, 36, 2.0, chu, chv, 1.04, 0.97);
  return 0;
}
