// Test helper: builds the host operator plan of libins2d_hip (ops1d.cpp) for
// a case read from stdin and prints every 1-D factor and boundary term as
// JSON, so the Python tests can rebuild the 2-D operators with kron() and
// compare them with the CPU oracle's assembled matrices.
//
// stdin: Nx Ny | x[0..Nx] | y[0..Ny] | 8 BC kinds | Re | 12 samples (n v...)
// With the argument --fdm it also prints the fast-diagonalization factors of
// the pressure matrix and its IC(0) factor (fdm_host.cpp).
#include <cstdio>
#include <iostream>
#include <string>
#include <vector>

#include "../../ins2d_amd/csrc/ops1d.hpp"

using namespace ins2d;

static void vec(const std::vector<double>& v) {
  printf("[");
  for (size_t i = 0; i < v.size(); ++i) printf(i ? ",%.17g" : "%.17g", v[i]);
  printf("]");
}
static void ivec(const std::vector<int>& v) {
  printf("[");
  for (size_t i = 0; i < v.size(); ++i) printf(i ? ",%d" : "%d", v[i]);
  printf("]");
}
static void op(const HOp1& o) {
  printf("{\"nrows\":%d,\"ncols\":%d,\"K\":%d,\"col\":", o.nrows, o.ncols, o.K);
  ivec(o.col);
  printf(",\"val\":");
  vec(o.val);
  printf("}");
}
static void kr(const char* name, const HKron& k) {
  printf("\"%s\":{\"X\":", name);
  op(k.X);
  printf(",\"Y\":");
  op(k.Y);
  printf("}");
}
static void yv(const char* name, const HYVec& Y, const BCSamples& S) {
  std::vector<std::vector<double>> ax, ay;
  eval_yvec(Y, S, ax, ay);
  printf("\"%s\":[", name);
  for (size_t t = 0; t < ax.size(); ++t) {
    printf(t ? ",[" : "[");
    vec(ax[t]);
    printf(",");
    vec(ay[t]);
    printf("]");
  }
  printf("]");
}
static void comp(const char* name, const HComp& C, const BCSamples& S) {
  printf("\"%s\":{\"nx\":%d,\"ny\":%d,\"conv\":[", name, C.nx, C.ny);
  for (size_t t = 0; t < C.conv.size(); ++t) {
    printf(t ? ",{" : "{");
    kr("D", C.conv[t].D); printf(",");
    kr("I", C.conv[t].I); printf(",");
    kr("A", C.conv[t].A); printf(",");
    yv("yI", C.conv[t].yI, S); printf(",");
    yv("yA", C.conv[t].yA, S);
    printf(",\"icomp\":%d,\"scale\":%.17g}", C.conv[t].icomp, C.conv[t].scale);
  }
  printf("],\"diff\":[");
  for (size_t t = 0; t < C.diff.size(); ++t) {
    printf(t ? ",{" : "{");
    kr("D", C.diff[t].D[0]); printf(",");
    kr("S", C.diff[t].S[0]); printf(",");
    yv("yS", C.diff[t].yS, S);
    printf(",\"sscale\":%.17g}", C.diff[t].sscale[0]);
  }
  printf("],\"ox\":");
  vec(C.ox);
  printf(",\"oy\":");
  vec(C.oy);
  printf(",");
  kr("G", C.G); printf(",");
  kr("M", C.M); printf(",");
  yv("yM", C.yM, S);
  printf("}");
}

int main(int argc, char** argv) {
  const bool fdm = argc > 1 && std::string(argv[1]) == "--fdm";
  CaseDesc cd;
  std::cin >> cd.Nx >> cd.Ny;
  cd.x.resize(cd.Nx + 1);
  cd.y.resize(cd.Ny + 1);
  for (auto& v : cd.x) std::cin >> v;
  for (auto& v : cd.y) std::cin >> v;
  std::cin >> cd.u_left >> cd.u_right >> cd.u_low >> cd.u_up >> cd.v_left >> cd.v_right >> cd.v_low >> cd.v_up;
  std::cin >> cd.Re;
  BCSamples S;
  for (int s = 0; s < S_NSRC; ++s) {
    int n;
    std::cin >> n;
    S.s[s].resize(n);
    for (auto& v : S.s[s]) std::cin >> v;
  }
  Plan P;
  try {
    build_plan(cd, P);
  } catch (const std::exception& e) {
    printf("{\"error\":\"%s\"}\n", e.what());
    return 0;
  }
  for (int s = 0; s < S_NSRC; ++s)
    if ((int)S.s[s].size() != sample_len(P, s)) {
      printf("{\"error\":\"sample %d has length %zu, expected %d\"}\n", s, S.s[s].size(), sample_len(P, s));
      return 0;
    }
  printf("{\"sizes\":");
  ivec({P.Nux_in, P.Nuy_in, P.Nvx_in, P.Nvy_in, P.Npx, P.Npy});
  printf(",\"xin\":");
  vec(P.xin);
  printf(",\"yin\":");
  vec(P.yin);
  printf(",");
  comp("u", P.cu, S);
  printf(",");
  comp("v", P.cv, S);
  printf(",");
  kr("L1", P.L1);
  printf(",");
  kr("L2", P.L2);
  printf(",\"Adiag\":");
  vec(P.Adiag);
  if (fdm) {
    FDMHost F;
    try {
      build_fdm_host(P, F);
    } catch (const std::exception& e) {
      printf(",\"fdm_error\":\"%s\"}\n", e.what());
      return 0;
    }
    printf(",\"fdm\":{\"nx0\":%d,\"ny0\":%d,\"Vx\":", F.nx0, F.ny0);
    vec(F.Vx);
    printf(",\"Vy\":");
    vec(F.Vy);
    printf(",\"lx\":");
    vec(F.lx);
    printf(",\"ly\":");
    vec(F.ly);
    printf("}");
    IC0Host I;
    try {
      build_ic0_host(P, I);
      printf(",\"ic0\":{\"D\":");
      vec(I.D);
      printf("}");
    } catch (const std::exception& e) {
      printf(",\"ic0_error\":\"%s\"", e.what());
    }
  }
  printf("}\n");
  return 0;
}
