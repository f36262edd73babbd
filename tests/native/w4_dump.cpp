// Test harness (host only): builds the fourth-order wall operators of
// ins2d_amd/csrc/ops4.cpp and hands them out in COO form, so that the CPU
// tests compare the assembly itself with the oracle, matrix by matrix.
#include <cstring>
#include <exception>
#include <string>

#include "../../ins2d_amd/csrc/ops4.hpp"

using namespace ins2d;

extern "C" {

void* w4h_build(int Nx, int Ny, const double* x, const double* y, const int* bcu, const int* bcv, double Re,
                double alfa, double beta, char* err, int errlen) {
  try {
    CaseDesc cd;
    cd.Nx = Nx;
    cd.Ny = Ny;
    cd.x.assign(x, x + Nx + 1);
    cd.y.assign(y, y + Ny + 1);
    cd.u_left = bcu[0]; cd.u_right = bcu[1]; cd.u_low = bcu[2]; cd.u_up = bcu[3];
    cd.v_left = bcv[0]; cd.v_right = bcv[1]; cd.v_low = bcv[2]; cd.v_up = bcv[3];
    cd.Re = Re;
    cd.order4 = 1;
    cd.alfa = alfa;
    cd.beta = beta;
    auto* W = new W4Ops();
    build_w4(cd, *W);
    return W;
  } catch (const std::exception& e) {
    if (err && errlen > 0) {
      std::strncpy(err, e.what(), errlen - 1);
      err[errlen - 1] = 0;
    }
    return nullptr;
  }
}

void w4h_free(void* p) { delete static_cast<W4Ops*>(p); }

// Nx Ny nux nuy nvx nvy Nu Nv NV Np Nf singular
void w4h_sizes(void* p, int* out) {
  const W4Ops& W = *static_cast<W4Ops*>(p);
  const int v[12] = {W.Nx, W.Ny, W.nux, W.nuy, W.nvx, W.nvy, W.Nu, W.Nv, W.NV, W.Np, W.Nf, W.singular ? 1 : 0};
  std::memcpy(out, v, sizeof(v));
}

// nonzeros of the named matrix; with rows non-null also its COO triplets
long w4h_matrix(void* p, const char* name, int* rows, int* cols, double* vals) {
  const W4Ops& W = *static_cast<W4Ops*>(p);
  const std::string n(name);
  const Sp* A = n == "Af" ? &W.Af : n == "If" ? &W.If : n == "Cw" ? &W.Cw : n == "Diff" ? &W.Diff
              : n == "G" ? &W.G : n == "M" ? &W.M : n == "A" ? &W.A : nullptr;
  if (!A) return -1;
  long k = 0;
  for (int i = 0; i < A->m; ++i)
    for (auto& kv : A->r[i]) {
      if (rows) { rows[k] = i; cols[k] = kv.first; vals[k] = kv.second; }
      ++k;
    }
  return k;
}

// Om (NV), xin, yin
long w4h_vector(void* p, const char* name, double* out) {
  const W4Ops& W = *static_cast<W4Ops*>(p);
  const std::string n(name);
  const std::vector<double>* v = n == "Om" ? &W.Om : n == "xin" ? &W.xin : n == "yin" ? &W.yin : nullptr;
  if (!v) return -1;
  if (out) std::memcpy(out, v->data(), v->size() * sizeof(double));
  return (long)v->size();
}

// boundary vectors at the twelve samples (lengths as the C ABI takes them)
int w4h_bc(void* p, const double* const* samples, double* yA, double* yI, double* yD, double* yM) {
  const W4Ops& W = *static_cast<W4Ops*>(p);
  try {
    BCSamples S;
    for (int s = 0; s < S_NSRC; ++s) S.s[s].assign(samples[s], samples[s] + w4_sample_len(W, s));
    std::vector<double> a, b, c, d;
    w4_bc_vectors(W, S, a, b, c, d);
    std::memcpy(yA, a.data(), a.size() * sizeof(double));
    std::memcpy(yI, b.data(), b.size() * sizeof(double));
    std::memcpy(yD, c.data(), c.size() * sizeof(double));
    std::memcpy(yM, d.data(), d.size() * sizeof(double));
    return 0;
  } catch (const std::exception&) {
    return 1;
  }
}

}  // extern "C"
