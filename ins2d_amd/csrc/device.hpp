// Device-side descriptors of the kron-factored operators and the kernel entry
// points shared by the translation units of libins2d_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace ins2d {

// 1-D operator, ELL layout: row r has K (col, val) pairs
// K: entries stored per row (the row stride); kmax: the most nonzero entries
// of any row (K >= kmax, the rest zero-padded)
struct DOp1 {
  int nrows, ncols, K, kmax;
  const int* col;
  const double* val;
};
// kron(Y, X): input field (X.ncols x Y.ncols, x fastest) -> (X.nrows x Y.nrows)
struct DKron {
  DOp1 X, Y;
};
// y(f, g) = sum_t ax[t][f] * ay[t][g]; term t is nonzero only on the box
// [x0[t], x1[t]) x [y0[t], y1[t]) (the nonzero ranges of ax[t] and ay[t],
// set when the vectors are uploaded), so the kernels skip its loads outside
struct DYV {
  int nt, pad;
  const double* ax[4];
  const double* ay[4];
  int x0[4], x1[4], y0[4], y1[4];
};

// [lo, hi): the range of the nonzero entries of v (empty: lo = hi = 0)
inline void nonzero_range(const double* v, size_t n, int& lo, int& hi) {
  lo = 0;
  hi = 0;
  size_t a = 0, b = n;
  while (a < n && v[a] == 0.0) ++a;
  while (b > a && v[b - 1] == 0.0) --b;
  if (a < b) {
    lo = (int)a;
    hi = (int)b;
  }
}
// D * [(I c + yI) .* (A own + yA)] * scale
struct DConv {
  DKron D, I, A;
  DYV yI, yA;
  int icomp, pad;
  double scale;
};
// D * (sscale * S own + yS)
struct DDiff {
  DKron D, S;
  DYV yS;
  double sscale;
};
// closed form of the generic path's interior points on uniform grids:
// point (i, j) of a component takes it where fast[i + nx j] is set (and no
// boundary term reaches it); u array column 0 is grid line ou, v array row 0
// grid line ov
struct UniForm {
  const unsigned char* fast;
  double hx, hy, rhx, rhy, invRe;
  int ou, ov, nux, nuy, nvx, nvy;
  // stretched tensor grids: cell widths hx (Nx), hy (Ny), line spacings gx
  // (Nx + 1), gy (Ny + 1) and their reciprocals; null on uniform grids
  const double *ahx, *ahy, *agx, *agy, *arhx, *arhy, *argx, *argy;
};
// a w x h block of points (w h <= 256) of one momentum component: comp bit
// 0 the component, bit 1 set for points that take the closed form (the
// interior rectangle), clear for the generic operators (the boundary ring);
// inst: the ensemble member (0 for a single case)
struct RTile {
  int inst, comp, i0, j0, w, h;
};
struct DComp {
  int nx, ny, nconv, ndiff;
  DConv conv[2];
  DDiff diff[2];
  const double* ox;   // volume = ox[i] * oy[j]
  const double* oy;
  // boundary masks: the point (i, j) reads a boundary term only where
  // bmx[i] & bmy[j] != 0 (interior points skip them all)
  const unsigned* bmx;
  const unsigned* bmy;
  UniForm uni;   // uni.fast null: no closed form
  DKron G;            // gradient of p onto this component
  DKron M;            // divergence of this component onto p
  DYV yM;
};
struct DLap {        // A = kron(L1.Y, L1.X) + kron(L2.Y, L2.X)
  DKron L1, L2;
  const double* dinv;   // 1 / diag(A)
  int nx, ny;
};

// ---- fourth-order operators on grids with walls, CSR (row pointers rp,
// column indices ci, values v)
struct DCsr {
  int m, n;
  const int* rp;
  const int* ci;
  const double* v;
};
struct W4Dev {
  int NV, Np, Nf;
  DCsr Af, If, Cw, Diff, G, M;
  const double *yA, *yI, *yD, *yM, *Om;
  double* w;   // the Nf face products
  double* K;   // Np x Np inverse of the pressure matrix (row-major)
};

// ---- stage combination: Vtemp = sum_j a[j] k[j] (+ aself * k_i)
struct StageCombine {
  int n;
  double aself;
  double a[8];
  const double* k[8];
};

// ---- FFT Poisson plan (periodic, uniform)
struct FFTPlan {
  int Nx, Ny, Nh;            // Nh = Nx/2 + 1
  bool pow2x, pow2y;
  double* twx;               // exp(-2 pi i k / Nx), k < Nx (interleaved re, im)
  double* twy;
  double* ax;                // symbol: Ahat(kx, ky) = c * (ax[kx] + ay[ky]), Ahat(0,0) = 1
  double* ay;
  double c;
  double* H;                 // Ny x Nh complex (row transforms)
  double* HT;                // Nh x Ny complex (transposed)
  // max |fhat|^2 of the solve in progress (as the bits of a non-negative
  // double): zeroed by the forward row transform, raised by the column
  // solve, read by the inverse row transform, which returns dp = 0 below
  // 1e-28 (the early exit of pressure_poisson.m:21-31)
  unsigned long long* fmax2;
};

}  // namespace ins2d
