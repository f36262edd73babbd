// Boundary masks of the generic momentum kernels (host side).
#pragma once

#include <vector>

#include "device.hpp"
#include "ops1d.hpp"

namespace ins2d {

// Bit b of mx[i] & my[j] is set when the b-th boundary outer-product term of
// a momentum component (the terms of the conv terms' yI and yA, then of the
// diff terms' yS, in that order) can be nonzero at one of the faces that the
// point (i, j) reads through the term's D operator: a face (f, g) of D's rows
// i and j, with f in the term's [x0, x1) and g in its [y0, y1).  Points with
// mx[i] & my[j] == 0 read no boundary term at all.  Past 32 terms every
// point is marked.
inline void bc_masks(const HComp& H, const DComp& D, std::vector<unsigned>& mx, std::vector<unsigned>& my) {
  mx.assign(D.nx, 0u);
  my.assign(D.ny, 0u);
  int bit = 0;
  bool all = false;
  auto rows_hit = [](const HOp1& O, int lo, int hi, unsigned m, std::vector<unsigned>& out) {
    for (int r = 0; r < O.nrows && r < (int)out.size(); ++r)
      for (int a = 0; a < O.K; ++a) {
        const int c = O.col[(size_t)r * O.K + a];
        if (c >= lo && c < hi) out[r] |= m;
      }
  };
  auto add = [&](const HKron& K, const DYV& Y) {
    for (int t = 0; t < Y.nt; ++t, ++bit) {
      if (bit >= 32) {
        all = true;
        return;
      }
      if (Y.x0[t] >= Y.x1[t] || Y.y0[t] >= Y.y1[t]) continue;
      rows_hit(K.X, Y.x0[t], Y.x1[t], 1u << bit, mx);
      rows_hit(K.Y, Y.y0[t], Y.y1[t], 1u << bit, my);
    }
  };
  for (int t = 0; t < D.nconv; ++t) {
    add(H.conv[t].D, D.conv[t].yI);
    add(H.conv[t].D, D.conv[t].yA);
  }
  for (int t = 0; t < D.ndiff; ++t) add(H.diff[t].D[0], D.diff[t].yS);
  if (all) {
    mx.assign(D.nx, ~0u);
    my.assign(D.ny, ~0u);
  }
}

}  // namespace ins2d
