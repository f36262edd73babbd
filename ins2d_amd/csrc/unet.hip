// Learned Poisson warm start: a 4-level U-shaped CNN (16/32/64/64 channels,
// 3x3 convolutions, BF16 weights and activations, FP32 accumulation on the
// MFMA units) with hand-written forward and backward kernels, a trainer step
// (MSE loss, Adam) and the inference entry used by the pressure solver.
//
// Tensors are NHWC (x fastest over pixels, channels innermost), H = Npy rows,
// W = Npx columns, so a pressure field in the solver's layout is a 1-channel
// tensor.  A 3x3 convolution is an implicit GEMM: M = pixels, N = output
// channels, K = 9 * input channels, computed with v_mfma_f32_16x16x16_bf16.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/ins2d.h"
#include "stencil.hpp"

namespace ins2d {
namespace nn {

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);   // round to nearest even
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float(((uint32_t)b) << 16); }

// Input of a convolution: channels [0, a.c) from a, [a.c, a.c + b.c) from b,
// zeros up to the layer's cin.  Each source is read on the convolution's grid
// (H x W) in one of four ways:
enum : int {
  SRC_DIRECT = 0,   // same grid
  SRC_POOL = 1,     // a 2H x 2W tensor, 2x2 average rounded to bf16
  SRC_UP = 2,       // an H/2 x W/2 tensor, nearest neighbour
  SRC_F64 = 3       // one fp64 field, times 1 / (div * s), rounded to bf16
};
struct CvSrc {
  const void* p;
  int c;        // channels
  int fp32;     // storage of DIRECT/POOL/UP sources: 1 float, 0 bf16
  int mode;     // SRC_*
  double div;   // SRC_F64 divisor (besides s)
};

struct ConvArgs {
  CvSrc a, b;
  int cin;              // input channels of the weights (>= a.c + b.c; the rest are zero)
  int H, W, cout, ks;   // ks 3 (padding 1) or 1
  const uint16_t* w;    // [taps][cout][cin] bf16
  const float* bias;    // cout (or null)
  void* out;            // [H][W][cout], bf16 or fp32 (null with the fused output layer)
  int out_fp32;
  int relu;
  // s = sqrt(sum(sq[0..nsq)) / npix_s): the input scale of SRC_F64 sources and of the fused output
  const double* sq;
  int nsq;
  long npix_s;
  // fused 1x1 output layer (cout 16): y = b1 + sum_c bf16(act_c) w1[c], out64 = base + oscale * s * y
  const uint16_t* w1;
  const float* b1;
  double* out64;
  const double* base;
  double oscale;
};

// Tiles: a block of 4 waves covers 4 rows x 16 MT pixels, wave w the row
// y0 + w (MT MFMA row tiles of 16 pixels) times 16 NT output channels
// (16 NT * blockIdx.z + [0, 16 NT)).  The input rows y0-1 .. y0+4 and columns
// x0-1 .. x0+16 MT are staged in LDS as bf16, 32 channels at a time, with a
// per-pixel stride of 36 channels (the 16 pixels a row tile reads land in
// distinct banks); the chunk's weights are staged beside them (per output
// channel the same stride).  Each thread issues all its staging loads before
// its LDS stores, so a chunk's loads are in flight together.
constexpr int CV_TY = 4, CV_KC = 32, CV_CP = 36;
__device__ __forceinline__ float load_ch(const CvSrc& s, long pix, int c) {
  if (s.fp32) return ((const float*)s.p)[pix * s.c + c];
  return bf2f(((const uint16_t*)s.p)[pix * s.c + c]);
}

__device__ __forceinline__ float src_val(const CvSrc& S, int H, int W, int yy, int xx, int c) {
  if (S.mode == SRC_POOL) {
    const int W2 = 2 * W;
    const long b = ((long)(2 * yy) * W2 + 2 * xx) * S.c + c;
    const long C = S.c, R = (long)W2 * S.c;
    float s;
    if (S.fp32) {
      const float* q = (const float*)S.p;
      s = q[b] + q[b + C] + q[b + R] + q[b + R + C];
    } else {
      const uint16_t* q = (const uint16_t*)S.p;
      s = bf2f(q[b]) + bf2f(q[b + C]) + bf2f(q[b + R]) + bf2f(q[b + R + C]);
    }
    return bf2f(f2bf(0.25f * s));
  }
  long pix;
  if (S.mode == SRC_UP) pix = (long)(yy >> 1) * (W >> 1) + (xx >> 1);
  else pix = (long)yy * W + xx;
  const long o = pix * S.c + c;
  return S.fp32 ? ((const float*)S.p)[o] : bf2f(((const uint16_t*)S.p)[o]);
}

// Staging kinds (a template parameter of the convolution, so that each kernel
// instance holds only the loads it needs):
enum : int {
  SK_ANY = 0,    // any sources (fp32 and odd channel counts: the backward pass, layer checks)
  SK_BF16 = 1,   // bf16 sources with 4k channels, DIRECT or UP
  SK_POOL = 2,   // one bf16 POOL source with 4k channels
  SK_F64 = 3     // SRC_F64 sources (the network input)
};

// the 4 channels c .. c+3 (c a multiple of 4) at (yy, xx), as bf16; zero outside
// the grid and beyond the sources' channels
template <int SK>
__device__ __forceinline__ short4_t stage_quad(const ConvArgs& A, int yy, int xx, int c, double inv) {
  short4_t v = short4_t{0, 0, 0, 0};
  if (yy < 0 || yy >= A.H || xx < 0 || xx >= A.W) return v;
  if (SK == SK_F64) {   // channel 0 from a, channel 1 from b (if any)
    if (c == 0) {
      const long pix = (long)yy * A.W + xx;
      v[0] = (short)f2bf((float)(((const double*)A.a.p)[pix] * (inv / A.a.div)));
      if (A.b.p) v[1] = (short)f2bf((float)(((const double*)A.b.p)[pix] * (inv / A.b.div)));
    }
    return v;
  }
  const int nc = A.a.c + A.b.c;
  if (c >= nc) return v;
  const bool ina = c < A.a.c;
  const CvSrc& S = ina ? A.a : A.b;
  const int cc = ina ? c : c - A.a.c;
  if (SK == SK_BF16) {
    const uint16_t* q = (const uint16_t*)S.p;
    const long pix = S.mode == SRC_UP ? (long)(yy >> 1) * (A.W >> 1) + (xx >> 1) : (long)yy * A.W + xx;
    return *(const short4_t*)(q + pix * S.c + cc);
  }
  if (SK == SK_POOL) {
    const uint16_t* q = (const uint16_t*)S.p;
    const long C = S.c, R = 2L * A.W * S.c;
    const long b = ((long)(2 * yy) * (2 * A.W) + 2 * xx) * S.c + cc;
    const short4_t v00 = *(const short4_t*)(q + b), v01 = *(const short4_t*)(q + b + C);
    const short4_t v10 = *(const short4_t*)(q + b + R), v11 = *(const short4_t*)(q + b + R + C);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float s = bf2f((uint16_t)v00[i]) + bf2f((uint16_t)v01[i]) + bf2f((uint16_t)v10[i]) + bf2f((uint16_t)v11[i]);
      v[i] = (short)f2bf(0.25f * s);
    }
    return v;
  }
  if (S.mode == SRC_DIRECT && S.fp32 && (S.c & 3) == 0) {
    const float4_t f = *(const float4_t*)((const float*)S.p + ((long)yy * A.W + xx) * S.c + cc);
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] = (short)f2bf(f[i]);
    return v;
  }
  for (int i = 0; i < 4; ++i) {
    const int ci = c + i;
    if (ci >= nc) break;
    const bool ia = ci < A.a.c;
    v[i] = (short)f2bf(src_val(ia ? A.a : A.b, A.H, A.W, yy, xx, ia ? ci : ci - A.a.c));
  }
  return v;
}

// out[y][x][co] = act(bias[co] + sum_t sum_ci in[y+dy][x+dx][ci] w[t][co][ci])
// grid (ceil(W / (16 MT)), ceil(H / 4), ceil(cout / (16 NT))), 256 threads
template <int NT, int MT, int SK>
__global__ void __launch_bounds__(256) k_conv(ConvArgs A) {
  constexpr int TW = 16 * MT;
  constexpr int TP = (CV_TY + 2) * (TW + 2);   // staged pixels (ks 3)
  constexpr int QA = (TP * 8 + 255) / 256;     // input quads per thread
  constexpr int WR = 9 * 16 * NT;              // staged weight rows (tap, output channel)
  constexpr int QW = (WR * 8 + 255) / 256;     // weight quads per thread
  __shared__ __align__(16) uint16_t tile[TP * CV_CP];
  __shared__ __align__(16) uint16_t wsh[WR * CV_CP];
  __shared__ double red[16];
  __shared__ double s_sh;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int x0 = blockIdx.x * TW, y0 = blockIdx.y * CV_TY;
  const int cob = 16 * NT * blockIdx.z;
  const int r = A.ks / 2, taps = A.ks * A.ks;
  const int tw = TW + 2 * r, th = CV_TY + 2 * r;
  double s = 1.0;
  if (A.sq) {
    const double t = block_sum((int)threadIdx.x < A.nsq ? A.sq[threadIdx.x] : 0.0, red);
    if (threadIdx.x == 0) s_sh = sqrt(t / (double)A.npix_s);
    __syncthreads();
    s = s_sh;
  }
  const double inv = 1.0 / fmax(s, 1e-300);
  const int kq = 4 * (lane >> 4);
  const int m = lane & 15;
  float4_t acc[MT][NT];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int n = 0; n < NT; ++n) acc[mt][n] = float4_t{0.f, 0.f, 0.f, 0.f};
  const bool wvec = (A.cin & 3) == 0;
  for (int kc = 0; kc < A.cin; kc += CV_KC) {
    const int nsub = min(2, (A.cin - kc + 15) / 16);
    const int lq = nsub + 1;                    // log2 of the quads staged per pixel (4 or 8)
    const int nq = (th * tw) << lq;
    const int nw = (taps * 16 * NT) << lq;
    short4_t ba[QA], bw[QW];
#pragma unroll
    for (int k = 0; k < QA; ++k) {
      const int q = threadIdx.x + 256 * k;
      ba[k] = short4_t{0, 0, 0, 0};
      if (q < nq) {
        const int px = q >> lq, cq = q & ((1 << lq) - 1);
        const int ty = px / tw, tx = px - ty * tw;
        ba[k] = stage_quad<SK>(A, y0 - r + ty, x0 - r + tx, kc + 4 * cq, inv);
      }
    }
#pragma unroll
    for (int k = 0; k < QW; ++k) {
      const int q = threadIdx.x + 256 * k;
      bw[k] = short4_t{0, 0, 0, 0};
      if (q < nw) {
        const int row = q >> lq, cq = q & ((1 << lq) - 1);
        const int t = row / (16 * NT), co = cob + row - t * (16 * NT);
        const int ci = kc + 4 * cq;
        if (co < A.cout && ci < A.cin) {
          const uint16_t* wp = A.w + ((long)t * A.cout + co) * A.cin + ci;
          if (wvec) bw[k] = *(const short4_t*)wp;
          else
#pragma unroll
            for (int i = 0; i < 4; ++i)
              if (ci + i < A.cin) bw[k][i] = (short)wp[i];
        }
      }
    }
    if (kc) __syncthreads();   // the previous chunk's reads are done
#pragma unroll
    for (int k = 0; k < QA; ++k) {
      const int q = threadIdx.x + 256 * k;
      if (q < nq) *(short4_t*)&tile[(q >> lq) * CV_CP + 4 * (q & ((1 << lq) - 1))] = ba[k];
    }
#pragma unroll
    for (int k = 0; k < QW; ++k) {
      const int q = threadIdx.x + 256 * k;
      if (q < nw) *(short4_t*)&wsh[(q >> lq) * CV_CP + 4 * (q & ((1 << lq) - 1))] = bw[k];
    }
    __syncthreads();
    for (int t = 0; t < taps; ++t) {
      const int dy = t / A.ks, dx = t - dy * A.ks;
      for (int sub = 0; sub < nsub; ++sub) {
        short4_t bv[NT];
#pragma unroll
        for (int n = 0; n < NT; ++n) bv[n] = *(const short4_t*)&wsh[(t * 16 * NT + 16 * n + m) * CV_CP + 16 * sub + kq];
        const uint16_t* ap = &tile[((wave + dy) * tw + m + dx) * CV_CP + 16 * sub + kq];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
          const short4_t av = *(const short4_t*)(ap + 16 * mt * CV_CP);
#pragma unroll
          for (int n = 0; n < NT; ++n) acc[mt][n] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv[n], acc[mt][n], 0, 0, 0);
        }
      }
    }
  }
  const int y = y0 + wave;
  if (A.w1) {   // fused output layer (NT 1, one z block): channel m of pixel x in lane (x, m)
    const float bias = A.bias ? A.bias[m] : 0.f;
    const float w1 = bf2f(A.w1[m]);
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        float v = acc[mt][0][i] + bias;
        if (A.relu) v = fmaxf(v, 0.f);
        float pr = bf2f(f2bf(v)) * w1;
        pr += __shfl_xor(pr, 1, 64);
        pr += __shfl_xor(pr, 2, 64);
        pr += __shfl_xor(pr, 4, 64);
        pr += __shfl_xor(pr, 8, 64);
        const int x = x0 + 16 * mt + kq + i;
        if (m == 0 && y < A.H && x < A.W) {
          const long o = (long)y * A.W + x;
          const float yv = pr + (A.b1 ? A.b1[0] : 0.f);
          A.out64[o] = (A.base ? A.base[o] : 0.0) + A.oscale * s * (double)yv;
        }
      }
    return;
  }
  if (y >= A.H) return;
#pragma unroll
  for (int n = 0; n < NT; ++n) {
    const int co = cob + 16 * n + m;
    if (co >= A.cout) continue;
    const float bias = A.bias ? A.bias[co] : 0.f;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int x = x0 + 16 * mt + kq + i;
        if (x >= A.W) continue;
        float v = acc[mt][n][i] + bias;
        if (A.relu) v = fmaxf(v, 0.f);
        const long o = ((long)y * A.W + x) * A.cout + co;
        if (A.out_fp32) ((float*)A.out)[o] = v;
        else ((uint16_t*)A.out)[o] = f2bf(v);
      }
  }
}

// dw[t][co][ci] += sum_p g[p][co] in[p + off_t][ci]   (fp32 atomics)
// grid (chunks, taps, tiles_co * tiles_ci); 64 threads
__global__ void __launch_bounds__(64) k_conv_dw(CvSrc a, CvSrc b, const float* __restrict__ g, int H, int W, int cout,
                                               int ks, int chunk, float* __restrict__ dw) {
  const int lane = threadIdx.x;
  const int cin = a.c + b.c;
  const int tci = (cin + 15) / 16;
  const int co0 = 16 * (blockIdx.z / tci), ci0 = 16 * (blockIdx.z % tci);
  const int t = blockIdx.y;
  const int r = ks / 2;
  const int dy = t / ks - r, dx = t % ks - r;
  const long npix = (long)H * W;
  const long p0 = (long)blockIdx.x * chunk;
  const long p1 = std::min<long>(p0 + chunk, npix);
  float4_t acc = float4_t{0.f, 0.f, 0.f, 0.f};
  const int com = co0 + (lane & 15);   // A row (output channel)
  const int cin_n = ci0 + (lane & 15); // B column (input channel)
  const int kq = 4 * (lane >> 4);
  for (long p = p0; p < p1; p += 16) {
    short4_t av, bv;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const long pp = p + kq + i;
      float ga = 0.f, xb = 0.f;
      if (pp < p1) {
        if (com < cout) ga = g[pp * cout + com];
        const int y = (int)(pp / W), x = (int)(pp % W);
        const int yy = y + dy, xx = x + dx;
        if (cin_n < cin && yy >= 0 && yy < H && xx >= 0 && xx < W) {
          const long sp = (long)yy * W + xx;
          xb = cin_n < a.c ? load_ch(a, sp, cin_n) : load_ch(b, sp, cin_n - a.c);
        }
      }
      av[i] = (short)f2bf(ga);
      bv[i] = (short)f2bf(xb);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int co = co0 + 4 * (lane >> 4) + i;
    const int ci = ci0 + (lane & 15);
    if (co < cout && ci < cin) atomicAdd(&dw[((long)t * cout + co) * cin + ci], acc[i]);
  }
}

// db[co] += sum_p g[p][co]
__global__ void k_bias_grad(const float* __restrict__ g, long npix, int cout, float* __restrict__ db) {
  __shared__ float red[256];
  const int co = blockIdx.y;
  float s = 0.f;
  for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < npix; p += (long)gridDim.x * blockDim.x)
    s += g[p * cout + co];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(&db[co], red[0]);
}

// g = dout .* (out > 0) (ReLU backward; out is the stored bf16 activation)
__global__ void k_relu_bwd(const float* __restrict__ dout, const uint16_t* __restrict__ out, long n,
                           float* __restrict__ g) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    g[i] = bf2f(out[i]) > 0.f ? dout[i] : 0.f;
}

// 2x2 average pool (H, W even), bf16 -> bf16
__global__ void k_pool(const uint16_t* __restrict__ in, int H, int W, int C, uint16_t* __restrict__ out) {
  const int Ho = H / 2, Wo = W / 2;
  const long n = (long)Ho * Wo * C;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int c = (int)(i % C);
    const long px = i / C;
    const int x = (int)(px % Wo), y = (int)(px / Wo);
    const long b = ((long)(2 * y) * W + 2 * x) * C + c;
    const float s = bf2f(in[b]) + bf2f(in[b + C]) + bf2f(in[b + (long)W * C]) + bf2f(in[b + (long)W * C + C]);
    out[i] = f2bf(0.25f * s);
  }
}
// dIn[y][x] += dOut[y/2][x/2] / 4
__global__ void k_pool_bwd(const float* __restrict__ dout, int H, int W, int C, float* __restrict__ din) {
  const long n = (long)H * W * C;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int c = (int)(i % C);
    const long px = i / C;
    const int x = (int)(px % W), y = (int)(px / W);
    din[i] += 0.25f * dout[((long)(y / 2) * (W / 2) + x / 2) * C + c];
  }
}
// nearest upsampling x2: out[y][x] = in[y/2][x/2]
__global__ void k_up(const uint16_t* __restrict__ in, int Hs, int Ws, int C, uint16_t* __restrict__ out) {
  const int H = 2 * Hs, W = 2 * Ws;
  const long n = (long)H * W * C;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int c = (int)(i % C);
    const long px = i / C;
    const int x = (int)(px % W), y = (int)(px / W);
    out[i] = in[((long)(y / 2) * Ws + x / 2) * C + c];
  }
}
// dIn[y][x] = sum of the 4 dOut children
__global__ void k_up_bwd(const float* __restrict__ dout, int Hs, int Ws, int C, float* __restrict__ din) {
  const int W = 2 * Ws;
  const long n = (long)Hs * Ws * C;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int c = (int)(i % C);
    const long px = i / C;
    const int x = (int)(px % Ws), y = (int)(px / Ws);
    const long b = ((long)(2 * y) * W + 2 * x) * C + c;
    din[i] = dout[b] + dout[b + C] + dout[b + (long)W * C] + dout[b + (long)W * C + C];
  }
}
// split a (ca + cb)-channel gradient: da (=) first ca channels, db (+=) rest
__global__ void k_split(const float* __restrict__ g, long npix, int ca, int cb, float* __restrict__ da,
                        float* __restrict__ db) {
  const int C = ca + cb;
  const long n = npix * C;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int c = (int)(i % C);
    const long px = i / C;
    if (c < ca) da[px * ca + c] = g[i];
    else db[px * cb + (c - ca)] += g[i];
  }
}
// weight transform for the backward-data convolution:
// wt[8 - t][ci][co] = w[t][co][ci] (taps flipped, channels swapped)
__global__ void k_wflip(const float* __restrict__ w, int taps, int cout, int cin, uint16_t* __restrict__ wt,
                        uint16_t* __restrict__ wb) {
  const long n = (long)taps * cout * cin;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int ci = (int)(i % cin);
    const int co = (int)((i / cin) % cout);
    const int t = (int)(i / ((long)cin * cout));
    wb[i] = f2bf(w[i]);
    wt[((long)(taps - 1 - t) * cin + ci) * cout + co] = f2bf(w[i]);
  }
}
// network input (fp64 -> 4-channel bf16): [f / s, xc / (scale s), 0, 0] with
// s = rms(f); xc (optional) a coarse-grid solution the output corrects
__global__ void k_input(const double* __restrict__ f, const double* __restrict__ xc, long n,
                        const double* __restrict__ s, double scale, uint16_t* __restrict__ x) {
  const double inv = 1.0 / fmax(s[0], 1e-300);
  const double invc = inv / scale;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    x[4 * i] = f2bf((float)(f[i] * inv));
    x[4 * i + 1] = xc ? f2bf((float)(xc[i] * invc)) : (uint16_t)0;
    x[4 * i + 2] = 0;
    x[4 * i + 3] = 0;
  }
}
// p = xc + scale * s * y (network output fp32 -> fp64; xc optional, may alias p)
__global__ void k_output(const float* __restrict__ y, const double* xc, long n, const double* __restrict__ s,
                         double scale, double* p) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    p[i] = (xc ? xc[i] : 0.0) + scale * s[0] * (double)y[i];
}
// partial sums of f^2 over kSq blocks of 256 threads
constexpr int kSq = 256;
__global__ void k_sumsq(const double* __restrict__ f, long n, double* __restrict__ part) {
  __shared__ double red[16];
  double a = 0.0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) a += f[i] * f[i];
  a = block_sum(a, red);
  if (threadIdx.x == 0) part[blockIdx.x] = a;
}
// s = sqrt(sum(part) / n) (one block of 256 threads: the same sum the
// convolution kernels form from the partials)
__global__ void k_rms(const double* __restrict__ part, int np, long n, double* __restrict__ s) {
  __shared__ double red[16];
  const double t = block_sum((int)threadIdx.x < np ? part[threadIdx.x] : 0.0, red);
  if (threadIdx.x == 0) s[0] = sqrt(t / (double)n);
}
// MSE loss of y against the target tt = (t - xc) / (scale s) (xc optional):
// dy = 2 (y - tt) / n; loss += sum (y - tt)^2 / n.  With tpart (partial sums
// of (t - xc)^2 over kSq blocks) the relative loss: dy = 2 (y - tt) / T,
// loss += sum (y - tt)^2 / T with T = sum tt^2.
__global__ void k_mse(const float* __restrict__ y, const double* __restrict__ t, const double* __restrict__ xc, long n,
                      const double* __restrict__ s, double scale, const double* __restrict__ tpart,
                      float* __restrict__ dy, double* __restrict__ loss) {
  __shared__ double red[16];
  __shared__ double den_sh;
  const double inv = 1.0 / (scale * fmax(s[0], 1e-300));
  double den = (double)n;
  if (tpart) {
    const double tt = block_sum((int)threadIdx.x < kSq ? tpart[threadIdx.x] : 0.0, red);
    if (threadIdx.x == 0) den_sh = fmax(tt * inv * inv, 1e-300);
    __syncthreads();
    den = den_sh;
  }
  double a = 0.0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const double d = (double)y[i] - (t[i] - (xc ? xc[i] : 0.0)) * inv;
    dy[i] = (float)(2.0 * d / den);
    a += d * d / den;
  }
  a = block_sum(a, red);
  if (threadIdx.x == 0) atomicAdd(loss, a);
}
// partial sums of (t - xc)^2 over kSq blocks of 256 threads (xc optional)
__global__ void k_sumsq_diff(const double* __restrict__ t, const double* __restrict__ xc, long n,
                             double* __restrict__ part) {
  __shared__ double red[16];
  double a = 0.0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const double d = t[i] - (xc ? xc[i] : 0.0);
    a += d * d;
  }
  a = block_sum(a, red);
  if (threadIdx.x == 0) part[blockIdx.x] = a;
}
// Adam (fp32 master parameters)
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       long n, float lr, float b1, float b2, float eps, float c1, float c2, float gscale) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const float gi = g[i] * gscale;
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}
// pair dst = ca f_i + cb f_j with ca = a / rms(f_i), cb = b / rms(f_j)
// (the same weights on p and xc: the three arrays of a pair are contiguous)
__global__ void k_combine(const double* __restrict__ si, const double* __restrict__ sj, long n3,
                          const double* __restrict__ rs, double a, double b, double* __restrict__ dst) {
  const double ca = a / fmax(rs[0], 1e-300), cb = b / fmax(rs[1], 1e-300);
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n3; i += (long)gridDim.x * blockDim.x)
    dst[i] = ca * si[i] + cb * sj[i];
}
__global__ void k_zero_f(float* x, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) x[i] = 0.f;
}
__global__ void k_zero_d(double* x, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) x[i] = 0.0;
}

// one 16x16x16 bf16 MFMA of fp32 inputs rounded to bf16: D = A (16x16) B (16x16)
// with A, B, D row-major; used to check the operand layout the kernels assume
__global__ void k_mfma_probe(const float* __restrict__ A, const float* __restrict__ B, float* __restrict__ D) {
  const int lane = threadIdx.x;
  short4_t av, bv;
  for (int i = 0; i < 4; ++i) {
    av[i] = (short)f2bf(A[(lane & 15) * 16 + 4 * (lane >> 4) + i]);
    bv[i] = (short)f2bf(B[(4 * (lane >> 4) + i) * 16 + (lane & 15)]);
  }
  float4_t acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, float4_t{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
  for (int i = 0; i < 4; ++i) D[(4 * (lane >> 4) + i) * 16 + (lane & 15)] = acc[i];
}

// ---------------------------------------------------------------------------
// network definition
// ---------------------------------------------------------------------------
struct LayerDef {
  int level, cin_a, cin_b, cout, ks, relu;
};
// c1a c1b | c2a c2b | c3a c3b | c4a c4b | d3a d3b | d2a d2b | d1a d1b | out
static const LayerDef kLayers[15] = {
    {0, 4, 0, 16, 3, 1},  {0, 16, 0, 16, 3, 1},  {1, 16, 0, 32, 3, 1}, {1, 32, 0, 32, 3, 1},
    {2, 32, 0, 64, 3, 1}, {2, 64, 0, 64, 3, 1},  {3, 64, 0, 64, 3, 1}, {3, 64, 0, 64, 3, 1},
    {2, 64, 64, 64, 3, 1}, {2, 64, 0, 64, 3, 1}, {1, 64, 32, 32, 3, 1}, {1, 32, 0, 32, 3, 1},
    {0, 32, 16, 16, 3, 1}, {0, 16, 0, 16, 3, 1}, {0, 16, 0, 1, 1, 0}};
constexpr int kNL = 15;

inline long wcount(const LayerDef& L) { return (long)L.ks * L.ks * L.cout * (L.cin_a + L.cin_b); }

}  // namespace nn
}  // namespace ins2d

using namespace ins2d;
using namespace ins2d::nn;

struct ins2d_nn {
  std::string err;
  int H = 0, W = 0;
  int Hl[4], Wl[4];
  long nparam = 0;
  std::vector<long> woff, boff;   // offsets of each layer's weights/bias in the flat vector
  std::vector<void*> allocs;
  float *param = nullptr, *grad = nullptr, *m = nullptr, *v = nullptr;
  uint16_t *wbf = nullptr, *wtbf = nullptr;   // bf16 weights and flipped weights
  uint16_t* act[kNL];                          // layer outputs (bf16)
  float* out = nullptr;                        // network output (fp32, 1 channel)
  uint16_t* x0 = nullptr;                      // input (4 channels)
  uint16_t *pool1 = nullptr, *pool2 = nullptr, *pool3 = nullptr, *up3 = nullptr, *up2 = nullptr, *up1 = nullptr;
  float *g0 = nullptr, *g1 = nullptr, *gs1 = nullptr, *gs2 = nullptr, *gs3 = nullptr, *gtmp = nullptr;
  double* scale = nullptr;   // rms of the current input
  double* sq = nullptr;      // partial sums of the input's squares
  double* tsq = nullptr;     // partial sums of the target's squares (relative loss)
  double* cscale = nullptr;  // rms of the two sources of a combined pair
  double* loss = nullptr;
  double out_scale = 1.0;    // p = out_scale * rms(f) * y
  double* data = nullptr;    // training pairs: [k][0] = f, [k][1] = p, [k][2] = xc (H*W each)
  int data_count = 0;
  long step = 0;
  bool weights_ready = false;
};

namespace {

#define NNCHK(x)                                                                                 \
  do {                                                                                           \
    hipError_t e_ = (x);                                                                         \
    if (e_ != hipSuccess) throw std::runtime_error(std::string(#x ": ") + hipGetErrorString(e_)); \
  } while (0)

template <typename T>
T* nalloc(ins2d_nn* n, size_t count) {
  void* p = nullptr;
  NNCHK(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)));
  n->allocs.push_back(p);
  return (T*)p;
}

inline unsigned nb(long n) { return (unsigned)std::min<long>((n + 255) / 256, 8192); }

// tile shape per launch: the largest pixel and channel tiles that still give
// about 1024 blocks (the coarse levels of the network are small grids)
template <int NT, int SK>
void launch_conv_nt(hipStream_t st, const ConvArgs& A, int mt) {
  const dim3 g((A.W + 16 * mt - 1) / (16 * mt), (A.H + CV_TY - 1) / CV_TY, (A.cout + 16 * NT - 1) / (16 * NT));
  if (mt == 4) hipLaunchKernelGGL((k_conv<NT, 4, SK>), g, dim3(256), 0, st, A);
  else if (mt == 2) hipLaunchKernelGGL((k_conv<NT, 2, SK>), g, dim3(256), 0, st, A);
  else hipLaunchKernelGGL((k_conv<NT, 1, SK>), g, dim3(256), 0, st, A);
}
template <int SK>
void launch_conv_sk(hipStream_t st, const ConvArgs& A, int nt, int mt) {
  if (nt == 1) launch_conv_nt<1, SK>(st, A, mt);
  else if (nt == 2) launch_conv_nt<2, SK>(st, A, mt);
  else launch_conv_nt<4, SK>(st, A, mt);
}
int staging_kind(const ConvArgs& A) {
  if (A.a.mode == SRC_F64) return SK_F64;
  auto bf = [](const CvSrc& s) { return s.c == 0 || (!s.fp32 && (s.c & 3) == 0); };
  if (!bf(A.a) || !bf(A.b)) return SK_ANY;
  if (A.a.mode == SRC_POOL && A.b.c == 0) return SK_POOL;
  if (A.a.mode != SRC_POOL && (A.b.c == 0 || A.b.mode != SRC_POOL)) return SK_BF16;
  return SK_ANY;
}
void launch_conv(hipStream_t st, const ConvArgs& A) {
  const long rows = (A.H + CV_TY - 1) / CV_TY;
  int nt = std::min(4, (A.cout + 15) / 16);
  if (nt == 3) nt = 4;
  int mt = 4;
  auto blocks = [&]() { return ((A.W + 16 * mt - 1) / (16 * mt)) * rows * ((A.cout + 16 * nt - 1) / (16 * nt)); };
  if (A.w1) nt = 1;   // the fused output layer reads all 16 channels of a pixel in one block
  while (blocks() < 1024 && nt > 1) nt /= 2;
  while (blocks() < 1024 && mt > 1) mt /= 2;
  switch (staging_kind(A)) {
    case SK_F64: launch_conv_sk<SK_F64>(st, A, nt, mt); break;
    case SK_BF16: launch_conv_sk<SK_BF16>(st, A, nt, mt); break;
    case SK_POOL: launch_conv_sk<SK_POOL>(st, A, nt, mt); break;
    default: launch_conv_sk<SK_ANY>(st, A, nt, mt); break;
  }
}

ConvArgs layer_args(ins2d_nn* N, int l, CvSrc a, CvSrc b, void* out, int out_fp32, const uint16_t* w, const float* bias,
                    int relu) {
  const LayerDef& L = kLayers[l];
  ConvArgs A{};
  A.a = a;
  A.b = b;
  A.cin = L.cin_a + L.cin_b;
  A.H = N->Hl[L.level];
  A.W = N->Wl[L.level];
  A.cout = L.cout;
  A.ks = L.ks;
  A.w = w;
  A.bias = bias;
  A.out = out;
  A.out_fp32 = out_fp32;
  A.relu = relu;
  return A;
}

void conv(hipStream_t st, ins2d_nn* N, int l, CvSrc a, CvSrc b, void* out, int out_fp32, const uint16_t* w,
          const float* bias, int relu) {
  launch_conv(st, layer_args(N, l, a, b, out, out_fp32, w, bias, relu));
}

CvSrc bfin(const uint16_t* p, int c, int mode = SRC_DIRECT) { return CvSrc{p, c, 0, mode, 1.0}; }
CvSrc f32in(const float* p, int c) { return CvSrc{p, c, 1, SRC_DIRECT, 1.0}; }
CvSrc f64in(const double* p, double div) { return CvSrc{p, p ? 1 : 0, 0, SRC_F64, div}; }
CvSrc none() { return CvSrc{nullptr, 0, 0, SRC_DIRECT, 1.0}; }

void refresh_weights(hipStream_t st, ins2d_nn* N) {
  for (int l = 0; l < kNL; ++l) {
    const LayerDef& L = kLayers[l];
    const long n = wcount(L);
    hipLaunchKernelGGL(k_wflip, dim3(nb(n)), dim3(256), 0, st, N->param + N->woff[l], L.ks * L.ks, L.cout,
                       L.cin_a + L.cin_b, N->wtbf + N->woff[l], N->wbf + N->woff[l]);
  }
  N->weights_ready = true;
}

// forward pass on a prepared input tensor N->x0; output in N->out (fp32); the
// pooled and upsampled tensors are stored for the backward pass
void forward(hipStream_t st, ins2d_nn* N) {
  uint16_t** A = N->act;
  auto W = [&](int l) { return N->wbf + N->woff[l]; };
  auto B = [&](int l) { return N->param + N->boff[l]; };
  conv(st, N, 0, bfin(N->x0, 4), none(), A[0], 0, W(0), B(0), 1);
  conv(st, N, 1, bfin(A[0], 16), none(), A[1], 0, W(1), B(1), 1);   // s1
  hipLaunchKernelGGL(k_pool, dim3(nb((long)N->Hl[1] * N->Wl[1] * 16)), dim3(256), 0, st, A[1], N->Hl[0], N->Wl[0], 16, N->pool1);
  conv(st, N, 2, bfin(N->pool1, 16), none(), A[2], 0, W(2), B(2), 1);
  conv(st, N, 3, bfin(A[2], 32), none(), A[3], 0, W(3), B(3), 1);   // s2
  hipLaunchKernelGGL(k_pool, dim3(nb((long)N->Hl[2] * N->Wl[2] * 32)), dim3(256), 0, st, A[3], N->Hl[1], N->Wl[1], 32, N->pool2);
  conv(st, N, 4, bfin(N->pool2, 32), none(), A[4], 0, W(4), B(4), 1);
  conv(st, N, 5, bfin(A[4], 64), none(), A[5], 0, W(5), B(5), 1);   // s3
  hipLaunchKernelGGL(k_pool, dim3(nb((long)N->Hl[3] * N->Wl[3] * 64)), dim3(256), 0, st, A[5], N->Hl[2], N->Wl[2], 64, N->pool3);
  conv(st, N, 6, bfin(N->pool3, 64), none(), A[6], 0, W(6), B(6), 1);
  conv(st, N, 7, bfin(A[6], 64), none(), A[7], 0, W(7), B(7), 1);   // bottleneck
  hipLaunchKernelGGL(k_up, dim3(nb((long)N->Hl[2] * N->Wl[2] * 64)), dim3(256), 0, st, A[7], N->Hl[3], N->Wl[3], 64, N->up3);
  conv(st, N, 8, bfin(N->up3, 64), bfin(A[5], 64), A[8], 0, W(8), B(8), 1);
  conv(st, N, 9, bfin(A[8], 64), none(), A[9], 0, W(9), B(9), 1);
  hipLaunchKernelGGL(k_up, dim3(nb((long)N->Hl[1] * N->Wl[1] * 64)), dim3(256), 0, st, A[9], N->Hl[2], N->Wl[2], 64, N->up2);
  conv(st, N, 10, bfin(N->up2, 64), bfin(A[3], 32), A[10], 0, W(10), B(10), 1);
  conv(st, N, 11, bfin(A[10], 32), none(), A[11], 0, W(11), B(11), 1);
  hipLaunchKernelGGL(k_up, dim3(nb((long)N->Hl[0] * N->Wl[0] * 32)), dim3(256), 0, st, A[11], N->Hl[1], N->Wl[1], 32, N->up1);
  conv(st, N, 12, bfin(N->up1, 32), bfin(A[1], 16), A[12], 0, W(12), B(12), 1);
  conv(st, N, 13, bfin(A[12], 16), none(), A[13], 0, W(13), B(13), 1);
  conv(st, N, 14, bfin(A[13], 16), none(), N->out, 1, W(14), B(14), 0);
}

// inference: p = xc + out_scale s U(f / s, xc / (out_scale s)) with s = rms(f),
// 14 launches: the input scaling is done by the first convolution's staging,
// pooling and upsampling by the staging of the convolutions that read them, and
// the 1x1 output layer and the fp64 output by the epilogue of the last 3x3 one
void forward_fused(hipStream_t st, ins2d_nn* N, const double* f, const double* xc, double* p) {
  uint16_t** A = N->act;
  auto W = [&](int l) { return N->wbf + N->woff[l]; };
  auto B = [&](int l) { return N->param + N->boff[l]; };
  const long n = (long)N->H * N->W;
  hipLaunchKernelGGL(k_sumsq, dim3(kSq), dim3(256), 0, st, f, n, N->sq);
  ConvArgs a0 = layer_args(N, 0, f64in(f, 1.0), f64in(xc, N->out_scale), A[0], 0, W(0), B(0), 1);
  a0.sq = N->sq;
  a0.nsq = kSq;
  a0.npix_s = n;
  launch_conv(st, a0);
  conv(st, N, 1, bfin(A[0], 16), none(), A[1], 0, W(1), B(1), 1);
  conv(st, N, 2, bfin(A[1], 16, SRC_POOL), none(), A[2], 0, W(2), B(2), 1);
  conv(st, N, 3, bfin(A[2], 32), none(), A[3], 0, W(3), B(3), 1);
  conv(st, N, 4, bfin(A[3], 32, SRC_POOL), none(), A[4], 0, W(4), B(4), 1);
  conv(st, N, 5, bfin(A[4], 64), none(), A[5], 0, W(5), B(5), 1);
  conv(st, N, 6, bfin(A[5], 64, SRC_POOL), none(), A[6], 0, W(6), B(6), 1);
  conv(st, N, 7, bfin(A[6], 64), none(), A[7], 0, W(7), B(7), 1);
  conv(st, N, 8, bfin(A[7], 64, SRC_UP), bfin(A[5], 64), A[8], 0, W(8), B(8), 1);
  conv(st, N, 9, bfin(A[8], 64), none(), A[9], 0, W(9), B(9), 1);
  conv(st, N, 10, bfin(A[9], 64, SRC_UP), bfin(A[3], 32), A[10], 0, W(10), B(10), 1);
  conv(st, N, 11, bfin(A[10], 32), none(), A[11], 0, W(11), B(11), 1);
  conv(st, N, 12, bfin(A[11], 32, SRC_UP), bfin(A[1], 16), A[12], 0, W(12), B(12), 1);
  ConvArgs a13 = layer_args(N, 13, bfin(A[12], 16), none(), nullptr, 0, W(13), B(13), 1);
  a13.sq = N->sq;
  a13.nsq = kSq;
  a13.npix_s = n;
  a13.w1 = W(14);
  a13.b1 = B(14);
  a13.out64 = p;
  a13.base = xc;
  a13.oscale = N->out_scale;
  launch_conv(st, a13);
}

// backward pass for the loss gradient dy (N->g0, fp32, 1 channel at level 0);
// accumulates parameter gradients into N->grad
void backward(hipStream_t st, ins2d_nn* N) {
  uint16_t** A = N->act;
  auto WT = [&](int l) { return N->wtbf + N->woff[l]; };
  auto GW = [&](int l) { return N->grad + N->woff[l]; };
  auto GB = [&](int l) { return N->grad + N->boff[l]; };
  const long npx[4] = {(long)N->Hl[0] * N->Wl[0], (long)N->Hl[1] * N->Wl[1], (long)N->Hl[2] * N->Wl[2],
                       (long)N->Hl[3] * N->Wl[3]};
  // dW, db of layer l given output gradient g (already through the activation)
  auto wgrad = [&](int l, CvSrc a, CvSrc b, const float* g) {
    const LayerDef& L = kLayers[l];
    const int lv = L.level;
    const int chunk = 4096;
    const int tci = (L.cin_a + L.cin_b + 15) / 16, tco = (L.cout + 15) / 16;
    hipLaunchKernelGGL(k_conv_dw, dim3((unsigned)((npx[lv] + chunk - 1) / chunk), L.ks * L.ks, tci * tco), dim3(64), 0,
                       st, a, b, g, N->Hl[lv], N->Wl[lv], L.cout, L.ks, chunk, GW(l));
    hipLaunchKernelGGL(k_bias_grad, dim3(std::min<long>(64, (npx[lv] + 255) / 256), L.cout), dim3(256), 0, st, g,
                       npx[lv], L.cout, GB(l));
  };
  // input gradient of layer l: din = conv(g, flipped W), no bias, fp32
  auto dgrad = [&](int l, const float* g, float* din) {
    const LayerDef& L = kLayers[l];
    ConvArgs C{};
    C.a = f32in(g, L.cout);
    C.b = none();
    C.cin = L.cout;
    C.H = N->Hl[L.level];
    C.W = N->Wl[L.level];
    C.cout = L.cin_a + L.cin_b;
    C.ks = L.ks;
    C.w = WT(l);
    C.out = din;
    C.out_fp32 = 1;
    launch_conv(st, C);
  };
  auto relub = [&](float* g, const uint16_t* out, long n) {
    hipLaunchKernelGGL(k_relu_bwd, dim3(nb(n)), dim3(256), 0, st, g, out, n, g);
  };
  float* g0 = N->g0;     // level-0 scratch (up to 48 channels)
  float* g1 = N->g1;
  // out layer (linear, 16 -> 1, ks 1)
  wgrad(14, bfin(A[13], 16), none(), g0);
  dgrad(14, g0, g1);                         // g1: d A13 (16 ch)
  relub(g1, A[13], npx[0] * 16);
  wgrad(13, bfin(A[12], 16), none(), g1);
  dgrad(13, g1, g0);                         // d A12 (16)
  relub(g0, A[12], npx[0] * 16);
  wgrad(12, bfin(N->up1, 32), bfin(A[1], 16), g0);
  dgrad(12, g0, g1);                         // d [up1, s1] (48 ch)
  // split: gtmp = d up1 (32 ch @ L0), gs1 += d s1 (16 ch @ L0)
  hipLaunchKernelGGL(k_zero_f, dim3(nb(npx[0] * 16)), dim3(256), 0, st, N->gs1, npx[0] * 16);
  hipLaunchKernelGGL(k_split, dim3(nb(npx[0] * 48)), dim3(256), 0, st, g1, npx[0], 32, 16, N->gtmp, N->gs1);
  // up1 backward -> d A11 (32 ch @ L1)
  hipLaunchKernelGGL(k_up_bwd, dim3(nb(npx[1] * 32)), dim3(256), 0, st, N->gtmp, N->Hl[1], N->Wl[1], 32, g0);
  relub(g0, A[11], npx[1] * 32);
  wgrad(11, bfin(A[10], 32), none(), g0);
  dgrad(11, g0, g1);                         // d A10 (32 @ L1)
  relub(g1, A[10], npx[1] * 32);
  wgrad(10, bfin(N->up2, 64), bfin(A[3], 32), g1);
  dgrad(10, g1, g0);                         // d [up2, s2] (96 @ L1)
  hipLaunchKernelGGL(k_zero_f, dim3(nb(npx[1] * 32)), dim3(256), 0, st, N->gs2, npx[1] * 32);
  hipLaunchKernelGGL(k_split, dim3(nb(npx[1] * 96)), dim3(256), 0, st, g0, npx[1], 64, 32, N->gtmp, N->gs2);
  hipLaunchKernelGGL(k_up_bwd, dim3(nb(npx[2] * 64)), dim3(256), 0, st, N->gtmp, N->Hl[2], N->Wl[2], 64, g1);
  relub(g1, A[9], npx[2] * 64);              // d A9 (64 @ L2)
  wgrad(9, bfin(A[8], 64), none(), g1);
  dgrad(9, g1, g0);
  relub(g0, A[8], npx[2] * 64);
  wgrad(8, bfin(N->up3, 64), bfin(A[5], 64), g0);
  dgrad(8, g0, g1);                          // d [up3, s3] (128 @ L2)
  hipLaunchKernelGGL(k_zero_f, dim3(nb(npx[2] * 64)), dim3(256), 0, st, N->gs3, npx[2] * 64);
  hipLaunchKernelGGL(k_split, dim3(nb(npx[2] * 128)), dim3(256), 0, st, g1, npx[2], 64, 64, N->gtmp, N->gs3);
  hipLaunchKernelGGL(k_up_bwd, dim3(nb(npx[3] * 64)), dim3(256), 0, st, N->gtmp, N->Hl[3], N->Wl[3], 64, g0);
  relub(g0, A[7], npx[3] * 64);              // d A7 (64 @ L3)
  wgrad(7, bfin(A[6], 64), none(), g0);
  dgrad(7, g0, g1);
  relub(g1, A[6], npx[3] * 64);
  wgrad(6, bfin(N->pool3, 64), none(), g1);
  dgrad(6, g1, g0);                          // d pool3 (64 @ L3)
  // pool3 backward adds into d s3 (gs3, 64 @ L2)
  hipLaunchKernelGGL(k_pool_bwd, dim3(nb(npx[2] * 64)), dim3(256), 0, st, g0, N->Hl[2], N->Wl[2], 64, N->gs3);
  relub(N->gs3, A[5], npx[2] * 64);          // d A5
  wgrad(5, bfin(A[4], 64), none(), N->gs3);
  dgrad(5, N->gs3, g0);
  relub(g0, A[4], npx[2] * 64);
  wgrad(4, bfin(N->pool2, 32), none(), g0);
  dgrad(4, g0, g1);                          // d pool2 (32 @ L2)
  hipLaunchKernelGGL(k_pool_bwd, dim3(nb(npx[1] * 32)), dim3(256), 0, st, g1, N->Hl[1], N->Wl[1], 32, N->gs2);
  relub(N->gs2, A[3], npx[1] * 32);
  wgrad(3, bfin(A[2], 32), none(), N->gs2);
  dgrad(3, N->gs2, g0);
  relub(g0, A[2], npx[1] * 32);
  wgrad(2, bfin(N->pool1, 16), none(), g0);
  dgrad(2, g0, g1);                          // d pool1 (16 @ L1)
  hipLaunchKernelGGL(k_pool_bwd, dim3(nb(npx[0] * 16)), dim3(256), 0, st, g1, N->Hl[0], N->Wl[0], 16, N->gs1);
  relub(N->gs1, A[1], npx[0] * 16);
  wgrad(1, bfin(A[0], 16), none(), N->gs1);
  dgrad(1, N->gs1, g0);
  relub(g0, A[0], npx[0] * 16);
  wgrad(0, bfin(N->x0, 4), none(), g0);
}

void prepare_input(hipStream_t st, ins2d_nn* N, const double* f, const double* xc) {
  const long n = (long)N->H * N->W;
  hipLaunchKernelGGL(k_sumsq, dim3(kSq), dim3(256), 0, st, f, n, N->sq);
  hipLaunchKernelGGL(k_rms, dim3(1), dim3(256), 0, st, N->sq, kSq, n, N->scale);
  hipLaunchKernelGGL(k_input, dim3(nb(n)), dim3(256), 0, st, f, xc, n, N->scale, N->out_scale, N->x0);
}

int nn_fail(ins2d_nn* N, const std::exception& e) {
  if (N) N->err = e.what();
  return INS2D_E_HIP;
}

std::string g_nn_err;

}  // namespace

extern "C" {

int ins2d_nn_create(int H, int W, ins2d_nn** out) {
  if (!out || H < 16 || W < 16 || H % 8 || W % 8) {
    g_nn_err = "network grid must be at least 16x16 with sides divisible by 8";
    return INS2D_E_ARG;
  }
  auto N = std::make_unique<ins2d_nn>();
  try {
    N->H = H;
    N->W = W;
    for (int l = 0; l < 4; ++l) { N->Hl[l] = H >> l; N->Wl[l] = W >> l; }
    long off = 0;
    for (int l = 0; l < kNL; ++l) {
      N->woff.push_back(off);
      off += wcount(kLayers[l]);
      N->boff.push_back(off);
      off += kLayers[l].cout;
    }
    N->nparam = off;
    N->param = nalloc<float>(N.get(), off);
    N->grad = nalloc<float>(N.get(), off);
    N->m = nalloc<float>(N.get(), off);
    N->v = nalloc<float>(N.get(), off);
    N->wbf = nalloc<uint16_t>(N.get(), off);
    N->wtbf = nalloc<uint16_t>(N.get(), off);
    const long p0 = (long)H * W;
    for (int l = 0; l < kNL; ++l) {
      const LayerDef& L = kLayers[l];
      N->act[l] = nalloc<uint16_t>(N.get(), (p0 >> (2 * L.level)) * L.cout);
    }
    N->out = nalloc<float>(N.get(), p0);
    N->x0 = nalloc<uint16_t>(N.get(), p0 * 4);
    N->pool1 = nalloc<uint16_t>(N.get(), (p0 >> 2) * 16);
    N->pool2 = nalloc<uint16_t>(N.get(), (p0 >> 4) * 32);
    N->pool3 = nalloc<uint16_t>(N.get(), (p0 >> 6) * 64);
    N->up3 = nalloc<uint16_t>(N.get(), (p0 >> 4) * 64);
    N->up2 = nalloc<uint16_t>(N.get(), (p0 >> 2) * 64);
    N->up1 = nalloc<uint16_t>(N.get(), p0 * 32);
    N->g0 = nalloc<float>(N.get(), p0 * 48);
    N->g1 = nalloc<float>(N.get(), p0 * 48);
    N->gtmp = nalloc<float>(N.get(), p0 * 32);
    N->gs1 = nalloc<float>(N.get(), p0 * 16);
    N->gs2 = nalloc<float>(N.get(), (p0 >> 2) * 32);
    N->gs3 = nalloc<float>(N.get(), (p0 >> 4) * 64);
    N->scale = nalloc<double>(N.get(), 1);
    N->sq = nalloc<double>(N.get(), kSq);
    N->tsq = nalloc<double>(N.get(), kSq);
    N->cscale = nalloc<double>(N.get(), 2);
    N->loss = nalloc<double>(N.get(), 1);
    NNCHK(hipMemset(N->param, 0, sizeof(float) * off));
    NNCHK(hipMemset(N->m, 0, sizeof(float) * off));
    NNCHK(hipMemset(N->v, 0, sizeof(float) * off));
    NNCHK(hipDeviceSynchronize());
  } catch (const std::exception& e) {
    g_nn_err = e.what();
    for (void* p : N->allocs) (void)hipFree(p);
    return INS2D_E_HIP;
  }
  *out = N.release();
  return INS2D_OK;
}

void ins2d_nn_destroy(ins2d_nn* N) {
  if (!N) return;
  (void)hipDeviceSynchronize();
  for (void* p : N->allocs) (void)hipFree(p);
  if (N->data) (void)hipFree(N->data);
  delete N;
}

const char* ins2d_nn_last_error(const ins2d_nn* N) { return N ? N->err.c_str() : g_nn_err.c_str(); }

long ins2d_nn_num_params(const ins2d_nn* N) { return N ? N->nparam : 0; }

int ins2d_nn_shape(const ins2d_nn* N, int* H, int* W) {
  if (!N) return INS2D_E_ARG;
  if (H) *H = N->H;
  if (W) *W = N->W;
  return INS2D_OK;
}

int ins2d_nn_set_params(ins2d_nn* N, const float* params, double out_scale) {
  if (!N || !params) return INS2D_E_ARG;
  try {
    NNCHK(hipMemcpy(N->param, params, sizeof(float) * N->nparam, hipMemcpyHostToDevice));
    N->out_scale = out_scale;
    refresh_weights(nullptr, N);
    NNCHK(hipDeviceSynchronize());
  } catch (const std::exception& e) {
    return nn_fail(N, e);
  }
  return INS2D_OK;
}

// He-normal weights (std sqrt(2 / fan_in) before a ReLU, sqrt(1 / fan_in)
// for the linear output layer), zero biases, fresh Adam state
int ins2d_nn_init(ins2d_nn* N, unsigned long long seed, double out_scale) {
  if (!N) return INS2D_E_ARG;
  std::vector<float> th(N->nparam, 0.f);
  uint64_t st = seed * 0x9E3779B97F4A7C15ull + 0x2545F4914F6CDD1Dull;
  auto next = [&]() {   // splitmix64
    uint64_t z = (st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  auto unif = [&]() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); };
  for (int l = 0; l < kNL; ++l) {
    const LayerDef& L = kLayers[l];
    const double fan_in = (double)L.ks * L.ks * (L.cin_a + L.cin_b);
    const double sd = std::sqrt((L.relu ? 2.0 : 1.0) / fan_in);
    const long n = wcount(L);
    for (long k = 0; k < n; ++k) {
      const double g = std::sqrt(-2.0 * std::log(unif())) * std::cos(6.283185307179586 * unif());   // Box-Muller
      th[N->woff[l] + k] = (float)(sd * g);
    }
  }
  try {
    NNCHK(hipMemset(N->m, 0, sizeof(float) * N->nparam));
    NNCHK(hipMemset(N->v, 0, sizeof(float) * N->nparam));
  } catch (const std::exception& e) {
    return nn_fail(N, e);
  }
  N->step = 0;
  return ins2d_nn_set_params(N, th.data(), out_scale);
}

int ins2d_nn_get_params(ins2d_nn* N, float* params) {
  if (!N || !params) return INS2D_E_ARG;
  try {
    NNCHK(hipMemcpy(params, N->param, sizeof(float) * N->nparam, hipMemcpyDeviceToHost));
  } catch (const std::exception& e) {
    return nn_fail(N, e);
  }
  return INS2D_OK;
}

// p = xc + out_scale * rms(f) * U(f / rms(f), xc / (out_scale rms(f))); f,
// xc (optional, may be p itself), p device arrays (H*W doubles)
int ins2d_nn_forward_dev(ins2d_nn* N, const double* f, const double* xc, double* p, void* stream) {
  if (!N || !f || !p) return INS2D_E_ARG;
  hipStream_t st = (hipStream_t)stream;
  try {
    forward_fused(st, N, f, xc, p);
    NNCHK(hipGetLastError());
  } catch (const std::exception& e) {
    return nn_fail(N, e);
  }
  return INS2D_OK;
}

// the same output through the training forward (stored input tensor, pooled
// and upsampled tensors, separate output layer)
int ins2d_nn_forward_unfused_dev(ins2d_nn* N, const double* f, const double* xc, double* p, void* stream) {
  if (!N || !f || !p) return INS2D_E_ARG;
  hipStream_t st = (hipStream_t)stream;
  try {
    const long n = (long)N->H * N->W;
    prepare_input(st, N, f, xc);
    forward(st, N);
    hipLaunchKernelGGL(k_output, dim3(nb(n)), dim3(256), 0, st, N->out, xc, n, N->scale, N->out_scale, p);
    NNCHK(hipGetLastError());
  } catch (const std::exception& e) {
    return nn_fail(N, e);
  }
  return INS2D_OK;
}

// host-array forward (tests); xc optional
static int nn_forward_host(ins2d_nn* N, const double* f, const double* xc, double* p, int unfused) {
  if (!N || !f || !p) return INS2D_E_ARG;
  try {
    const long n = (long)N->H * N->W;
    double *df = nullptr, *dp = nullptr, *dx = nullptr;
    NNCHK(hipMalloc(&df, n * sizeof(double)));
    NNCHK(hipMalloc(&dp, n * sizeof(double)));
    NNCHK(hipMemcpy(df, f, n * sizeof(double), hipMemcpyHostToDevice));
    if (xc) {
      NNCHK(hipMalloc(&dx, n * sizeof(double)));
      NNCHK(hipMemcpy(dx, xc, n * sizeof(double), hipMemcpyHostToDevice));
    }
    int rc = unfused ? ins2d_nn_forward_unfused_dev(N, df, dx, dp, nullptr) : ins2d_nn_forward_dev(N, df, dx, dp, nullptr);
    NNCHK(hipDeviceSynchronize());
    NNCHK(hipMemcpy(p, dp, n * sizeof(double), hipMemcpyDeviceToHost));
    (void)hipFree(df);
    (void)hipFree(dp);
    if (dx) (void)hipFree(dx);
    return rc;
  } catch (const std::exception& e) {
    return nn_fail(N, e);
  }
}

int ins2d_nn_forward(ins2d_nn* N, const double* f, const double* xc, double* p) {
  return nn_forward_host(N, f, xc, p, 0);
}
int ins2d_nn_forward_unfused(ins2d_nn* N, const double* f, const double* xc, double* p) {
  return nn_forward_host(N, f, xc, p, 1);
}

// loss and parameter gradients (accumulated over a batch of device samples
// f[b], p[b]); mode 0: gradient only (into grad, zeroed first); mode 1: also
// one Adam step with learning rate lr.  loss_out: mean loss over the batch.
int ins2d_nn_train_step(ins2d_nn* N, const double* const* f, const double* const* p, const double* const* xc,
                        int batch, float lr, int mode, double* loss_out, void* stream) {
  if (!N || !f || !p || batch < 1) return INS2D_E_ARG;
  hipStream_t st = (hipStream_t)stream;
  try {
    const long n = (long)N->H * N->W;
    hipLaunchKernelGGL(k_zero_f, dim3(nb(N->nparam)), dim3(256), 0, st, N->grad, N->nparam);
    hipLaunchKernelGGL(k_zero_d, dim3(1), dim3(64), 0, st, N->loss, 1L);
    for (int b = 0; b < batch; ++b) {
      const double* x = xc ? xc[b] : nullptr;
      prepare_input(st, N, f[b], x);
      forward(st, N);
      if (mode & 2) hipLaunchKernelGGL(k_sumsq_diff, dim3(kSq), dim3(256), 0, st, p[b], x, n, N->tsq);
      hipLaunchKernelGGL(k_mse, dim3(nb(n)), dim3(256), 0, st, N->out, p[b], x, n, N->scale, N->out_scale,
                         (mode & 2) ? (const double*)N->tsq : nullptr, N->g0,
                         N->loss);
      backward(st, N);
    }
    if (mode & 1) {
      N->step += 1;
      const float b1 = 0.9f, b2 = 0.999f;
      const float c1 = 1.f - powf(b1, (float)N->step), c2 = 1.f - powf(b2, (float)N->step);
      hipLaunchKernelGGL(k_adam, dim3(nb(N->nparam)), dim3(256), 0, st, N->param, N->grad, N->m, N->v, N->nparam, lr,
                         b1, b2, 1e-8f, c1, c2, 1.0f / batch);
      refresh_weights(st, N);
    }
    double l = 0.0;
    NNCHK(hipMemcpyAsync(&l, N->loss, sizeof(double), hipMemcpyDeviceToHost, st));
    NNCHK(hipStreamSynchronize(st));
    if (loss_out) *loss_out = l / batch;
  } catch (const std::exception& e) {
    return nn_fail(N, e);
  }
  return INS2D_OK;
}

int ins2d_nn_data_reserve(ins2d_nn* N, int count) {
  if (!N || count < 0) return INS2D_E_ARG;
  try {
    NNCHK(hipDeviceSynchronize());
    if (N->data) NNCHK(hipFree(N->data));
    N->data = nullptr;
    N->data_count = 0;
    if (count > 0) {
      NNCHK(hipMalloc(&N->data, sizeof(double) * 3 * (size_t)count * N->H * N->W));
      N->data_count = count;
    }
  } catch (const std::exception& e) {
    return nn_fail(N, e);
  }
  return INS2D_OK;
}

int ins2d_nn_data_set(ins2d_nn* N, int k, const double* f, const double* p, const double* xc) {
  if (!N || !f || !p) return INS2D_E_ARG;
  if (k < 0 || k >= N->data_count) {
    N->err = "training pair index out of range";
    return INS2D_E_ARG;
  }
  try {
    const size_t n = (size_t)N->H * N->W;
    double* d = N->data + 3 * n * k;
    NNCHK(hipMemcpy(d, f, n * sizeof(double), hipMemcpyHostToDevice));
    NNCHK(hipMemcpy(d + n, p, n * sizeof(double), hipMemcpyHostToDevice));
    if (xc) NNCHK(hipMemcpy(d + 2 * n, xc, n * sizeof(double), hipMemcpyHostToDevice));
    else NNCHK(hipMemset(d + 2 * n, 0, n * sizeof(double)));
  } catch (const std::exception& e) {
    return nn_fail(N, e);
  }
  return INS2D_OK;
}

double* ins2d_nn_data_ptr(ins2d_nn* N, int k, int which) {
  if (!N || k < 0 || k >= N->data_count || which < 0 || which > 2) return nullptr;
  const size_t n = (size_t)N->H * N->W;
  return N->data + 3 * n * k + n * which;
}

int ins2d_nn_data_get(ins2d_nn* N, int k, double* f, double* p, double* xc) {
  if (!N) return INS2D_E_ARG;
  if (k < 0 || k >= N->data_count) {
    N->err = "training pair index out of range";
    return INS2D_E_ARG;
  }
  try {
    const size_t n = (size_t)N->H * N->W;
    const double* d = N->data + 3 * n * k;
    NNCHK(hipDeviceSynchronize());
    if (f) NNCHK(hipMemcpy(f, d, n * sizeof(double), hipMemcpyDeviceToHost));
    if (p) NNCHK(hipMemcpy(p, d + n, n * sizeof(double), hipMemcpyDeviceToHost));
    if (xc) NNCHK(hipMemcpy(xc, d + 2 * n, n * sizeof(double), hipMemcpyDeviceToHost));
  } catch (const std::exception& e) {
    return nn_fail(N, e);
  }
  return INS2D_OK;
}

int ins2d_nn_data_combine(ins2d_nn* N, int dst, int i, int j, double a, double b) {
  if (!N) return INS2D_E_ARG;
  if (dst < 0 || dst >= N->data_count || i < 0 || i >= N->data_count || j < 0 || j >= N->data_count || dst == i ||
      dst == j) {
    N->err = "training pair index out of range (or the destination is a source)";
    return INS2D_E_ARG;
  }
  try {
    const long n = (long)N->H * N->W;
    const double* si = N->data + 3 * n * i;
    const double* sj = N->data + 3 * n * j;
    hipLaunchKernelGGL(k_sumsq, dim3(kSq), dim3(256), 0, nullptr, si, n, N->sq);
    hipLaunchKernelGGL(k_rms, dim3(1), dim3(256), 0, nullptr, N->sq, kSq, n, N->cscale);
    hipLaunchKernelGGL(k_sumsq, dim3(kSq), dim3(256), 0, nullptr, sj, n, N->sq);
    hipLaunchKernelGGL(k_rms, dim3(1), dim3(256), 0, nullptr, N->sq, kSq, n, N->cscale + 1);
    hipLaunchKernelGGL(k_combine, dim3(nb(3 * n)), dim3(256), 0, nullptr, si, sj, 3 * n, (const double*)N->cscale, a,
                       b, N->data + 3 * n * dst);
    NNCHK(hipGetLastError());
  } catch (const std::exception& e) {
    return nn_fail(N, e);
  }
  return INS2D_OK;
}

int ins2d_nn_train(ins2d_nn* N, const int* idx, int batch, float lr, int mode, double* loss_out) {
  if (!N || !idx || batch < 1) return INS2D_E_ARG;
  const size_t n = (size_t)N->H * N->W;
  std::vector<const double*> f(batch), p(batch), xc(batch);
  for (int b = 0; b < batch; ++b) {
    if (idx[b] < 0 || idx[b] >= N->data_count) {
      N->err = "training pair index out of range";
      return INS2D_E_ARG;
    }
    f[b] = N->data + 3 * n * idx[b];
    p[b] = f[b] + n;
    xc[b] = f[b] + 2 * n;
  }
  return ins2d_nn_train_step(N, f.data(), p.data(), xc.data(), batch, lr, mode, loss_out, nullptr);
}

int ins2d_nn_get_grad(ins2d_nn* N, float* grad) {
  if (!N || !grad) return INS2D_E_ARG;
  try {
    NNCHK(hipMemcpy(grad, N->grad, sizeof(float) * N->nparam, hipMemcpyDeviceToHost));
  } catch (const std::exception& e) {
    return nn_fail(N, e);
  }
  return INS2D_OK;
}

// One convolution layer's kernels on host data (FP32 arrays rounded to BF16
// where the kernels read BF16): forward out = conv(x) (+ bias, ReLU), and for
// an output gradient g: din = conv^T(g), dw, db.  xa: H*W*ca, xb: H*W*cb,
// w: ks*ks*cout*(ca+cb), g/out: H*W*cout, din: H*W*(ca+cb).
int ins2d_nn_layer_check(int H, int W, int ca, int cb, int cout, int ks, int relu, const float* xa, const float* xb,
                         const float* w, const float* bias, const float* g, float* out, float* din, float* dw,
                         float* db) {
  if (H < 1 || W < 1 || ca < 1 || cb < 0 || cout < 1 || (ks != 1 && ks != 3) || !xa || (cb && !xb) || !w || !bias ||
      !g || !out || !din || !dw || !db)
    return INS2D_E_ARG;
  const long np_ = (long)H * W, cin = ca + cb, nw = (long)ks * ks * cout * cin;
  std::vector<void*> mem;
  auto dev = [&](size_t bytes) {
    void* p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(bytes, 4)) != hipSuccess) throw std::runtime_error("hipMalloc");
    mem.push_back(p);
    return p;
  };
  int rc = INS2D_OK;
  try {
    float* dxa = (float*)dev(np_ * ca * 4);
    float* dxb = (float*)dev(np_ * std::max(cb, 1) * 4);
    float* dwf = (float*)dev(nw * 4);
    uint16_t* dwb = (uint16_t*)dev(nw * 2);
    uint16_t* dwt = (uint16_t*)dev(nw * 2);
    float* dbias = (float*)dev(cout * 4);
    float* dg = (float*)dev(np_ * cout * 4);
    float* dout = (float*)dev(np_ * cout * 4);
    float* ddin = (float*)dev(np_ * cin * 4);
    float* ddw = (float*)dev(nw * 4);
    float* ddb = (float*)dev(cout * 4);
    NNCHK(hipMemcpy(dxa, xa, np_ * ca * 4, hipMemcpyHostToDevice));
    if (cb) NNCHK(hipMemcpy(dxb, xb, np_ * cb * 4, hipMemcpyHostToDevice));
    NNCHK(hipMemcpy(dwf, w, nw * 4, hipMemcpyHostToDevice));
    NNCHK(hipMemcpy(dbias, bias, cout * 4, hipMemcpyHostToDevice));
    NNCHK(hipMemcpy(dg, g, np_ * cout * 4, hipMemcpyHostToDevice));
    NNCHK(hipMemset(ddw, 0, nw * 4));
    NNCHK(hipMemset(ddb, 0, cout * 4));
    hipLaunchKernelGGL(k_wflip, dim3(nb(nw)), dim3(256), 0, nullptr, dwf, ks * ks, cout, (int)cin, dwt, dwb);
    ConvArgs A{};
    A.a = f32in(dxa, ca);
    A.b = cb ? f32in(dxb, cb) : none();
    A.cin = (int)cin;
    A.H = H;
    A.W = W;
    A.cout = cout;
    A.ks = ks;
    A.w = dwb;
    A.bias = dbias;
    A.out = dout;
    A.out_fp32 = 1;
    A.relu = relu;
    launch_conv(nullptr, A);
    ConvArgs C{};
    C.a = f32in(dg, cout);
    C.b = none();
    C.cin = cout;
    C.H = H;
    C.W = W;
    C.cout = (int)cin;
    C.ks = ks;
    C.w = dwt;
    C.out = ddin;
    C.out_fp32 = 1;
    launch_conv(nullptr, C);
    const int chunk = 4096, tci = ((int)cin + 15) / 16, tco = (cout + 15) / 16;
    hipLaunchKernelGGL(k_conv_dw, dim3((unsigned)((np_ + chunk - 1) / chunk), ks * ks, tci * tco), dim3(64), 0, nullptr,
                       A.a, A.b, dg, H, W, cout, ks, chunk, ddw);
    hipLaunchKernelGGL(k_bias_grad, dim3(std::min<long>(64, (np_ + 255) / 256), cout), dim3(256), 0, nullptr, dg, np_,
                       cout, ddb);
    NNCHK(hipGetLastError());
    NNCHK(hipDeviceSynchronize());
    NNCHK(hipMemcpy(out, dout, np_ * cout * 4, hipMemcpyDeviceToHost));
    NNCHK(hipMemcpy(din, ddin, np_ * cin * 4, hipMemcpyDeviceToHost));
    NNCHK(hipMemcpy(dw, ddw, nw * 4, hipMemcpyDeviceToHost));
    NNCHK(hipMemcpy(db, ddb, cout * 4, hipMemcpyDeviceToHost));
  } catch (const std::exception& e) {
    g_nn_err = e.what();
    rc = INS2D_E_HIP;
  }
  for (void* p : mem) (void)hipFree(p);
  return rc;
}

int ins2d_nn_mfma_probe(const float* A, const float* B, float* D) {
  if (!A || !B || !D) return INS2D_E_ARG;
  float *dA = nullptr, *dB = nullptr, *dD = nullptr;
  if (hipMalloc(&dA, 256 * 4) != hipSuccess || hipMalloc(&dB, 256 * 4) != hipSuccess ||
      hipMalloc(&dD, 256 * 4) != hipSuccess)
    return INS2D_E_HIP;
  (void)hipMemcpy(dA, A, 256 * 4, hipMemcpyHostToDevice);
  (void)hipMemcpy(dB, B, 256 * 4, hipMemcpyHostToDevice);
  hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, nullptr, dA, dB, dD);
  hipError_t e = hipMemcpy(D, dD, 256 * 4, hipMemcpyDeviceToHost);
  (void)hipFree(dA);
  (void)hipFree(dB);
  (void)hipFree(dD);
  return e == hipSuccess ? INS2D_OK : INS2D_E_HIP;
}

}  // extern "C"
