// The periodic FFT pressure solve without a handle and without allocation:
// for callers that cannot call hipMalloc (work captured into HIP graphs,
// tight serving loops).  Twiddles, symbol and the half spectrum live in a
// file-scope device array of fixed size; each call fills what it needs and
// runs the register-transform solve of the handle-based path on the caller's
// stream.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>

#include "../../include/ins2d.h"
#include "device.hpp"
#include "launch.hpp"

namespace ins2d {
void set_handle_free_error(const std::string& msg);
bool fft_register_lengths(int Nx, int Ny);
}  // namespace ins2d

namespace {

using namespace ins2d;

// capacity: the half spectrum of up to 2^20 cells (Nh * Ny <= 525312 complex
// values, e.g. 1024 x 1024 or 4096 x 256) plus twiddles and symbols of
// lengths up to 4096
constexpr size_t kSpecCap = 2 * (size_t)513 * 1024;    // doubles
constexpr size_t kTableCap = 32768;                      // doubles
constexpr size_t kScratchDoubles = kSpecCap + kTableCap;
constexpr int kMaxDevices = 64;

// the solve's scratch memory on every GPU the library runs on (held until
// the process ends); host code reaches it only through the address that
// hipGetSymbolAddress returns on each device
__device__ __align__(256) double scratch[kScratchDoubles];

std::mutex g_mu;
double* g_scratch[kMaxDevices] = {};

// the scratch address on the current device (looked up once per device)
hipError_t scratch_ptr(double** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[dev] = static_cast<double*>(p);
  }
  *out = g_scratch[dev];
  return hipSuccess;
}

// twiddles exp(-2 pi i k / N) of both lengths and the 1-D symbol factors
// (the same formulas as the handle's FFT plan: operator_divergence.m:283 for
// the second-order symbol, or the exact symbol of the fourth-order A)
__global__ void k_noalloc_tables(int Nx, int Ny, int Nh, double hx, double hy, int exact4, double* twx, double* twy,
                                 double* ax, double* ay) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = Nx + Ny + Nh + Ny;
  if (i >= n) return;
  if (i < Nx) {
    double s, c;
    sincospi(-2.0 * (double)i / (double)Nx, &s, &c);
    twx[2 * i] = c;
    twx[2 * i + 1] = s;
  } else if (i < Nx + Ny) {
    const int k = i - Nx;
    double s, c;
    sincospi(-2.0 * (double)k / (double)Ny, &s, &c);
    twy[2 * k] = c;
    twy[2 * k + 1] = s;
  } else {
    const bool xdir = i < Nx + Ny + Nh;
    const int k = xdir ? i - Nx - Ny : i - Nx - Ny - Nh;
    const int N = xdir ? Nx : Ny;
    const double d = xdir ? hx : hy, o = xdir ? hy : hx;
    const double t = (double)k / (double)N;
    double v;
    if (exact4) {
      const double m = 81.0 * sinpi(t) - 3.0 * sinpi(3.0 * t);
      v = 4.0 * o / (72.0 * d) * m * m;
    } else {
      const double sn = sinpi(t);
      v = (1.0 / (d * d)) * (sn * sn);
    }
    (xdir ? ax : ay)[k] = v;
  }
}

}  // namespace

extern "C" {

int ins2d_poisson_fft_noalloc(int Nx, int Ny, double hx, double hy, int order4, int fft_symbol, const double* f,
                              double* dp, void* stream) {
  auto fail = [](int code, const std::string& msg) {
    set_handle_free_error("ins2d_poisson_fft_noalloc: " + msg);
    return code;
  };
  if (!f || !dp) return fail(INS2D_E_ARG, "null array");
  if (!(hx > 0.0) || !(hy > 0.0)) return fail(INS2D_E_ARG, "grid spacings must be positive");
  if (!fft_register_lengths(Nx, Ny))
    return fail(INS2D_E_UNSUPPORTED, "Nx and Ny must be powers of two from 16 to 4096");
  const int Nh = Nx / 2 + 1;
  const size_t spec = 2 * (size_t)Nh * (size_t)Ny;
  const size_t tables = 2 * (size_t)Nx + 2 * (size_t)Ny + (size_t)Nh + (size_t)Ny;
  if (spec > kSpecCap || tables > kTableCap - 64)
    return fail(INS2D_E_ARG, "grid " + std::to_string(Nx) + "x" + std::to_string(Ny) +
                                 " exceeds the scratch capacity (half spectrum (Nx/2+1)*Ny at most 525312 complex "
                                 "values, e.g. 1024x1024)");
  double* base = nullptr;
  hipError_t e = scratch_ptr(&base);
  if (e != hipSuccess) return fail(INS2D_E_HIP, std::string("hipGetSymbolAddress: ") + hipGetErrorString(e));
  hipStream_t st = static_cast<hipStream_t>(stream);
  FFTPlan P{};
  P.Nx = Nx;
  P.Ny = Ny;
  P.Nh = Nh;
  P.pow2x = P.pow2y = true;
  P.twx = base;
  P.twy = P.twx + 2 * Nx;
  P.ax = P.twy + 2 * Ny;
  P.ay = P.ax + Nh;
  P.HT = base + kTableCap;   // 256-byte aligned (kTableCap doubles)
  P.H = nullptr;             // register transforms store the spectra transposed
  // the early-exit maxima: the last 64 slots of the table region (the
  // tables take at most 22529 of its 32768 doubles)
  P.fmax2 = reinterpret_cast<unsigned long long*>(base + kTableCap - 64);
  const bool exact4 = order4 && fft_symbol;
  P.c = exact4 ? 1.0 : 4.0 * hx * hy;
  const int n = Nx + Ny + Nh + Ny;
  hipLaunchKernelGGL(k_noalloc_tables, dim3((n + 255) / 256), dim3(256), 0, st, Nx, Ny, Nh, hx, hy, exact4 ? 1 : 0,
                     P.twx, P.twy, P.ax, P.ay);
  launch_fft_poisson(st, P, f, dp);
  e = hipGetLastError();
  if (e != hipSuccess) return fail(INS2D_E_HIP, std::string("launch: ") + hipGetErrorString(e));
  return INS2D_OK;
}

}  // extern "C"
