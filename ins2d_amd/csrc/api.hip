// C-ABI of libins2d_hip: handle management, operator upload, the explicit RK
// step (time_ERK.m), the Poisson solvers and the diagnostics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <exception>
#include <memory>
#include <string>
#include <vector>

#include "../../include/ins2d.h"
#include "bcmask.hpp"
#include "uniform.hpp"
#include "device.hpp"
#include "launch.hpp"
#include "ops1d.hpp"
#include "ops4.hpp"

using namespace ins2d;

namespace {

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct ArgError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct NoConv : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct Unsupported : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define HIPCHK(x)                                                                    \
  do {                                                                               \
    hipError_t e_ = (x);                                                             \
    if (e_ != hipSuccess) throw HipError(std::string(#x ": ") + hipGetErrorString(e_)); \
  } while (0)

}  // namespace

struct ins2d_handle {
  std::string err;
  hipStream_t st = nullptr;
  int path = 0;              // 0 generic, 2/4 periodic uniform order
  int Nx = 0, Ny = 0, Nu = 0, Nv = 0, NV = 0, Np = 0;
  int nux = 0, nuy = 0, nvx = 0, nvy = 0;
  double hx = 0, hy = 0, Re = 1, invRe = 0;
  std::unique_ptr<Plan> plan;
  // fourth-order operators on grids with walls (CSR on the device) and the
  // dense inverse of their pressure matrix; null on the other paths
  std::unique_ptr<W4Ops> w4;
  W4Dev w4d{};
  double* one = nullptr;   // a device 1.0 (unit weights of the reductions)
  int bc_u[4] = {0, 0, 0, 0}, bc_v[4] = {0, 0, 0, 0};
  std::vector<void*> allocs;
  // operators (generic)
  DComp hCu{}, hCv{};
  DComp *dCu = nullptr, *dCv = nullptr;
  // split tiles of the momentum right-hand side (closed form inside, generic
  // ring); none when the grid has no closed-form interior
  RTile* split = nullptr;
  int nsplit = 0;
  DLap lap{};
  std::vector<void*> yv_allocs;   // boundary vectors (re-uploaded by set_bc)
  // FFT
  FFTPlan fft{};
  bool has_fft = false;
  // fast-diagonalization direct solve (poisson = direct), or the
  // preconditioner of PCG (cg_precond = 3) with its two work arrays
  FDMPlan fdm{};
  bool fdm_pcg = false;
  // device-iteration mode of the fast-diagonalization PCG inside ins2d_step:
  // each solve runs one iteration (the rule with the exact inverse as
  // preconditioner) and counts a missed stopping rule into d_missed; the
  // call checks once at its end and redoes its steps from the saved state
  // on the host-checked path when any solve missed
  bool fdm_dev = false;
  // the first stage's k_1 and Vtemp of the next step are already in kv[0]
  // and Vt (formed with the additional solve's F of the same velocity)
  bool k1_ready = false;
  int* d_missed = nullptr;
  double *Vsave = nullptr, *psave = nullptr;
  double *fw1 = nullptr, *fw2 = nullptr;
  // IC(0) preconditioner of the reference (cg_precond = 1)
  bool ic0 = false;
  DIC dic{};
  // accelerated PCG: 1 learned warm start, 2 two-level preconditioner
  int accel = 0;
  int ws_sweeps = 4;   // damped Jacobi sweep pairs after a start with the coarse-solution input
  ins2d_nn* nn = nullptr;
  TLPlan tl{};
  long long fallbacks = 0;
  // chunks of PCG iterations captured once into a graph and replayed (the
  // iteration kernels test the device-side done flag, so a chunk may run
  // past convergence harmlessly); keyed by the solution array
  hipGraphExec_t cg_exec = nullptr, tl_exec = nullptr;
  hipGraph_t cg_graph = nullptr, tl_graph = nullptr;
  const double *cg_x = nullptr, *tl_x = nullptr;
  // state and work buffers
  double *V = nullptr, *Vb = nullptr, *p = nullptr, *Vt = nullptr, *f = nullptr, *dp = nullptr, *tV = nullptr;
  std::vector<double*> kv;
  double *cx = nullptr, *cr = nullptr, *cz = nullptr, *cp = nullptr, *cq = nullptr;
  double* cp2 = nullptr;   // second search-direction buffer of the two-launch PCG iteration
  CGStateH cg{};
  int* d_it = nullptr;
  double* d_rel = nullptr;
  double* part = nullptr;
  int nred = 0;
  double* ee_part = nullptr;   // partial maxima of |f| (early exit of the matrix solvers)
  double* ox = nullptr;   // volume factors (all paths)
  double* oy = nullptr;
  double* ovx = nullptr;
  double* ovy = nullptr;
  // time integration
  int s = 0;
  std::vector<double> At, ct;   // shifted tableau (time_ERK.m:29,32)
  std::vector<bool> keep;        // kV_j needed by a later stage
  int poisson = 6;
  int fft_symbol = 0;
  double tol = 1e-8;
  int maxit = 1000;
  int p_add = 1;
  int order4 = 0;
  // boundary samples (for the momentum/energy boundary terms)
  BCSamples samp;
  bool have_bc = false;
  // time derivatives of the boundary samples (dudtLe_i, dudtRi_i, dvdtLo_i,
  // dvdtUp_i) and the operator records whose yM holds ydM, built from them
  std::vector<double> rate[4];
  bool have_rate = false;
  DComp *dCu_rate = nullptr, *dCv_rate = nullptr;
  std::vector<void*> rate_allocs;
  // boundary vectors of time-dependent boundary values, one version per
  // stage time of a chunk of steps: all arrays of all versions in one device
  // block (filled from pinned staging by one copy), and per version a (u, v)
  // pair of operator records pointing into it; a stage selects its version by
  // pointing dCu/dCv at that pair, so the step allocates nothing
  size_t ver_len = 0;
  int ver_cap = 0;
  double* ver_dev = nullptr;
  double* ver_pin = nullptr;
  DComp* ver_comp = nullptr;
  hipEvent_t ver_ev = nullptr;
  unsigned* ver_ones = nullptr;
  // stats
  int last_it = 0;
  double last_rel = 0.0;
  long long total_it = 0;
  // profiling
  bool prof = false;
  // kinds 0-3: momentum RHS, divergence, Poisson solve, projection of the RK
  // stages; 4-6: RHS, divergence and Poisson solve of the additional solve
  // 8-11: the forward row transform, column solve, inverse row transform and
  // inverse row transform with p += dp of the register-transform FFT solve
  // (every solve, kernel timestamps)
  static constexpr int kProfKinds = 12;
  double prof_ms[kProfKinds] = {};
  long long prof_n[kProfKinds] = {};
  int prof_off = 0;   // 4 inside the additional solve
  std::vector<hipEvent_t> ev_free;
  struct Pending {
    int kind;
    hipEvent_t a, b;
  };
  std::vector<Pending> pend;
};

namespace {

template <typename T>
T* dalloc(ins2d_handle* h, size_t n, std::vector<void*>* list = nullptr) {
  void* p = nullptr;
  HIPCHK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
  (list ? *list : h->allocs).push_back(p);
  return (T*)p;
}

template <typename T>
T* dupload(ins2d_handle* h, const std::vector<T>& v, std::vector<void*>* list = nullptr) {
  T* p = dalloc<T>(h, v.size(), list);
  if (!v.empty()) HIPCHK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return p;
}

DOp1 up_op(ins2d_handle* h, const HOp1& o) {
  DOp1 d{};
  d.nrows = o.nrows;
  d.ncols = o.ncols;
  d.K = o.K;
  d.kmax = o.kmax;
  d.col = dupload(h, o.col);
  d.val = dupload(h, o.val);
  return d;
}
DKron up_kron(ins2d_handle* h, const HKron& k) {
  DKron d;
  d.X = up_op(h, k.X);
  d.Y = up_op(h, k.Y);
  return d;
}

void up_yv(ins2d_handle* h, const HYVec& Y, DYV& d) {
  std::vector<std::vector<double>> ax, ay;
  eval_yvec(Y, h->samp, ax, ay);
  if (ax.size() > 4) throw std::runtime_error("too many boundary terms");
  d.nt = (int)ax.size();
  for (int t = 0; t < d.nt; ++t) {
    d.ax[t] = dupload(h, ax[t], &h->yv_allocs);
    d.ay[t] = dupload(h, ay[t], &h->yv_allocs);
    nonzero_range(ax[t].data(), ax[t].size(), d.x0[t], d.x1[t]);
    nonzero_range(ay[t].data(), ay[t].size(), d.y0[t], d.y1[t]);
  }
  for (int t = d.nt; t < 4; ++t) {
    d.ax[t] = d.ay[t] = nullptr;
    d.x0[t] = d.x1[t] = d.y0[t] = d.y1[t] = 0;
  }
}

void upload_comp_ops(ins2d_handle* h, const HComp& C, DComp& D) {
  D.nx = C.nx;
  D.ny = C.ny;
  D.nconv = (int)C.conv.size();
  D.ndiff = (int)C.diff.size();
  if (D.nconv > 2 || D.ndiff > 2) throw std::runtime_error("operator term count");
  for (int t = 0; t < D.nconv; ++t) {
    D.conv[t].D = up_kron(h, C.conv[t].D);
    D.conv[t].I = up_kron(h, C.conv[t].I);
    D.conv[t].A = up_kron(h, C.conv[t].A);
    D.conv[t].icomp = C.conv[t].icomp;
    D.conv[t].scale = C.conv[t].scale;
    D.conv[t].yI.nt = D.conv[t].yA.nt = 0;
  }
  for (int t = 0; t < D.ndiff; ++t) {
    D.diff[t].D = up_kron(h, C.diff[t].D[0]);
    D.diff[t].S = up_kron(h, C.diff[t].S[0]);
    D.diff[t].sscale = C.diff[t].sscale[0];
    D.diff[t].yS.nt = 0;
  }
  D.ox = dupload(h, C.ox);
  D.oy = dupload(h, C.oy);
  D.G = up_kron(h, C.G);
  D.M = up_kron(h, C.M);
  D.yM.nt = 0;
}

void upload_comp_bc(ins2d_handle* h, const HComp& C, DComp& D) {
  for (int t = 0; t < D.nconv; ++t) {
    up_yv(h, C.conv[t].yI, D.conv[t].yI);
    up_yv(h, C.conv[t].yA, D.conv[t].yA);
  }
  for (int t = 0; t < D.ndiff; ++t) up_yv(h, C.diff[t].yS, D.diff[t].yS);
  up_yv(h, C.yM, D.yM);
  std::vector<unsigned> mx, my;
  bc_masks(C, D, mx, my);
  D.bmx = dupload(h, mx, &h->yv_allocs);
  D.bmy = dupload(h, my, &h->yv_allocs);
}

void push_comps(ins2d_handle* h) {
  if (!h->dCu) {
    h->dCu = dalloc<DComp>(h, 1);
    h->dCv = dalloc<DComp>(h, 1);
  }
  HIPCHK(hipMemcpy(h->dCu, &h->hCu, sizeof(DComp), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(h->dCv, &h->hCv, sizeof(DComp), hipMemcpyHostToDevice));
}

bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

// the dense inverse of the fourth-order pressure matrix with walls holds
// Np^2 doubles (8192 cells: 512 MiB)
constexpr long kW4MaxNp = 8192;

DCsr up_csr(ins2d_handle* h, const Sp& A) {
  std::vector<int> rp(A.m + 1, 0), ci;
  std::vector<double> v;
  for (int i = 0; i < A.m; ++i) {
    for (auto& kv : A.r[i]) {
      ci.push_back(kv.first);
      v.push_back(kv.second);
    }
    rp[i + 1] = (int)ci.size();
  }
  DCsr d{};
  d.m = A.m;
  d.n = A.n;
  d.rp = dupload(h, rp);
  d.ci = dupload(h, ci);
  d.v = dupload(h, v);
  return d;
}

// boundary vectors of the fourth-order wall operators at the current samples
void w4_upload_bc(ins2d_handle* h) {
  std::vector<double> yA, yI, yD, yM;
  w4_bc_vectors(*h->w4, h->samp, yA, yI, yD, yM);
  const W4Dev& D = h->w4d;
  HIPCHK(hipMemcpy(const_cast<double*>(D.yA), yA.data(), yA.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(const_cast<double*>(D.yI), yI.data(), yI.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(const_cast<double*>(D.yD), yD.data(), yD.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(const_cast<double*>(D.yM), yM.data(), yM.size() * sizeof(double), hipMemcpyHostToDevice));
}

// timing with HIP events on the handle's stream.  KERNEL scopes wrap one
// launcher of the stencil kernels: its kernels record their own start and end
// timestamps into the two events (launch_timed), which measures the kernels
// alone; other scopes record the events on the stream around the work.
struct ProfScope {
  ins2d_handle* h;
  int kind;
  bool kernel;
  hipEvent_t a{}, b{};
  KernelTimer kt;
  ProfScope(ins2d_handle* h_, int k, bool kernel_ts = false) : h(h_), kind(k + h_->prof_off), kernel(kernel_ts) {
    if (!h->prof) return;
    a = get_ev();
    b = get_ev();
    if (kernel) {
      kt.start = a;
      kt.stop = b;
      g_ktimer = &kt;
    } else {
      HIPCHK(hipEventRecord(a, h->st));
    }
  }
  hipEvent_t get_ev() {
    if (!h->ev_free.empty()) {
      hipEvent_t e = h->ev_free.back();
      h->ev_free.pop_back();
      return e;
    }
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    return e;
  }
  ~ProfScope() {
    if (!h->prof) return;
    if (kernel) {
      g_ktimer = nullptr;
      if (kt.launches > 0) {
        h->pend.push_back({kind, a, b});
      } else {
        h->ev_free.push_back(a);
        h->ev_free.push_back(b);
      }
      return;
    }
    if (hipEventRecord(b, h->st) == hipSuccess) h->pend.push_back({kind, a, b});
  }
};

// kernel timestamps of the FFT solve's kernels (profile kinds 8-11)
struct FftProf {
  ins2d_handle* h;
  KernelTimer kt[4];
  explicit FftProf(ins2d_handle* h_) : h(h_) {
    if (!h->prof) return;
    for (int k = 0; k < 4; ++k) {
      kt[k].start = get_ev();
      kt[k].stop = get_ev();
      g_fft_timer[k] = &kt[k];
    }
  }
  hipEvent_t get_ev() {
    if (!h->ev_free.empty()) {
      hipEvent_t e = h->ev_free.back();
      h->ev_free.pop_back();
      return e;
    }
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    return e;
  }
  ~FftProf() {
    if (!h->prof) return;
    for (int k = 0; k < 4; ++k) {
      g_fft_timer[k] = nullptr;
      if (kt[k].launches > 0) {
        h->pend.push_back({8 + k, kt[k].start, kt[k].stop});
      } else {
        h->ev_free.push_back(kt[k].start);
        h->ev_free.push_back(kt[k].stop);
      }
    }
  }
};

void prof_collect(ins2d_handle* h) {
  for (auto& pe : h->pend) {
    float ms = 0.f;
    if (hipEventSynchronize(pe.b) == hipSuccess && hipEventElapsedTime(&ms, pe.a, pe.b) == hipSuccess) {
      h->prof_ms[pe.kind] += ms;
      h->prof_n[pe.kind] += 1;
    }
    h->ev_free.push_back(pe.a);
    h->ev_free.push_back(pe.b);
  }
  h->pend.clear();
}

// the samples with the four divergence-side sources replaced by their time
// derivatives and all others zero: the yM of these samples is ydM
// (set_bc_vectors.m:73-79,120-146)
BCSamples rate_samples(const ins2d_handle* h, const double* const* rate) {
  BCSamples S;
  for (int q = 0; q < S_NSRC; ++q) S.s[q].assign(h->samp.s[q].size(), 0.0);
  const int src[4] = {S_uLe_i, S_uRi_i, S_vLo_i, S_vUp_i};
  for (int q = 0; q < 4; ++q) std::copy(rate[q], rate[q] + S.s[src[q]].size(), S.s[src[q]].begin());
  return S;
}

// operator records whose yM is ydM of the stored rates (regular allocation)
void upload_rate(ins2d_handle* h) {
  for (void* p : h->rate_allocs) (void)hipFree(p);
  h->rate_allocs.clear();
  if (!h->plan) return;
  if (!h->dCu_rate) {
    h->dCu_rate = dalloc<DComp>(h, 1);
    h->dCv_rate = dalloc<DComp>(h, 1);
  }
  const double* r[4] = {h->rate[0].data(), h->rate[1].data(), h->rate[2].data(), h->rate[3].data()};
  const BCSamples S = rate_samples(h, r);
  DComp cu = h->hCu, cv = h->hCv;
  for (int c = 0; c < 2; ++c) {
    std::vector<std::vector<double>> ax, ay;
    eval_yvec(c ? h->plan->cv.yM : h->plan->cu.yM, S, ax, ay);
    DYV& d = c ? cv.yM : cu.yM;
    for (int t = 0; t < d.nt; ++t) {
      d.ax[t] = dupload(h, ax[t], &h->rate_allocs);
      d.ay[t] = dupload(h, ay[t], &h->rate_allocs);
      nonzero_range(ax[t].data(), ax[t].size(), d.x0[t], d.x1[t]);
      nonzero_range(ay[t].data(), ay[t].size(), d.y0[t], d.y1[t]);
    }
  }
  HIPCHK(hipMemcpy(h->dCu_rate, &cu, sizeof(DComp), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(h->dCv_rate, &cv, sizeof(DComp), hipMemcpyHostToDevice));
}

// writes the arrays of one boundary vector at hb + off (when hb is given) and
// points d at the same offsets from db; the boxes span the whole arrays, as
// the values change from version to version
size_t put_yv(const HYVec& Y, const BCSamples& S, double* hb, double* db, size_t off, DYV& d) {
  std::vector<std::vector<double>> ax, ay;
  eval_yvec(Y, S, ax, ay);
  if (ax.size() > 4) throw std::runtime_error("too many boundary terms");
  d.nt = (int)ax.size();
  for (int t = 0; t < 4; ++t) {
    if (t >= d.nt) {
      d.ax[t] = d.ay[t] = nullptr;
      d.x0[t] = d.x1[t] = d.y0[t] = d.y1[t] = 0;
      continue;
    }
    if (hb) std::copy(ax[t].begin(), ax[t].end(), hb + off);
    d.ax[t] = db + off;
    d.x0[t] = 0;
    d.x1[t] = (int)ax[t].size();
    off += ax[t].size();
    if (hb) std::copy(ay[t].begin(), ay[t].end(), hb + off);
    d.ay[t] = db + off;
    d.y0[t] = 0;
    d.y1[t] = (int)ay[t].size();
    off += ay[t].size();
  }
  return off;
}

size_t put_comp(const HComp& C, const BCSamples& S, double* hb, double* db, size_t off, DComp& D) {
  for (int t = 0; t < D.nconv; ++t) {
    off = put_yv(C.conv[t].yI, S, hb, db, off, D.conv[t].yI);
    off = put_yv(C.conv[t].yA, S, hb, db, off, D.conv[t].yA);
  }
  for (int t = 0; t < D.ndiff; ++t) off = put_yv(C.diff[t].yS, S, hb, db, off, D.diff[t].yS);
  return put_yv(C.yM, S, hb, db, off, D.yM);
}

// room for n versions; the operator records of every slot are written once
// here (they depend on the slot, not on the values)
void ver_reserve(ins2d_handle* h, int n) {
  if (n <= h->ver_cap) return;
  HIPCHK(hipStreamSynchronize(h->st));
  const Plan& P = *h->plan;
  if (!h->ver_len) {
    DComp cu = h->hCu, cv = h->hCv;
    h->ver_len = put_comp(P.cv, h->samp, nullptr, nullptr, put_comp(P.cu, h->samp, nullptr, nullptr, 0, cu), cv);
    HIPCHK(hipEventCreateWithFlags(&h->ver_ev, hipEventDisableTiming));
  }
  if (h->ver_dev) {
    (void)hipFree(h->ver_dev);
    (void)hipFree(h->ver_comp);
    (void)hipHostFree(h->ver_pin);
    h->ver_dev = h->ver_pin = nullptr;
    h->ver_comp = nullptr;
  }
  const size_t len = std::max<size_t>(h->ver_len, 1);
  HIPCHK(hipMalloc(&h->ver_dev, sizeof(double) * len * n));
  HIPCHK(hipMalloc(&h->ver_comp, sizeof(DComp) * 2 * n));
  if (!h->ver_ones) {
    // versions mark every point as reading boundary terms (their boxes span
    // the whole arrays)
    std::vector<unsigned> ones((size_t)std::max(std::max(h->nux, h->nvx), std::max(h->nuy, h->nvy)), ~0u);
    h->ver_ones = dupload(h, ones);
  }
  HIPCHK(hipHostMalloc(&h->ver_pin, sizeof(double) * len * n, hipHostMallocDefault));
  h->ver_cap = n;
  std::vector<DComp> recs(2 * (size_t)n);
  for (int v = 0; v < n; ++v) {
    DComp cu = h->hCu, cv = h->hCv;
    double* db = h->ver_dev + (size_t)v * len;
    put_comp(P.cv, h->samp, nullptr, db, put_comp(P.cu, h->samp, nullptr, db, 0, cu), cv);
    cu.bmx = cu.bmy = cv.bmx = cv.bmy = h->ver_ones;
    recs[2 * v] = cu;
    recs[2 * v + 1] = cv;
  }
  HIPCHK(hipMemcpy(h->ver_comp, recs.data(), sizeof(DComp) * recs.size(), hipMemcpyHostToDevice));
}

// the values of version v from samples S (staged; copied by the caller)
void ver_fill(ins2d_handle* h, int v, const BCSamples& S) {
  const Plan& P = *h->plan;
  const size_t len = std::max<size_t>(h->ver_len, 1);
  double* hb = h->ver_pin + (size_t)v * len;
  DComp cu = h->hCu, cv = h->hCv;
  put_comp(P.cv, S, hb, hb, put_comp(P.cu, S, hb, hb, 0, cu), cv);
}

// one sample set of the C ABI's layout (the 12 arrays concatenated)
void unpack_samples(const ins2d_handle* h, const double* src, BCSamples& S) {
  for (int q = 0; q < S_NSRC; ++q) {
    const size_t n = h->samp.s[q].size();
    S.s[q].assign(src, src + n);
    src += n;
  }
}

// ------------------------------------------------------------------ kernels
const StageCombine kNoCombine{0, 0.0, {0}, {nullptr}};

void rhs(ins2d_handle* h, const double* V, const double* p, int mode, double* out, const StageCombine& sc, double* vt) {
  ProfScope ps(h, 0, true);
  if (h->w4) {
    launch_w4_rhs(h->st, h->w4d, V, p, mode, out, sc, vt);
  } else if (h->path) {
    launch_rhs_per(h->st, h->path, h->Nx, h->Ny, h->hx, h->hy, h->invRe, V, V + h->Nu, p, mode, out, sc, vt);
  } else {
    StageCombine scv = sc;
    for (int q = 0; q < sc.n; ++q) scv.k[q] = sc.k[q] + h->Nu;
    if (h->nsplit)
      launch_rhs_split(h->st, h->dCu, h->dCv, h->split, h->nsplit, V, V + h->Nu, p, mode, out,
                       out ? out + h->Nu : nullptr, sc, scv, vt, vt ? vt + h->Nu : nullptr);
    else
      launch_rhs_generic(h->st, h->dCu, h->dCv, h->nux, h->nuy, h->nvx, h->nvy, V, V + h->Nu, p, mode, out,
                         out ? out + h->Nu : nullptr, sc, scv, vt, vt ? vt + h->Nu : nullptr);
  }
}

// f = oscale * (M (ca A + cb B) + ycoef yM)
void div(ins2d_handle* h, const double* A, const double* B, double ca, double cb, double ycoef, double oscale,
         double* f) {
  ProfScope ps(h, 1, true);
  if (h->w4)
    launch_w4_div(h->st, h->w4d, A, B, ca, cb, ycoef, oscale, f);
  else if (h->path)
    launch_div_per(h->st, h->path, h->Nx, h->Ny, h->hx, h->hy, A, B, ca, cb, oscale, f);
  else
    launch_div_generic(h->st, h->dCu, h->dCv, h->Nx, h->Ny, A, B, ca, cb, ycoef, oscale, h->Nu, f);
}

void proj(ins2d_handle* h, const double* Vn, const double* Vt, const double* dp, double dt, double c, double* V) {
  ProfScope ps(h, 3, true);
  if (h->w4)
    launch_w4_proj(h->st, h->w4d, Vn, Vt, dp, dt, c, V);
  else if (h->path)
    launch_proj_per(h->st, h->path, h->Nx, h->Ny, h->hx, h->hy, Vn, Vt, dp, dt, c, V);
  else
    launch_proj_generic(h->st, h->dCu, h->dCv, h->nux, h->nuy, h->nvx, h->nvy, h->Nu, Vn, Vt, dp, dt, c, V);
}

// PCG with the learned start vector and/or the two-level preconditioner;
// the stopping rule and tolerance are the plain solver's (cg.F:247)
void accel_poisson(ins2d_handle* h, const double* f, double* dp) {
  const double tol2 = h->tol * h->tol;
  const bool net = (h->accel & 1) != 0, xc_in = (h->accel & 4) != 0, coarse_start = (h->accel & 8) != 0;
  int x0_zero = 1;
  if (net || coarse_start) {
    // the coarse-grid solution P Ac^-1 P^T f: the network's second input
    // (the network then predicts its correction), or the start itself
    if (xc_in || coarse_start) launch_tl_coarse_solution(h->st, h->tl, f, dp, h->cr);
    if (net) {
      if (ins2d_nn_forward_dev(h->nn, f, xc_in ? dp : nullptr, dp, h->st) != INS2D_OK)
        throw HipError(std::string("warm start: ") + ins2d_nn_last_error(h->nn));
      // the network's BF16 rounding noise is rough and A amplifies rough
      // error most: damped Jacobi sweeps remove it before CG starts; without
      // the coarse input the start's smooth (global) error is also corrected
      // on the coarse grid, between the sweeps
      if (xc_in) {
        if (h->ws_sweeps > 0) launch_tl_smooth(h->st, h->tl, f, dp, h->cz, h->ws_sweeps, 0.6);
      } else {
        launch_tl_smooth(h->st, h->tl, f, dp, h->cz, 12, 0.6);
        if (h->tl.coarse) {
          launch_tl_coarse_correct(h->st, h->tl, f, dp, h->cr);
          launch_tl_smooth(h->st, h->tl, f, dp, h->cz, 12, 0.6);
        }
      }
    }
    x0_zero = 0;
  } else {
    HIPCHK(hipMemsetAsync(dp, 0, sizeof(double) * h->Np, h->st));
  }
  launch_tl_init(h->st, h->tl, f, dp, h->cr, h->cp, x0_zero, tol2);
  if (!x0_zero) {
    // keep the start only when it is closer to the solution than x0 = 0 in
    // the norm CG minimises: phi(x0) = x0 (f + r0) / 2 < phi(0) = 0, with
    // phi(x) - phi(A^-1 f) = |x - A^-1 f|^2_(-A) / 2
    launch_tl_energy(h->st, h->tl, f, dp, h->cr, h->tl.en);
    const int nblk = h->tl.C.ncx * h->tl.C.ncy;
    std::vector<double> en(nblk);
    HIPCHK(hipMemcpyAsync(en.data(), h->tl.en, nblk * sizeof(double), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    double e = 0.0;
    for (double v : en) e += v;
    if (!(e < 0.0)) {
      ++h->fallbacks;
      HIPCHK(hipMemsetAsync(dp, 0, sizeof(double) * h->Np, h->st));
      launch_tl_init(h->st, h->tl, f, dp, h->cr, h->cp, 1, tol2);
    }
  }
  int flag[3] = {0, 0, 0};
  HIPCHK(hipMemcpyAsync(flag, h->tl.flag, 3 * sizeof(int), hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  if (!flag[0] && (!h->tl_exec || h->tl_x != dp)) {
    if (h->tl_exec) { HIPCHK(hipGraphExecDestroy(h->tl_exec)); HIPCHK(hipGraphDestroy(h->tl_graph)); }
    HIPCHK(hipStreamBeginCapture(h->st, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < 16; ++k) launch_tl_iter(h->st, h->tl, dp, h->cr, h->cp, h->cq);
    HIPCHK(hipStreamEndCapture(h->st, &h->tl_graph));
    HIPCHK(hipGraphInstantiate(&h->tl_exec, h->tl_graph, nullptr, nullptr, 0));
    h->tl_x = dp;
  }
  while (!flag[0]) {
    HIPCHK(hipGraphLaunch(h->tl_exec, h->st));
    HIPCHK(hipMemcpyAsync(flag, h->tl.flag, 3 * sizeof(int), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
  }
  // report the true relative residual ||f - A x|| / ||f|| of the result
  launch_tl_true_residual(h->st, h->tl, f, dp, h->cr);
  double rel = 0;
  HIPCHK(hipMemcpyAsync(&rel, h->tl.s + 7, sizeof(double), hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  h->last_it = flag[1];
  h->last_rel = std::sqrt(rel);
  h->total_it += flag[1];
  if (flag[2]) throw NoConv("CG did not converge in maxit iterations");
}

void poisson_matrix(ins2d_handle* h, const double* f, double* dp);

// A dp = f, x0 = 0 (pressure_poisson.m:12); dp = 0 when the right-hand side
// is negligible (pressure_poisson.m:21-39: the FFT kernels test max |fft2(f)|
// themselves, the matrix solvers max |f|, both against 1e-14)
void poisson(ins2d_handle* h, const double* f, double* dp) {
  ProfScope ps(h, 2);
  if (h->poisson == INS2D_POISSON_FFT) {
    FftProf fp(h);
    launch_fft_poisson(h->st, h->fft, f, dp);
    return;
  }
  if (h->poisson == INS2D_POISSON_DIRECT && !h->w4) {
    // the early exit inside the first and last GEMM launches
    launch_fdm_poisson(h->st, h->fdm, f, dp, h->cr, h->cz, nullptr, h->ee_part);
    h->last_it = 0;
    h->last_rel = 0.0;
    return;
  }
  launch_absmax_part(h->st, f, (size_t)h->Np, h->ee_part);
  poisson_matrix(h, f, dp);
  launch_zero_if_small(h->st, h->ee_part, 1e-14, dp, (size_t)h->Np);
}

// the direct and CG solvers of pressure_poisson.m:42-90
void poisson_matrix(ins2d_handle* h, const double* f, double* dp) {
  if (h->poisson == INS2D_POISSON_DIRECT) {
    if (h->w4)
      launch_w4_solve(h->st, h->w4d, f, dp);
    else
      launch_fdm_poisson(h->st, h->fdm, f, dp, h->cr, h->cz);
    h->last_it = 0;
    h->last_rel = 0.0;
    return;
  }
  const double tol2 = h->tol * h->tol;
  if (h->accel) {
    accel_poisson(h, f, dp);
    return;
  }
  if (h->ic0) {
    launch_cg_ic0(h->st, h->lap, h->dic, f, dp, h->cr, h->cz, h->cp, h->cq, tol2, h->maxit, 1, h->d_it, h->d_rel);
    int it = 0;
    double rel = 0;
    HIPCHK(hipMemcpyAsync(&it, h->d_it, sizeof(int), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipMemcpyAsync(&rel, h->d_rel, sizeof(double), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    h->last_it = it;
    h->last_rel = std::sqrt(rel);
    h->total_it += it;
    if (it >= h->maxit && rel > tol2) throw NoConv("CG did not converge in maxit iterations");
    return;
  }
  if (h->fdm_pcg) {
    // PCG preconditioned by the fast-diagonalization inverse (exact on this
    // tensor-product grid up to rounding): the plain stopping rule, normally
    // met after one iteration; the convergence test precedes each further
    // preconditioner application
    HIPCHK(hipMemsetAsync(dp, 0, sizeof(double) * h->Np, h->st));
    launch_cgx(h->st, h->lap, h->cg, 0, f, h->cr, nullptr, nullptr, tol2);
    launch_fdm_poisson(h->st, h->fdm, h->cr, h->cz, h->fw1, h->fw2);
    launch_cgx(h->st, h->lap, h->cg, 1, h->cr, h->cz, h->cp, nullptr, tol2);
    if (h->fdm_dev) {
      // one iteration, its outcome counted on the device (no host wait)
      launch_cgx(h->st, h->lap, h->cg, 2, h->cp, h->cq, dp, h->cr, tol2);
      launch_cgx_missed(h->st, h->cg, h->d_missed);
      return;
    }
    // the host checks after each convergence test (one iteration is the
    // rule with the exact inverse as preconditioner): the preconditioner
    // apply and the direction update run only when another iteration follows
    int flag[3] = {0, 0, 0};
    double rel = 0;
    while (true) {
      launch_cgx(h->st, h->lap, h->cg, 2, h->cp, h->cq, dp, h->cr, tol2);
      HIPCHK(hipMemcpyAsync(flag, h->cg.flag, 3 * sizeof(int), hipMemcpyDeviceToHost, h->st));
      HIPCHK(hipMemcpyAsync(&rel, h->cg.s + 4, sizeof(double), hipMemcpyDeviceToHost, h->st));
      HIPCHK(hipStreamSynchronize(h->st));
      if (flag[0]) break;
      launch_fdm_poisson(h->st, h->fdm, h->cr, h->cz, h->fw1, h->fw2, h->cg.flag);
      launch_cgx(h->st, h->lap, h->cg, 3, h->cr, h->cz, h->cp, nullptr, tol2);
    }
    h->last_it = flag[1];
    h->last_rel = std::sqrt(rel);
    h->total_it += flag[1];
    if (flag[2]) throw NoConv("CG did not converge in maxit iterations");
    return;
  }
  if (h->Np <= 32768) {
    launch_cg_single(h->st, h->lap, f, dp, h->cr, h->cz, h->cp, h->cq, tol2, h->maxit, 1, h->d_it, h->d_rel);
    int it = 0;
    double rel = 0;
    HIPCHK(hipMemcpyAsync(&it, h->d_it, sizeof(int), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipMemcpyAsync(&rel, h->d_rel, sizeof(double), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    h->last_it = it;
    h->last_rel = std::sqrt(rel);
    h->total_it += it;
    if (it >= h->maxit && rel > tol2) throw NoConv("CG did not converge in maxit iterations");
    return;
  }
  HIPCHK(hipMemsetAsync(dp, 0, sizeof(double) * h->Np, h->st));
  launch_cg_init(h->st, h->lap, h->cg, f, dp, h->cr, h->cz, h->cp, 1, tol2);
  int flag[3] = {0, 0, 0};
  if (!h->cg_exec || h->cg_x != dp) {
    if (h->cg_exec) { HIPCHK(hipGraphExecDestroy(h->cg_exec)); HIPCHK(hipGraphDestroy(h->cg_graph)); }
    HIPCHK(hipStreamBeginCapture(h->st, hipStreamCaptureModeThreadLocal));
    // 16 iterations of two launches each; the direction alternates between
    // cp and cp2 and is back in cp after the chunk
    for (int k = 0; k < 16; ++k)
      launch_cg_iter2(h->st, h->lap, h->cg, dp, h->cr, h->cz, (k & 1) ? h->cp2 : h->cp, (k & 1) ? h->cp : h->cp2,
                      h->cq);
    launch_cg2_check(h->st, h->cg);
    HIPCHK(hipStreamEndCapture(h->st, &h->cg_graph));
    HIPCHK(hipGraphInstantiate(&h->cg_exec, h->cg_graph, nullptr, nullptr, 0));
    h->cg_x = dp;
  }
  while (true) {
    HIPCHK(hipGraphLaunch(h->cg_exec, h->st));
    HIPCHK(hipMemcpyAsync(flag, h->cg.flag, 3 * sizeof(int), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    if (flag[0]) break;
  }
  double rel = 0;
  HIPCHK(hipMemcpy(&rel, h->cg.s + 4, sizeof(double), hipMemcpyDeviceToHost));
  h->last_it = flag[1];
  h->last_rel = std::sqrt(rel);
  h->total_it += flag[1];
  if (flag[2]) throw NoConv("CG did not converge in maxit iterations");
}

// A dp = f with f = oscale (M (ca A + cb B) + ycoef yM); acc: out += dp,
// else out = dp.  With the register-transform FFT solve the update is done
// by the last transform, and on the second-order periodic kernels f is
// computed inside the first transform (never stored); h->f and h->dp then
// keep their earlier contents.
void div_poisson(ins2d_handle* h, const double* A, const double* B, double ca, double cb, double ycoef, double oscale,
                 double* out, bool acc) {
  const bool reg = h->poisson == INS2D_POISSON_FFT && fft_register_lengths(h->Nx, h->Ny);
  if (reg && (h->path == 2 || (h->path == 4 && fft_fused_div4(h->Nx, h->Ny)))) {
    ProfScope ps(h, 2);
    FftProf fp(h);
    launch_fft_poisson_fused(h->st, h->fft, nullptr, h->hx, h->hy, A, B, ca, cb, oscale, out, acc, h->path);
    return;
  }
  div(h, A, B, ca, cb, ycoef, oscale, h->f);
  if (reg && acc) {
    ProfScope ps(h, 2);
    FftProf fp(h);
    launch_fft_poisson_fused(h->st, h->fft, h->f, h->hx, h->hy, nullptr, nullptr, 0.0, 0.0, 0.0, out, true);
    return;
  }
  if (acc && h->poisson == INS2D_POISSON_DIRECT && !h->w4) {
    // out += dp in the last GEMM of the solve (left as it is on the early exit)
    ProfScope ps(h, 2);
    launch_fdm_poisson(h->st, h->fdm, h->f, out, h->cr, h->cz, nullptr, h->ee_part, true);
    h->last_it = 0;
    h->last_rel = 0.0;
    return;
  }
  poisson(h, h->f, acc ? h->dp : out);
  if (acc) launch_axpy(h->st, (size_t)h->Np, 1.0, h->dp, out);
}

// p <- p + A^-1 (M Om^-1 F(V, p) + ydM)   (pressure_additional_solve.m:19-25);
// rate_cu/rate_cv: operator records whose yM is ydM (null: ydM = 0)
void additional_solve(ins2d_handle* h, const double* V, DComp* rate_cu = nullptr, DComp* rate_cv = nullptr) {
  struct Off {
    ins2d_handle* h;
    explicit Off(ins2d_handle* h_) : h(h_) { h->prof_off = 4; }
    ~Off() { h->prof_off = 0; }
  } off(h);
  rhs(h, V, h->p, RHS_WITH_P | RHS_SCALE_OM, h->tV, kNoCombine, nullptr);
  if (rate_cu && !h->path) {
    DComp *cu = h->dCu, *cv = h->dCv;
    h->dCu = rate_cu;
    h->dCv = rate_cv;
    div(h, h->tV, nullptr, 1.0, 0.0, 1.0, 1.0, h->f);
    h->dCu = cu;
    h->dCv = cv;
    poisson(h, h->f, h->dp);
    launch_axpy(h->st, (size_t)h->Np, 1.0, h->dp, h->p);
  } else {
    div_poisson(h, h->tV, nullptr, 1.0, 0.0, 0.0, 1.0, h->p, true);
  }
}

// one step of time_ERK.m.  ver: for time-dependent boundary values, the
// operator records of s + 2 consecutive versions (boundary values at tn, at
// the s stage times tn + c_i dt, and the ydM record at tn + dt); null for
// steady boundary values
// F of V_{n+1} once at the end of a step that another step follows in the
// same call (periodic fourth order, rhs_per_dual_ok): the additional solve's
// (F - G p) / Omega into tV and the next step's first-stage k_1 = F / Omega
// (into kv[0] when kept) with its stage combination into Vt, both as their
// own launches would form them; the next erk_step then skips its first
// right-hand side
void additional_solve_dual(ins2d_handle* h) {
  StageCombine sc0{};
  sc0.aself = h->At[0];
  sc0.n = 0;
  if (h->path) {
    launch_rhs_per(h->st, h->path, h->Nx, h->Ny, h->hx, h->hy, h->invRe, h->V, h->V + h->Nu, h->p, RHS_SCALE_OM,
                   h->keep[0] ? h->kv[0] : nullptr, sc0, h->Vt, h->tV);
  } else if (h->w4) {
    // fourth order with walls: k1 - Om^-1 G p by the projection kernel
    // (Vn = k1, no Vtemp, dt = c = 1)
    rhs(h, h->V, h->p, RHS_SCALE_OM, h->kv[0], sc0, h->Vt);
    launch_w4_proj(h->st, h->w4d, h->kv[0], nullptr, h->p, 1.0, 1.0, h->tV);
  } else {
    // generic operators: the first stage's right-hand side as its own launch
    // would form it, then (F - G p) / Omega = k1 - Om^-1 G p from it
    rhs(h, h->V, h->p, RHS_SCALE_OM, h->kv[0], sc0, h->Vt);
    launch_sub_grad_generic(h->st, h->dCu, h->dCv, h->Nu, h->NV - h->Nu, h->kv[0], h->p, h->tV);
  }
  div_poisson(h, h->tV, nullptr, 1.0, 0.0, 0.0, 1.0, h->p, true);
  h->k1_ready = true;
}

// the steps of an ins2d_step call whose boundary F can be shared (steady
// boundary values): the periodic kernels with two cells per thread (one
// instance writes both), the generic operators (second order, and fourth
// order with walls) when k1 is kept (the additional solve's right-hand side
// follows from it)
bool dual_capable(const ins2d_handle* h) {
  if (h->prof) return false;
  if (h->path) return rhs_per_dual_ok(h->path, h->Nx);
  return (h->w4 || h->plan) && h->keep[0];
}

void erk_step(ins2d_handle* h, double tn, double dt, DComp* ver = nullptr, bool dual_end = false) {
  (void)tn;
  double* Vn = h->V;
  double* Vs = h->Vb;   // stage velocity, V_{n+1} at the end
  DComp *cu0 = h->dCu, *cv0 = h->dCv;
  auto use = [&](int k) {
    h->dCu = ver + 2 * k;
    h->dCv = ver + 2 * k + 1;
  };
  if (ver) use(0);   // time_ERK.m:62-64
  for (int i = 0; i < h->s; ++i) {
    const double* Vsrc = (i == 0) ? Vn : Vs;
    StageCombine sc{};
    sc.aself = h->At[i * h->s + i];
    sc.n = 0;
    for (int j = 0; j < i; ++j) {
      const double a = h->At[i * h->s + j];
      if (a != 0.0) {
        sc.a[sc.n] = a;
        sc.k[sc.n] = h->kv[j];
        ++sc.n;
      }
    }
    // kV_i = Om^-1 (F_i + G p) = Om^-1 (-conv + diff - y_p)   (time_ERK.m:90-95)
    // Vtemp = sum_j At(i,j) kV_j                              (time_ERK.m:101)
    if (i == 0 && h->k1_ready)
      h->k1_ready = false;   // formed at the end of the previous step
    else
      rhs(h, Vsrc, h->p, RHS_SCALE_OM, h->keep[i] ? h->kv[i] : nullptr, sc, h->Vt);
    const double c = h->ct[i];
    const double* dpp;
    // unsteady BC: yM at tn + c_i dt (time_ERK.m:106-109), and a pressure
    // solve at every stage (time_ERK.m:121-123)
    if (ver) use(i + 1);
    if (ver || i > 0) {
      // f = (M (Vn/dt + Vtemp) + yM/dt) / c_i                 (time_ERK.m:117)
      div_poisson(h, Vn, h->Vt, 1.0 / dt, 1.0, 1.0 / dt, 1.0 / c, h->dp, false);
      dpp = h->dp;
    } else {
      dpp = h->p;   // steady BC, first stage: dp = pn (time_ERK.m:125-126)
    }
    // V = Vn + dt (Vtemp - c_i Om^-1 G dp)                    (time_ERK.m:132)
    proj(h, Vn, h->Vt, dpp, dt, c, Vs);
  }
  std::swap(h->V, h->Vb);
  if (!ver) {
    // steady BC: always the additional solve (time_ERK.m:160-163)
    if (dual_end)
      additional_solve_dual(h);
    else
      additional_solve(h, h->V);
  } else if (h->p_add) {
    // at tn + dt with ydM (time_ERK.m:147-148)
    additional_solve(h, h->V, ver + 2 * (h->s + 1), ver + 2 * (h->s + 1) + 1);
  } else {
    // p = kp(:, end) (time_ERK.m:162)
    HIPCHK(hipMemcpyAsync(h->p, h->dp, sizeof(double) * h->Np, hipMemcpyDeviceToDevice, h->st));
  }
  h->dCu = cu0;
  h->dCv = cv0;
}

void fill_samples_default(ins2d_handle* h) {
  const Plan* P = h->plan.get();
  for (int s = 0; s < S_NSRC; ++s) {
    int n = 0;
    switch (s) {
      case S_uLo: case S_uUp: n = h->Nx + 1; break;
      case S_uLo_i: case S_uUp_i: n = h->nux; break;
      case S_uLe_i: case S_uRi_i: n = h->Ny; break;
      case S_vLe: case S_vRi: n = h->Ny + 1; break;
      case S_vLo_i: case S_vUp_i: n = h->Nx; break;
      default: n = h->nvy; break;
    }
    (void)P;
    h->samp.s[s].assign(n, 0.0);
  }
}

void build_fft(ins2d_handle* h, int order) {
  FFTPlan& F = h->fft;
  F.Nx = h->Nx;
  F.Ny = h->Ny;
  F.Nh = h->Nx / 2 + 1;
  F.pow2x = is_pow2(h->Nx);
  F.pow2y = is_pow2(h->Ny);
  if ((F.pow2x && h->Nx > 4096) || (F.pow2y && h->Ny > 4096) || (!F.pow2x && h->Nx > 2048) ||
      (!F.pow2y && h->Ny > 2048))
    throw Unsupported("FFT Poisson: grid dimension too large (power-of-two up to 4096, other sizes up to 2048)");
  auto twid = [](int N) {
    std::vector<double> t(2 * (size_t)N);
    for (int k = 0; k < N; ++k) {
      const double a = -2.0 * M_PI * (double)k / (double)N;
      t[2 * k] = std::cos(a);
      t[2 * k + 1] = std::sin(a);
    }
    return t;
  };
  F.twx = dupload(h, twid(h->Nx));
  F.twy = dupload(h, twid(h->Ny));
  std::vector<double> ax(F.Nh), ay(h->Ny);
  const double dx = h->hx, dy = h->hy;
  if (order == 2 || h->fft_symbol == 0) {
    // operator_divergence.m:283: Ahat = (dx*dy)*4*((1/dx^2) sin^2(pi I/Npx) + (1/dy^2) sin^2(pi J/Npy))
    F.c = (dx * dy) * 4.0;
    for (int k = 0; k < F.Nh; ++k) {
      const double sn = std::sin((double)k * M_PI / (double)h->Nx);
      ax[k] = (1.0 / (dx * dx)) * (sn * sn);
    }
    for (int k = 0; k < h->Ny; ++k) {
      const double sn = std::sin((double)k * M_PI / (double)h->Ny);
      ay[k] = (1.0 / (dy * dy)) * (sn * sn);
    }
  } else {
    // exact symbol of the fourth-order A = (alfa M1 - M3) Om^-1 G with
    // Om = 72 hx hy: -A hat = 4 hy/(72 hx) (81 s1 - 3 s3)^2 + (same in y)
    F.c = 1.0;
    for (int k = 0; k < F.Nh; ++k) {
      const double t = (double)k * M_PI / (double)h->Nx;
      const double m = 81.0 * std::sin(t) - 3.0 * std::sin(3.0 * t);
      ax[k] = 4.0 * dy / (72.0 * dx) * m * m;
    }
    for (int k = 0; k < h->Ny; ++k) {
      const double t = (double)k * M_PI / (double)h->Ny;
      const double m = 81.0 * std::sin(t) - 3.0 * std::sin(3.0 * t);
      ay[k] = 4.0 * dx / (72.0 * dy) * m * m;
    }
  }
  F.ax = dupload(h, ax);
  F.ay = dupload(h, ay);
  // the row-major spectra H are needed only where a direct DFT (length not a
  // power of two in [16, 4096]) sits on one side of the transposes
  // the row-major spectra H are needed only where a direct DFT (length not a
  // power of two in [16, 4096]) sits on one side of the transposes
  F.H = fft_register_lengths(F.Nx, F.Ny) ? nullptr : dalloc<double>(h, 2 * (size_t)F.Nh * h->Ny);
  F.HT = dalloc<double>(h, 2 * (size_t)F.Nh * h->Ny);
  F.fmax2 = dalloc<unsigned long long>(h, 64);   // the early-exit maxima (kFmaxSlots)
  h->has_fft = true;
}

// max |x|, sum w x, sum w x^2 (w = wx[i] wy[j], or 1): partials of one
// launch into slot `slot` of h->part (kStatSlots slots), read back together
// by stats_collect (one host wait for several reductions)
constexpr int kStatSlots = 5;
void stats_launch(ins2d_handle* h, int slot, const double* x, int nx, int ny, const double* wx, const double* wy) {
  launch_reduce_stats(h->st, x, nx, ny, wx, wy, h->part + 4 * (size_t)h->nred * slot, h->nred);
}
void stats_collect(ins2d_handle* h, int nslots, double (*out)[3]) {
  std::vector<double> part(4 * (size_t)h->nred * nslots);
  HIPCHK(hipMemcpyAsync(part.data(), h->part, part.size() * sizeof(double), hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  for (int s = 0; s < nslots; ++s) {
    const double* q = part.data() + 4 * (size_t)h->nred * s;
    out[s][0] = out[s][1] = out[s][2] = 0.0;
    for (int b = 0; b < h->nred; ++b) {
      out[s][0] = std::max(out[s][0], q[4 * b]);
      out[s][1] += q[4 * b + 1];
      out[s][2] += q[4 * b + 2];
    }
  }
}
void stats_reduce(ins2d_handle* h, const double* x, int nx, int ny, const double* wx, const double* wy, double out[3]) {
  stats_launch(h, 0, x, nx, ny, wx, wy);
  double o[1][3];
  stats_collect(h, 1, o);
  out[0] = o[0][0];
  out[1] = o[0][1];
  out[2] = o[0][2];
}

int fail(ins2d_handle* h, const std::exception& e, int code) {
  if (h) h->err = e.what();
  return code;
}

#define API_TRY try {
#define API_CATCH(h)                                    \
  }                                                     \
  catch (const NoConv& e) { return fail(h, e, INS2D_E_NOCONV); } \
  catch (const Unsupported& e) { return fail(h, e, INS2D_E_UNSUPPORTED); } \
  catch (const ArgError& e) { return fail(h, e, INS2D_E_ARG); } \
  catch (const HipError& e) { return fail(h, e, INS2D_E_HIP); } \
  catch (const std::exception& e) { return fail(h, e, INS2D_E_ARG); }

}  // namespace

static std::string g_create_err;

namespace ins2d {
// the error text of the handle-free entry points (ins2d_last_error(NULL))
void set_handle_free_error(const std::string& msg) { g_create_err = msg; }
}  // namespace ins2d

extern "C" {

const char* ins2d_last_error(const ins2d_handle* h) { return h ? h->err.c_str() : g_create_err.c_str(); }

int ins2d_create(const ins2d_case_desc* d, ins2d_handle** out) {
  if (!d || !out) return INS2D_E_ARG;
  *out = nullptr;
  auto h = std::make_unique<ins2d_handle>();
  try {
    if (d->Nx < 2 || d->Ny < 2 || !d->x || !d->y) throw ArgError("bad grid");
    if (d->rk_stages < 1 || d->rk_stages > 8 || !d->rk_A || !d->rk_b || !d->rk_c) throw ArgError("bad tableau");
    for (int k = 0; k < 4; ++k)
      if (d->bc_u[k] < 0 || d->bc_u[k] > 3 || d->bc_v[k] < 0 || d->bc_v[k] > 3) throw ArgError("bad boundary kind");
    h->Nx = d->Nx;
    h->Ny = d->Ny;
    h->Re = d->Re;
    h->invRe = 1.0 / d->Re;
    h->poisson = d->poisson;
    h->fft_symbol = d->fft_symbol;
    h->tol = d->cg_tol;
    h->maxit = d->cg_maxit;
    h->p_add = d->p_add_solve;
    h->order4 = d->order4;
    HIPCHK(hipStreamCreate(&h->st));
    bool allper = true;
    for (int k = 0; k < 4; ++k) allper = allper && d->bc_u[k] == INS2D_BC_PER && d->bc_v[k] == INS2D_BC_PER;
    double hxmin = 1e300, hxmax = 0, hymin = 1e300, hymax = 0;
    for (int i = 0; i < d->Nx; ++i) { double t = d->x[i + 1] - d->x[i]; hxmin = std::min(hxmin, t); hxmax = std::max(hxmax, t); }
    for (int j = 0; j < d->Ny; ++j) { double t = d->y[j + 1] - d->y[j]; hymin = std::min(hymin, t); hymax = std::max(hymax, t); }
    const bool uniform = (hxmax - hxmin) <= 1e-14 && (hymax - hymin) <= 1e-14;
    if (allper && (d->Nx < 8 || d->Ny < 8)) throw Unsupported("periodic grids need at least 8 cells per direction");
    for (int k = 0; k < 4; ++k) { h->bc_u[k] = d->bc_u[k]; h->bc_v[k] = d->bc_v[k]; }
    // fourth order: the periodic uniform kernels with the FFT solve, else the
    // assembled operators with the direct solve
    bool w4 = false;
    if (d->order4) {
      if (allper && uniform && d->kernels != 1 && h->poisson == INS2D_POISSON_FFT) {
        if (d->alfa != 81.0 || d->beta != 9.0 / 8.0)
          throw Unsupported("the periodic fourth-order kernels need alfa = 81, beta = 9/8");
        h->path = 4;
      } else {
        if (h->poisson != INS2D_POISSON_DIRECT)
          throw Unsupported("fourth order solves the pressure by FFT (periodic uniform grids) or directly (poisson = 1)");
        w4 = true;
      }
    } else if (allper && uniform && d->kernels != 1) {
      h->path = 2;
    }
    if (h->poisson == INS2D_POISSON_FFT && !(allper && uniform))
      throw Unsupported("FFT Poisson needs periodic boundaries and a uniform grid (operator_divergence.m:268-290)");
    if (h->poisson != INS2D_POISSON_FFT && h->poisson != INS2D_POISSON_CG && h->poisson != INS2D_POISSON_CG_IC &&
        h->poisson != INS2D_POISSON_DIRECT && h->poisson != INS2D_POISSON_PCG_ILU)
      throw ArgError("unknown poisson mode");
    if (!d->order4) {
      CaseDesc cd;
      cd.Nx = d->Nx;
      cd.Ny = d->Ny;
      cd.x.assign(d->x, d->x + d->Nx + 1);
      cd.y.assign(d->y, d->y + d->Ny + 1);
      cd.u_left = d->bc_u[0]; cd.u_right = d->bc_u[1]; cd.u_low = d->bc_u[2]; cd.u_up = d->bc_u[3];
      cd.v_left = d->bc_v[0]; cd.v_right = d->bc_v[1]; cd.v_low = d->bc_v[2]; cd.v_up = d->bc_v[3];
      cd.Re = d->Re;
      cd.order4 = 0;
      h->plan.reset(new Plan());
      build_plan(cd, *h->plan);
      const Plan& P = *h->plan;
      h->nux = P.Nux_in; h->nuy = P.Nuy_in; h->nvx = P.Nvx_in; h->nvy = P.Nvy_in;
      upload_comp_ops(h.get(), P.cu, h->hCu);
      upload_comp_ops(h.get(), P.cv, h->hCv);
      {
        // closed form at the interior points of uniform grids
        UniForm g{};
        UniArrays ua;
        std::vector<unsigned char> fu, fv;
        uniform_maps(P, g, ua, fu, fv);
        if (g.ahx) {   // stretched grid: the spacing arrays on the device
          g.ahx = dupload(h.get(), ua.hx); g.ahy = dupload(h.get(), ua.hy);
          g.agx = dupload(h.get(), ua.gx); g.agy = dupload(h.get(), ua.gy);
          g.arhx = dupload(h.get(), ua.rhx); g.arhy = dupload(h.get(), ua.rhy);
          g.argx = dupload(h.get(), ua.rgx); g.argy = dupload(h.get(), ua.rgy);
        }
        h->hCu.uni = h->hCv.uni = g;
        h->hCu.uni.fast = fu.empty() ? nullptr : dupload(h.get(), fu);
        h->hCv.uni.fast = fv.empty() ? nullptr : dupload(h.get(), fv);
        std::vector<RTile> tiles;
        if (split_tiles(0, 0, P.Nux_in, P.Nuy_in, fu, 64, tiles, rhs_split_rows()) &&
            split_tiles(0, 1, P.Nvx_in, P.Nvy_in, fv, 64, tiles, rhs_split_rows())) {
          h->split = dupload(h.get(), tiles);
          h->nsplit = (int)tiles.size();
        }
      }
      fill_samples_default(h.get());
      upload_comp_bc(h.get(), P.cu, h->hCu);
      upload_comp_bc(h.get(), P.cv, h->hCv);
      push_comps(h.get());
      // Laplacian and Jacobi preconditioner
      h->lap.L1 = up_kron(h.get(), P.L1);
      h->lap.L2 = up_kron(h.get(), P.L2);
      std::vector<double> dinv(P.Adiag.size());
      for (size_t k = 0; k < dinv.size(); ++k) dinv[k] = 1.0 / P.Adiag[k];
      h->lap.dinv = dupload(h.get(), dinv);
      h->lap.nx = P.Npx;
      h->lap.ny = P.Npy;
      h->ox = dupload(h.get(), P.cu.ox);
      h->oy = dupload(h.get(), P.cu.oy);
      h->ovx = dupload(h.get(), P.cv.ox);
      h->ovy = dupload(h.get(), P.cv.oy);
      h->hx = P.hx[0];
      h->hy = P.hy[0];
    } else if (w4) {
      if ((long)d->Nx * d->Ny > kW4MaxNp)
        throw Unsupported("fourth order with the direct solve: at most " + std::to_string(kW4MaxNp) +
                          " pressure cells (dense inverse)");
      CaseDesc cd;
      cd.Nx = d->Nx;
      cd.Ny = d->Ny;
      cd.x.assign(d->x, d->x + d->Nx + 1);
      cd.y.assign(d->y, d->y + d->Ny + 1);
      cd.u_left = d->bc_u[0]; cd.u_right = d->bc_u[1]; cd.u_low = d->bc_u[2]; cd.u_up = d->bc_u[3];
      cd.v_left = d->bc_v[0]; cd.v_right = d->bc_v[1]; cd.v_low = d->bc_v[2]; cd.v_up = d->bc_v[3];
      cd.Re = d->Re;
      cd.order4 = 1;
      cd.alfa = d->alfa;
      cd.beta = d->beta;
      h->w4.reset(new W4Ops());
      try {
        build_w4(cd, *h->w4);
      } catch (const std::runtime_error& e) {
        throw Unsupported(e.what());
      }
      const W4Ops& W = *h->w4;
      h->nux = W.nux; h->nuy = W.nuy; h->nvx = W.nvx; h->nvy = W.nvy;
      W4Dev& D = h->w4d;
      D.NV = W.NV;
      D.Np = W.Np;
      D.Nf = W.Nf;
      D.Af = up_csr(h.get(), W.Af);
      D.If = up_csr(h.get(), W.If);
      D.Cw = up_csr(h.get(), W.Cw);
      D.Diff = up_csr(h.get(), W.Diff);
      D.G = up_csr(h.get(), W.G);
      D.M = up_csr(h.get(), W.M);
      D.Om = dupload(h.get(), W.Om);
      D.yA = dalloc<double>(h.get(), W.Nf);
      D.yI = dalloc<double>(h.get(), W.Nf);
      D.yD = dalloc<double>(h.get(), W.NV);
      D.yM = dalloc<double>(h.get(), W.Np);
      D.w = dalloc<double>(h.get(), W.Nf);
      fill_samples_default(h.get());
      w4_upload_bc(h.get());
    } else {
      h->nux = h->Nx; h->nuy = h->Ny; h->nvx = h->Nx; h->nvy = h->Ny;
    }
    h->one = dupload(h.get(), std::vector<double>{1.0});
    if (h->path) {
      h->nux = h->Nx; h->nuy = h->Ny; h->nvx = h->Nx; h->nvy = h->Ny;
      h->hx = (d->x[d->Nx] - d->x[0]) / d->Nx;
      h->hy = (d->y[d->Ny] - d->y[0]) / d->Ny;
      h->hx = d->x[1] - d->x[0];
      h->hy = d->y[1] - d->y[0];
      const double f = (h->path == 4) ? 72.0 : 1.0;
      std::vector<double> ox(h->Nx, f * h->hx), oy(h->Ny, h->hy);
      h->ox = dupload(h.get(), ox);
      h->oy = dupload(h.get(), oy);
      h->ovx = h->ox;
      h->ovy = h->oy;
      if (!h->plan) fill_samples_default(h.get());
    }
    h->Nu = h->nux * h->nuy;
    h->Nv = h->nvx * h->nvy;
    h->NV = h->Nu + h->Nv;
    h->Np = h->Nx * h->Ny;
    if (h->poisson == INS2D_POISSON_FFT) build_fft(h.get(), h->path == 4 ? 4 : 2);
    // tableau (time_ERK.m:21-32)
    const int s = d->rk_stages;
    h->s = s;
    h->At.assign((size_t)s * s, 0.0);
    h->ct.assign(s, 0.0);
    for (int i = 0; i < s; ++i) {
      for (int j = 0; j < s; ++j) h->At[i * s + j] = (i < s - 1) ? d->rk_A[(i + 1) * s + j] : d->rk_b[j];
      h->ct[i] = (i < s - 1) ? d->rk_c[i + 1] : 1.0;
    }
    for (int i = 0; i < s; ++i)
      for (int j = i + 1; j < s; ++j)
        if (h->At[i * s + j] != 0.0) throw ArgError("tableau is not explicit");
    h->keep.assign(s, false);
    for (int j = 0; j < s; ++j)
      for (int i = j + 1; i < s; ++i)
        if (h->At[i * s + j] != 0.0) h->keep[j] = true;
    for (int i = 1; i < s; ++i)
      if (h->ct[i] == 0.0) throw Unsupported("zero stage time increment");
    // buffers
    h->V = dalloc<double>(h.get(), h->NV);
    h->Vb = dalloc<double>(h.get(), h->NV);
    h->Vt = dalloc<double>(h.get(), h->NV);
    h->tV = dalloc<double>(h.get(), h->NV);
    h->kv.assign(s, nullptr);
    for (int j = 0; j < s; ++j)
      if (h->keep[j]) h->kv[j] = dalloc<double>(h.get(), h->NV);
    h->p = dalloc<double>(h.get(), h->Np);
    h->f = dalloc<double>(h.get(), h->Np);
    h->dp = dalloc<double>(h.get(), h->Np);
    HIPCHK(hipMemset(h->V, 0, sizeof(double) * h->NV));
    HIPCHK(hipMemset(h->p, 0, sizeof(double) * h->Np));
    if (h->poisson != INS2D_POISSON_FFT) {
      h->cr = dalloc<double>(h.get(), h->Np);
      h->cz = dalloc<double>(h.get(), h->Np);
      h->cp = dalloc<double>(h.get(), h->Np);
      h->cq = dalloc<double>(h.get(), h->Np);
      h->cp2 = dalloc<double>(h.get(), h->Np);
      h->d_it = dalloc<int>(h.get(), 1);
      h->d_rel = dalloc<double>(h.get(), 1);
      h->cg.nblk = (int)std::min<long>(2048, std::max<long>(1, (h->Np + 255) / 256));
      h->cg.s = dalloc<double>(h.get(), 8);
      h->cg.flag = dalloc<int>(h.get(), 4);
      h->cg.part = dalloc<double>(h.get(), 2 * (size_t)h->cg.nblk);
      h->cg.e = dalloc<double>(h.get(), 4 + (size_t)h->cg.nblk);
      h->cg.maxit = h->maxit;
      if ((d->cg_precond == 1 && (h->poisson == INS2D_POISSON_CG || h->poisson == INS2D_POISSON_CG_IC)) ||
          h->poisson == INS2D_POISSON_PCG_ILU) {
        if (h->Np > (1 << 20)) throw Unsupported("IC(0) / ILU(0) preconditioner: at most 2^20 pressure cells");
        IC0Host I;
        build_ic0_host(*h->plan, I);
        h->dic.nx = h->Nx;
        h->dic.ny = h->Ny;
        h->dic.D = dupload(h.get(), I.D);
        h->dic.cw = dupload(h.get(), I.cw);
        h->dic.cs = dupload(h.get(), I.cs);
        h->dic.cwx = I.cwx.empty() ? nullptr : dupload(h.get(), I.cwx);
        h->dic.cwy = I.cwy.empty() ? nullptr : dupload(h.get(), I.cwy);
        // the ILU(0) pivots of a symmetric 5-point matrix are those of IC(0)
        // (the same recurrence); mode 2 uses M = L1 L1^T of the unit factor
        h->dic.lu = h->poisson == INS2D_POISSON_PCG_ILU ? 1 : 0;
        h->ic0 = true;
      } else if (d->cg_precond == 3 && (h->poisson == INS2D_POISSON_CG || h->poisson == INS2D_POISSON_CG_IC)) {
        h->fdm_pcg = true;
        h->fw1 = dalloc<double>(h.get(), h->Np);
        h->fw2 = dalloc<double>(h.get(), h->Np);
      } else if (d->cg_precond < 0 || d->cg_precond > 3) {
        throw ArgError("cg_precond must be 0 (Jacobi), 1 (IC(0)), 2 (two-level) or 3 (fast diagonalization)");
      }
      if (h->w4) {
        // dense inverse of the pressure matrix: Gauss-Jordan on C = -A, plus
        // e e^T / Np when A has the constant null vector; the solve then
        // returns the zero-mean solution (that of the bordered system
        // [A e; e^T 0])
        const int n = h->Np;
        const double shift = h->w4->singular ? 1.0 / n : 0.0;
        std::vector<double> C((size_t)n * n, shift);
        for (int i = 0; i < n; ++i)
          for (auto& kv : h->w4->A.r[i]) C[(size_t)i * n + kv.first] -= kv.second;
        h->w4d.K = dupload(h.get(), C);
        w4_invert(h->st, h->w4d.K, n, h->cr, h->cz, shift);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(h->st));
      } else if (h->poisson == INS2D_POISSON_DIRECT || h->fdm_pcg) {
        FDMHost F;
        build_fdm_host(*h->plan, F);
        h->fdm.nx = F.nx;
        h->fdm.ny = F.ny;
        h->fdm.nx0 = F.nx0;
        h->fdm.ny0 = F.ny0;
        h->fdm.Vx = dupload(h.get(), F.Vx);
        h->fdm.VxT = dupload(h.get(), F.VxT);
        h->fdm.Vy = dupload(h.get(), F.Vy);
        h->fdm.VyT = dupload(h.get(), F.VyT);
        h->fdm.lx = dupload(h.get(), F.lx);
        h->fdm.ly = dupload(h.get(), F.ly);
      }
    }
    h->nred = (int)std::min<long>(1024, std::max<long>(1, ((long)std::max(h->NV, h->Np) + 255) / 256));
    h->part = dalloc<double>(h.get(), 4 * (size_t)h->nred * kStatSlots);
    h->ee_part = dalloc<double>(h.get(), std::max(64, (h->Ny + 15) / 16));
    HIPCHK(hipDeviceSynchronize());
  } catch (const std::exception& e) {
    g_create_err = e.what();
    int code = INS2D_E_ARG;
    if (dynamic_cast<const HipError*>(&e)) code = INS2D_E_HIP;
    if (dynamic_cast<const Unsupported*>(&e)) code = INS2D_E_UNSUPPORTED;
    for (void* p : h->allocs) (void)hipFree(p);
    for (void* p : h->yv_allocs) (void)hipFree(p);
    h->allocs.clear();
    h->yv_allocs.clear();
    if (h->st) (void)hipStreamDestroy(h->st);
    h->st = nullptr;
    return code;
  }
  ins2d_handle* hp = h.release();
  if (d->cg_precond == 2 && (hp->poisson == INS2D_POISSON_CG || hp->poisson == INS2D_POISSON_CG_IC)) {
    const int rc = ins2d_set_poisson_accel(hp, nullptr, 2, 0);
    if (rc != INS2D_OK) {
      g_create_err = hp->err;
      ins2d_destroy(hp);
      return rc;
    }
  }
  *out = hp;
  return INS2D_OK;
}

void ins2d_destroy(ins2d_handle* h) {
  if (!h) return;
  if (h->st) (void)hipStreamSynchronize(h->st);
  if (h->cg_exec) { (void)hipGraphExecDestroy(h->cg_exec); (void)hipGraphDestroy(h->cg_graph); }
  if (h->tl_exec) { (void)hipGraphExecDestroy(h->tl_exec); (void)hipGraphDestroy(h->tl_graph); }
  for (auto& pe : h->pend) { h->ev_free.push_back(pe.a); h->ev_free.push_back(pe.b); }
  for (auto e : h->ev_free) (void)hipEventDestroy(e);
  for (void* p : h->allocs) (void)hipFree(p);
  for (void* p : h->yv_allocs) (void)hipFree(p);
  for (void* p : h->rate_allocs) (void)hipFree(p);
  if (h->ver_dev) (void)hipFree(h->ver_dev);
  if (h->ver_comp) (void)hipFree(h->ver_comp);
  if (h->ver_pin) (void)hipHostFree(h->ver_pin);
  if (h->ver_ev) (void)hipEventDestroy(h->ver_ev);
  if (h->st) (void)hipStreamDestroy(h->st);
  delete h;
}

int ins2d_grid_info(const ins2d_handle* h, int* out, double* xin, double* yin) {
  if (!h || !out) return INS2D_E_ARG;
  out[0] = h->nux; out[1] = h->nuy; out[2] = h->nvx; out[3] = h->nvy;
  out[4] = h->Nx; out[5] = h->Ny; out[6] = h->path;
  if (h->plan) {
    if (xin) std::copy(h->plan->xin.begin(), h->plan->xin.end(), xin);
    if (yin) std::copy(h->plan->yin.begin(), h->plan->yin.end(), yin);
  } else if (h->w4) {
    if (xin) std::copy(h->w4->xin.begin(), h->w4->xin.end(), xin);
    if (yin) std::copy(h->w4->yin.begin(), h->w4->yin.end(), yin);
  }
  return INS2D_OK;
}

int ins2d_set_bc(ins2d_handle* h, double t, const double* const* samples) {
  if (!h || !samples) return INS2D_E_ARG;
  (void)t;
  API_TRY
  HIPCHK(hipStreamSynchronize(h->st));
  for (int s = 0; s < S_NSRC; ++s) {
    if (!samples[s]) throw ArgError("missing boundary sample");
    std::copy(samples[s], samples[s] + h->samp.s[s].size(), h->samp.s[s].begin());
  }
  h->have_bc = true;
  if (h->w4) w4_upload_bc(h);
  if (h->plan) {
    for (void* p : h->yv_allocs) (void)hipFree(p);
    h->yv_allocs.clear();
    upload_comp_bc(h, h->plan->cu, h->hCu);
    upload_comp_bc(h, h->plan->cv, h->hCv);
    push_comps(h);
    if (h->have_rate) upload_rate(h);
  }
  return INS2D_OK;
  API_CATCH(h)
}

int ins2d_set_bc_rate(ins2d_handle* h, const double* const* rates) {
  if (!h) return INS2D_E_ARG;
  API_TRY
  if (h->w4 && rates) throw Unsupported("time-dependent boundary values are not provided with fourth-order walls");
  HIPCHK(hipStreamSynchronize(h->st));
  if (!rates) {
    h->have_rate = false;
    for (void* p : h->rate_allocs) (void)hipFree(p);
    h->rate_allocs.clear();
    return INS2D_OK;
  }
  const int src[4] = {S_uLe_i, S_uRi_i, S_vLo_i, S_vUp_i};
  for (int q = 0; q < 4; ++q) {
    if (!rates[q]) throw ArgError("missing boundary rate sample");
    h->rate[q].assign(rates[q], rates[q] + h->samp.s[src[q]].size());
  }
  h->have_rate = true;
  upload_rate(h);
  return INS2D_OK;
  API_CATCH(h)
}

int ins2d_set_state(ins2d_handle* h, const double* V, const double* p) {
  if (!h) return INS2D_E_ARG;
  API_TRY
  if (V) HIPCHK(hipMemcpyAsync(h->V, V, sizeof(double) * h->NV, hipMemcpyHostToDevice, h->st));
  if (p) HIPCHK(hipMemcpyAsync(h->p, p, sizeof(double) * h->Np, hipMemcpyHostToDevice, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  return INS2D_OK;
  API_CATCH(h)
}

int ins2d_get_state(ins2d_handle* h, double* V, double* p) {
  if (!h) return INS2D_E_ARG;
  API_TRY
  if (V) HIPCHK(hipMemcpyAsync(V, h->V, sizeof(double) * h->NV, hipMemcpyDeviceToHost, h->st));
  if (p) HIPCHK(hipMemcpyAsync(p, h->p, sizeof(double) * h->Np, hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  return INS2D_OK;
  API_CATCH(h)
}

int ins2d_rhs(ins2d_handle* h, const double* V, const double* p, double t, double* F, double* maxres) {
  if (!h || !V || !p) return INS2D_E_ARG;
  (void)t;
  API_TRY
  HIPCHK(hipMemcpyAsync(h->Vb, V, sizeof(double) * h->NV, hipMemcpyHostToDevice, h->st));
  HIPCHK(hipMemcpyAsync(h->dp, p, sizeof(double) * h->Np, hipMemcpyHostToDevice, h->st));
  rhs(h, h->Vb, h->dp, RHS_WITH_P, h->tV, kNoCombine, nullptr);
  if (F) HIPCHK(hipMemcpyAsync(F, h->tV, sizeof(double) * h->NV, hipMemcpyDeviceToHost, h->st));
  if (maxres) {
    double a[3], b[3];
    stats_reduce(h, h->tV, h->nux, h->nuy, nullptr, nullptr, a);
    stats_reduce(h, h->tV + h->Nu, h->nvx, h->nvy, nullptr, nullptr, b);
    *maxres = std::max(a[0], b[0]);
  }
  HIPCHK(hipStreamSynchronize(h->st));
  return INS2D_OK;
  API_CATCH(h)
}

int ins2d_poisson(ins2d_handle* h, const double* f, double* dp, int* iters, double* relres) {
  if (!h || !f || !dp) return INS2D_E_ARG;
  API_TRY
  HIPCHK(hipMemcpyAsync(h->f, f, sizeof(double) * h->Np, hipMemcpyHostToDevice, h->st));
  poisson(h, h->f, h->dp);
  HIPCHK(hipMemcpyAsync(dp, h->dp, sizeof(double) * h->Np, hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  if (iters) *iters = h->poisson == INS2D_POISSON_FFT ? 0 : h->last_it;
  if (relres) *relres = h->poisson == INS2D_POISSON_FFT ? 0.0 : h->last_rel;
  return INS2D_OK;
  API_CATCH(h)
}

int ins2d_input_check(ins2d_handle* h, double t, int p_initial) {
  if (!h) return INS2D_E_ARG;
  (void)t;
  API_TRY
  // maxdiv = max|M V + yM| (check_conservation.m:37)
  div(h, h->V, nullptr, 1.0, 0.0, 1.0, 1.0, h->f);
  double st[3];
  stats_reduce(h, h->f, h->Nx, h->Ny, nullptr, nullptr, st);
  if (st[0] > 1e-12) {
    // input_check.m:104-107: V = V - Om^-1 G A^-1 (M V + yM)
    poisson(h, h->f, h->dp);
    HIPCHK(hipMemcpyAsync(h->Vb, h->V, sizeof(double) * h->NV, hipMemcpyDeviceToDevice, h->st));
    proj(h, h->Vb, nullptr, h->dp, 1.0, 1.0, h->V);
  }
  if (p_initial) {
    if (h->have_rate) additional_solve(h, h->V, h->dCu_rate, h->dCv_rate);
    else additional_solve(h, h->V);
  }
  HIPCHK(hipStreamSynchronize(h->st));
  return INS2D_OK;
  API_CATCH(h)
}

int ins2d_step(ins2d_handle* h, double tn, double dt, int nsteps) {
  if (!h || nsteps < 0 || !(dt > 0)) return INS2D_E_ARG;
  API_TRY
  if (h->fdm_pcg && !h->accel && !h->prof && nsteps > 0) {
    // fast-diagonalization PCG: one device iteration per solve and one host
    // wait per call (see fdm_dev); the saved state lets the call start over
    // on the host-checked path should a solve need a second iteration
    if (!h->d_missed) {
      h->d_missed = dalloc<int>(h, 1);
      h->Vsave = dalloc<double>(h, h->NV);
      h->psave = dalloc<double>(h, h->Np);
    }
    double* V0 = h->V;
    HIPCHK(hipMemcpyAsync(h->Vsave, h->V, sizeof(double) * h->NV, hipMemcpyDeviceToDevice, h->st));
    HIPCHK(hipMemcpyAsync(h->psave, h->p, sizeof(double) * h->Np, hipMemcpyDeviceToDevice, h->st));
    HIPCHK(hipMemsetAsync(h->d_missed, 0, sizeof(int), h->st));
    h->fdm_dev = true;
    h->k1_ready = false;
    const bool dual_fdm = dual_capable(h);
    try {
      for (int n = 0; n < nsteps; ++n) erk_step(h, tn + n * dt, dt, nullptr, dual_fdm && n + 1 < nsteps);
    } catch (...) {
      h->fdm_dev = false;
      h->k1_ready = false;
      throw;
    }
    h->fdm_dev = false;
    h->k1_ready = false;
    int missed = 0;
    double rel = 0.0;
    HIPCHK(hipMemcpyAsync(&missed, h->d_missed, sizeof(int), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipMemcpyAsync(&rel, h->cg.s + 4, sizeof(double), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    if (!missed) {
      const long long solves = (long long)nsteps * h->s;   // stages 2..s and the additional solve
      h->last_it = 1;
      h->last_rel = std::sqrt(rel);
      h->total_it += solves;
      return INS2D_OK;
    }
    if (h->V != V0) std::swap(h->V, h->Vb);
    HIPCHK(hipMemcpyAsync(h->V, h->Vsave, sizeof(double) * h->NV, hipMemcpyDeviceToDevice, h->st));
    HIPCHK(hipMemcpyAsync(h->p, h->psave, sizeof(double) * h->Np, hipMemcpyDeviceToDevice, h->st));
  }
  h->k1_ready = false;
  // the F of each step's end velocity formed once for the additional solve
  // and the next step's first stage (profiling off: the profile times the
  // launches one by one)
  const bool dual = dual_capable(h);
  for (int n = 0; n < nsteps; ++n) erk_step(h, tn + n * dt, dt, nullptr, dual && n + 1 < nsteps);
  h->k1_ready = false;
  if (h->prof) prof_collect(h);
  return INS2D_OK;
  API_CATCH(h)
}

int ins2d_step_unsteady(ins2d_handle* h, double tn, double dt, int nsteps, const double* samples,
                        const double* rates) {
  if (!h || nsteps < 0 || !(dt > 0) || (nsteps > 0 && (!samples || !rates))) return INS2D_E_ARG;
  API_TRY
  if (h->path || !h->plan)
    throw Unsupported("time-dependent boundary values need the second-order kernels with boundary terms");
  if (nsteps == 0) return INS2D_OK;
  const int s = h->s, nv = s + 2;
  size_t set_len = 0, rate_len = 0;
  for (int q = 0; q < S_NSRC; ++q) set_len += h->samp.s[q].size();
  const int src[4] = {S_uLe_i, S_uRi_i, S_vLo_i, S_vUp_i};
  for (int q = 0; q < 4; ++q) rate_len += h->samp.s[src[q]].size();
  // versions are staged for up to kChunk steps at a time
  constexpr int kChunk = 64;
  ver_reserve(h, nv * std::min(nsteps, kChunk));
  const size_t len = std::max<size_t>(h->ver_len, 1);
  BCSamples S;
  for (int n0 = 0; n0 < nsteps; n0 += kChunk) {
    const int m = std::min(kChunk, nsteps - n0);
    // the staging buffer is free once the previous chunk's copy has ended
    HIPCHK(hipEventSynchronize(h->ver_ev));
    for (int n = 0; n < m; ++n) {
      const double* smp = samples + (size_t)(n0 + n) * (s + 1) * set_len;
      for (int k = 0; k <= s; ++k) {
        unpack_samples(h, smp + k * set_len, S);
        ver_fill(h, n * nv + k, S);
      }
      const double* r = rates + (size_t)(n0 + n) * rate_len;
      const double* rq[4];
      for (int q = 0; q < 4; ++q) {
        rq[q] = r;
        r += h->samp.s[src[q]].size();
      }
      ver_fill(h, n * nv + s + 1, rate_samples(h, rq));
    }
    HIPCHK(hipMemcpyAsync(h->ver_dev, h->ver_pin, sizeof(double) * len * nv * m, hipMemcpyHostToDevice, h->st));
    HIPCHK(hipEventRecord(h->ver_ev, h->st));
    for (int n = 0; n < m; ++n) erk_step(h, tn + (n0 + n) * dt, dt, h->ver_comp + 2 * (size_t)n * nv);
    if (n0 + m < nsteps) {
      // the next chunk rewrites the device versions: this chunk must be done
      HIPCHK(hipStreamSynchronize(h->st));
    }
  }
  // the handle's own boundary state becomes that of tn + nsteps dt
  const double* last = samples + ((size_t)nsteps * (s + 1) - 1) * set_len;
  std::vector<const double*> sp(S_NSRC);
  for (int q = 0; q < S_NSRC; ++q) {
    sp[q] = last;
    last += h->samp.s[q].size();
  }
  const double* r = rates + (size_t)(nsteps - 1) * rate_len;
  const double* rq[4];
  for (int q = 0; q < 4; ++q) {
    rq[q] = r;
    r += h->samp.s[src[q]].size();
  }
  HIPCHK(hipStreamSynchronize(h->st));
  for (int q = 0; q < S_NSRC; ++q) std::copy(sp[q], sp[q] + h->samp.s[q].size(), h->samp.s[q].begin());
  for (void* p : h->yv_allocs) (void)hipFree(p);
  h->yv_allocs.clear();
  upload_comp_bc(h, h->plan->cu, h->hCu);
  upload_comp_bc(h, h->plan->cv, h->hCv);
  push_comps(h);
  for (int q = 0; q < 4; ++q) h->rate[q].assign(rq[q], rq[q] + h->samp.s[src[q]].size());
  h->have_rate = true;
  upload_rate(h);
  if (h->prof) prof_collect(h);
  return INS2D_OK;
  API_CATCH(h)
}

int ins2d_diagnostics(ins2d_handle* h, double t, double* out) {
  if (!h || !out) return INS2D_E_ARG;
  (void)t;
  API_TRY
  // the five reductions queued behind the F and divergence launches, read
  // back with one host wait
  rhs(h, h->V, h->p, RHS_WITH_P, h->tV, kNoCombine, nullptr);
  stats_launch(h, 0, h->tV, h->nux, h->nuy, nullptr, nullptr);
  stats_launch(h, 1, h->tV + h->Nu, h->nvx, h->nvy, nullptr, nullptr);
  div(h, h->V, nullptr, 1.0, 0.0, 1.0, 1.0, h->f);
  stats_launch(h, 2, h->f, h->Nx, h->Ny, nullptr, nullptr);
  if (h->w4) {
    // volumes Om = alfa Om1 - Om3 do not factor: full weight arrays as one row
    stats_launch(h, 3, h->V, h->Nu, 1, h->w4d.Om, h->one);
    stats_launch(h, 4, h->V + h->Nu, h->Nv, 1, h->w4d.Om + h->Nu, h->one);
  } else {
    stats_launch(h, 3, h->V, h->nux, h->nuy, h->ox, h->oy);
    stats_launch(h, 4, h->V + h->Nu, h->nvx, h->nvy, h->ovx, h->ovy);
  }
  double S5[kStatSlots][3];
  stats_collect(h, kStatSlots, S5);
  out[0] = std::max(S5[0][0], S5[1][0]);
  out[1] = S5[2][0];
  double umom = S5[3][1], vmom = S5[4][1], k = 0.5 * S5[3][2] + 0.5 * S5[4][2];
  // Dirichlet boundary contributions with the second-order widths
  // (check_conservation.m:45-76)
  if (h->plan || h->w4) {
    const std::vector<double>& hx = h->plan ? h->plan->hx : h->w4->hx;
    const std::vector<double>& hy = h->plan ? h->plan->hy : h->w4->hy;
    const std::vector<double>& gx = h->plan ? h->plan->gx : h->w4->gx;
    const std::vector<double>& gy = h->plan ? h->plan->gy : h->w4->gy;
    const auto& S = h->samp.s;
    if (h->bc_u[0] == BC_DIR)
      for (int j = 0; j < h->Ny; ++j) { umom += S[S_uLe_i][j] * hy[j] * gx[0]; k += 0.5 * S[S_uLe_i][j] * S[S_uLe_i][j] * hy[j] * gx[0]; }
    if (h->bc_u[1] == BC_DIR)
      for (int j = 0; j < h->Ny; ++j) { umom += S[S_uRi_i][j] * hy[j] * gx[h->Nx]; k += 0.5 * S[S_uRi_i][j] * S[S_uRi_i][j] * hy[j] * gx[h->Nx]; }
    if (h->bc_v[2] == BC_DIR)
      for (int i = 0; i < h->Nx; ++i) { vmom += S[S_vLo_i][i] * hx[i] * gy[0]; k += 0.5 * S[S_vLo_i][i] * S[S_vLo_i][i] * hx[i] * gy[0]; }
    if (h->bc_v[3] == BC_DIR)
      for (int i = 0; i < h->Nx; ++i) { vmom += S[S_vUp_i][i] * hx[i] * gy[h->Ny]; k += 0.5 * S[S_vUp_i][i] * S[S_vUp_i][i] * hx[i] * gy[h->Ny]; }
  }
  out[2] = umom;
  out[3] = vmom;
  out[4] = k;
  return INS2D_OK;
  API_CATCH(h)
}

int ins2d_set_poisson_accel(ins2d_handle* h, ins2d_nn* nn, int flags, int agg) {
  if (!h || flags < 0 || (flags & ~0xff0f)) return INS2D_E_ARG;
  API_TRY
  const int sweeps = (flags >> 8) & 0xff;
  flags &= 0xf;
  if (((flags & 4) && (flags & 3) != 3) || ((flags & 8) && ((flags & 1) || !(flags & 2))))
    throw ArgError("flag 4 needs flags 1 and 2; flag 8 needs flag 2 and excludes flag 1");
  if (flags == 0) {
    h->accel = 0;
    h->nn = nullptr;
    return INS2D_OK;
  }
  if (h->poisson != INS2D_POISSON_CG && h->poisson != INS2D_POISSON_CG_IC)
    throw Unsupported("Poisson acceleration applies to the PCG modes");
  if ((flags & 1) && !nn) throw ArgError("warm start needs a network");
  if (flags & 1) {
    int H = 0, W = 0;
    ins2d_nn_shape(nn, &H, &W);
    if (H != h->Ny || W != h->Nx) throw ArgError("network grid does not match the pressure grid");
  }
  const int nx = h->lap.nx, ny = h->lap.ny;
  auto ncoarse = [&](int a) { return (long)((nx + a - 1) / a) * ((ny + a - 1) / a); };
  if (agg == 0) {
    for (int a : {8, 16, 32})
      if (ncoarse(a) <= kTLMaxCoarse) { agg = a; break; }
    if (agg == 0) throw Unsupported("grid too large for the one-workgroup coarse solve");
  }
  if (agg != 4 && agg != 8 && agg != 16 && agg != 32) throw ArgError("aggregate size must be 4, 8, 16 or 32");
  if (ncoarse(agg) > kTLMaxCoarse) throw Unsupported("too many coarse unknowns for this aggregate size");
  TLPlan T{};
  T.L = h->lap;
  T.maxit = h->maxit;
  T.coarse = (flags & 2) ? 1 : 0;
  T.C.agg = agg;
  T.C.ncx = (nx + agg - 1) / agg;
  T.C.ncy = (ny + agg - 1) / agg;
  T.C.nx0 = T.C.ny0 = -1;
  if (flags & 2) {
    FDMHost C;
    build_coarse_host(*h->plan, agg, C);
    if (C.nx != T.C.ncx || C.ny != T.C.ncy) throw std::runtime_error("coarse grid mismatch");
    T.C.Vx = dupload(h, C.Vx);
    T.C.Vy = dupload(h, C.Vy);
    T.C.lx = dupload(h, C.lx);
    T.C.ly = dupload(h, C.ly);
    T.C.nx0 = C.nx0;
    T.C.ny0 = C.ny0;
  }
  const size_t nc = (size_t)T.C.ncx * T.C.ncy;
  T.C.rc = dalloc<double>(h, nc);
  T.C.ec = dalloc<double>(h, nc);
  T.s = dalloc<double>(h, 8);
  T.flag = dalloc<int>(h, 4);
  T.part = dalloc<double>(h, 3 * nc);
  T.en = dalloc<double>(h, nc);
  HIPCHK(hipMemset(T.s, 0, 8 * sizeof(double)));
  HIPCHK(hipMemset(T.flag, 0, 4 * sizeof(int)));
  h->tl = T;
  if (h->tl_exec) {
    HIPCHK(hipGraphExecDestroy(h->tl_exec));
    HIPCHK(hipGraphDestroy(h->tl_graph));
    h->tl_exec = nullptr;
    h->tl_graph = nullptr;
  }
  h->accel = flags;
  h->nn = (flags & 1) ? nn : nullptr;
  h->ws_sweeps = sweeps ? sweeps : 4;
  return INS2D_OK;
  API_CATCH(h)
}

int ins2d_pressure_rhs_dev(ins2d_handle* h, double t, double* f) {
  if (!h || !f) return INS2D_E_ARG;
  (void)t;
  API_TRY
  if (h->order4) throw Unsupported("pressure right-hand side is built for second order");
  rhs(h, h->V, h->p, RHS_SCALE_OM, h->tV, kNoCombine, nullptr);
  div(h, h->tV, nullptr, 1.0, 0.0, 0.0, 1.0, f);
  HIPCHK(hipStreamSynchronize(h->st));
  return INS2D_OK;
  API_CATCH(h)
}

int ins2d_poisson_dev(ins2d_handle* h, const double* f, double* dp) {
  if (!h || !f || !dp) return INS2D_E_ARG;
  API_TRY
  poisson(h, f, dp);
  HIPCHK(hipStreamSynchronize(h->st));
  return INS2D_OK;
  API_CATCH(h)
}

int ins2d_poisson_coarse_dev(ins2d_handle* h, const double* f, double* xc) {
  if (!h || !f || !xc) return INS2D_E_ARG;
  API_TRY
  if (!(h->accel & 2)) throw Unsupported("the coarse-grid solution needs the two-level preconditioner (flag 2)");
  launch_tl_coarse_solution(h->st, h->tl, f, xc, h->cr);
  HIPCHK(hipStreamSynchronize(h->st));
  return INS2D_OK;
  API_CATCH(h)
}

int ins2d_poisson_coarse(ins2d_handle* h, const double* f, double* xc) {
  if (!h || !f || !xc) return INS2D_E_ARG;
  API_TRY
  if (!(h->accel & 2)) throw Unsupported("the coarse-grid solution needs the two-level preconditioner (flag 2)");
  HIPCHK(hipMemcpyAsync(h->f, f, sizeof(double) * h->Np, hipMemcpyHostToDevice, h->st));
  launch_tl_coarse_solution(h->st, h->tl, h->f, h->dp, h->cr);
  HIPCHK(hipMemcpyAsync(xc, h->dp, sizeof(double) * h->Np, hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  return INS2D_OK;
  API_CATCH(h)
}

int ins2d_poisson_accel_stats(const ins2d_handle* h, long long* fallbacks, int* agg) {
  if (!h) return INS2D_E_ARG;
  if (fallbacks) *fallbacks = h->fallbacks;
  if (agg) *agg = h->accel ? h->tl.C.agg : 0;
  return INS2D_OK;
}

int ins2d_residual(ins2d_handle* h, const double* f, const double* x, double* relres) {
  if (!h || !f || !x || !relres) return INS2D_E_ARG;
  API_TRY
  if (!h->plan) throw Unsupported("the residual needs the second-order pressure matrix");
  HIPCHK(hipMemcpyAsync(h->f, f, sizeof(double) * h->Np, hipMemcpyHostToDevice, h->st));
  HIPCHK(hipMemcpyAsync(h->dp, x, sizeof(double) * h->Np, hipMemcpyHostToDevice, h->st));
  launch_lap(h->st, h->lap, h->dp, h->tV);
  launch_axpy(h->st, (size_t)h->Np, -1.0, h->f, h->tV);
  double a[3], b[3];
  stats_reduce(h, h->tV, h->Nx, h->Ny, nullptr, nullptr, a);
  stats_reduce(h, h->f, h->Nx, h->Ny, nullptr, nullptr, b);
  *relres = b[2] > 0.0 ? std::sqrt(a[2] / b[2]) : std::sqrt(a[2]);
  return INS2D_OK;
  API_CATCH(h)
}

int ins2d_pressure_rhs(ins2d_handle* h, double t, double* f) {
  if (!h || !f) return INS2D_E_ARG;
  (void)t;
  API_TRY
  if (h->order4) throw Unsupported("pressure right-hand side is built for second order");
  rhs(h, h->V, h->p, RHS_SCALE_OM, h->tV, kNoCombine, nullptr);
  div(h, h->tV, nullptr, 1.0, 0.0, 0.0, 1.0, h->f);
  HIPCHK(hipMemcpyAsync(f, h->f, sizeof(double) * h->Np, hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  return INS2D_OK;
  API_CATCH(h)
}

int ins2d_solver_stats(const ins2d_handle* h, int* iters, double* relres, long long* total) {
  if (!h) return INS2D_E_ARG;
  if (iters) *iters = h->last_it;
  if (relres) *relres = h->last_rel;
  if (total) *total = h->total_it;
  return INS2D_OK;
}

int ins2d_profile(ins2d_handle* h, int enable) {
  if (!h) return INS2D_E_ARG;
  API_TRY
  prof_collect(h);
  h->prof = enable != 0;
  for (int k = 0; k < ins2d_handle::kProfKinds; ++k) { h->prof_ms[k] = 0.0; h->prof_n[k] = 0; }
  return INS2D_OK;
  API_CATCH(h)
}

int ins2d_profile_read(const ins2d_handle* h, int kind, double* ms, long long* count) {
  if (!h || kind < 0 || kind >= ins2d_handle::kProfKinds) return INS2D_E_ARG;
  if (ms) *ms = h->prof_ms[kind];
  if (count) *count = h->prof_n[kind];
  return INS2D_OK;
}

int ins2d_sync(ins2d_handle* h) {
  if (!h) return INS2D_E_ARG;
  API_TRY
  HIPCHK(hipStreamSynchronize(h->st));
  if (h->prof) prof_collect(h);
  return INS2D_OK;
  API_CATCH(h)
}

}  // extern "C"
