// smx_api.hip -- host side of libsmectic: handle, device buffers, step chain,
// FP64 diagnostics and the C ABI declared in include/smectic.h.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/smectic.h"
#include "smx_launch.h"

using namespace smx;

// ---------------------------------------------------------------------------
// errors
static thread_local std::string g_err;
static int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}
#define HIPCHK(x)                                                                                  \
  do {                                                                                             \
    hipError_t e_ = (x);                                                                           \
    if (e_ != hipSuccess) return fail(SMX_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)
#define CHK(x)                   \
  do {                           \
    int r_ = (x);                \
    if (r_ != SMX_OK) return r_; \
  } while (0)

// ---------------------------------------------------------------------------
// diagnostics kernels (FP64)
namespace {

// Per-block FP64 Kahan partials of a k-space reduction over the normalised
// half spectrum psi_hat, layout [kz][kx][ky]:
//   out[3b+0] = sum w (alpha/2 (q0^2 - k^2)^2 - eps/2) |psi_hat|^2
//   out[3b+1] = sum_{k != 0} w |psi_hat|^2
//   out[3b+2] = sum_{k != 0} w |k| |psi_hat|^2
// w = 1 on the kz = 0 and kz = nz/2 planes, 2 elsewhere (Hermitian half).
struct Kahan {
  double s = 0.0, c = 0.0;
  __device__ __forceinline__ void add(double x) {
    double y = x - c, t = s + y;
    c = (t - s) - y;
    s = t;
  }
};

__global__ void k_reduce_spec(const cplx<double>* __restrict__ a, int64_t n, int nx, int ny, int nz,
                              const double* __restrict__ kx, const double* __restrict__ ky,
                              const double* __restrict__ kz, double eps, double alpha, double q02,
                              double* __restrict__ out) {
  __shared__ double red[64];
  Kahan s0, s1, s2;
  const int64_t nxy = (int64_t)nx * ny;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int iz = (int)(i / nxy);
    const int64_t r = i - iz * nxy;
    const int ix = (int)(r / ny), iy = (int)(r - (int64_t)ix * ny);
    const double k2 = kx[ix] * kx[ix] + ky[iy] * ky[iy] + kz[iz] * kz[iz];
    const double w = (iz == 0 || 2 * iz == nz) ? 1.0 : 2.0;
    const double p2 = w * (a[i].x * a[i].x + a[i].y * a[i].y);
    const double d = q02 - k2;
    s0.add((0.5 * alpha * d * d - 0.5 * eps) * p2);
    if (i != 0) {
      s1.add(p2);
      s2.add(sqrt(k2) * p2);
    }
  }
  double t0 = block_sum(s0.s, red);
  double t1 = block_sum(s1.s, red);
  double t2 = block_sum(s2.s, red);
  if (threadIdx.x == 0) {
    out[3 * blockIdx.x + 0] = t0;
    out[3 * blockIdx.x + 1] = t1;
    out[3 * blockIdx.x + 2] = t2;
  }
}

// Per-block partial of sum f_local(psi) over a real field.
__global__ void k_reduce_local(const double* __restrict__ psi, int64_t n, int nl_model, double beta, double gamma,
                               double* __restrict__ out) {
  __shared__ double red[64];
  Kahan s;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double p = psi[i], p2 = p * p;
    // energyPsi.m:225-228 (quintic) or the cubic model
    const double f = (nl_model == SMX_NL_CUBIC) ? (-beta / 3.0 * p2 * p + gamma / 4.0 * p2 * p2)
                                                : (-beta / 4.0 * p2 * p2 + gamma / 6.0 * p2 * p2 * p2);
    s.add(f);
  }
  double t = block_sum(s.s, red);
  if (threadIdx.x == 0) out[blockIdx.x] = t;
}

// Per-block partials of sum |a - b| and sum |a| (pf3d.cpp:494-502).
__global__ void k_reduce_l1(const double* __restrict__ a, const double* __restrict__ b, int64_t n,
                            double* __restrict__ out) {
  __shared__ double red[64];
  Kahan s0, s1;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    s0.add(fabs(a[i] - b[i]));
    s1.add(fabs(a[i]));
  }
  double t0 = block_sum(s0.s, red);
  double t1 = block_sum(s1.s, red);
  if (threadIdx.x == 0) {
    out[2 * blockIdx.x] = t0;
    out[2 * blockIdx.x + 1] = t1;
  }
}

// Non-finite check over a complex array: out[0] += count.
template <typename T>
__global__ void k_count_nonfinite(const cplx<T>* __restrict__ a, int64_t n, unsigned int* __restrict__ out) {
  unsigned int c = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    c += (!isfinite((double)a[i].x) || !isfinite((double)a[i].y)) ? 1u : 0u;
  if (c) atomicAdd(out, c);
}

__global__ void k_f32_to_f64(const cplx<float>* __restrict__ a, cplx<double>* __restrict__ b, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = mk((double)a[i].x, (double)a[i].y);
}

// dpsi/dt in k-space: out = aLin(k) psi_hat + N_hat (both normalised), layout [kz][kx][ky]
// (cosNoAdvMorphCurv.cpp:744, surfTrack.cpp:611).
__global__ void k_dtpsi_combine(const cplx<double>* __restrict__ psi, cplx<double>* __restrict__ nq, int64_t n,
                                int nx, int ny, const double* __restrict__ kx, const double* __restrict__ ky,
                                const double* __restrict__ kz, double eps, double alpha, double q02) {
  const int64_t nxy = (int64_t)nx * ny;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int iz = (int)(i / nxy);
    const int64_t r = i - iz * nxy;
    const int ix = (int)(r / ny), iy = (int)(r - (int64_t)ix * ny);
    const double k2 = kx[ix] * kx[ix] + ky[iy] * ky[iy] + kz[iz] * kz[iz];
    const double d = k2 - q02;
    const double aLin = eps - alpha * d * d;
    nq[i] = mk(aLin * psi[i].x + nq[i].x, aLin * psi[i].y + nq[i].y);
  }
}

__global__ void k_copy_f4(const float4* __restrict__ a, float4* __restrict__ b, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}

constexpr int kRedBlocks = 1024;
constexpr int kRedThreads = 256;

}  // namespace

// ---------------------------------------------------------------------------
// handle
template <typename T>
struct DevTables {
  T* kx = nullptr;
  T* ky = nullptr;
  T* kz = nullptr;
  cplx<T>* twx = nullptr;
  cplx<T>* twy = nullptr;
  cplx<T>* twz = nullptr;
};

struct smx_handle {
  smx_config cfg;
  hipStream_t st = nullptr;
  bool fp32 = false;
  int nzh = 0;
  int64_t N = 0, Nh = 0, nxy = 0, npencils = 0;
  double dx = 0, dy = 0, dz = 0;
  void* A = nullptr;  // stage buffer, working precision
  void* R = nullptr;  // CNAB2 carry R = C1 psi_hat - dt/2 N_hat_past
  void* P = nullptr;  // bootstrap spectrum
  cplx<double>* D1 = nullptr;  // FP64 diagnostics
  cplx<double>* D2 = nullptr;
  double* real_io = nullptr;
  double* real_prev = nullptr;
  double* partial = nullptr;  // z-pass FP64 partial sums (mixed path mean mode)
  double* mean_state = nullptr;
  double* red = nullptr;      // reduction partials
  unsigned int* flag = nullptr;
  DevTables<float> t32;
  DevTables<double> t64;
  std::vector<void*> allocs;
  bool has_state = false, has_prev = false;
  int64_t steps = 0;
};

static int dalloc(smx_handle* h, void** p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) return fail(SMX_ERR_HIP, "hipMalloc(" + std::to_string(bytes) + "): " + hipGetErrorString(e));
  h->allocs.push_back(*p);
  return SMX_OK;
}

// Periodic wavenumbers in FFT order with the Nyquist entry at +pi/dx (pf3d.cpp:162-190).
static std::vector<double> wavenumbers_periodic(int n, double d) {
  std::vector<double> v(n);
  for (int i = 0; i < n; ++i) v[i] = (i < n / 2 + 1) ? 2.0 * M_PI * i / (d * n) : -(double)(n - i) * 2.0 * M_PI / (d * n);
  return v;
}

// e^{-2 pi i m / n}, m < n, evaluated in extended precision and rounded once.
static std::vector<cplx<double>> twiddles(int n) {
  std::vector<cplx<double>> t(n);
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int m = 0; m < n; ++m) {
    long double a = -2.0L * pi * (long double)m / (long double)n;
    t[m].x = (double)cosl(a);
    t[m].y = (double)sinl(a);
  }
  return t;
}

template <typename T>
static int upload_tables(smx_handle* h, DevTables<T>& t) {
  const smx_config& c = h->cfg;
  auto put_real = [&](T** dst, const std::vector<double>& v) -> int {
    std::vector<T> tmp(v.begin(), v.end());
    CHK(dalloc(h, (void**)dst, sizeof(T) * tmp.size()));
    HIPCHK(hipMemcpy(*dst, tmp.data(), sizeof(T) * tmp.size(), hipMemcpyHostToDevice));
    return SMX_OK;
  };
  auto put_tw = [&](cplx<T>** dst, int n) -> int {
    auto tw = twiddles(n);
    std::vector<cplx<T>> tmp(n);
    for (int i = 0; i < n; ++i) {
      tmp[i].x = (T)tw[i].x;
      tmp[i].y = (T)tw[i].y;
    }
    CHK(dalloc(h, (void**)dst, sizeof(cplx<T>) * n));
    HIPCHK(hipMemcpy(*dst, tmp.data(), sizeof(cplx<T>) * n, hipMemcpyHostToDevice));
    return SMX_OK;
  };
  CHK(put_real(&t.kx, wavenumbers_periodic(c.nx, h->dx)));
  CHK(put_real(&t.ky, wavenumbers_periodic(c.ny, h->dy)));
  CHK(put_real(&t.kz, wavenumbers_periodic(c.nz, h->dz)));
  CHK(put_tw(&t.twx, c.nx));
  CHK(put_tw(&t.twy, c.ny));
  CHK(put_tw(&t.twz, c.nz));
  return SMX_OK;
}

template <typename T>
static PParams<T> params(const smx_handle* h, const DevTables<T>& t, bool mixed_mean) {
  const smx_config& c = h->cfg;
  PParams<T> p;
  p.nx = c.nx;
  p.ny = c.ny;
  p.nz = c.nz;
  p.nzh = h->nzh;
  p.nxy = h->nxy;
  p.scale = (T)(1.0 / (double)h->N);
  p.eps = (T)c.eps;
  p.alpha = (T)c.alpha;
  p.q02 = (T)(c.q0 * c.q0);
  p.beta = (T)c.beta;
  p.gamma = (T)c.gamma;
  p.dtd2 = (T)(0.5 * c.dt);
  p.nl_model = c.nl_model;
  p.mid = (c.w_mid != 0.0) ? c.mid_plane : -1;
  p.pen = (T)(2.0 * c.w_mid / (h->dz * h->dz));
  p.kx = t.kx;
  p.ky = t.ky;
  p.kz = t.kz;
  p.twx = t.twx;
  p.twy = t.twy;
  p.twz = t.twz;
  p.partial = mixed_mean ? h->partial : nullptr;
  p.npartial = PeriodicOps<T>::zpass_blocks(p);
  p.mean_state = h->mean_state;
  const double q02 = c.q0 * c.q0, aLin0 = c.eps - c.alpha * q02 * q02;
  p.m_scale = 1.0 / (double)h->N;
  p.m_dtd2 = 0.5 * c.dt;
  p.m_c1 = 1.0 + p.m_dtd2 * aLin0;
  p.m_c2 = 1.0 - p.m_dtd2 * aLin0;
  return p;
}

#define HIPERR(x)                                                                                  \
  do {                                                                                             \
    hipError_t e_ = (x);                                                                           \
    if (e_ != hipSuccess) return fail(SMX_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

// ---- working-precision chains ----
template <typename T>
static DevTables<T>& tables(smx_handle* h);
template <>
DevTables<float>& tables<float>(smx_handle* h) { return h->t32; }
template <>
DevTables<double>& tables<double>(smx_handle* h) { return h->t64; }

// Bootstrap from real psi in h->real_io: P = psi_hat (normalised), R = C1 psi_hat - dt/2 N_hat(psi),
// A = x-inverse(psi_hat).  pf3d.cpp:363-374, pf3dMorph.cpp:451-472.
template <typename T>
static int bootstrap(smx_handle* h) {
  PParams<T> p = params<T>(h, tables<T>(h), h->fp32);
  cplx<T>* A = (cplx<T>*)h->A;
  cplx<T>* R = (cplx<T>*)h->R;
  cplx<T>* P = (cplx<T>*)h->P;
  using Ops = PeriodicOps<T>;
  HIPERR(Ops::zpass(Z_IN_REAL, Z_OUT_SPEC, Z_OP_ID, nullptr, A, h->real_io, nullptr, p, h->st));
  HIPERR(Ops::ypass(A, A, p.ny, h->npencils, -1, p.twy, h->st));
  HIPERR(Ops::xpass(X_FWD_SCALED, A, P, nullptr, nullptr, p, h->st));
  HIPERR(Ops::zpass(Z_IN_REAL, Z_OUT_SPEC, Z_OP_NL, nullptr, A, h->real_io, nullptr, p, h->st));
  HIPERR(Ops::ypass(A, A, p.ny, h->npencils, -1, p.twy, h->st));
  HIPERR(Ops::xpass(X_BOOT, A, A, R, P, p, h->st));
  return SMX_OK;
}

// Kernel slots of one step, in launch order.
static const char* kSlotNames[] = {"k_ypass_inv", "k_zpass_nl", "k_ypass_fwd", "k_xpass_step"};
constexpr int kNumSlots = 4;

template <typename T>
static int step_once(smx_handle* h, const PParams<T>& p, hipEvent_t* ev) {
  cplx<T>* A = (cplx<T>*)h->A;
  cplx<T>* R = (cplx<T>*)h->R;
  using Ops = PeriodicOps<T>;
  if (ev) HIPERR(hipEventRecord(ev[0], h->st));
  HIPERR(Ops::ypass(A, A, p.ny, h->npencils, +1, p.twy, h->st));
  if (ev) HIPERR(hipEventRecord(ev[1], h->st));
  HIPERR(Ops::zpass(Z_IN_SPEC, Z_OUT_SPEC, Z_OP_NL, A, A, nullptr, nullptr, p, h->st));
  if (ev) HIPERR(hipEventRecord(ev[2], h->st));
  HIPERR(Ops::ypass(A, A, p.ny, h->npencils, -1, p.twy, h->st));
  if (ev) HIPERR(hipEventRecord(ev[3], h->st));
  HIPERR(Ops::xpass(X_STEP, A, A, R, nullptr, p, h->st));
  if (ev) HIPERR(hipEventRecord(ev[4], h->st));
  return SMX_OK;
}

// ---- FP64 diagnostics on a copy of the state ----
static int state_to_d1(smx_handle* h) {
  if (h->fp32) {
    hipLaunchKernelGGL(k_f32_to_f64, dim3(kRedBlocks), dim3(kRedThreads), 0, h->st, (const cplx<float>*)h->A, h->D1,
                       h->Nh);
    HIPERR(hipGetLastError());
  } else {
    HIPERR(hipMemcpyAsync(h->D1, h->A, sizeof(cplx<double>) * h->Nh, hipMemcpyDeviceToDevice, h->st));
  }
  return SMX_OK;
}

// Real psi of the current state into `out` (dense [x][y][z], FP64).
static int c2r_state(smx_handle* h, double* out) {
  PParams<double> p = params<double>(h, h->t64, false);
  CHK(state_to_d1(h));
  HIPERR(PeriodicOps<double>::ypass(h->D1, h->D1, p.ny, h->npencils, +1, p.twy, h->st));
  HIPERR(PeriodicOps<double>::zpass(Z_IN_SPEC, Z_OUT_REAL, Z_OP_ID, h->D1, nullptr, nullptr, out, p, h->st));
  return SMX_OK;
}

// Normalised full half spectrum psi_hat of the current state into D2.
static int spectrum_state(smx_handle* h) {
  PParams<double> p = params<double>(h, h->t64, false);
  p.scale = 1.0 / (double)h->cfg.nx;  // the stage buffer holds the x-inverse of psi_hat
  CHK(state_to_d1(h));
  HIPERR(PeriodicOps<double>::xpass(X_FWD_SCALED, h->D1, h->D2, nullptr, nullptr, p, h->st));
  return SMX_OK;
}

static int host_sum(smx_handle* h, int nper, int nblocks, std::vector<double>& out) {
  std::vector<double> tmp((size_t)nper * nblocks);
  HIPERR(hipMemcpyAsync(tmp.data(), h->red, sizeof(double) * tmp.size(), hipMemcpyDeviceToHost, h->st));
  HIPERR(hipStreamSynchronize(h->st));
  out.assign(nper, 0.0);
  std::vector<double> c(nper, 0.0);
  for (int b = 0; b < nblocks; ++b)
    for (int j = 0; j < nper; ++j) {  // Kahan over the block partials, fixed order
      double y = tmp[(size_t)b * nper + j] - c[j], t = out[j] + y;
      c[j] = (t - out[j]) - y;
      out[j] = t;
    }
  return SMX_OK;
}

static int check_finite(smx_handle* h) {
  HIPERR(hipMemsetAsync(h->flag, 0, sizeof(unsigned int), h->st));
  if (h->fp32)
    hipLaunchKernelGGL(k_count_nonfinite<float>, dim3(kRedBlocks), dim3(kRedThreads), 0, h->st,
                       (const cplx<float>*)h->R, h->Nh, h->flag);
  else
    hipLaunchKernelGGL(k_count_nonfinite<double>, dim3(kRedBlocks), dim3(kRedThreads), 0, h->st,
                       (const cplx<double>*)h->R, h->Nh, h->flag);
  HIPERR(hipGetLastError());
  unsigned int bad = 0;
  HIPERR(hipMemcpyAsync(&bad, h->flag, sizeof(bad), hipMemcpyDeviceToHost, h->st));
  HIPERR(hipStreamSynchronize(h->st));
  if (bad) return fail(SMX_ERR_NONFINITE, std::to_string(bad) + " non-finite spectral entries in the state");
  return SMX_OK;
}

template <typename T>
static int run_steps(smx_handle* h, int64_t nsteps) {
  PParams<T> p = params<T>(h, tables<T>(h), h->fp32);
  for (int64_t n = 0; n < nsteps; ++n) {
    if (n == nsteps - 1) {
      CHK(c2r_state(h, h->real_prev));  // psi^n for the L1 criterion of the last step
    }
    CHK(step_once<T>(h, p, nullptr));
  }
  return SMX_OK;
}

template <typename T>
static int profile_T(smx_handle* h, int64_t nsteps, std::vector<double>& acc) {
  PParams<T> p = params<T>(h, tables<T>(h), h->fp32);
  std::vector<hipEvent_t> ev(kNumSlots + 1);
  for (auto& e : ev) HIPERR(hipEventCreate(&e));
  acc.assign(kNumSlots, 0.0);
  int rc = SMX_OK;
  for (int64_t n = 0; n < nsteps && rc == SMX_OK; ++n) {
    rc = step_once<T>(h, p, ev.data());
    if (rc) break;
    if (hipEventSynchronize(ev[kNumSlots]) != hipSuccess) { rc = fail(SMX_ERR_HIP, "event sync"); break; }
    for (int s = 0; s < kNumSlots; ++s) {
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, ev[s], ev[s + 1]);
      acc[s] += ms;
    }
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  for (auto& a : acc) a /= (double)(nsteps > 0 ? nsteps : 1);
  return rc;
}

// ---------------------------------------------------------------------------
// C ABI
extern "C" {

const char* smx_last_error(void) { return g_err.c_str(); }

void smx_config_default(smx_config* c, int n) {
  // periodicFFT/pf3d.cpp:35-75
  std::memset(c, 0, sizeof(*c));
  c->abi_version = SMX_ABI_VERSION;
  c->bc = SMX_BC_PERIODIC;
  c->precision = SMX_FP64;
  c->nx = c->ny = c->nz = n;
  c->q0 = 1.0;
  const int nw = 16;
  c->lx = c->ly = c->lz = n * 2.0 * M_PI / (c->q0 * nw);
  c->eps = -0.7;
  c->alpha = 1.0;
  c->beta = 2.0;
  c->gamma = 1.0;
  c->dt = 0.001;
  c->nl_model = SMX_NL_QUINTIC;
  c->advect = 1;
  c->eta = 1.0;
  c->rho = 1.0;
  c->force_form = SMX_FORCE_SH;
  c->w_mid = 0.0;
  c->mid_plane = -1;
  c->device = 0;
}

int smx_create(const smx_config* cfg, smx_handle** out) {
  if (!cfg || !out) return fail(SMX_ERR_INVALID, "null argument");
  *out = nullptr;
  const smx_config& c = *cfg;
  if (c.abi_version != SMX_ABI_VERSION) return fail(SMX_ERR_INVALID, "abi_version mismatch");
  if (c.bc != SMX_BC_PERIODIC && c.bc != SMX_BC_NEUMANN) return fail(SMX_ERR_INVALID, "bad bc");
  if (c.precision != SMX_FP64 && c.precision != SMX_FP32_MIXED) return fail(SMX_ERR_INVALID, "bad precision");
  if (c.nl_model != SMX_NL_QUINTIC && c.nl_model != SMX_NL_CUBIC) return fail(SMX_ERR_INVALID, "bad nl_model");
  if (!(c.lx > 0 && c.ly > 0 && c.lz > 0 && c.dt > 0)) return fail(SMX_ERR_INVALID, "lengths and dt must be > 0");
  if (c.bc == SMX_BC_NEUMANN) return fail(SMX_ERR_UNSUPPORTED, "Neumann (DCT) path not built yet");
  if (c.advect) return fail(SMX_ERR_UNSUPPORTED, "advection not built yet");
  for (int n : {c.nx, c.ny, c.nz})
    if (!length_supported(n))
      return fail(SMX_ERR_UNSUPPORTED, "grid length " + std::to_string(n) +
                                           " has no radix plan (supported: 16 24 32 48 64 96 128 192 256 384 512)");
  if (c.ny % kNC) return fail(SMX_ERR_UNSUPPORTED, "ny must be a multiple of 16");
  if (c.w_mid != 0.0 && (c.mid_plane < 1 || c.mid_plane > c.nz - 2))
    return fail(SMX_ERR_INVALID, "mid_plane must lie in [1, nz-2] when w_mid != 0");

  smx_handle* h = new smx_handle();
  h->cfg = c;
  h->fp32 = (c.precision == SMX_FP32_MIXED);
  h->nzh = c.nz / 2 + 1;
  h->nxy = (int64_t)c.nx * c.ny;
  h->N = h->nxy * c.nz;
  h->Nh = h->nxy * h->nzh;
  h->npencils = (int64_t)h->nzh * c.nx;
  h->dx = c.lx / c.nx;
  h->dy = c.ly / c.ny;
  h->dz = c.lz / c.nz;
  auto bail = [&](int code) {
    smx_destroy(h);
    return code;
  };
  hipError_t e = hipSetDevice(c.device);
  if (e != hipSuccess) return bail(fail(SMX_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e)));
  e = hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking);
  if (e != hipSuccess) return bail(fail(SMX_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e)));
  const size_t wsz = h->fp32 ? sizeof(cplx<float>) : sizeof(cplx<double>);
  int rc = SMX_OK;
  if ((rc = dalloc(h, &h->A, wsz * h->Nh)) || (rc = dalloc(h, &h->R, wsz * h->Nh)) ||
      (rc = dalloc(h, &h->P, wsz * h->Nh)) || (rc = dalloc(h, (void**)&h->D1, sizeof(cplx<double>) * h->Nh)) ||
      (rc = dalloc(h, (void**)&h->D2, sizeof(cplx<double>) * h->Nh)) ||
      (rc = dalloc(h, (void**)&h->real_io, sizeof(double) * h->N)) ||
      (rc = dalloc(h, (void**)&h->real_prev, sizeof(double) * h->N)) ||
      (rc = dalloc(h, (void**)&h->partial, sizeof(double) * (h->nxy / 8 + 64))) ||
      (rc = dalloc(h, (void**)&h->mean_state, sizeof(double) * 8)) ||
      (rc = dalloc(h, (void**)&h->red, sizeof(double) * 4 * kRedBlocks)) ||
      (rc = dalloc(h, (void**)&h->flag, sizeof(unsigned int) * 4)))
    return bail(rc);
  if ((rc = upload_tables<double>(h, h->t64))) return bail(rc);
  if (h->fp32 && (rc = upload_tables<float>(h, h->t32))) return bail(rc);
  e = hipMemset(h->mean_state, 0, sizeof(double) * 8);
  if (e != hipSuccess) return bail(fail(SMX_ERR_HIP, "hipMemset"));
  *out = h;
  return SMX_OK;
}

void smx_destroy(smx_handle* h) {
  if (!h) return;
  if (h->st) (void)hipStreamSynchronize(h->st);
  for (void* p : h->allocs) (void)hipFree(p);
  if (h->st) (void)hipStreamDestroy(h->st);
  delete h;
}

int smx_get_config(const smx_handle* h, smx_config* cfg) {
  if (!h || !cfg) return fail(SMX_ERR_INVALID, "null argument");
  *cfg = h->cfg;
  return SMX_OK;
}

int smx_set_psi(smx_handle* h, const double* psi, int64_t row_stride) {
  if (!h || !psi) return fail(SMX_ERR_INVALID, "null argument");
  const smx_config& c = h->cfg;
  if (row_stride < c.nz) return fail(SMX_ERR_INVALID, "row_stride < nz");
  HIPCHK(hipSetDevice(c.device));
  HIPCHK(hipMemcpy2DAsync(h->real_io, sizeof(double) * c.nz, psi, sizeof(double) * row_stride,
                          sizeof(double) * c.nz, (size_t)h->nxy, hipMemcpyHostToDevice, h->st));
  CHK(h->fp32 ? bootstrap<float>(h) : bootstrap<double>(h));
  HIPCHK(hipStreamSynchronize(h->st));
  h->has_state = true;
  h->has_prev = false;
  h->steps = 0;
  return SMX_OK;
}

int smx_get_psi(smx_handle* h, double* psi, int64_t row_stride) {
  if (!h || !psi) return fail(SMX_ERR_INVALID, "null argument");
  if (!h->has_state) return fail(SMX_ERR_STATE, "no state: call smx_set_psi first");
  const smx_config& c = h->cfg;
  if (row_stride < c.nz) return fail(SMX_ERR_INVALID, "row_stride < nz");
  HIPCHK(hipSetDevice(c.device));
  CHK(c2r_state(h, h->real_io));
  HIPCHK(hipMemcpy2DAsync(psi, sizeof(double) * row_stride, h->real_io, sizeof(double) * c.nz,
                          sizeof(double) * c.nz, (size_t)h->nxy, hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  return SMX_OK;
}

int smx_step(smx_handle* h, int64_t nsteps) {
  if (!h) return fail(SMX_ERR_INVALID, "null handle");
  if (!h->has_state) return fail(SMX_ERR_STATE, "no state: call smx_set_psi first");
  if (nsteps < 0) return fail(SMX_ERR_INVALID, "nsteps < 0");
  if (nsteps == 0) return SMX_OK;
  HIPCHK(hipSetDevice(h->cfg.device));
  CHK(h->fp32 ? run_steps<float>(h, nsteps) : run_steps<double>(h, nsteps));
  h->steps += nsteps;
  h->has_prev = true;
  return SMX_OK;
}

int smx_sync(smx_handle* h) {
  if (!h) return fail(SMX_ERR_INVALID, "null handle");
  HIPCHK(hipStreamSynchronize(h->st));
  return SMX_OK;
}

int smx_step_count(const smx_handle* h, int64_t* n) {
  if (!h || !n) return fail(SMX_ERR_INVALID, "null argument");
  *n = h->steps;
  return SMX_OK;
}

int smx_l1(smx_handle* h, double* l1) {
  if (!h || !l1) return fail(SMX_ERR_INVALID, "null argument");
  if (!h->has_prev) return fail(SMX_ERR_STATE, "no step taken since smx_set_psi");
  HIPCHK(hipSetDevice(h->cfg.device));
  CHK(check_finite(h));
  CHK(c2r_state(h, h->real_io));
  hipLaunchKernelGGL(k_reduce_l1, dim3(kRedBlocks), dim3(kRedThreads), 0, h->st, h->real_io, h->real_prev, h->N,
                     h->red);
  HIPCHK(hipGetLastError());
  std::vector<double> s;
  CHK(host_sum(h, 2, kRedBlocks, s));
  *l1 = s[0] / (h->cfg.dt * s[1]);  // pf3d.cpp:509
  return SMX_OK;
}

static int spectral_sums(smx_handle* h, std::vector<double>& s) {
  CHK(spectrum_state(h));
  const smx_config& c = h->cfg;
  hipLaunchKernelGGL(k_reduce_spec, dim3(kRedBlocks), dim3(kRedThreads), 0, h->st, h->D2, h->Nh, c.nx, c.ny, c.nz,
                     h->t64.kx, h->t64.ky, h->t64.kz, c.eps, c.alpha, c.q0 * c.q0, h->red);
  HIPCHK(hipGetLastError());
  return host_sum(h, 3, kRedBlocks, s);
}

int smx_energy(smx_handle* h, double* F) {
  if (!h || !F) return fail(SMX_ERR_INVALID, "null argument");
  if (!h->has_state) return fail(SMX_ERR_STATE, "no state: call smx_set_psi first");
  HIPCHK(hipSetDevice(h->cfg.device));
  const smx_config& c = h->cfg;
  std::vector<double> s, l;
  CHK(spectral_sums(h, s));
  CHK(c2r_state(h, h->real_io));
  hipLaunchKernelGGL(k_reduce_local, dim3(kRedBlocks), dim3(kRedThreads), 0, h->st, h->real_io, h->N, c.nl_model,
                     c.beta, c.gamma, h->red);
  HIPCHK(hipGetLastError());
  CHK(host_sum(h, 1, kRedBlocks, l));
  const double V = c.lx * c.ly * c.lz, dV = h->dx * h->dy * h->dz;
  *F = V * s[0] + dV * l[0];  // Parseval: int psi^2 = V sum_k |psi_hat_k|^2
  return SMX_OK;
}

int smx_dominant_wavenumber(smx_handle* h, double* q) {
  if (!h || !q) return fail(SMX_ERR_INVALID, "null argument");
  if (!h->has_state) return fail(SMX_ERR_STATE, "no state: call smx_set_psi first");
  HIPCHK(hipSetDevice(h->cfg.device));
  std::vector<double> s;
  CHK(spectral_sums(h, s));
  *q = s[1] > 0 ? s[2] / s[1] : 0.0;
  return SMX_OK;
}

int smx_dtpsi(smx_handle* h, double* out, int64_t row_stride) {
  if (!h || !out) return fail(SMX_ERR_INVALID, "null argument");
  if (!h->has_state) return fail(SMX_ERR_STATE, "no state: call smx_set_psi first");
  const smx_config& c = h->cfg;
  if (row_stride < c.nz) return fail(SMX_ERR_INVALID, "row_stride < nz");
  HIPCHK(hipSetDevice(c.device));
  PParams<double> p = params<double>(h, h->t64, false);
  using Ops = PeriodicOps<double>;
  CHK(spectrum_state(h));                     // D2 = psi_hat
  CHK(c2r_state(h, h->real_io));              // real psi
  HIPCHK(Ops::zpass(Z_IN_REAL, Z_OUT_SPEC, Z_OP_NL, nullptr, h->D1, h->real_io, nullptr, p, h->st));
  HIPCHK(Ops::ypass(h->D1, h->D1, p.ny, h->npencils, -1, p.twy, h->st));
  HIPCHK(Ops::xpass(X_FWD_SCALED, h->D1, h->D1, nullptr, nullptr, p, h->st));  // N_hat (normalised)
  hipLaunchKernelGGL(k_dtpsi_combine, dim3(kRedBlocks), dim3(kRedThreads), 0, h->st, h->D2, h->D1, h->Nh, c.nx, c.ny,
                     h->t64.kx, h->t64.ky, h->t64.kz, c.eps, c.alpha, c.q0 * c.q0);
  HIPCHK(hipGetLastError());
  HIPCHK(Ops::xpass(X_INV, h->D1, h->D1, nullptr, nullptr, p, h->st));
  HIPCHK(Ops::ypass(h->D1, h->D1, p.ny, h->npencils, +1, p.twy, h->st));
  HIPCHK(Ops::zpass(Z_IN_SPEC, Z_OUT_REAL, Z_OP_ID, h->D1, nullptr, nullptr, h->real_io, p, h->st));
  HIPCHK(hipMemcpy2DAsync(out, sizeof(double) * row_stride, h->real_io, sizeof(double) * c.nz, sizeof(double) * c.nz,
                          (size_t)h->nxy, hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  return SMX_OK;
}

int smx_surftrack(smx_handle* h, const smx_surf_opts* opts, double* surfz, double* vel, double* H, double* K,
                  int64_t* found) {
  (void)opts; (void)surfz; (void)vel; (void)H; (void)K; (void)found;
  if (!h) return fail(SMX_ERR_INVALID, "null handle");
  return fail(SMX_ERR_UNSUPPORTED, "surfTrack not built yet");
}

int smx_profile_step(smx_handle* h, int64_t nsteps, int max_slots, int* nslots, double* avg_ms, double* alg_bytes,
                     char* names, int names_cap) {
  if (!h || !nslots || !avg_ms || !alg_bytes) return fail(SMX_ERR_INVALID, "null argument");
  if (!h->has_state) return fail(SMX_ERR_STATE, "no state: call smx_set_psi first");
  if (max_slots < kNumSlots) return fail(SMX_ERR_INVALID, "max_slots too small");
  HIPCHK(hipSetDevice(h->cfg.device));
  std::vector<double> acc;
  CHK(h->fp32 ? profile_T<float>(h, nsteps, acc) : profile_T<double>(h, nsteps, acc));
  const double sc = (h->fp32 ? 8.0 : 16.0) * (double)h->Nh;  // bytes of one spectral-stage array
  const double bytes[kNumSlots] = {2 * sc, 2 * sc, 2 * sc, 4 * sc};
  *nslots = kNumSlots;
  std::string nm;
  for (int s = 0; s < kNumSlots; ++s) {
    avg_ms[s] = acc[s];
    alg_bytes[s] = bytes[s];
    nm += kSlotNames[s];
    nm += "\n";
  }
  if (names && names_cap > 0) {
    std::strncpy(names, nm.c_str(), (size_t)names_cap - 1);
    names[names_cap - 1] = 0;
  }
  h->steps += nsteps;
  h->has_prev = false;
  return SMX_OK;
}

int smx_calibration_copy(smx_handle* h, int64_t nbytes, int reps, double* avg_ms) {
  if (!h || !avg_ms || nbytes <= 0 || reps <= 0) return fail(SMX_ERR_INVALID, "bad argument");
  HIPCHK(hipSetDevice(h->cfg.device));
  const int64_t n4 = nbytes / 16;
  void *a = nullptr, *b = nullptr;
  HIPCHK(hipMalloc(&a, n4 * 16));
  hipError_t e = hipMalloc(&b, n4 * 16);
  if (e != hipSuccess) {
    (void)hipFree(a);
    return fail(SMX_ERR_HIP, "hipMalloc");
  }
  (void)hipMemsetAsync(a, 0, n4 * 16, h->st);
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  hipLaunchKernelGGL(k_copy_f4, dim3(2048), dim3(256), 0, h->st, (const float4*)a, (float4*)b, n4);
  (void)hipEventRecord(e0, h->st);
  for (int r = 0; r < reps; ++r)
    hipLaunchKernelGGL(k_copy_f4, dim3(2048), dim3(256), 0, h->st, (const float4*)a, (float4*)b, n4);
  (void)hipEventRecord(e1, h->st);
  (void)hipEventSynchronize(e1);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  *avg_ms = ms / reps;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  HIPCHK(hipGetLastError());
  return SMX_OK;
}

}  // extern "C"
