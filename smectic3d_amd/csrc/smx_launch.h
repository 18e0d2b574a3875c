// smx_launch.h -- runtime-length -> template-instance dispatch for the pass kernels.
// Instantiated once per precision in smx_inst_f32.hip / smx_inst_f64.hip.
#pragma once
#include <hip/hip_ext.h>

#include <algorithm>
#include <type_traits>

#include "smx_adjoint.h"
#include "smx_neuadv.h"

namespace smx {

// Kernel timing: when the calling thread has set a start/stop event pair, launches go
// through hipExtLaunchKernelGGL, which stamps the events with the kernel's own start and
// end (no extra packets in the stream).  Used to time each kernel inside a real step chain.
struct LaunchEvents {
  hipEvent_t start = nullptr, stop = nullptr;
};
inline LaunchEvents& launch_events() {
  static thread_local LaunchEvents ev;
  return ev;
}
#define SMX_LAUNCH(K, G, B, SH, ST, ...)                                                   \
  do {                                                                                   \
    const LaunchEvents& ev_ = launch_events();                                           \
    if (ev_.start) hipExtLaunchKernelGGL(K, G, B, SH, ST, ev_.start, ev_.stop, 0, __VA_ARGS__); \
    else hipLaunchKernelGGL(K, G, B, SH, ST, __VA_ARGS__);                               \
  } while (0)

constexpr int kNC = 16;  // columns (pencils) per tile

// Grid lengths with a radix plan.  Every axis length must be one of these.
#define SMX_FOR_EACH_LENGTH(X) \
  X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512)

inline bool length_supported(int n) {
#define SMX_CASE(L) if (n == L) return true;
  SMX_FOR_EACH_LENGTH(SMX_CASE)
#undef SMX_CASE
  return false;
}

// Tile schedule of one pass kernel: columns per tile (NC) and plan variant (V, PlanV;
// in the x slot v >> 1 also selects the carry order of k_xpass).
// The defaults are the fixed schedule; an autotuner may pick others per pass.
struct PassSched {
  int nc = kNC, v = 0;
};
// Schedule slots of the periodic step chain, in launch order.
enum SchedSlot { SLOT_YINV = 0, SLOT_Z = 1, SLOT_YFWD = 2, SLOT_X = 3, NSLOTS = 4 };
constexpr int kSchedNC[3] = {8, 16, 32};
constexpr int kSchedV[2] = {0, 1};
constexpr int kSchedVX[8] = {0, 1, 2, 3, 4, 5, 6, 7};  // the x pass adds the carry order (v >> 1)
constexpr size_t kMaxLds = 160 * 1024 - 1024;  // LDS of one workgroup on gfx950, less the reduction scratch
// (each tile check adds the L-entry twiddle copy every pass kernel keeps in LDS)

template <int N> using IC = std::integral_constant<int, N>;

// Calls f(IC<NC>, IC<V>) for a schedule of length L; invalid schedules give hipErrorInvalidValue.
template <int L, typename F>
hipError_t sched_dispatch(const PassSched& s, F&& f) {
  if (s.v == 0) {
    switch (s.nc) {
      case 8: return f(IC<8>{}, IC<0>{});
      case 16: return f(IC<16>{}, IC<0>{});
      case 32: return f(IC<32>{}, IC<0>{});
      default: break;
    }
  } else if (s.v == 1) {
    if constexpr (HasPlanV<L, 1>::value) {
      switch (s.nc) {
        case 8: return f(IC<8>{}, IC<1>{});
        case 16: return f(IC<16>{}, IC<1>{});
        case 32: return f(IC<32>{}, IC<1>{});
        default: break;
      }
    }
  }
  return hipErrorInvalidValue;
}

template <typename T>
struct PeriodicOps {
  // y pass, in place allowed.  sign -1 forward, +1 inverse.  kmul: k_y table of a multiplier
  // applied to the input (kop KM_IKY: i k_y, KM_DYY: -k_y^2).  dry: only report whether the
  // schedule is valid for this length (no launch).
  static hipError_t ypass(const cplx<T>* in, cplx<T>* out, int ny, int64_t npencils, int sign,
                          const cplx<T>* tw, hipStream_t st, const T* kmul = nullptr, PassSched s = PassSched{},
                          bool dry = false, int kop = KM_IKY);
  // x pass (mode: XMode), tiles over (y, kz).  Only X_STEP takes a non-default schedule.
  static hipError_t xpass(int mode, const cplx<T>* in, cplx<T>* out, cplx<T>* rbuf, const cplx<T>* psi0,
                          const PParams<T>& p, hipStream_t st, PassSched s = PassSched{}, bool dry = false);
  // z pass.  in_mode: ZIn, out_mode: ZOut bit set, op: ZOp.  Only the spectral-to-spectral
  // form takes a non-default schedule; its grid has nx*ny/s.nc blocks (FP64 partials).
  static hipError_t zpass(int in_mode, int out_mode, int op, const cplx<T>* in, cplx<T>* out,
                          const double* real_in, double* real_out, const PParams<T>& p, hipStream_t st,
                          PassSched s = PassSched{}, bool dry = false);
  static int zpass_blocks(const PParams<T>& p, int nc = kNC) { return (int)((p.nxy + nc - 1) / nc); }
  // One CNAB2 step of nbox boxes that share the grid of p and the tile schedules s[NSLOTS]:
  // the four passes, one grouped launch each (GrpArgs in device memory, one per box, with the
  // z pass's partial count for its schedule).  dry: only check the schedules.
  static hipError_t step_grp(const GrpArgs<T>* g, int nbox, const PParams<T>& p, int64_t npencils,
                             const PassSched* s, hipStream_t st, bool dry = false);
  // The grouped passes one by one (in place on each box's stage array g[b].a).
  static hipError_t ypass_grp(const GrpArgs<T>* g, int nbox, const PParams<T>& p, int64_t npencils, int sign,
                              PassSched s, hipStream_t st, bool dry = false);
  static hipError_t zpass_grp(const GrpArgs<T>* g, int nbox, const PParams<T>& p, int op, PassSched s,
                              hipStream_t st, bool dry = false);
  static hipError_t xpass_grp(const GrpArgs<T>* g, int nbox, const PParams<T>& p, PassSched s, hipStream_t st,
                              bool dry = false);
};

template <typename T>
hipError_t PeriodicOps<T>::ypass(const cplx<T>* in, cplx<T>* out, int ny, int64_t npencils, int sign,
                                 const cplx<T>* tw, hipStream_t st, const T* kmul, PassSched s, bool dry,
                                 int kop) {
  if (kmul != nullptr && kop != KM_IKY && kop != KM_DYY) return hipErrorInvalidValue;
#define SMX_CASE(L)                                                                                      \
  if (ny == L)                                                                                           \
    return sched_dispatch<L>(s, [&](auto nc, auto v) -> hipError_t {                                     \
      constexpr int NC = decltype(nc)::value, V = decltype(v)::value;                                    \
      constexpr int NT = NC * (L / PlanV<L, V>::E);                                                      \
      if constexpr (NT > 1024 || (NC * ColAddr<L>::LP + L) * sizeof(cplx<T>) + L * sizeof(T) > kMaxLds) {    \
        return hipErrorInvalidValue;                                                                     \
      } else {                                                                                           \
        if (dry) return hipSuccess;                                                                      \
        const dim3 grid((unsigned)((npencils + NC - 1) / NC));                                           \
        if (sign < 0) SMX_LAUNCH((k_ypass<T, L, -1, NC, V>), grid, dim3(NT), 0, st, in, out, npencils, tw, kmul, kop); \
        else SMX_LAUNCH((k_ypass<T, L, +1, NC, V>), grid, dim3(NT), 0, st, in, out, npencils, tw, kmul, kop);          \
        return hipGetLastError();                                                                        \
      }                                                                                                  \
    });
  SMX_FOR_EACH_LENGTH(SMX_CASE)
#undef SMX_CASE
  return hipErrorInvalidValue;
}

template <typename T>
hipError_t PeriodicOps<T>::xpass(int mode, const cplx<T>* in, cplx<T>* out, cplx<T>* rbuf, const cplx<T>* psi0,
                                 const PParams<T>& p, hipStream_t st, PassSched s, bool dry) {
  if (mode != X_STEP) {
    if (s.nc != kNC || s.v != 0) return hipErrorInvalidValue;
    if (dry) return hipSuccess;
    const dim3 grid((unsigned)(p.ny / kNC), (unsigned)p.nzh);
#define SMX_CASE(L)                                                                                   \
  if (p.nx == L) {                                                                                    \
    const dim3 block(kNC * (L / Plan<L>::E));                                                         \
    switch (mode) {                                                                                   \
      case X_FWD_SCALED: SMX_LAUNCH((k_xpass<T, L, kNC, X_FWD_SCALED>), grid, block, 0, st, in, out, rbuf, psi0, p); break; \
      case X_BOOT: SMX_LAUNCH((k_xpass<T, L, kNC, X_BOOT>), grid, block, 0, st, in, out, rbuf, psi0, p); break; \
      case X_INV: SMX_LAUNCH((k_xpass<T, L, kNC, X_INV>), grid, block, 0, st, in, out, rbuf, psi0, p); break; \
      case X_MUL2:                                                                                    \
        if constexpr (std::is_same<T, double>::value)                                                 \
          SMX_LAUNCH((k_xpass<T, L, kNC, X_MUL2>), grid, block, 0, st, in, out, rbuf, psi0, p);      \
        else                                                                                          \
          return hipErrorInvalidValue;                                                                \
        break;                                                                                        \
      default: return hipErrorInvalidValue;                                                           \
    }                                                                                                 \
    return hipGetLastError();                                                                         \
  }
    SMX_FOR_EACH_LENGTH(SMX_CASE)
#undef SMX_CASE
    return hipErrorInvalidValue;
  }
  if (p.ny % s.nc != 0) return hipErrorInvalidValue;  // whole tiles along y
  // variant bit 0: radix plan; v >> 1: carry order (k_xpass ORD 0..3)
  if (s.v < 0 || s.v > 7) return hipErrorInvalidValue;
  const int ord = s.v >> 1;
  const PassSched sp{s.nc, s.v & 1};
#define SMX_CASE(L)                                                                                      \
  if (p.nx == L)                                                                                         \
    return sched_dispatch<L>(sp, [&](auto nc, auto v) -> hipError_t {                                    \
      constexpr int NC = decltype(nc)::value, V = decltype(v)::value;                                    \
      constexpr int NT = NC * (L / PlanV<L, V>::E);                                                      \
      if constexpr (NT > 1024 || (size_t)L * (NC + 2) * sizeof(cplx<T>) > kMaxLds) {                     \
        return hipErrorInvalidValue;                                                                     \
      } else {                                                                                           \
        if (dry) return hipSuccess;                                                                      \
        const dim3 grid((unsigned)(p.ny / NC), (unsigned)p.nzh);                                         \
        if (ord == 0) SMX_LAUNCH((k_xpass<T, L, NC, X_STEP, V, 0>), grid, dim3(NT), 0, st, in, out, rbuf, psi0, p); \
        else if (ord == 1) SMX_LAUNCH((k_xpass<T, L, NC, X_STEP, V, 1>), grid, dim3(NT), 0, st, in, out, rbuf, psi0, p); \
        else if (ord == 2) SMX_LAUNCH((k_xpass<T, L, NC, X_STEP, V, 2>), grid, dim3(NT), 0, st, in, out, rbuf, psi0, p); \
        else SMX_LAUNCH((k_xpass<T, L, NC, X_STEP, V, 3>), grid, dim3(NT), 0, st, in, out, rbuf, psi0, p);          \
        return hipGetLastError();                                                                        \
      }                                                                                                  \
    });
  SMX_FOR_EACH_LENGTH(SMX_CASE)
#undef SMX_CASE
  return hipErrorInvalidValue;
}

template <typename T>
hipError_t PeriodicOps<T>::zpass(int in_mode, int out_mode, int op, const cplx<T>* in, cplx<T>* out,
                                 const double* real_in, double* real_out, const PParams<T>& p, hipStream_t st,
                                 PassSched s, bool dry) {
  if (p.nxy % s.nc != 0) return hipErrorInvalidValue;  // the kernel assumes whole tiles
  if (in_mode != Z_IN_SPEC || out_mode != Z_OUT_SPEC) {
    if (s.nc != kNC || s.v != 0) return hipErrorInvalidValue;
    if (dry) return hipSuccess;
    const dim3 grid((unsigned)zpass_blocks(p));
#define SMX_CASE(L)                                                                                  \
  if (p.nz == L) {                                                                                   \
    const dim3 block((kNC / 2) * (L / Plan<L>::E));                                                  \
    if (in_mode == Z_IN_REAL && out_mode == Z_OUT_SPEC)                                              \
      SMX_LAUNCH((k_zpass<T, L, kNC, Z_IN_REAL, Z_OUT_SPEC>), grid, block, 0, st, in, out, real_in, real_out, op, p); \
    else if (in_mode == Z_IN_SPEC && out_mode == Z_OUT_REAL)                                         \
      SMX_LAUNCH((k_zpass<T, L, kNC, Z_IN_SPEC, Z_OUT_REAL>), grid, block, 0, st, in, out, real_in, real_out, op, p); \
    else                                                                                             \
      return hipErrorInvalidValue;                                                                   \
    return hipGetLastError();                                                                        \
  }
    SMX_FOR_EACH_LENGTH(SMX_CASE)
#undef SMX_CASE
    return hipErrorInvalidValue;
  }
#define SMX_CASE(L)                                                                                      \
  if (p.nz == L)                                                                                         \
    return sched_dispatch<L>(s, [&](auto nc, auto v) -> hipError_t {                                     \
      constexpr int NC = decltype(nc)::value, V = decltype(v)::value;                                    \
      constexpr int NT = (NC / 2) * (L / PlanV<L, V>::E);                                                \
      if constexpr (NT > 1024 || (size_t)L * (NC / 2 + 2) * sizeof(cplx<T>) > kMaxLds) {                 \
        return hipErrorInvalidValue;                                                                     \
      } else {                                                                                           \
        if (dry) return hipSuccess;                                                                      \
        SMX_LAUNCH((k_zpass<T, L, NC, Z_IN_SPEC, Z_OUT_SPEC, V>), dim3((unsigned)(p.nxy / NC)), dim3(NT), 0, st, \
                           in, out, real_in, real_out, op, p);                                           \
        return hipGetLastError();                                                                        \
      }                                                                                                  \
    });
  SMX_FOR_EACH_LENGTH(SMX_CASE)
#undef SMX_CASE
  return hipErrorInvalidValue;
}


template <typename T>
hipError_t PeriodicOps<T>::ypass_grp(const GrpArgs<T>* g, int nbox, const PParams<T>& p, int64_t npencils, int sign,
                                     PassSched ps, hipStream_t st, bool dry) {
  if (nbox < 1 || nbox > 65535) return hipErrorInvalidValue;
  const unsigned nb = (unsigned)nbox;
#define SMX_CASE(L)                                                                                      \
  if (p.ny == L)                                                                                         \
    return sched_dispatch<L>(ps, [&](auto nc, auto v) -> hipError_t {                                    \
      constexpr int NC = decltype(nc)::value, V = decltype(v)::value;                                    \
      constexpr int NT = NC * (L / PlanV<L, V>::E);                                                      \
      if constexpr (NT > 1024 || (NC * ColAddr<L>::LP + L) * sizeof(cplx<T>) + L * sizeof(T) > kMaxLds) { \
        return hipErrorInvalidValue;                                                                     \
      } else {                                                                                           \
        if (dry) return hipSuccess;                                                                      \
        const dim3 grid((unsigned)((npencils + NC - 1) / NC), nb);                                       \
        if (sign < 0) SMX_LAUNCH((k_ypass_grp<T, L, -1, NC, V>), grid, dim3(NT), 0, st, g, npencils);   \
        else SMX_LAUNCH((k_ypass_grp<T, L, +1, NC, V>), grid, dim3(NT), 0, st, g, npencils);            \
        return hipGetLastError();                                                                        \
      }                                                                                                  \
    });
  SMX_FOR_EACH_LENGTH(SMX_CASE)
#undef SMX_CASE
  return hipErrorInvalidValue;
}

template <typename T>
hipError_t PeriodicOps<T>::zpass_grp(const GrpArgs<T>* g, int nbox, const PParams<T>& p, int op, PassSched ps,
                                     hipStream_t st, bool dry) {
  if (nbox < 1 || nbox > 65535 || p.nxy % ps.nc != 0) return hipErrorInvalidValue;
  const unsigned nb = (unsigned)nbox;
#define SMX_CASE(L)                                                                                      \
  if (p.nz == L)                                                                                         \
    return sched_dispatch<L>(ps, [&](auto nc, auto v) -> hipError_t {                                    \
      constexpr int NC = decltype(nc)::value, V = decltype(v)::value;                                    \
      constexpr int NT = (NC / 2) * (L / PlanV<L, V>::E);                                                \
      if constexpr (NT > 1024 || (size_t)L * (NC / 2 + 2) * sizeof(cplx<T>) > kMaxLds) {                 \
        return hipErrorInvalidValue;                                                                     \
      } else {                                                                                           \
        if (dry) return hipSuccess;                                                                      \
        SMX_LAUNCH((k_zpass_grp<T, L, NC, V>), dim3((unsigned)(p.nxy / NC), nb), dim3(NT), 0, st, g, op); \
        return hipGetLastError();                                                                        \
      }                                                                                                  \
    });
  SMX_FOR_EACH_LENGTH(SMX_CASE)
#undef SMX_CASE
  return hipErrorInvalidValue;
}

template <typename T>
hipError_t PeriodicOps<T>::xpass_grp(const GrpArgs<T>* g, int nbox, const PParams<T>& p, PassSched ps,
                                     hipStream_t st, bool dry) {
  if (nbox < 1 || nbox > 65535 || p.ny % ps.nc != 0 || ps.v < 0 || ps.v > 7) return hipErrorInvalidValue;
  const unsigned nb = (unsigned)nbox;
  const int ord = ps.v >> 1;
  const PassSched sp{ps.nc, ps.v & 1};
#define SMX_CASE(L)                                                                                      \
  if (p.nx == L)                                                                                         \
    return sched_dispatch<L>(sp, [&](auto nc, auto v) -> hipError_t {                                    \
      constexpr int NC = decltype(nc)::value, V = decltype(v)::value;                                    \
      constexpr int NT = NC * (L / PlanV<L, V>::E);                                                      \
      if constexpr (NT > 1024 || (size_t)L * (NC + 2) * sizeof(cplx<T>) > kMaxLds) {                     \
        return hipErrorInvalidValue;                                                                     \
      } else {                                                                                           \
        if (dry) return hipSuccess;                                                                      \
        const dim3 grid((unsigned)(p.ny / NC), (unsigned)p.nzh, nb);                                     \
        if (ord == 0) SMX_LAUNCH((k_xpass_grp<T, L, NC, V, 0>), grid, dim3(NT), 0, st, g);              \
        else if (ord == 1) SMX_LAUNCH((k_xpass_grp<T, L, NC, V, 1>), grid, dim3(NT), 0, st, g);         \
        else if (ord == 2) SMX_LAUNCH((k_xpass_grp<T, L, NC, V, 2>), grid, dim3(NT), 0, st, g);         \
        else SMX_LAUNCH((k_xpass_grp<T, L, NC, V, 3>), grid, dim3(NT), 0, st, g);                       \
        return hipGetLastError();                                                                        \
      }                                                                                                  \
    });
  SMX_FOR_EACH_LENGTH(SMX_CASE)
#undef SMX_CASE
  return hipErrorInvalidValue;
}

template <typename T>
hipError_t PeriodicOps<T>::step_grp(const GrpArgs<T>* g, int nbox, const PParams<T>& p, int64_t npencils,
                                    const PassSched* s, hipStream_t st, bool dry) {
  hipError_t e = ypass_grp(g, nbox, p, npencils, +1, s[SLOT_YINV], st, dry);
  if (e == hipSuccess) e = zpass_grp(g, nbox, p, Z_OP_NL, s[SLOT_Z], st, dry);
  if (e == hipSuccess) e = ypass_grp(g, nbox, p, npencils, -1, s[SLOT_YFWD], st, dry);
  if (e == hipSuccess) e = xpass_grp(g, nbox, p, s[SLOT_X], st, dry);
  return e;
}

// Tile width (columns, in pairs of adjacent lines) of the Neumann passes: rows of 128 bytes.
template <typename T> constexpr int neumann_nc() { return sizeof(T) == 4 ? 32 : 16; }
// The second width of the plain Neumann step chain's kernels (rows of 64 bytes), one of the two
// candidates of the autotuner; diagnostics and the advected step keep the first.
template <typename T> constexpr int neumann_nc_alt() { return sizeof(T) == 4 ? 16 : 8; }

// nc: columns per tile, 0 for neumann_nc<T>(); neumann_nc_alt<T>() exists for the step chain's
// forms (cosine y passes, the z pass spectrum -> spectrum, every x pass).  dry: only check that
// the kernel exists for this grid.
template <typename T>
struct NeumannOps {
  // y pass on real pencils (npencils = nz*nx), dir -1: DCT-II, +1: DCT-III.
  // par: parity of the pencils along y (DPar; DP_DSIN with dir +1 and kq = k_y).
  static hipError_t ypass(const T* in, T* out, int ny, int64_t npencils, int dir, const cplx<T>* tw,
                          const cplx<T>* qtw, hipStream_t st, int par = DP_COS, const T* kq = nullptr, int nc = 0,
                          bool dry = false);
  static hipError_t xpass(int mode, const T* in, T* out, T* rbuf, const T* psi0, const PParams<T>& p,
                          const cplx<T>* qtw, hipStream_t st, int nc = 0, bool dry = false);
  static hipError_t zpass(int in_mode, int out_mode, int op, const T* in, T* out, const double* real_in,
                          double* real_out, const PParams<T>& p, const cplx<T>* qtw, hipStream_t st, int nc = 0,
                          bool dry = false);
  static int zpass_blocks(const PParams<T>& p, int nc = 0) { return (int)(p.nxy / (nc > 0 ? nc : neumann_nc<T>())); }
};

template <typename T, int L, int NC>
static hipError_t dct_ypass_launch(const T* in, T* out, int64_t npencils, int dir, const cplx<T>* tw,
                                   const cplx<T>* qtw, hipStream_t st, int par, const T* kq) {
  const dim3 grid((unsigned)(npencils / NC)), block((NC / 2) * (L / Plan<L>::E));
  if (dir < 0 && par == DP_COS) SMX_LAUNCH((k_dct_ypass<T, L, -1, NC>), grid, block, 0, st, in, out, tw, qtw, kq);
  else if (par == DP_COS) SMX_LAUNCH((k_dct_ypass<T, L, +1, NC>), grid, block, 0, st, in, out, tw, qtw, kq);
  else if constexpr (NC == neumann_nc<T>()) {
    if (dir < 0) SMX_LAUNCH((k_dct_ypass<T, L, -1, NC, DP_SIN>), grid, block, 0, st, in, out, tw, qtw, kq);
    else if (par == DP_SIN) SMX_LAUNCH((k_dct_ypass<T, L, +1, NC, DP_SIN>), grid, block, 0, st, in, out, tw, qtw, kq);
    else SMX_LAUNCH((k_dct_ypass<T, L, +1, NC, DP_DSIN>), grid, block, 0, st, in, out, tw, qtw, kq);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename T>
hipError_t NeumannOps<T>::ypass(const T* in, T* out, int ny, int64_t npencils, int dir, const cplx<T>* tw,
                                const cplx<T>* qtw, hipStream_t st, int par, const T* kq, int nc, bool dry) {
  constexpr int NCD = neumann_nc<T>(), NCA = neumann_nc_alt<T>();
  if (nc == 0) nc = NCD;
  if (nc != NCD && !(nc == NCA && par == DP_COS)) return hipErrorInvalidValue;
  if (npencils % nc != 0) return hipErrorInvalidValue;  // whole tiles of pencil pairs
  if (par == DP_DSIN && (dir < 0 || kq == nullptr)) return hipErrorInvalidValue;
  if (par == DP_COS && dir < 0 && kq != nullptr) return hipErrorInvalidValue;  // -k^2 is an inverse input
#define SMX_CASE(L)                                                                                   \
  if (ny == L) {                                                                                      \
    if (dry) return hipSuccess;                                                                       \
    return nc == NCD ? dct_ypass_launch<T, L, NCD>(in, out, npencils, dir, tw, qtw, st, par, kq)     \
                     : dct_ypass_launch<T, L, NCA>(in, out, npencils, dir, tw, qtw, st, par, kq);    \
  }
  SMX_FOR_EACH_LENGTH(SMX_CASE)
#undef SMX_CASE
  return hipErrorInvalidValue;
}

template <typename T, int L, int NC>
static hipError_t dct_xpass_launch(int mode, const T* in, T* out, T* rbuf, const T* psi0, const PParams<T>& p,
                                   const cplx<T>* qtw, hipStream_t st) {
  const dim3 grid((unsigned)(p.ny / NC), (unsigned)p.nzh), block((NC / 2) * (L / Plan<L>::E));
  switch (mode) {
    case X_STEP: SMX_LAUNCH((k_dct_xpass<T, L, NC, X_STEP>), grid, block, 0, st, in, out, rbuf, psi0, p, qtw); break;
    case X_FWD_SCALED: SMX_LAUNCH((k_dct_xpass<T, L, NC, X_FWD_SCALED>), grid, block, 0, st, in, out, rbuf, psi0, p, qtw); break;
    case X_BOOT: SMX_LAUNCH((k_dct_xpass<T, L, NC, X_BOOT>), grid, block, 0, st, in, out, rbuf, psi0, p, qtw); break;
    case X_INV: SMX_LAUNCH((k_dct_xpass<T, L, NC, X_INV>), grid, block, 0, st, in, out, rbuf, psi0, p, qtw); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename T>
hipError_t NeumannOps<T>::xpass(int mode, const T* in, T* out, T* rbuf, const T* psi0, const PParams<T>& p,
                                const cplx<T>* qtw, hipStream_t st, int nc, bool dry) {
  // 128-byte rows where ny allows it (ny is a multiple of 16), else the 16-column tile
  constexpr int NCD = neumann_nc<T>(), NCA = neumann_nc_alt<T>();
  if (p.ny % 16 != 0) return hipErrorInvalidValue;
  if (nc == 0) nc = (p.ny % NCD == 0) ? NCD : 16;
  if (nc != NCD && nc != NCA && nc != 16) return hipErrorInvalidValue;
  if (p.ny % nc != 0) return hipErrorInvalidValue;
#define SMX_CASE(L)                                                                          \
  if (p.nx == L) {                                                                           \
    if (dry) return hipSuccess;                                                              \
    if (nc == NCD) return dct_xpass_launch<T, L, NCD>(mode, in, out, rbuf, psi0, p, qtw, st); \
    if (nc == NCA) return dct_xpass_launch<T, L, NCA>(mode, in, out, rbuf, psi0, p, qtw, st); \
    return dct_xpass_launch<T, L, 16>(mode, in, out, rbuf, psi0, p, qtw, st);                \
  }
  SMX_FOR_EACH_LENGTH(SMX_CASE)
#undef SMX_CASE
  return hipErrorInvalidValue;
}

template <typename T, int L, int NC>
static hipError_t dct_zpass_launch(int in_mode, int out_mode, int op, const T* in, T* out, const double* real_in,
                                   double* real_out, const PParams<T>& p, const cplx<T>* qtw, hipStream_t st) {
  const dim3 grid((unsigned)(p.nxy / NC)), block((NC / 2) * (L / Plan<L>::E));
  if (in_mode == Z_IN_SPEC && out_mode == Z_OUT_SPEC)
    SMX_LAUNCH((k_dct_zpass<T, L, NC, Z_IN_SPEC, Z_OUT_SPEC>), grid, block, 0, st, in, out, real_in, real_out, op, p, qtw);
  else if constexpr (NC == neumann_nc<T>()) {
    if (in_mode == Z_IN_REAL && out_mode == Z_OUT_SPEC)
      SMX_LAUNCH((k_dct_zpass<T, L, NC, Z_IN_REAL, Z_OUT_SPEC>), grid, block, 0, st, in, out, real_in, real_out, op, p, qtw);
    else if (in_mode == Z_IN_SPEC && out_mode == Z_OUT_REAL)
      SMX_LAUNCH((k_dct_zpass<T, L, NC, Z_IN_SPEC, Z_OUT_REAL>), grid, block, 0, st, in, out, real_in, real_out, op, p, qtw);
    else
      return hipErrorInvalidValue;
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename T>
hipError_t NeumannOps<T>::zpass(int in_mode, int out_mode, int op, const T* in, T* out, const double* real_in,
                                double* real_out, const PParams<T>& p, const cplx<T>* qtw, hipStream_t st, int nc,
                                bool dry) {
  constexpr int NCD = neumann_nc<T>(), NCA = neumann_nc_alt<T>();
  if (nc == 0) nc = NCD;
  if (nc != NCD && !(nc == NCA && in_mode == Z_IN_SPEC && out_mode == Z_OUT_SPEC)) return hipErrorInvalidValue;
  if (p.nxy % nc != 0) return hipErrorInvalidValue;  // whole tiles
#define SMX_CASE(L)                                                                                           \
  if (p.nz == L) {                                                                                            \
    if (dry) return hipSuccess;                                                                               \
    return nc == NCD ? dct_zpass_launch<T, L, NCD>(in_mode, out_mode, op, in, out, real_in, real_out, p, qtw, st) \
                     : dct_zpass_launch<T, L, NCA>(in_mode, out_mode, op, in, out, real_in, real_out, p, qtw, st); \
  }
  SMX_FOR_EACH_LENGTH(SMX_CASE)
#undef SMX_CASE
  return hipErrorInvalidValue;
}



// Plan of the adjoint x pass: the E = 8 variant where there is one and its block stays at
// <= 512 threads (half the registers per thread of the default plan).
template <int L>
constexpr int adj_variant() {
  if constexpr (HasPlanV<L, 1>::value) return (kNC * (L / PlanV<L, 1>::E) <= 512) ? 1 : 0;
  else return 0;
}

template <typename T>
hipError_t AdjointOps<T>::xpass_adj(cplx<T>* stage, cplx<T>* carry, const cplx<T>* psin, const PParams<T>& p,
                                    T cnext, int init, int tail, double q0d, double alphad, double* acc,
                                    hipStream_t st) {
  const dim3 grid((unsigned)(p.ny / kNC), (unsigned)p.nzh);
#define SMX_CASE(L)                                                                                      \
  if (p.nx == L) {                                                                                       \
    constexpr int V = adj_variant<L>();                                                                  \
    SMX_LAUNCH((k_xpass_adj<T, L, kNC, V>), grid, dim3(kNC * (L / PlanV<L, V>::E)), 0, st, stage, carry, psin, \
               p, cnext, init, tail, q0d, alphad, acc);                                                  \
    return hipGetLastError();                                                                            \
  }
  SMX_FOR_EACH_LENGTH(SMX_CASE)
#undef SMX_CASE
  return hipErrorInvalidValue;
}

template <typename T>
hipError_t AdjointOps<T>::adj_step_grp(const GrpArgs<T>* gz, const AdjGrpArgs<T>* gx, int nbox, const PParams<T>& p,
                                       int64_t npencils, const PassSched* s, T cnext, int init, int tail,
                                       hipStream_t st) {
  using PO = PeriodicOps<T>;
  hipError_t e = PO::ypass_grp(gz, nbox, p, npencils, +1, s[SLOT_YINV], st);
  if (e == hipSuccess) e = PO::zpass_grp(gz, nbox, p, Z_OP_MUL, s[SLOT_Z], st);
  if (e == hipSuccess) e = PO::ypass_grp(gz, nbox, p, npencils, -1, s[SLOT_YFWD], st);
  if (e != hipSuccess) return e;
  return xpass_adj_grp(gx, nbox, p, cnext, init, tail, st);
}

template <typename T>
hipError_t AdjointOps<T>::xpass_adj_grp(const AdjGrpArgs<T>* gx, int nbox, const PParams<T>& p, T cnext, int init,
                                        int tail, hipStream_t st) {
  if (nbox < 1 || nbox > 65535) return hipErrorInvalidValue;
  const dim3 grid((unsigned)(p.ny / kNC), (unsigned)p.nzh, (unsigned)nbox);
#define SMX_CASE(L)                                                                                      \
  if (p.nx == L) {                                                                                       \
    constexpr int V = adj_variant<L>();                                                                  \
    SMX_LAUNCH((k_xpass_adj_grp<T, L, kNC, V>), grid, dim3(kNC * (L / PlanV<L, V>::E)), 0, st, gx, cnext, init, \
               tail);                                                                                    \
    return hipGetLastError();                                                                            \
  }
  SMX_FOR_EACH_LENGTH(SMX_CASE)
#undef SMX_CASE
  return hipErrorInvalidValue;
}

// Plan of a pass kernel that keeps several fields in registers: the E = 8 variant where
// there is one and the block stays at <= 512 threads.
template <int L, int LANES>
constexpr int light_variant() {
  if constexpr (HasPlanV<L, 1>::value) return (LANES * (L / PlanV<L, 1>::E) <= 512) ? 1 : 0;
  else return 0;
}

// Tiles of the advected-step kernels (measured at 256^3): the z kernels take 128-byte rows and,
// in FP32, the E = 16 plan (fewer exchanges; FP64 keeps the E = 8 plan for registers); the
// Stokes x pass, which holds three fields, takes 8-column tiles and the E = 8 plan; the
// CNAB2 x pass takes 16 columns and, in FP32, the E = 16 plan.
template <typename T>
struct AdvectOps {
  static constexpr int NCX = 16;                         // x pass with the CNAB2 update
  static constexpr int NCS = 8;                          // Stokes x pass
  static constexpr int NCZ = sizeof(T) == 4 ? 32 : 16;   // z kernels
  template <int L> static constexpr int vx() { return sizeof(T) == 4 ? 0 : light_variant<L, NCX>(); }
  template <int L> static constexpr int vz() { return sizeof(T) == 4 ? 0 : light_variant<L, NCZ / 2>(); }
  static hipError_t xpass_adv(int mode, const cplx<T>* in, cplx<T>* rbuf, const cplx<T>* psi0, cplx<T>* oa,
                              cplx<T>* ox, cplx<T>* os, const PParams<T>& p, hipStream_t st) {
    if (p.ny % NCX) return hipErrorInvalidValue;
    const dim3 grid((unsigned)(p.ny / NCX), (unsigned)p.nzh);
#define SMX_CASE(L)                                                                                       \
  if (p.nx == L) {                                                                                        \
    constexpr int V = vx<L>();                                                                            \
    const dim3 block(NCX * (L / PlanV<L, V>::E));                                                         \
    if (mode == X_STEP)                                                                                   \
      SMX_LAUNCH((k_xpass_adv<T, L, NCX, X_STEP, V>), grid, block, 0, st, in, rbuf, psi0, oa, ox, os, p);  \
    else                                                                                                  \
      SMX_LAUNCH((k_xpass_adv<T, L, NCX, X_BOOT, V>), grid, block, 0, st, in, rbuf, psi0, oa, ox, os, p);  \
    return hipGetLastError();                                                                             \
  }
    SMX_FOR_EACH_LENGTH(SMX_CASE)
#undef SMX_CASE
    return hipErrorInvalidValue;
  }
  static hipError_t stokes(cplx<T>* s0, cplx<T>* s1, cplx<T>* s2, const AParams<T>& ap, hipStream_t st) {
    const PParams<T>& p = ap.p;
    if (p.ny % NCS) return hipErrorInvalidValue;
    const dim3 grid((unsigned)(p.ny / NCS), (unsigned)p.nzh);
#define SMX_CASE(L)                                                                                \
  if (p.nx == L) {                                                                                 \
    constexpr int V = light_variant<L, NCS>();                                                     \
    SMX_LAUNCH((k_xpass_stokes<T, L, NCS, V>), grid, dim3(NCS * (L / PlanV<L, V>::E)), 0, st, s0, s1, s2, ap); \
    return hipGetLastError();                                                                      \
  }
    SMX_FOR_EACH_LENGTH(SMX_CASE)
#undef SMX_CASE
    return hipErrorInvalidValue;
  }
  static hipError_t zforce(const cplx<T>* bpsi, const cplx<T>* bx, const cplx<T>* by, const cplx<T>* bs,
                           cplx<T>* o0, cplx<T>* o1, cplx<T>* o2, const AParams<T>& ap, hipStream_t st) {
    const PParams<T>& p = ap.p;
    if (p.nxy % NCZ) return hipErrorInvalidValue;
    const dim3 grid((unsigned)(p.nxy / NCZ));
#define SMX_CASE(L)                                                                                     \
  if (p.nz == L) {                                                                                      \
    constexpr int V = vz<L>();                                                                          \
    SMX_LAUNCH((k_zpass_force<T, L, NCZ, V>), grid, dim3((NCZ / 2) * (L / PlanV<L, V>::E)), 0, st, bpsi, bx, by, \
               bs, o0, o1, o2, ap);                                                                     \
    return hipGetLastError();                                                                           \
  }
    SMX_FOR_EACH_LENGTH(SMX_CASE)
#undef SMX_CASE
    return hipErrorInvalidValue;
  }
  static hipError_t zadvnl(const cplx<T>* u0, const cplx<T>* u1, const cplx<T>* u2, const cplx<T>* bpsi,
                           const cplx<T>* bx, const cplx<T>* by, cplx<T>* nout, const AParams<T>& ap, hipStream_t st) {
    const PParams<T>& p = ap.p;
    if (p.nxy % NCZ) return hipErrorInvalidValue;
    const dim3 grid((unsigned)(p.nxy / NCZ));
#define SMX_CASE(L)                                                                                     \
  if (p.nz == L) {                                                                                      \
    constexpr int V = vz<L>();                                                                          \
    SMX_LAUNCH((k_zpass_advnl<T, L, NCZ, V>), grid, dim3((NCZ / 2) * (L / PlanV<L, V>::E)), 0, st, u0, u1, u2, \
               bpsi, bx, by, nout, ap);                                                                 \
    return hipGetLastError();                                                                           \
  }
    SMX_FOR_EACH_LENGTH(SMX_CASE)
#undef SMX_CASE
    return hipErrorInvalidValue;
  }
  // partial sums written by zadvnl: one per z tile
  static int zpass_blocks(const PParams<T>& p) { return (int)(p.nxy / NCZ); }
};

// Kernels of the advected Neumann step (smx_neuadv.h): 128-byte rows (x kernels where ny
// allows it); the z kernels take the Neumann z pass's tiles.
template <typename T>
struct NeuAdvOps {
  static constexpr int NCW = neumann_nc<T>();  // 128-byte rows
  template <int L> static constexpr int ncz() { return NCW; }
  template <int L, int NC>
  static hipError_t xadv_launch(int mode, const T* in, T* rbuf, const T* psi0, T* oa, T* ox, T* os,
                                const PParams<T>& p, const cplx<T>* qtw, hipStream_t st) {
    const dim3 grid((unsigned)(p.ny / NC), (unsigned)p.nzh), block((NC / 2) * (L / Plan<L>::E));
    if (mode == X_STEP)
      SMX_LAUNCH((k_dctx_adv<T, L, NC, X_STEP>), grid, block, 0, st, in, rbuf, psi0, oa, ox, os, p, qtw);
    else if (mode == X_BOOT)
      SMX_LAUNCH((k_dctx_adv<T, L, NC, X_BOOT>), grid, block, 0, st, in, rbuf, psi0, oa, ox, os, p, qtw);
    else
      return hipErrorInvalidValue;
    return hipGetLastError();
  }
  static hipError_t xpass_adv(int mode, const T* in, T* rbuf, const T* psi0, T* oa, T* ox, T* os,
                              const PParams<T>& p, const cplx<T>* qtw, hipStream_t st) {
    if (p.ny % 16) return hipErrorInvalidValue;
    const bool wide = (p.ny % NCW == 0);
#define SMX_CASE(L)                                                                                      \
  if (p.nx == L) {                                                                                       \
    if (wide) return xadv_launch<L, NCW>(mode, in, rbuf, psi0, oa, ox, os, p, qtw, st);                  \
    return xadv_launch<L, 16>(mode, in, rbuf, psi0, oa, ox, os, p, qtw, st);                             \
  }
    SMX_FOR_EACH_LENGTH(SMX_CASE)
#undef SMX_CASE
    return hipErrorInvalidValue;
  }
  // Stokes pass: the E = 8 plan where there is one (it holds every component's spectrum of its
  // slots in registers), blocks of at most 256 threads
  template <int L> static constexpr int vs() { return HasPlanV<L, 1>::value ? 1 : 0; }
  template <int L> static constexpr int ncs() { return std::min(NCW, 512 / (L / PlanV<L, vs<L>()>::E)); }
  template <int L, int NC>
  static hipError_t stokes_launch(T* s0, T* s1, T* s2, const AParams<T>& ap, const cplx<T>* qtw, T qnx, T qny,
                                  T qnz, hipStream_t st) {
    constexpr int V = vs<L>();
    static_assert(((size_t)L * (NC / 2 + 1) + 2 * L) * sizeof(cplx<T>) + L * sizeof(T) <= kMaxLds,
                  "Stokes tile exceeds LDS");
    const dim3 grid((unsigned)(ap.p.ny / NC), (unsigned)ap.p.nzh);
    SMX_LAUNCH((k_dctx_stokes<T, L, NC, V>), grid, dim3((NC / 2) * (L / PlanV<L, V>::E)), 0, st, s0, s1, s2, ap, qtw,
               qnx, qny, qnz);
    return hipGetLastError();
  }
  static hipError_t stokes(T* s0, T* s1, T* s2, const AParams<T>& ap, const cplx<T>* qtw, T qnx, T qny, T qnz,
                           hipStream_t st) {
    const PParams<T>& p = ap.p;
    if (p.ny % 16) return hipErrorInvalidValue;
#define SMX_CASE(L)                                                                           \
  if (p.nx == L) {                                                                            \
    constexpr int NC = ncs<L>();                                                              \
    if (p.ny % NC == 0) return stokes_launch<L, NC>(s0, s1, s2, ap, qtw, qnx, qny, qnz, st);  \
    return stokes_launch<L, 16>(s0, s1, s2, ap, qtw, qnx, qny, qnz, st);                      \
  }
    SMX_FOR_EACH_LENGTH(SMX_CASE)
#undef SMX_CASE
    return hipErrorInvalidValue;
  }
  static hipError_t zforce(const T* bpsi, const T* bx, const T* by, const T* bs, T* o0, T* o1, T* o2,
                           const AParams<T>& ap, const cplx<T>* qtw, hipStream_t st) {
    const PParams<T>& p = ap.p;
#define SMX_CASE(L)                                                                                     \
  if (p.nz == L) {                                                                                      \
    constexpr int NC = ncz<L>();                                                                        \
    if (p.nxy % NC) return hipErrorInvalidValue;                                                        \
    SMX_LAUNCH((k_dctz_force<T, L, NC>), grid_of(p, NC), dim3((NC / 2) * (L / Plan<L>::E)), 0, st, bpsi, bx, by, \
               bs, o0, o1, o2, ap, qtw);                                                                \
    return hipGetLastError();                                                                           \
  }
    SMX_FOR_EACH_LENGTH(SMX_CASE)
#undef SMX_CASE
    return hipErrorInvalidValue;
  }
  static hipError_t zadvnl(const T* u0, const T* u1, const T* u2, const T* bpsi, const T* bx, const T* by, T* nout,
                           const AParams<T>& ap, const cplx<T>* qtw, hipStream_t st) {
    const PParams<T>& p = ap.p;
#define SMX_CASE(L)                                                                                     \
  if (p.nz == L) {                                                                                      \
    constexpr int NC = ncz<L>();                                                                        \
    if (p.nxy % NC) return hipErrorInvalidValue;                                                        \
    SMX_LAUNCH((k_dctz_advnl<T, L, NC>), grid_of(p, NC), dim3((NC / 2) * (L / Plan<L>::E)), 0, st, u0, u1, u2,   \
               bpsi, bx, by, nout, ap, qtw);                                                            \
    return hipGetLastError();                                                                           \
  }
    SMX_FOR_EACH_LENGTH(SMX_CASE)
#undef SMX_CASE
    return hipErrorInvalidValue;
  }
  static dim3 grid_of(const PParams<T>& p, int nc) { return dim3((unsigned)(p.nxy / nc)); }
  // partial sums written by zadvnl: one per z tile
  static int zpass_blocks(const PParams<T>& p) {
#define SMX_CASE(L) \
  if (p.nz == L) return (int)(p.nxy / ncz<L>());
    SMX_FOR_EACH_LENGTH(SMX_CASE)
#undef SMX_CASE
    return 0;
  }
};

}  // namespace smx
