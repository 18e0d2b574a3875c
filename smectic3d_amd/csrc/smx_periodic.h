// smx_periodic.h -- periodic (R2C/C2R) pass kernels of the Swift-Hohenberg step.
//
// HBM layout of every spectral-stage array: [kz][x or kx][y or ky], y fastest,
// kz in [0, nz/2] (nz/2+1 planes).  A step is four in-place launches on the
// stage buffer A plus the CNAB2 carry R:
//   k_ypass<+1>  (x,ky,kz) -> (x,y,kz)       inverse y
//   k_zpass      (x,y,kz)  -> (x,y,kz)       z C2R, N(psi), z R2C
//   k_ypass<-1>  (x,y,kz)  -> (x,ky,kz)      forward y
//   k_xpass      (x,ky,kz) -> (x,ky,kz)      forward x, CNAB2 update, inverse x
// Real fields (host I/O, diagnostics) are dense [x][y][z] in FP64.
#pragma once
#include <type_traits>

#include "smx_dev.h"

namespace smx {

// Runtime parameters shared by the periodic kernels.
template <typename T>
struct PParams {
  int nx, ny, nz, nzh;
  int64_t nxy;      // nx*ny (columns of a kz plane)
  T scale;          // 1/(nx ny nz), pf3d.cpp:62
  T eps, alpha, q02, beta, gamma, dtd2;
  int nl_model;     // 0 quintic, 1 cubic
  int mid;          // mid-plane penalty z index, <0 off
  T pen;            // 2 w_mid / dz^2
  const T* kx;      // wavenumbers, FFT order, Nyquist +pi/dx (pf3d.cpp:162-190)
  const T* ky;
  const T* kz;
  const cplx<T>* twx;  // e^{-2 pi i m/n} tables, m < n
  const cplx<T>* twy;
  const cplx<T>* twz;
  // FP64 mean (k = 0) mode of the mixed path: partial sums written by the z
  // pass, consumed by the x pass.  Unused (nullptr) in FP64 mode.
  double* partial;
  int npartial;
  double* mean_state;  // [0] = R(k=0) in FP64, [1] = psi_hat(0) in FP64
  double m_scale, m_dtd2, m_c1, m_c2;  // FP64 scale, dt/2, C1(0), C2(0) for the mean mode
  // Recording for the adjoint sweep (nullptr: off): N'(psi_n) by the z pass, layout
  // [z][x][y]; psi_hat_{n+1} by the x pass, layout of R.
  T* rec_np = nullptr;
  cplx<T>* rec_psi = nullptr;
  // Multiplier of the z pass in the Z_OP_MUL form (the adjoint sweep: the recorded N'(psi_n)),
  // layout [z][x][y].
  const T* mul_np = nullptr;
  // k-space multiplier (KMode, < 0: none) applied to the spectral input of the inverse x pass
  // (X_INV) as it is loaded; the two multipliers of X_MUL2 (kop, kop2)
  int kop = -1;
  int kop2 = -1;
};

// dN/dpsi: 3 beta psi^2 - 5 gamma psi^4 (quintic) or 2 beta psi - 3 gamma psi^2 (cubic).
template <typename T>
__device__ __forceinline__ T nonlin_prime(const PParams<T>& p, T s) {
  if (p.nl_model == 1) return s * ((T)2 * p.beta - (T)3 * p.gamma * s);
  const T s2 = s * s;
  return s2 * ((T)3 * p.beta - (T)5 * p.gamma * s2);
}

// Local nonlinearity: beta psi^3 - gamma psi^5 (pf3d.cpp:460-461) or the cubic model
// beta psi^2 - gamma psi^3.
template <typename T>
__device__ __forceinline__ T nonlin3(const PParams<T>& p, T s) {
  return s * s * (p.beta - p.gamma * s);
}
template <typename T>
__device__ __forceinline__ T nonlin5(const PParams<T>& p, T s) {
  const T s2 = s * s;
  return s2 * s * (p.beta - p.gamma * s2);
}
template <typename T>
__device__ __forceinline__ T nonlin(const PParams<T>& p, T s) {
  return (p.nl_model == 1) ? nonlin3(p, s) : nonlin5(p, s);
}

// Factor of the k-space multiplier `op` at (qx, qy, qz): the multiplier itself, or its
// imaginary part for the first derivatives (i q).
template <typename T>
__device__ __forceinline__ T kfactor(const PParams<T>& p, int op, T qx, T qy, T qz) {
  switch (op) {
    case KM_ALIN: {
      const T d = qx * qx + qy * qy + qz * qz - p.q02;
      return p.eps - p.alpha * d * d;
    }
    case KM_IKX: return qx;
    case KM_IKY: return qy;
    case KM_IKZ: return qz;
    case KM_DXX: return -qx * qx;
    case KM_DYY: return -qy * qy;
    case KM_DZZ: return -qz * qz;
    case KM_DXY: return -qx * qy;
    default: return -(qx * qx) - qy * qy;  // KM_LAP_XY
  }
}
// The multiplier `op` applied to one complex coefficient.
template <typename T>
__device__ __forceinline__ cplx<T> kapply(const PParams<T>& p, int op, T qx, T qy, T qz, cplx<T> a) {
  const T m = kfactor(p, op, qx, qy, qz);
  return (op == KM_IKX || op == KM_IKY || op == KM_IKZ) ? mk(-m * a.y, m * a.x) : m * a;
}

// ---------------------------------------------------------------------------
// y pass: contiguous pencils of length L = ny; pencils g = kz*nx + x.  One tile (NC pencils
// from bx*NC) per block: the body of k_ypass and of the grouped k_ypass_grp.
template <typename T, int L, int SIGN, int NC, int V>
__device__ __forceinline__ void ypass_tile(const cplx<T>* in, cplx<T>* out, int64_t npencils,
                                           const cplx<T>* __restrict__ tw, const T* __restrict__ kmul, int kop,
                                           int bx) {
  // lanes run along each pencil (tc fastest), so the first-pass loads and last-pass
  // stores go straight between HBM and registers (16 contiguous complex per lane group)
  constexpr int E = PlanV<L, V>::E, TPC = L / E;
  using AD = ColAddr<L>;
  using TL = TileA<T, L, E, AD>;
  __shared__ cplx<T> S[NC * AD::LP];
  const int tid = threadIdx.x, tc = tid % TPC, col = tid / TPC;
  __shared__ cplx<T> TW[L];
  const cplx<T>* twl = tw_lds<L, NC * TPC>(TW, tw, tid);
  const int64_t pen = (int64_t)bx * NC + col;
  const bool valid = pen < npencils;
  const cplx<T>* src = in + pen * L;
  // LDS copy of the multiplier's k_y (global loads behind the row loads would double the
  // vector memory instructions)
  __shared__ T KM[L];
  if (kmul != nullptr) {
    for (int i = tid; i < L; i += NC * TPC) KM[i] = kmul[i];
    __syncthreads();
  }
  cplx<T> v[E];
  using FFT = BlockFFTA<T, L, E, AD, SIGN,
                        typename std::conditional<(SIGN < 0), typename PlanV<L, V>::Fwd, typename PlanV<L, V>::Inv>::type>;
  constexpr int RF = (SIGN < 0) ? fwd_first<L, V>() : inv_first<L, V>();
  constexpr int RL = (SIGN < 0) ? fwd_last<L, V>() : inv_last<L, V>();
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = valid ? src[TL::template pos_in<RF>(tc, e)] : mk<T>(0, 0);
  if (kmul != nullptr) {  // multiply by i k_y (spectral d/dy, pf3d.cpp:405) or by -k_y^2 (kop KM_DYY)
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const T k = KM[TL::template pos_in<RF>(tc, e)];
      v[e] = (kop == KM_DYY) ? (-k * k) * v[e] : mk(-k * v[e].y, k * v[e].x);
    }
  }
  FFT::run(v, S, col, tc, twl);
  if (valid) {
    cplx<T>* dst = out + pen * L;
#pragma unroll
    for (int e = 0; e < E; ++e) dst[TL::template pos_in<RL>(tc, e)] = v[e];
  }
}
template <typename T, int L, int SIGN, int NC, int V = 0>
__global__ __launch_bounds__(NC*(L / PlanV<L, V>::E)) void k_ypass(const cplx<T>* in, cplx<T>* out,
                                                               int64_t npencils,
                                                               const cplx<T>* __restrict__ tw,
                                                               const T* __restrict__ kmul, int kop) {
  ypass_tile<T, L, SIGN, NC, V>(in, out, npencils, tw, kmul, kop, blockIdx.x);
}

// ---------------------------------------------------------------------------
// x pass.  Tile: plane kz = by, columns y0 = bx*NC .. y0+NC-1, rows x (the body of k_xpass,
// bx = blockIdx.x, by = blockIdx.y, and of the grouped k_xpass_grp).
// X_MUL2: F^-1[m_kop F in] into out and F^-1[m_kop2 F in] into rbuf from one forward transform
// (surfTrack's x-dependent multipliers; FP64).
enum XMode { X_STEP = 0, X_FWD_SCALED = 1, X_BOOT = 2, X_INV = 3, X_MUL2 = 4 };

// ORD (X_STEP only) orders the carry traffic.  0: R is read and rewritten element by
// element between the two transforms.  1: all of R (and k_x) is read first, the new R is
// kept in registers through the inverse transform and every global store is issued last.
// 2: all of R is read first and each new R is stored as soon as it is formed.  3: the new R
// goes to the LDS tile (as on the self-conjugate planes) and to HBM after the loop, so the
// loop issues no stores and keeps no extra registers.  A wave's loads and stores retire
// through one in-order counter, so a load issued behind a store waits for that store;
// which order hides more latency depends on the precision and tile shape, and the
// autotuner measures all four.  k_x comes from an LDS copy in every order.  The tile's
// exchanges order LDS only (lds_barrier), so R stores issued before the inverse transform
// drain while it runs: every global access of a thread is to entries no other thread of the
// block reads or writes, except the in-place rows, whose loads are consumed before the first
// exchange.
template <typename T, int L, int NC, int MODE, int V, int ORD>
__device__ __forceinline__ void xpass_tile(const cplx<T>* in, cplx<T>* out, cplx<T>* __restrict__ rbuf,
                                           const cplx<T>* __restrict__ psi0, const PParams<T> p, int bx, int by) {
  static_assert(ORD == 0 || MODE == X_STEP, "the late-store order is a form of the CNAB2 step");
  constexpr int E = PlanV<L, V>::E, TPC = L / E, NT = NC * TPC, NCP = NC + 1;
  constexpr uint32_t CB = sizeof(cplx<T>);
  __shared__ cplx<T> S[L * NCP];
  __shared__ double red[64];
  const int tid = threadIdx.x, col = tid % NC, tc = tid / NC;
  __shared__ cplx<T> TW[L];
  const cplx<T>* twl = tw_lds<L, NT>(TW, p.twx, tid);
  __shared__ T KX[(MODE == X_STEP || MODE == X_BOOT || MODE == X_INV || MODE == X_MUL2) ? L : 1];
  if constexpr (MODE == X_STEP || MODE == X_BOOT || MODE == X_MUL2) {
    for (int i = tid; i < L; i += NT) KX[i] = p.kx[i];
    __syncthreads();  // a one-pass plan has no barrier of its own before the update reads KX
  } else if constexpr (MODE == X_INV) {
    if (p.kop >= 0) {  // k_x of the multiplier
      for (int i = tid; i < L; i += NT) KX[i] = p.kx[i];
      __syncthreads();
    }
  }
  const int kz = by, y0 = bx * NC;
  const int64_t plane = (int64_t)kz * p.nxy;
  const uint32_t rowb = (uint32_t)p.ny * CB;  // bytes between kx rows
  // On the self-conjugate planes kz = 0 and kz = nz/2 a block takes NC/2 columns ky and
  // their partners -ky, so that R can be projected onto Hermitian symmetry in place
  // (the carried state otherwise accumulates an anti-Hermitian part there that the
  // C2R never sees -- it reads only real parts on those planes -- and that grows by
  // (1 + dt L/2)/(1 - dt L/2) per step in the unstable band L(k) > 0; the reference
  // re-derives psi_hat from real psi each step, pf3d.cpp:395, and never carries it).
  constexpr int H = NC / 2;
  const bool herm = (MODE == X_STEP || MODE == X_BOOT) && (kz == 0 || 2 * kz == p.nz);
  int ky = y0 + col, pcol = col;
  if (herm) {
    const int b = bx * H;
    if (col < H) {
      ky = b + col;
      pcol = (b + col == 0) ? col : col + H;
    } else {
      ky = (b == 0 && col == H) ? p.ny / 2 : p.ny - (b + col - H);
      pcol = (b == 0 && col == H) ? col : col - H;
    }
  }
  // wave-uniform plane bases, per-lane 32-bit byte offsets (row kx, column ky)
  const cplx<T>* src = in + plane;
  cplx<T>* rb = rbuf + plane;
  const uint32_t cofs = (uint32_t)ky * CB;
  cplx<T> v[E];
  // first-pass inputs straight from HBM: lanes run along y (16 contiguous complex per row)
  {
    constexpr int R0 = (MODE == X_INV) ? inv_first<L, V>() : fwd_first<L, V>();
#pragma unroll
    for (int e = 0; e < E; ++e)
      v[e] = ld_off(src, (uint32_t)Tile<T, L, E, NCP>::template pos_in<R0>(tc, e) * rowb + cofs);
    if constexpr (MODE == X_INV) {
      if (p.kop >= 0) {
        const T kyv = p.ky[ky], kzv = p.kz[kz];
#pragma unroll
        for (int e = 0; e < E; ++e)
          v[e] = kapply(p, p.kop, KX[Tile<T, L, E, NCP>::template pos_in<R0>(tc, e)], kyv, kzv, v[e]);
      }
    }
  }
  // FP64 mean mode (mixed path): the block owning k = 0 sums the z-pass partials.
  double nsum = 0.0;
  const bool owner0 = (p.partial != nullptr) && kz == 0 && y0 == 0;
  if (owner0) {
    double a = 0.0;
    for (int i = tid; i < p.npartial; i += NT) a += p.partial[i];
    nsum = block_sum(a, red);  // valid in thread 0 (= col 0, tc 0)
  }
  cplx<T> cr[(ORD == 1 || ORD == 2) ? E : 1];  // ORD 1, 2: R (ORD 1: then the new R)
  if constexpr (MODE != X_INV) {
    FwdFFT<T, L, NCP, V, true>::run(v, S, col, tc, twl);
    const T kyv = p.ky[ky], kzv = p.kz[kz];
    if constexpr (MODE == X_MUL2) {
      // F^-1[m_kop F in] into out and F^-1[m_kop2 F in] into rbuf from this one forward
      // transform (p.scale undoes the round trip), the scaled spectrum kept in registers between
      // the two: the arithmetic of X_FWD_SCALED followed by X_INV with each multiplier
      cplx<T> sp[E];
#pragma unroll
      for (int e = 0; e < E; ++e) sp[e] = p.scale * v[e];
#pragma unroll
      for (int pass = 0; pass < 2; ++pass) {
#pragma unroll
        for (int e = 0; e < E; ++e)
          v[e] = kapply(p, pass == 0 ? p.kop : p.kop2, KX[Tile<T, L, E, NCP>::template pos_in<fwd_last<L, V>()>(tc, e)],
                        kyv, kzv, sp[e]);
        InvFFT<T, L, NCP, V, true>::run(v, S, col, tc, twl);
        cplx<T>* dst = (pass == 0 ? out : rbuf) + plane;
#pragma unroll
        for (int e = 0; e < E; ++e)
          st_off(dst, (uint32_t)Tile<T, L, E, NCP>::template pos_in<inv_last<L, V>()>(tc, e) * rowb + cofs, v[e]);
      }
      return;
    }
    if constexpr (ORD == 1 || ORD == 2) {
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int kxi = Tile<T, L, E, NCP>::template pos_in<fwd_last<L, V>()>(tc, e);
        cr[e] = ld_off(rb, (uint32_t)kxi * rowb + cofs);
      }
    }
    constexpr bool kStage = (ORD == 3);  // every plane's R' through the LDS tile
    if (herm || kStage) lds_barrier();  // S receives R' below; the FFT's last reads must be done
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int kxi = Tile<T, L, E, NCP>::template pos_in<fwd_last<L, V>()>(tc, e);
      const uint32_t o = (uint32_t)kxi * rowb + cofs;
      if constexpr (MODE == X_FWD_SCALED) {
        v[e] = p.scale * v[e];
        if (owner0 && tid == 0 && kxi == 0) {
          p.mean_state[1] = p.m_scale * nsum;  // FP64 mean of the transformed field
          v[e] = mk((T)p.mean_state[1], (T)0);
        }
      } else {
        const T kxv = KX[kxi];
        const T k2 = kxv * kxv + kyv * kyv + kzv * kzv;
        const T d = k2 - p.q02;
        const T aLin = p.eps - p.alpha * d * d;  // pf3d.cpp:205-206
        const T C1 = (T)1 + p.dtd2 * aLin, C2 = (T)1 - p.dtd2 * aLin;  // :214-215
        const cplx<T> nn = p.scale * v[e];
        if constexpr (MODE == X_STEP) {
          // psi_new = (C1 psi + dt/2 (3 N - N_past))/C2 carried as R = C1 psi - dt/2 N_past
          // (pf3d.cpp:476-478): psi_new = (R + 3 dt/2 N)/C2, R' = C1 psi_new - dt/2 N.
          const cplx<T> r = (ORD == 0 || ORD == 3) ? ld_off<cplx<T>>(rb, o) : cr[e];
          const T ic2 = recip(C2);
          cplx<T> ps = mk(ic2 * (r.x + (T)3 * p.dtd2 * nn.x), ic2 * (r.y + (T)3 * p.dtd2 * nn.y));
          cplx<T> rn = mk(C1 * ps.x - p.dtd2 * nn.x, C1 * ps.y - p.dtd2 * nn.y);
          if (owner0 && tid == 0 && kxi == 0) {
            // k = 0 in FP64: N_hat(0) = sum of N over the grid.
            const double dd = p.m_dtd2, n0 = p.m_scale * nsum;
            const double ps0 = (p.mean_state[0] + 3.0 * dd * n0) / p.m_c2;
            p.mean_state[0] = p.m_c1 * ps0 - dd * n0;
            p.mean_state[1] = ps0;
            ps = mk((T)ps0, (T)0);
            rn = mk((T)p.mean_state[0], (T)0);
          }
          if (herm || kStage) S[kxi * NCP + col] = rn;
          else if constexpr (ORD != 1) st_off(rb, o, rn);
          if constexpr (ORD == 1) cr[e] = rn;
          if (p.rec_psi != nullptr) st_off(p.rec_psi + plane, o, ps);
          v[e] = ps;
        } else {  // X_BOOT: R = C1 psi0 - dt/2 N_past with N_past = N(psi0) (pf3d.cpp:363-374)
          cplx<T> ps = ld_off(psi0 + plane, o);
          cplx<T> rn = mk(C1 * ps.x - p.dtd2 * nn.x, C1 * ps.y - p.dtd2 * nn.y);
          if (owner0 && tid == 0 && kxi == 0) {
            const double ps0 = p.mean_state[1];  // set from the FP64 sum of psi0
            p.mean_state[0] = p.m_c1 * ps0 - p.m_dtd2 * p.m_scale * nsum;
            ps = mk((T)ps0, (T)0);
            rn = mk((T)p.mean_state[0], (T)0);
          }
          if (herm) S[kxi * NCP + col] = rn;
          else st_off(rb, o, rn);
          v[e] = ps;
        }
      }
    }
    if (herm) {  // R(k) <- (R(k) + conj R(-k))/2 with -k = (nx - kx, partner column)
      lds_barrier();
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int kxi = Tile<T, L, E, NCP>::template pos_in<fwd_last<L, V>()>(tc, e);
        const cplx<T> a = S[kxi * NCP + col], b = S[((L - kxi) % L) * NCP + pcol];
        const cplx<T> h = mk((T)0.5 * (a.x + b.x), (T)0.5 * (a.y - b.y));
        if constexpr (ORD != 1) st_off(rb, (uint32_t)kxi * rowb + cofs, h);
        else cr[e] = h;
      }
    } else if (kStage) {  // each thread reads back its own entries
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int kxi = Tile<T, L, E, NCP>::template pos_in<fwd_last<L, V>()>(tc, e);
        st_off(rb, (uint32_t)kxi * rowb + cofs, S[kxi * NCP + col]);
      }
    }
  }
  cplx<T>* dst = out + plane;
  if constexpr (MODE == X_FWD_SCALED) {
#pragma unroll
    for (int e = 0; e < E; ++e)
      st_off(dst, (uint32_t)Tile<T, L, E, NCP>::template pos_in<fwd_last<L, V>()>(tc, e) * rowb + cofs, v[e]);
  } else {
    InvFFT<T, L, NCP, V, true>::run(v, S, col, tc, twl);
#pragma unroll
    for (int e = 0; e < E; ++e)
      st_off(dst, (uint32_t)Tile<T, L, E, NCP>::template pos_in<inv_last<L, V>()>(tc, e) * rowb + cofs, v[e]);
  }
  if constexpr (ORD == 1) {
#pragma unroll
    for (int e = 0; e < E; ++e)
      st_off(rb, (uint32_t)Tile<T, L, E, NCP>::template pos_in<fwd_last<L, V>()>(tc, e) * rowb + cofs, cr[e]);
  }
}
template <typename T, int L, int NC, int MODE, int V = 0, int ORD = 0>
__global__ __launch_bounds__(NC*(L / PlanV<L, V>::E)) void k_xpass(const cplx<T>* in, cplx<T>* out,
                                                               cplx<T>* __restrict__ rbuf,
                                                               const cplx<T>* __restrict__ psi0,
                                                               PParams<T> p) {
  xpass_tile<T, L, NC, MODE, V, ORD>(in, out, rbuf, psi0, p, blockIdx.x, blockIdx.y);
}

// ---------------------------------------------------------------------------
// z pass.  Tile: columns f0 .. f0+NC-1 of the (x,y) plane; adjacent columns
// (2q, 2q+1) share one length-nz complex FFT (z = psi_a + i psi_b), so that a lane
// moves both with one 2-complex access per row (a full 128-byte row per 8 lanes).
enum ZIn { Z_IN_SPEC = 0, Z_IN_REAL = 1 };
enum ZOut { Z_OUT_SPEC = 1, Z_OUT_REAL = 2 };
enum ZOp { Z_OP_ID = 0, Z_OP_NL = 1, Z_OP_MUL = 2 };

// One tile (columns bx*NC ..) per block: the body of k_zpass and of the grouped k_zpass_grp.
template <typename T, int L, int NC, int IN, int OUT, int V>
__device__ __forceinline__ void zpass_tile(const cplx<T>* in, cplx<T>* out, const double* __restrict__ real_in,
                                           double* __restrict__ real_out, int op, const PParams<T> p, int bx) {
  // The grid covers whole tiles: nx*ny is a multiple of NC (every plan length is a
  // multiple of 8 and the launcher checks it).
  constexpr int E = PlanV<L, V>::E, TPC = L / E, NPR = NC / 2, NT = NPR * TPC;
  // Tile row pitch: in the spectrum-to-spectrum form (the step's), rows of exactly 128 bytes
  // (8 pairs in FP64, 16 in FP32) unpadded, others padded by one element.  Measured on the FP64 z
  // pass at 256^3: the 144-byte padded pitch took three times the LDS bank-conflict cycles (3.8 M
  // against 1.2 M per launch) and, at 42 KB of LDS per block, one block per CU fewer (46.7 against
  // 44.7 us).  (The x pass is no faster unpadded.)  The real-field forms keep the pad: their
  // lanes run down consecutive rows of one column, which an unpadded 128-byte pitch puts in
  // the same banks.
  constexpr bool UNPAD = IN == Z_IN_SPEC && OUT == Z_OUT_SPEC && NPR * sizeof(cplx<T>) == 128;
  constexpr int NZH = L / 2 + 1, FP = UNPAD ? NPR : NPR + 1;
  constexpr uint32_t CB = sizeof(cplx<T>);
  __shared__ cplx<T> F[L * FP];
  __shared__ double red[64];
  const int tid = threadIdx.x, q = tid % NPR, tc = tid / NPR;
  __shared__ cplx<T> TW[L];
  const cplx<T>* twl = tw_lds<L, NT>(TW, p.twz, tid);
  const int64_t f0 = (int64_t)bx * NC;
  const uint32_t rowb = (uint32_t)p.nxy * CB;  // bytes between kz rows of one column
  const uint32_t qofs = (uint32_t)q * 2 * CB;  // the lane's column pair within a row
  cplx<T> v[E];
  if constexpr (IN == Z_IN_SPEC) {
    // Hermitian extension of the column pair read straight from HBM; only the real
    // parts of kz = 0 and kz = nz/2 enter (the half-complex stage of an FFTW c2r).
    // Element e sits at n = tc + c_e (c_e a compile-time constant once unrolled) and reads row
    // m = min(n, L - n): its byte offset is tc*rowb + c_e*rowb or (L - c_e)*rowb - tc*rowb, one
    // add on a uniform product, decided at compile time except for the one element that can
    // straddle L/2.
    const cplx2<T>* col = reinterpret_cast<const cplx2<T>*>(in + f0);
    const uint32_t tcr = (uint32_t)tc * rowb + qofs;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int c = Tile<T, L, E, FP>::template pos_in<inv_first<L, V>()>(0, e);
      const int n = tc + c;
      const bool lo = (c + TPC - 1 <= L / 2) || (c <= L / 2 && n <= L / 2);  // n <= L/2
      const uint32_t off = lo ? tcr + (uint32_t)c * rowb : (uint32_t)(L - c) * rowb - tcr + 2 * qofs;
      const cplx2<T> ab = ld_off(col, off);
      cplx<T> A = ab.a, B = ab.b;
      if (n == 0 || n == L / 2) {
        A.y = 0;
        B.y = 0;
      }
      v[e] = lo ? mk(A.x - B.y, A.y + B.x) : mk(A.x + B.y, B.x - A.y);
    }
    InvFFT<T, L, FP, V>::run(v, F, q, tc, twl);
  } else {
    // a lane reads one z of both columns of a pair (two coalesced rows) and writes the pair as
    // one complex: consecutive lanes hit consecutive LDS rows, without bank conflicts
    for (int i = tid; i < NPR * L; i += NT) {
      const int qq = i / L, n = i - qq * L;
      F[n * FP + qq] = mk((T)real_in[(f0 + 2 * qq) * L + n], (T)real_in[(f0 + 2 * qq + 1) * L + n]);
    }
    __syncthreads();
    Tile<T, L, E, FP>::template load<fwd_first<L, V>()>(v, F, q, tc);
  }
  // v now holds (psi_a, psi_b) at positions pos_in<fwd_first>.
  constexpr int RP = fwd_first<L, V>();
  static_assert(fwd_first<L, V>() == inv_last<L, V>(), "plan symmetry");
  if constexpr ((OUT & Z_OUT_REAL) != 0) {
    __syncthreads();
    Tile<T, L, E, FP>::template store<RP>(v, F, q, tc);
    __syncthreads();
    for (int i = tid; i < NPR * L; i += NT) {
      const int qq = i / L, n = i - qq * L;
      const cplx<T> w = F[n * FP + qq];
      real_out[(f0 + 2 * qq) * L + n] = (double)w.x;
      real_out[(f0 + 2 * qq + 1) * L + n] = (double)w.y;
    }
  }
  if constexpr ((OUT & Z_OUT_SPEC) != 0) {
    const bool pen = (op == Z_OP_NL) && p.mid > 0 && p.mid < L - 1;
    if (pen) {
      __syncthreads();
      Tile<T, L, E, FP>::template store<RP>(v, F, q, tc);
      __syncthreads();
    }
    if (p.rec_np != nullptr) {  // N'(psi_n) for the adjoint sweep
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int n = Tile<T, L, E, FP>::template pos_in<RP>(tc, e);
        T* row = p.rec_np + (int64_t)n * p.nxy + f0 + 2 * q;
        row[0] = nonlin_prime(p, v[e].x);
        row[1] = nonlin_prime(p, v[e].y);
      }
    }
    // the op and model branches are uniform: one straight loop per case
    if (op == Z_OP_MUL) {
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int n = Tile<T, L, E, FP>::template pos_in<RP>(tc, e);
        const T* row = p.mul_np + (int64_t)n * p.nxy + f0 + 2 * q;
        v[e] = mk(row[0] * v[e].x, row[1] * v[e].y);
      }
    } else if (op == Z_OP_NL && !pen) {
      if (p.nl_model == 1) {
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = mk(nonlin3(p, v[e].x), nonlin3(p, v[e].y));
      } else {
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = mk(nonlin5(p, v[e].x), nonlin5(p, v[e].y));
      }
    } else if (op == Z_OP_NL) {
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int n = Tile<T, L, E, FP>::template pos_in<RP>(tc, e);
        T a = nonlin(p, v[e].x), b = nonlin(p, v[e].y);
        if (n == p.mid) {
          // 2 wMid (psi_{k-1} - 2 psi_k + psi_{k+1})/dz^2 at k = mid (pf3dMorph.cpp:572-581)
          const cplx<T> lo = F[(n - 1) * FP + q], hi = F[(n + 1) * FP + q];
          a += p.pen * (lo.x - (T)2 * v[e].x + hi.x);
          b += p.pen * (lo.y - (T)2 * v[e].y + hi.y);
        }
        v[e] = mk(a, b);
      }
    }
    // FP64 partial of sum N over the tile; stored after the split (its barriers order it)
    if (p.partial != nullptr) wave_partial<E, NT>(v, red, tid);
    FwdFFT<T, L, FP, V>::run(v, F, q, tc, twl);
    __syncthreads();
    Tile<T, L, E, FP>::template store<fwd_last<L, V>()>(v, F, q, tc);
    __syncthreads();
    // split W = X + iY:  X = (W_k + conj W_{-k})/2,  Y = (W_k - conj W_{-k})/(2i); straight to HBM.
    // Thread (q, tc) takes rows k = tc, tc + TPC, ... of its own column pair q.
    cplx2<T>* dst = reinterpret_cast<cplx2<T>*>(out + f0);
#pragma unroll
    for (int it = 0; it < (NZH + TPC - 1) / TPC; ++it) {
      const int k = tc + it * TPC;
      if (k < NZH) {
        const cplx<T> W = F[k * FP + q], Wm = F[(k == 0 ? 0 : L - k) * FP + q];
        cplx2<T> xy;
        xy.a = mk((T)0.5 * (W.x + Wm.x), (T)0.5 * (W.y - Wm.y));
        xy.b = mk((T)0.5 * (W.y + Wm.y), (T)0.5 * (Wm.x - W.x));
        st_off(dst, (uint32_t)tc * rowb + (uint32_t)(it * TPC) * rowb + qofs, xy);
      }
    }
    if (p.partial != nullptr) store_partial<NT>(red, p.partial + bx, tid);
  }
}
template <typename T, int L, int NC, int IN, int OUT, int V = 0>
__global__ __launch_bounds__((NC / 2) * (L / PlanV<L, V>::E)) void k_zpass(const cplx<T>* in, cplx<T>* out,
                                                                       const double* __restrict__ real_in,
                                                                       double* __restrict__ real_out,
                                                                       int op, PParams<T> p) {
  zpass_tile<T, L, NC, IN, OUT, V>(in, out, real_in, real_out, op, p, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Grouped step kernels: boxes of one batch that share grid and tile schedules advance with one
// launch per pass instead of one per box (small boxes are bound by launch latency, not by
// HBM).  Box b of the group is blockIdx.y (y, z passes) or blockIdx.z (x pass); its stage
// array a (in place), carry r, y twiddles and parameters (with the z pass's partial count)
// come from g[b].  Same tile bodies as the single-box kernels, so results are bitwise equal.
template <typename T>
struct GrpArgs {
  cplx<T>* a;
  cplx<T>* r;
  const cplx<T>* twy;
  PParams<T> p;
};
template <typename T, int L, int SIGN, int NC, int V>
__global__ __launch_bounds__(NC*(L / PlanV<L, V>::E)) void k_ypass_grp(const GrpArgs<T>* __restrict__ g,
                                                                   int64_t npencils) {
  const GrpArgs<T>& b = g[blockIdx.y];
  ypass_tile<T, L, SIGN, NC, V>(b.a, b.a, npencils, b.twy, nullptr, KM_IKY, blockIdx.x);
}
template <typename T, int L, int NC, int V>
__global__ __launch_bounds__((NC / 2) * (L / PlanV<L, V>::E)) void k_zpass_grp(const GrpArgs<T>* __restrict__ g,
                                                                           int op) {
  const GrpArgs<T>& b = g[blockIdx.y];
  zpass_tile<T, L, NC, Z_IN_SPEC, Z_OUT_SPEC, V>(b.a, b.a, nullptr, nullptr, op, b.p, blockIdx.x);
}
template <typename T, int L, int NC, int V, int ORD>
__global__ __launch_bounds__(NC*(L / PlanV<L, V>::E)) void k_xpass_grp(const GrpArgs<T>* __restrict__ g) {
  const GrpArgs<T>& b = g[blockIdx.z];
  xpass_tile<T, L, NC, X_STEP, V, ORD>(b.a, b.a, b.r, nullptr, b.p, blockIdx.x, blockIdx.y);
}

}  // namespace smx
