// smx_advect.h -- kernels of the advected periodic step (pf3d.cpp:399-465):
// spectral gradient and psiSH, force S = f grad(psi), Stokes projection
// U = s CM1 (S - CM2 (k.S)), and N = N(psi) - U.grad(psi).
//
// Per step (all arrays spectral-stage [kz][x|kx][y|ky], in place per tile):
//   y inverse:  A -> By (i ky), A, Ax, As                 (psi, d_x psi, d_y psi, psiSH at (x,y,kz))
//   k_zpass_force:  A, Ax, By, As -> S0, S1, S2           (z C2R, f grad psi, z R2C)
//   y forward x3;  k_xpass_stokes: S -> U;  y inverse x3
//   k_zpass_advnl:  U0..2, A, Ax, By -> N                 (N(psi) - U.grad psi, z R2C)
//   y forward;  k_xpass_adv: N, R -> R, A, Ax, As          (forward x, CNAB2, three inverse x)
// 15 launches.  d_z psi is formed inside the z passes from A with i kz (Nyquist +pi/dz,
// pf3d.cpp:184).  Every kernel loads its inputs straight from HBM into registers and stores
// its outputs straight back; LDS carries only the exchanges between radix passes.
#pragma once
#include "smx_periodic.h"

namespace smx {

enum ForceForm { F_SH = 0, F_MU = 1 };

template <typename T>
struct AParams {
  PParams<T> p;
  T cm;          // 1/(rho eta), pf3d.cpp:210
  int force;     // ForceForm
};

// opSH = alpha (|k|^2 - q0^2)^2 (pf3dMorph.cpp:237)
template <typename T>
__device__ __forceinline__ T op_sh(const PParams<T>& p, T k2) {
  const T d = k2 - p.q02;
  return p.alpha * d * d;
}

// ---------------------------------------------------------------------------
// x pass with three inverse outputs.  MODE X_STEP: in = N stage, CNAB2 update;
// MODE X_BOOT: psi_hat from psi0, R = C1 psi_hat - dt/2 N_hat(psi0).  Inputs come straight
// from HBM into registers, outputs go straight back; one LDS tile serves every exchange.
// On the self-conjugate planes kz = 0, nz/2 the block pairs columns ky / -ky (as k_xpass)
// and projects R onto Hermitian symmetry before storing it.
template <typename T, int L, int NC, int MODE, int V>
__global__ __launch_bounds__(NC*(L / PlanV<L, V>::E)) void k_xpass_adv(const cplx<T>* in, cplx<T>* __restrict__ rbuf,
                                                                       const cplx<T>* __restrict__ psi0, cplx<T>* outA,
                                                                       cplx<T>* outX, cplx<T>* outS, PParams<T> p) {
  constexpr int E = PlanV<L, V>::E, TPC = L / E, NT = NC * TPC, NCP = NC + 1, H = NC / 2;
  constexpr uint32_t CB = sizeof(cplx<T>);
  __shared__ cplx<T> S[L * NCP];
  __shared__ double red[64];
  const int tid = threadIdx.x, col = tid % NC, tc = tid / NC;
  __shared__ cplx<T> TW[L];
  const cplx<T>* twl = tw_lds<L, NT>(TW, p.twx, tid);
  __shared__ T KX[L];  // k_x for the loops after the forward transform's barriers
  for (int i = tid; i < L; i += NT) KX[i] = p.kx[i];
  const int kz = blockIdx.y, y0 = blockIdx.x * NC;
  const int64_t plane = (int64_t)kz * p.nxy;
  const uint32_t rowb = (uint32_t)p.ny * CB;
  const bool herm = (kz == 0 || 2 * kz == p.nz);
  int ky = y0 + col, pcol = col;
  if (herm) {
    const int b = blockIdx.x * H;
    if (col < H) {
      ky = b + col;
      pcol = (b + col == 0) ? col : col + H;
    } else {
      ky = (b == 0 && col == H) ? p.ny / 2 : p.ny - (b + col - H);
      pcol = (b == 0 && col == H) ? col : col - H;
    }
  }
  const uint32_t cofs = (uint32_t)ky * CB;
  cplx<T>* rb = rbuf + plane;
  cplx<T> v[E], ps[E];
#pragma unroll
  for (int e = 0; e < E; ++e)
    v[e] = ld_off(in + plane, (uint32_t)Tile<T, L, E, NCP>::template pos_in<fwd_first<L, V>()>(tc, e) * rowb + cofs);
  double nsum = 0.0;
  const bool owner0 = (p.partial != nullptr) && kz == 0 && y0 == 0;
  if (owner0) {
    double a = 0.0;
    for (int i = tid; i < p.npartial; i += NT) a += p.partial[i];
    nsum = block_sum(a, red);
  }
  FwdFFT<T, L, NCP, V>::run(v, S, col, tc, twl);
  const T kyv = p.ky[ky], kzv = p.kz[kz];
  // The new R goes to the LDS tile first and to HBM after the loop, so that the loop's R
  // reads are not issued behind R writes (a wave's loads and stores retire in order).
  __syncthreads();  // S receives R' below; the transform's last reads must be done
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int kxi = Tile<T, L, E, NCP>::template pos_in<fwd_last<L, V>()>(tc, e);
    const uint32_t o = (uint32_t)kxi * rowb + cofs;
    const T kxv = KX[kxi];
    const T k2 = kxv * kxv + kyv * kyv + kzv * kzv;
    const T aLin = p.eps - op_sh(p, k2);
    const T C1 = (T)1 + p.dtd2 * aLin, C2 = (T)1 - p.dtd2 * aLin;
    const cplx<T> nn = p.scale * v[e];
    cplx<T> s, rn;
    if constexpr (MODE == X_STEP) {  // pf3d.cpp:476-478 carried as R = C1 psi - dt/2 N_past
      const cplx<T> r = ld_off<cplx<T>>(rb, o);
      const T ic2 = recip(C2);
      s = mk(ic2 * (r.x + (T)3 * p.dtd2 * nn.x), ic2 * (r.y + (T)3 * p.dtd2 * nn.y));
      rn = mk(C1 * s.x - p.dtd2 * nn.x, C1 * s.y - p.dtd2 * nn.y);
      if (owner0 && tid == 0 && kxi == 0) {
        const double dd = p.m_dtd2, n0 = p.m_scale * nsum;
        const double ps0 = (p.mean_state[0] + 3.0 * dd * n0) / p.m_c2;
        p.mean_state[0] = p.m_c1 * ps0 - dd * n0;
        p.mean_state[1] = ps0;
        s = mk((T)ps0, (T)0);
        rn = mk((T)p.mean_state[0], (T)0);
      }
    } else {
      s = ld_off(psi0 + plane, o);
      rn = mk(C1 * s.x - p.dtd2 * nn.x, C1 * s.y - p.dtd2 * nn.y);
      if (owner0 && tid == 0 && kxi == 0) {
        const double ps0 = p.mean_state[1];
        p.mean_state[0] = p.m_c1 * ps0 - p.m_dtd2 * p.m_scale * nsum;
        s = mk((T)ps0, (T)0);
        rn = mk((T)p.mean_state[0], (T)0);
      }
    }
    S[kxi * NCP + col] = rn;
    ps[e] = s;
  }
  if (herm) {  // R(k) <- (R(k) + conj R(-k))/2 with -k = (nx - kx, partner column)
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int kxi = Tile<T, L, E, NCP>::template pos_in<fwd_last<L, V>()>(tc, e);
      const cplx<T> a = S[kxi * NCP + col], b = S[((L - kxi) % L) * NCP + pcol];
      st_off(rb, (uint32_t)kxi * rowb + cofs, mk((T)0.5 * (a.x + b.x), (T)0.5 * (a.y - b.y)));
    }
  } else {  // each thread reads back its own entries
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int kxi = Tile<T, L, E, NCP>::template pos_in<fwd_last<L, V>()>(tc, e);
      st_off(rb, (uint32_t)kxi * rowb + cofs, S[kxi * NCP + col]);
    }
  }
  cplx<T>* outs[3] = {outA + plane, outX + plane, outS + plane};
#pragma unroll 1
  for (int o = 0; o < 3; ++o) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int kxi = Tile<T, L, E, NCP>::template pos_in<fwd_last<L, V>()>(tc, e);
      const T kxv = KX[kxi];
      if (o == 0) v[e] = ps[e];
      else if (o == 1) v[e] = mk(-kxv * ps[e].y, kxv * ps[e].x);  // i kx psi_hat (pf3d.cpp:404)
      else v[e] = op_sh(p, kxv * kxv + kyv * kyv + kzv * kzv) * ps[e];  // opSH psi_hat (pf3d.cpp:407)
    }
    InvFFT<T, L, NCP, V>::run(v, S, col, tc, twl);
#pragma unroll
    for (int e = 0; e < E; ++e)
      st_off(outs[o], (uint32_t)Tile<T, L, E, NCP>::template pos_in<inv_last<L, V>()>(tc, e) * rowb + cofs, v[e]);
  }
}

// ---------------------------------------------------------------------------
// Stokes projection in full k-space (pf3d.cpp:433-447):
//   U = s CM1 (S - CM2 (k.S)),  CM1 = 1/(rho eta |k|^2),  CM2 = k/|k|^2,  U(k = 0) = 0.
// Three components: forward x, project (all three in one thread's registers), inverse x;
// in place per tile; one LDS tile serves the six transforms.
template <typename T, int L, int NC, int V>
__global__ __launch_bounds__(NC*(L / PlanV<L, V>::E)) void k_xpass_stokes(cplx<T>* s0, cplx<T>* s1, cplx<T>* s2,
                                                                          AParams<T> ap) {
  constexpr int E = PlanV<L, V>::E, NCP = NC + 1;
  constexpr uint32_t CB = sizeof(cplx<T>);
  __shared__ cplx<T> Sx[L * NCP];
  const PParams<T>& p = ap.p;
  const int tid = threadIdx.x, col = tid % NC, tc = tid / NC;
  __shared__ cplx<T> TW[L];
  const cplx<T>* twl = tw_lds<L, NC * (L / E)>(TW, p.twx, tid);
  const int kz = blockIdx.y, y0 = blockIdx.x * NC;
  const int64_t plane = (int64_t)kz * p.nxy;
  const uint32_t rowb = (uint32_t)p.ny * CB, cofs = (uint32_t)(y0 + col) * CB;
  cplx<T>* arr[3] = {s0 + plane, s1 + plane, s2 + plane};
  cplx<T> v[3][E];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
#pragma unroll
    for (int e = 0; e < E; ++e)
      v[a][e] = ld_off<cplx<T>>(arr[a], (uint32_t)Tile<T, L, E, NCP>::template pos_in<fwd_first<L, V>()>(tc, e) * rowb + cofs);
    FwdFFT<T, L, NCP, V>::run(v[a], Sx, col, tc, twl);
  }
  const T kyv = p.ky[y0 + col], kzv = p.kz[kz];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int kxi = Tile<T, L, E, NCP>::template pos_in<fwd_last<L, V>()>(tc, e);
    const T kxv = p.kx[kxi];
    const T k2 = kxv * kxv + kyv * kyv + kzv * kzv;
    const cplx<T> a0 = v[0][e], a1 = v[1][e], a2 = v[2][e];
    cplx<T> u0 = mk<T>(0, 0), u1 = mk<T>(0, 0), u2 = mk<T>(0, 0);
    if (!(kxi == 0 && y0 + col == 0 && kz == 0)) {  // CM1[0] = CM2[0] = 0 (pf3d.cpp:219-222)
      const cplx<T> dot = mk(kxv * a0.x + kyv * a1.x + kzv * a2.x, kxv * a0.y + kyv * a1.y + kzv * a2.y);
      const T ik2 = recip(k2);
      const T cm1 = p.scale * ap.cm * ik2;
      const T c2x = kxv * ik2, c2y = kyv * ik2, c2z = kzv * ik2;
      u0 = mk(cm1 * (a0.x - c2x * dot.x), cm1 * (a0.y - c2x * dot.y));
      u1 = mk(cm1 * (a1.x - c2y * dot.x), cm1 * (a1.y - c2y * dot.y));
      u2 = mk(cm1 * (a2.x - c2z * dot.x), cm1 * (a2.y - c2z * dot.y));
    }
    v[0][e] = u0;
    v[1][e] = u1;
    v[2][e] = u2;
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    InvFFT<T, L, NCP, V>::run(v[a], Sx, col, tc, twl);
#pragma unroll
    for (int e = 0; e < E; ++e)
      st_off(arr[a], (uint32_t)Tile<T, L, E, NCP>::template pos_in<inv_last<L, V>()>(tc, e) * rowb + cofs, v[a][e]);
  }
}

// ---------------------------------------------------------------------------
// z-pass helpers over one tile of NC columns: adjacent columns (2q, 2q+1) share one complex
// FFT and move with one 2-complex access per row (as k_zpass).
template <typename T, int L, int NC, int V>
struct ZPair {
  static constexpr int E = PlanV<L, V>::E, TPC = L / E, NPR = NC / 2, NT = NPR * TPC;
  static constexpr int NZH = L / 2 + 1, FP = NPR + 1;
  static constexpr uint32_t CB = sizeof(cplx<T>);

  // The rows one thread reads of a z-stage field (its column pair, the Hermitian extension's
  // rows min(n, L - n)).
  __device__ static void load(const cplx<T>* in, int64_t f0, uint32_t rowb, int q, int tc, cplx2<T>* raw) {
    const cplx2<T>* col = reinterpret_cast<const cplx2<T>*>(in + f0);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int n = Tile<T, L, E, FP>::template pos_in<inv_first<L, V>()>(tc, e);
      const int m = (n <= L / 2) ? n : L - n;
      raw[e] = ld_off(col, (uint32_t)m * rowb + (uint32_t)q * 2 * CB);
    }
  }
  // C2R of one z-stage field straight from HBM: v <- (psi_a, psi_b) at pos_in<fwd_first>.
  // kzmul: multiply by i kz first (spectral d/dz, Nyquist +pi/dz, pf3d.cpp:184).
  __device__ static void c2r(const cplx<T>* in, int64_t f0, uint32_t rowb, int q, int tc, const T* kzmul,
                             cplx<T>* F, const cplx<T>* tw, cplx<T>* v) {
    cplx2<T> raw[E];
    load(in, f0, rowb, q, tc, raw);
    c2r_loaded(raw, tc, kzmul, F, tw, q, v);
  }
  // C2R from rows already in registers (load()).
  __device__ static void c2r_loaded(const cplx2<T>* raw, int tc, const T* kzmul, cplx<T>* F, const cplx<T>* tw, int q,
                                    cplx<T>* v) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int n = Tile<T, L, E, FP>::template pos_in<inv_first<L, V>()>(tc, e);
      const int m = (n <= L / 2) ? n : L - n;
      const cplx2<T> ab = raw[e];
      cplx<T> A = ab.a, B = ab.b;
      if (kzmul != nullptr) {
        const T k = kzmul[m];
        A = mk(-k * A.y, k * A.x);
        B = mk(-k * B.y, k * B.x);
      }
      if (n == 0 || n == L / 2) {
        A.y = 0;
        B.y = 0;
      }
      v[e] = (n <= L / 2) ? mk(A.x - B.y, A.y + B.x) : mk(A.x + B.y, B.x - A.y);
    }
    InvFFT<T, L, FP, V>::run(v, F, q, tc, tw);
  }

  // R2C of the (a, b) pairs in v and store of the z-stage spectra straight to HBM.
  __device__ static void r2c_store(cplx<T>* v, cplx<T>* out, int64_t f0, uint32_t rowb, int q, int tc, int tid,
                                   cplx<T>* F, const cplx<T>* tw) {
    FwdFFT<T, L, FP, V>::run(v, F, q, tc, tw);
    __syncthreads();
    Tile<T, L, E, FP>::template store<fwd_last<L, V>()>(v, F, q, tc);
    __syncthreads();
    cplx2<T>* dst = reinterpret_cast<cplx2<T>*>(out + f0);
    for (int i = tid; i < NZH * NPR; i += NT) {
      const int k = i / NPR, qq = i - k * NPR;
      const cplx<T> W = F[k * FP + qq], Wm = F[((L - k) % L) * FP + qq];
      cplx2<T> xy;
      xy.a = mk((T)0.5 * (W.x + Wm.x), (T)0.5 * (W.y - Wm.y));
      xy.b = mk((T)0.5 * (W.y + Wm.y), (T)0.5 * (Wm.x - W.x));
      st_off(dst, (uint32_t)k * rowb + (uint32_t)qq * 2 * CB, xy);
    }
  }
};

// Force S = f grad(psi) (pf3d.cpp:417-425), f = psiSH (SH form) or
// mu = psiSH - eps psi - N(psi) (variational derivative of F).  Whole tiles (nx*ny % NC == 0).
template <typename T, int L, int NC, int V>
__global__ __launch_bounds__((NC / 2) * (L / PlanV<L, V>::E)) void k_zpass_force(const cplx<T>* bpsi,
                                                                                 const cplx<T>* bx,
                                                                                 const cplx<T>* by,
                                                                                 const cplx<T>* bs, cplx<T>* o0,
                                                                                 cplx<T>* o1, cplx<T>* o2,
                                                                                 AParams<T> ap) {
  using Z = ZPair<T, L, NC, V>;
  constexpr int E = Z::E;
  __shared__ cplx<T> F[L * Z::FP];
  const PParams<T>& p = ap.p;
  const int tid = threadIdx.x, q = tid % Z::NPR, tc = tid / Z::NPR;
  __shared__ cplx<T> TW[L];
  const cplx<T>* twl = tw_lds<L, Z::NT>(TW, p.twz, tid);
  __shared__ T KZ[L];  // k_z for the d/dz field
  for (int i = tid; i < L; i += Z::NT) KZ[i] = p.kz[i];
  __syncthreads();  // one-pass plans have no transform barrier before KZ is read
  const int64_t f0 = (int64_t)blockIdx.x * NC;
  const uint32_t rowb = (uint32_t)p.nxy * Z::CB;
  cplx<T> v[E], f[E];
  Z::c2r(bs, f0, rowb, q, tc, nullptr, F, twl, f);  // psiSH
  if (ap.force == F_MU) {
    Z::c2r(bpsi, f0, rowb, q, tc, nullptr, F, twl, v);  // psi
#pragma unroll
    for (int e = 0; e < E; ++e)
      f[e] = mk(f[e].x - p.eps * v[e].x - nonlin(p, v[e].x), f[e].y - p.eps * v[e].y - nonlin(p, v[e].y));
  }
  const cplx<T>* src[3] = {bx, by, bpsi};
  cplx<T>* dst[3] = {o0, o1, o2};
  // the next component's rows are read before this component's stores are issued (a
  // wave's loads and stores retire in order: a read behind a store would wait for it)
  cplx2<T> raw[E];
  Z::load(src[0], f0, rowb, q, tc, raw);
#pragma unroll 1
  for (int a = 0; a < 3; ++a) {
    Z::c2r_loaded(raw, tc, a == 2 ? KZ : nullptr, F, twl, q, v);
    if (a < 2) Z::load(src[a + 1], f0, rowb, q, tc, raw);
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = mk(f[e].x * v[e].x, f[e].y * v[e].y);
    Z::r2c_store(v, dst[a], f0, rowb, q, tc, tid, F, twl);
  }
}

// N = N(psi) - U.grad(psi) [+ mid-plane penalty] (pf3d.cpp:455-465, pf3dMorph.cpp:572-590).
template <typename T, int L, int NC, int V>
__global__ __launch_bounds__((NC / 2) * (L / PlanV<L, V>::E)) void k_zpass_advnl(const cplx<T>* u0,
                                                                                 const cplx<T>* u1,
                                                                                 const cplx<T>* u2,
                                                                                 const cplx<T>* bpsi,
                                                                                 const cplx<T>* bx,
                                                                                 const cplx<T>* by, cplx<T>* nout,
                                                                                 AParams<T> ap) {
  using Z = ZPair<T, L, NC, V>;
  constexpr int E = Z::E, FP = Z::FP;
  __shared__ cplx<T> F[L * FP];
  __shared__ double red[64];
  const PParams<T>& p = ap.p;
  const int tid = threadIdx.x, q = tid % Z::NPR, tc = tid / Z::NPR;
  __shared__ cplx<T> TW[L];
  const cplx<T>* twl = tw_lds<L, Z::NT>(TW, p.twz, tid);
  __shared__ T KZ[L];  // k_z for the d/dz field
  for (int i = tid; i < L; i += Z::NT) KZ[i] = p.kz[i];
  __syncthreads();  // one-pass plans have no transform barrier before KZ is read
  const int64_t f0 = (int64_t)blockIdx.x * NC;
  const uint32_t rowb = (uint32_t)p.nxy * Z::CB;
  constexpr int RP = fwd_first<L, V>();
  cplx<T> v[E], nacc[E], u[E];
  Z::c2r(bpsi, f0, rowb, q, tc, nullptr, F, twl, v);
  const bool pen = p.mid > 0 && p.mid < L - 1;
  if (pen) {
    __syncthreads();
    Tile<T, L, E, FP>::template store<RP>(v, F, q, tc);
    __syncthreads();
  }
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int n = Tile<T, L, E, FP>::template pos_in<RP>(tc, e);
    T a = nonlin(p, v[e].x), b = nonlin(p, v[e].y);
    if (pen && n == p.mid) {
      const cplx<T> lo = F[(n - 1) * FP + q], hi = F[(n + 1) * FP + q];
      a += p.pen * (lo.x - (T)2 * v[e].x + hi.x);
      b += p.pen * (lo.y - (T)2 * v[e].y + hi.y);
    }
    nacc[e] = mk(a, b);
  }
  const cplx<T>* us[3] = {u0, u1, u2};
  const cplx<T>* gs[3] = {bx, by, bpsi};
#pragma unroll 1
  for (int a = 0; a < 3; ++a) {
    Z::c2r(us[a], f0, rowb, q, tc, nullptr, F, twl, u);
    Z::c2r(gs[a], f0, rowb, q, tc, a == 2 ? KZ : nullptr, F, twl, v);
#pragma unroll
    for (int e = 0; e < E; ++e) nacc[e] = mk(nacc[e].x - u[e].x * v[e].x, nacc[e].y - u[e].y * v[e].y);
  }
  if (p.partial != nullptr) wave_partial<E, Z::NT>(nacc, red, tid);
  Z::r2c_store(nacc, nout, f0, rowb, q, tc, tid, F, twl);  // its barriers order red[] for:
  if (p.partial != nullptr) store_partial<Z::NT>(red, p.partial + blockIdx.x, tid);
}

}  // namespace smx
