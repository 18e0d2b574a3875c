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
// d_z psi is formed inside the z passes from A with i kz (Nyquist +pi/dz, pf3d.cpp:184).
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
// MODE X_BOOT: psi_hat from psi0, R = C1 psi_hat - dt/2 N_hat(psi0).
template <typename T, int L, int NC, int MODE>
__global__ __launch_bounds__(NC*(L / Plan<L>::E)) void k_xpass_adv(const cplx<T>* in, cplx<T>* __restrict__ rbuf,
                                                                   const cplx<T>* __restrict__ psi0, cplx<T>* outA,
                                                                   cplx<T>* outX, cplx<T>* outS, PParams<T> p) {
  constexpr int E = Plan<L>::E, TPC = L / E, NT = NC * TPC, NCP = NC + 1;
  __shared__ cplx<T> S[L * NCP];
  __shared__ cplx<T> K[L * NCP];
  __shared__ double red[64];
  const int tid = threadIdx.x, col = tid % NC, tc = tid / NC;
  const int kz = blockIdx.y, y0 = blockIdx.x * NC;
  const int64_t plane = (int64_t)kz * p.nxy;
  for (int i = tid; i < NC * L; i += NT) {
    const int pos = i / NC, c = i - pos * NC;
    S[pos * NCP + c] = in[plane + (int64_t)pos * p.ny + y0 + c];
  }
  double nsum = 0.0;
  const bool owner0 = (p.partial != nullptr) && kz == 0 && y0 == 0;
  if (owner0) {
    double a = 0.0;
    for (int i = tid; i < p.npartial; i += NT) a += p.partial[i];
    nsum = block_sum(a, red);
  }
  __syncthreads();
  cplx<T> v[E];
  Tile<T, L, E, NCP>::template load<fwd_first<L>()>(v, S, col, tc);
  FwdFFT<T, L, NCP>::run(v, S, col, tc, p.twx);
  const T kyv = p.ky[y0 + col], kzv = p.kz[kz];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int kxi = Tile<T, L, E, NCP>::template pos_in<fwd_last<L>()>(tc, e);
    const int64_t idx = plane + (int64_t)kxi * p.ny + y0 + col;
    const T kxv = p.kx[kxi];
    const T k2 = kxv * kxv + kyv * kyv + kzv * kzv;
    const T aLin = p.eps - op_sh(p, k2);
    const T C1 = (T)1 + p.dtd2 * aLin, C2 = (T)1 - p.dtd2 * aLin;
    const cplx<T> nn = p.scale * v[e];
    cplx<T> ps, rn;
    if constexpr (MODE == X_STEP) {
      const cplx<T> r = rbuf[idx];
      const T ic2 = (T)1 / C2;
      ps = mk(ic2 * (r.x + (T)3 * p.dtd2 * nn.x), ic2 * (r.y + (T)3 * p.dtd2 * nn.y));
      rn = mk(C1 * ps.x - p.dtd2 * nn.x, C1 * ps.y - p.dtd2 * nn.y);
      if (owner0 && tid == 0 && kxi == 0) {
        const double dd = p.m_dtd2, n0 = p.m_scale * nsum;
        const double ps0 = (p.mean_state[0] + 3.0 * dd * n0) / p.m_c2;
        p.mean_state[0] = p.m_c1 * ps0 - dd * n0;
        p.mean_state[1] = ps0;
        ps = mk((T)ps0, (T)0);
        rn = mk((T)p.mean_state[0], (T)0);
      }
    } else {
      ps = psi0[idx];
      rn = mk(C1 * ps.x - p.dtd2 * nn.x, C1 * ps.y - p.dtd2 * nn.y);
      if (owner0 && tid == 0 && kxi == 0) {
        const double ps0 = p.mean_state[1];
        p.mean_state[0] = p.m_c1 * ps0 - p.m_dtd2 * p.m_scale * nsum;
        ps = mk((T)ps0, (T)0);
        rn = mk((T)p.mean_state[0], (T)0);
      }
    }
    rbuf[idx] = rn;
    v[e] = ps;
  }
  Tile<T, L, E, NCP>::template store<fwd_last<L>()>(v, K, col, tc);  // own positions, read back below
  cplx<T>* outs[3] = {outA, outX, outS};
#pragma unroll 1
  for (int o = 0; o < 3; ++o) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int kxi = Tile<T, L, E, NCP>::template pos_in<fwd_last<L>()>(tc, e);
      const cplx<T> a = K[kxi * NCP + col];
      if (o == 0) {
        v[e] = a;
      } else if (o == 1) {  // i kx psi_hat (pf3d.cpp:404)
        const T k = p.kx[kxi];
        v[e] = mk(-k * a.y, k * a.x);
      } else {  // opSH psi_hat (pf3d.cpp:407)
        const T kxv = p.kx[kxi];
        v[e] = op_sh(p, kxv * kxv + kyv * kyv + kzv * kzv) * a;
      }
    }
    InvFFT<T, L, NCP>::run(v, S, col, tc, p.twx);
    __syncthreads();
    Tile<T, L, E, NCP>::template store<inv_last<L>()>(v, S, col, tc);
    __syncthreads();
    cplx<T>* out = outs[o];
    for (int i = tid; i < NC * L; i += NT) {
      const int pos = i / NC, c = i - pos * NC;
      out[plane + (int64_t)pos * p.ny + y0 + c] = S[pos * NCP + c];
    }
  }
}

// ---------------------------------------------------------------------------
// Stokes projection in full k-space (pf3d.cpp:433-447):
//   U = s CM1 (S - CM2 (k.S)),  CM1 = 1/(rho eta |k|^2),  CM2 = k/|k|^2,  U(k = 0) = 0.
// Three components: forward x, project, inverse x; in place per tile.
template <typename T, int L, int NC>
__global__ __launch_bounds__(NC*(L / Plan<L>::E)) void k_xpass_stokes(cplx<T>* s0, cplx<T>* s1, cplx<T>* s2,
                                                                      AParams<T> ap) {
  constexpr int E = Plan<L>::E, TPC = L / E, NT = NC * TPC, NCP = NC + 1;
  __shared__ cplx<T> Tl[3][L * NCP];
  const PParams<T>& p = ap.p;
  const int tid = threadIdx.x, col = tid % NC, tc = tid / NC;
  const int kz = blockIdx.y, y0 = blockIdx.x * NC;
  const int64_t plane = (int64_t)kz * p.nxy;
  cplx<T>* arr[3] = {s0, s1, s2};
#pragma unroll 1
  for (int a = 0; a < 3; ++a)
    for (int i = tid; i < NC * L; i += NT) {
      const int pos = i / NC, c = i - pos * NC;
      Tl[a][pos * NCP + c] = arr[a][plane + (int64_t)pos * p.ny + y0 + c];
    }
  __syncthreads();
  cplx<T> v[E];
#pragma unroll 1
  for (int a = 0; a < 3; ++a) {
    Tile<T, L, E, NCP>::template load<fwd_first<L>()>(v, Tl[a], col, tc);
    FwdFFT<T, L, NCP>::run(v, Tl[a], col, tc, p.twx);
    __syncthreads();
    Tile<T, L, E, NCP>::template store<fwd_last<L>()>(v, Tl[a], col, tc);
  }
  __syncthreads();
  const T kyv = p.ky[y0 + col], kzv = p.kz[kz];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int kxi = Tile<T, L, E, NCP>::template pos_in<fwd_last<L>()>(tc, e);
    const T kxv = p.kx[kxi];
    const T k2 = kxv * kxv + kyv * kyv + kzv * kzv;
    const int o = kxi * NCP + col;
    const cplx<T> a0 = Tl[0][o], a1 = Tl[1][o], a2 = Tl[2][o];
    cplx<T> u0 = mk<T>(0, 0), u1 = mk<T>(0, 0), u2 = mk<T>(0, 0);
    if (!(kxi == 0 && y0 + col == 0 && kz == 0)) {  // CM1[0] = CM2[0] = 0 (pf3d.cpp:219-222)
      const cplx<T> dot = mk(kxv * a0.x + kyv * a1.x + kzv * a2.x, kxv * a0.y + kyv * a1.y + kzv * a2.y);
      const T cm1 = p.scale * ap.cm / k2;
      const T c2x = kxv / k2, c2y = kyv / k2, c2z = kzv / k2;
      u0 = mk(cm1 * (a0.x - c2x * dot.x), cm1 * (a0.y - c2x * dot.y));
      u1 = mk(cm1 * (a1.x - c2y * dot.x), cm1 * (a1.y - c2y * dot.y));
      u2 = mk(cm1 * (a2.x - c2z * dot.x), cm1 * (a2.y - c2z * dot.y));
    }
    Tl[0][o] = u0;
    Tl[1][o] = u1;
    Tl[2][o] = u2;
  }
#pragma unroll 1
  for (int a = 0; a < 3; ++a) {
    Tile<T, L, E, NCP>::template load<inv_first<L>()>(v, Tl[a], col, tc);
    InvFFT<T, L, NCP>::run(v, Tl[a], col, tc, p.twx);
    __syncthreads();
    Tile<T, L, E, NCP>::template store<inv_last<L>()>(v, Tl[a], col, tc);
    __syncthreads();
    for (int i = tid; i < NC * L; i += NT) {
      const int pos = i / NC, c = i - pos * NC;
      arr[a][plane + (int64_t)pos * p.ny + y0 + c] = Tl[a][pos * NCP + c];
    }
  }
}

// ---------------------------------------------------------------------------
// z-pass helpers over one tile of NC columns (pairs q, q + NC/2).
template <typename T, int L, int NC>
struct ZTile {
  static constexpr int E = Plan<L>::E, TPC = L / E, NPR = NC / 2, NT = NPR * TPC;
  static constexpr int NZH = L / 2 + 1, SP = NC + 1, FP = NPR + 1;

  // C2R of one z-stage field: v <- (psi_a, psi_b) at pos_in<fwd_first>.  kzmul: multiply by i kz first.
  __device__ static void c2r(const cplx<T>* in, int64_t f0, int nvalid, int64_t nxy, cplx<T>* S, cplx<T>* F, int q,
                             int tc, int tid, const cplx<T>* tw, const T* kzmul, cplx<T>* v) {
    __syncthreads();
    for (int i = tid; i < NZH * NC; i += NT) {
      const int k = i / NC, c = i - k * NC;
      cplx<T> a = (c < nvalid) ? in[(int64_t)k * nxy + f0 + c] : mk<T>(0, 0);
      if (kzmul != nullptr) a = mk(-kzmul[k] * a.y, kzmul[k] * a.x);
      S[k * SP + c] = a;
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int n = Tile<T, L, E, FP>::template pos_in<inv_first<L>()>(tc, e);
      if (n <= L / 2) {
        cplx<T> A = S[n * SP + q], B = S[n * SP + q + NPR];
        if (n == 0 || n == L / 2) {
          A.y = 0;
          B.y = 0;
        }
        v[e] = mk(A.x - B.y, A.y + B.x);
      } else {
        const int m = L - n;
        const cplx<T> A = S[m * SP + q], B = S[m * SP + q + NPR];
        v[e] = mk(A.x + B.y, B.x - A.y);
      }
    }
    InvFFT<T, L, FP>::run(v, F, q, tc, tw);
  }

  // R2C of (a, b) pairs in v and store the z-stage spectra of the NC columns.
  __device__ static void r2c_store(cplx<T>* v, cplx<T>* out, int64_t f0, int nvalid, int64_t nxy, cplx<T>* S,
                                   cplx<T>* F, int q, int tc, int tid, const cplx<T>* tw) {
    FwdFFT<T, L, FP>::run(v, F, q, tc, tw);
    __syncthreads();
    Tile<T, L, E, FP>::template store<fwd_last<L>()>(v, F, q, tc);
    __syncthreads();
    for (int i = tid; i < NZH * NPR; i += NT) {
      const int k = i / NPR, qq = i - k * NPR;
      const cplx<T> W = F[k * FP + qq], Wm = F[((L - k) % L) * FP + qq];
      S[k * SP + qq] = mk((T)0.5 * (W.x + Wm.x), (T)0.5 * (W.y - Wm.y));
      S[k * SP + qq + NPR] = mk((T)0.5 * (W.y + Wm.y), (T)0.5 * (Wm.x - W.x));
    }
    __syncthreads();
    for (int i = tid; i < NZH * NC; i += NT) {
      const int k = i / NC, c = i - k * NC;
      if (c < nvalid) out[(int64_t)k * nxy + f0 + c] = S[k * SP + c];
    }
  }
};

// Force S = f grad(psi) (pf3d.cpp:417-425), f = psiSH (SH form) or
// mu = psiSH - eps psi - N(psi) (variational derivative of F).
template <typename T, int L, int NC>
__global__ __launch_bounds__((NC / 2) * (L / Plan<L>::E)) void k_zpass_force(const cplx<T>* bpsi, const cplx<T>* bx,
                                                                             const cplx<T>* by, const cplx<T>* bs,
                                                                             cplx<T>* o0, cplx<T>* o1, cplx<T>* o2,
                                                                             AParams<T> ap) {
  using Z = ZTile<T, L, NC>;
  constexpr int E = Z::E;
  __shared__ cplx<T> S[Z::NZH * Z::SP];
  __shared__ cplx<T> F[L * Z::FP];
  const PParams<T>& p = ap.p;
  const int tid = threadIdx.x, q = tid % Z::NPR, tc = tid / Z::NPR;
  const int64_t f0 = (int64_t)blockIdx.x * NC;
  const int nvalid = (int)((p.nxy - f0) < NC ? (p.nxy - f0) : NC);
  cplx<T> v[E], f[E];
  Z::c2r(bs, f0, nvalid, p.nxy, S, F, q, tc, tid, p.twz, nullptr, f);  // psiSH
  if (ap.force == F_MU) {
    Z::c2r(bpsi, f0, nvalid, p.nxy, S, F, q, tc, tid, p.twz, nullptr, v);  // psi
#pragma unroll
    for (int e = 0; e < E; ++e)
      f[e] = mk(f[e].x - p.eps * v[e].x - nonlin(p, v[e].x), f[e].y - p.eps * v[e].y - nonlin(p, v[e].y));
  }
  const cplx<T>* src[3] = {bx, by, bpsi};
  cplx<T>* dst[3] = {o0, o1, o2};
#pragma unroll 1
  for (int a = 0; a < 3; ++a) {
    Z::c2r(src[a], f0, nvalid, p.nxy, S, F, q, tc, tid, p.twz, a == 2 ? p.kz : nullptr, v);
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = mk(f[e].x * v[e].x, f[e].y * v[e].y);
    Z::r2c_store(v, dst[a], f0, nvalid, p.nxy, S, F, q, tc, tid, p.twz);
  }
}

// N = N(psi) - U.grad(psi) [+ mid-plane penalty] (pf3d.cpp:455-465, pf3dMorph.cpp:572-590).
template <typename T, int L, int NC>
__global__ __launch_bounds__((NC / 2) * (L / Plan<L>::E)) void k_zpass_advnl(const cplx<T>* u0, const cplx<T>* u1,
                                                                             const cplx<T>* u2, const cplx<T>* bpsi,
                                                                             const cplx<T>* bx, const cplx<T>* by,
                                                                             cplx<T>* nout, AParams<T> ap) {
  using Z = ZTile<T, L, NC>;
  constexpr int E = Z::E, FP = Z::FP;
  __shared__ cplx<T> S[Z::NZH * Z::SP];
  __shared__ cplx<T> F[L * FP];
  __shared__ double red[64];
  const PParams<T>& p = ap.p;
  const int tid = threadIdx.x, q = tid % Z::NPR, tc = tid / Z::NPR;
  const int64_t f0 = (int64_t)blockIdx.x * NC;
  const int nvalid = (int)((p.nxy - f0) < NC ? (p.nxy - f0) : NC);
  constexpr int RP = fwd_first<L>();
  cplx<T> v[E], nacc[E], u[E];
  Z::c2r(bpsi, f0, nvalid, p.nxy, S, F, q, tc, tid, p.twz, nullptr, v);
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
    Z::c2r(us[a], f0, nvalid, p.nxy, S, F, q, tc, tid, p.twz, nullptr, u);
    Z::c2r(gs[a], f0, nvalid, p.nxy, S, F, q, tc, tid, p.twz, a == 2 ? p.kz : nullptr, v);
#pragma unroll
    for (int e = 0; e < E; ++e) nacc[e] = mk(nacc[e].x - u[e].x * v[e].x, nacc[e].y - u[e].y * v[e].y);
  }
  if (p.partial != nullptr) {
    double acc = 0.0;
#pragma unroll
    for (int e = 0; e < E; ++e) acc += (double)nacc[e].x + (double)nacc[e].y;
    const double s = block_sum(acc, red);
    if (tid == 0) p.partial[blockIdx.x] = s;
  }
  Z::r2c_store(nacc, nout, f0, nvalid, p.nxy, S, F, q, tc, tid, p.twz);
}

}  // namespace smx
