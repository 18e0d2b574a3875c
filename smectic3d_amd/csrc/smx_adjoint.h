// smx_adjoint.h -- discrete adjoint of the periodic CNAB2 step (pure SH).
//
// Forward map (pf3d.cpp:471-481, bootstrap :363-374), real space:
//   psi_{n+1} = A psi_n + B (c_n N(psi_n) - N(psi_{n-1})),  c_0 = 2, c_n = 3,
//   A = F^-1 (C1/C2) F,  B = F^-1 (dt/2 / C2) F   (F normalised; both symmetric).
// Adjoint of J = 1/2 |psi_T - g|^2, alpha_n = F a_n:
//   alpha_n = (C1/C2) alpha_{n+1} + F[ N'(psi_n) F^-1[ dt/2/C2 (c_n alpha_{n+1} - alpha_{n+2}) ] ].
// One adjoint step is the forward chain's shape: y inverse, z (C2R, x N'(psi_n), R2C),
// y forward, x (forward, update, next w, inverse) -- with the carry alpha in place of R.
#pragma once
#include "smx_advect.h"

namespace smx {

// x pass of the adjoint: stage holds the y,z-forward transform of v_n = N'(psi_n) w_n.
//   alpha_n = (C1/C2) alpha_{n+1} + s V,  w_{n-1}(k) = dt/2/C2 (cnext alpha_n - alpha_{n+1}),
// and the parameter sums  sum_k w_kz Re(conj alpha_{n+1} dt/2 dL/dtheta (psi_hat_n + psi_hat_{n+1}) / C2)
// for theta = eps (dL = 1) and q0 (dL = 4 alpha q0 (k^2 - q0^2)), added to acc[2b], acc[2b+1].
// init != 0: alpha_{n+1} = 0 and no parameter sums (sets alpha_T from a_T).
template <typename T, int L, int NC>
__global__ __launch_bounds__(NC*(L / Plan<L>::E)) void k_xpass_adj(cplx<T>* stage, cplx<T>* __restrict__ carry,
                                                                   const cplx<T>* __restrict__ psin,
                                                                   const cplx<T>* __restrict__ psin1, PParams<T> p,
                                                                   T cnext, int init, double q0d, double alphad,
                                                                   double* __restrict__ acc) {
  constexpr int E = Plan<L>::E, TPC = L / E, NT = NC * TPC, NCP = NC + 1;
  __shared__ cplx<T> S[L * NCP];
  __shared__ double red[64];
  const int tid = threadIdx.x, col = tid % NC, tc = tid / NC;
  const int kz = blockIdx.y, y0 = blockIdx.x * NC;
  const int64_t plane = (int64_t)kz * p.nxy;
  for (int i = tid; i < NC * L; i += NT) {
    const int pos = i / NC, c = i - pos * NC;
    S[pos * NCP + c] = stage[plane + (int64_t)pos * p.ny + y0 + c];
  }
  __syncthreads();
  cplx<T> v[E];
  Tile<T, L, E, NCP>::template load<fwd_first<L>()>(v, S, col, tc);
  FwdFFT<T, L, NCP>::run(v, S, col, tc, p.twx);
  const T kyv = p.ky[y0 + col], kzv = p.kz[kz];
  const double wk = (kz == 0 || 2 * kz == p.nz) ? 1.0 : 2.0;
  double ge = 0.0, gq = 0.0;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int kxi = Tile<T, L, E, NCP>::template pos_in<fwd_last<L>()>(tc, e);
    const int64_t idx = plane + (int64_t)kxi * p.ny + y0 + col;
    const T kxv = p.kx[kxi];
    const T k2 = kxv * kxv + kyv * kyv + kzv * kzv;
    const T d = k2 - p.q02;
    const T aLin = p.eps - p.alpha * d * d;
    const T C1 = (T)1 + p.dtd2 * aLin, C2 = (T)1 - p.dtd2 * aLin;
    cplx<T> a1 = mk<T>(0, 0);
    if (!init) {
      a1 = carry[idx];
      const cplx<T> s0 = psin[idx], s1 = psin1[idx];
      const double re = (double)a1.x * ((double)s0.x + (double)s1.x) + (double)a1.y * ((double)s0.y + (double)s1.y);
      const double base = wk * re * (double)p.dtd2 / (double)C2;
      ge += base;
      gq += base * 4.0 * alphad * q0d * ((double)k2 - q0d * q0d);
    }
    const T g = C1 / C2;
    const cplx<T> a0 = mk(g * a1.x + p.scale * v[e].x, g * a1.y + p.scale * v[e].y);
    carry[idx] = a0;
    const T m = p.dtd2 / C2;
    v[e] = mk(m * (cnext * a0.x - a1.x), m * (cnext * a0.y - a1.y));
  }
  if (acc != nullptr && !init) {
    const double be = block_sum(ge, red), bq = block_sum(gq, red);
    if (tid == 0) {
      const int b = blockIdx.y * gridDim.x + blockIdx.x;
      acc[2 * b] += be;
      acc[2 * b + 1] += bq;
    }
  }
  InvFFT<T, L, NCP>::run(v, S, col, tc, p.twx);
  __syncthreads();
  Tile<T, L, E, NCP>::template store<inv_last<L>()>(v, S, col, tc);
  __syncthreads();
  for (int i = tid; i < NC * L; i += NT) {
    const int pos = i / NC, c = i - pos * NC;
    stage[plane + (int64_t)pos * p.ny + y0 + c] = S[pos * NCP + c];
  }
}

// z pass of the adjoint: C2R, multiply by N'(psi_n) (recorded, layout [z][x][y]), R2C; in place.
template <typename T, int L, int NC>
__global__ __launch_bounds__((NC / 2) * (L / Plan<L>::E)) void k_zpass_adj(cplx<T>* stage,
                                                                           const T* __restrict__ nprime,
                                                                           PParams<T> p) {
  using Z = ZTile<T, L, NC>;
  constexpr int E = Z::E, FP = Z::FP, NPR = Z::NPR;
  __shared__ cplx<T> S[Z::NZH * Z::SP];
  __shared__ cplx<T> F[L * FP];
  const int tid = threadIdx.x, q = tid % NPR, tc = tid / NPR;
  const int64_t f0 = (int64_t)blockIdx.x * NC;
  const int nvalid = (int)((p.nxy - f0) < NC ? (p.nxy - f0) : NC);
  cplx<T> v[E];
  Z::c2r(stage, f0, nvalid, p.nxy, S, F, q, tc, tid, p.twz, nullptr, v);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int n = Tile<T, L, E, FP>::template pos_in<fwd_first<L>()>(tc, e);
    const T* row = nprime + (int64_t)n * p.nxy + f0;
    const T na = (q < nvalid) ? row[q] : (T)0, nb = (q + NPR < nvalid) ? row[q + NPR] : (T)0;
    v[e] = mk(na * v[e].x, nb * v[e].y);
  }
  Z::r2c_store(v, stage, f0, nvalid, p.nxy, S, F, q, tc, tid, p.twz);
}

template <typename T>
struct AdjointOps {
  static hipError_t xpass_adj(cplx<T>* stage, cplx<T>* carry, const cplx<T>* psin, const cplx<T>* psin1,
                              const PParams<T>& p, T cnext, int init, double q0d, double alphad, double* acc,
                              hipStream_t st);
  static hipError_t zpass_adj(cplx<T>* stage, const T* nprime, const PParams<T>& p, hipStream_t st);
};

}  // namespace smx
