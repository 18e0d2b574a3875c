// smx_adjoint.h -- discrete adjoint of the periodic CNAB2 step (pure SH).
//
// Forward map (pf3d.cpp:471-481, bootstrap :363-374), real space:
//   psi_{n+1} = A psi_n + B (c_n N(psi_n) - N(psi_{n-1})),  c_0 = 2, c_n = 3,
//   A = F^-1 (C1/C2) F,  B = F^-1 (dt/2 / C2) F   (F normalised; both symmetric).
// Adjoint of J = 1/2 |psi_T - g|^2, alpha_n = F a_n:
//   alpha_n = (C1/C2) alpha_{n+1} + F[ N'(psi_n) F^-1[ dt/2/C2 (c_n alpha_{n+1} - alpha_{n+2}) ] ].
// One adjoint step is the forward chain's shape: y inverse, z (C2R, x N'(psi_n), R2C: the
// periodic z pass in its Z_OP_MUL form), y forward, x (forward, update, next w, inverse) --
// with the carry alpha in place of R.
#pragma once
#include "smx_advect.h"

namespace smx {

// x pass of the adjoint step n: stage holds the y,z-forward transform of v_n = N'(psi_n) w_n,
// carry holds alpha_{n+1}.
//   alpha_n = (C1/C2) alpha_{n+1} + s V,  w_{n-1}(k) = dt/2/C2 (cnext alpha_n - alpha_{n+1}).
// Parameter sums: step n contributes  w_kz Re(conj alpha_{n+1} dt/2 dL/dtheta (Psi_n + Psi_{n+1})/C2)
// for theta = eps (dL = 1) and q0 (dL = 4 alpha q0 (k^2 - q0^2)).  Only one spectrum, psin, is
// read per call: the call for n adds the Psi_n half of step n (conj alpha_{n+1} Psi_n) and, when
// tail != 0, the Psi_n half of step n-1 (conj alpha_n Psi_n).  init != 0 sets alpha_T from a_T
// (alpha_{n+1} = 0); with psin = Psi_T and tail = 1 it adds the Psi_T half of step T-1.  Block
// sums are added to acc[2b], acc[2b+1] (psin == nullptr: no sums).
// The self-conjugate planes kz = 0, nz/2 pair columns ky / -ky as in k_xpass, and alpha_n is
// projected onto Hermitian symmetry there before it is stored.
// Tile (plane by, columns bx*NC ..; gx tiles per plane) of k_xpass_adj and of the grouped
// k_xpass_adj_grp.
template <typename T, int L, int NC, int V>
__device__ __forceinline__ void xpass_adj_tile(cplx<T>* stage, cplx<T>* __restrict__ carry,
                                               const cplx<T>* __restrict__ psin, const PParams<T> p, T cnext,
                                               int init, int tail, double q0d, double alphad,
                                               double* __restrict__ acc, int bx, int by, int gx) {
  constexpr int E = PlanV<L, V>::E, TPC = L / E, NT = NC * TPC, NCP = NC + 1;
  constexpr uint32_t CB = sizeof(cplx<T>);
  __shared__ cplx<T> S[L * NCP];
  __shared__ double red[64];
  const int tid = threadIdx.x, col = tid % NC, tc = tid / NC;
  __shared__ cplx<T> TW[L];
  const cplx<T>* twl = tw_lds<L, NT>(TW, p.twx, tid);
  __shared__ T KX[L];  // k_x for the loop after the forward transform's barriers
  for (int i = tid; i < L; i += NT) KX[i] = p.kx[i];
  const int kz = by, y0 = bx * NC;
  const int64_t plane = (int64_t)kz * p.nxy;
  const uint32_t rowb = (uint32_t)p.ny * CB;
  constexpr int H = NC / 2;
  const bool herm = (kz == 0 || 2 * kz == p.nz);
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
  const uint32_t cofs = (uint32_t)ky * CB;
  cplx<T>* sb = stage + plane;
  cplx<T>* cb = carry + plane;
  const bool sums = psin != nullptr && acc != nullptr;
  // Every read is issued first (the stage rows, then alpha_{n+1} and Psi_n at the entries the
  // forward transform leaves in this thread) and every store comes last (the new stage, then
  // alpha_n from registers): a wave's loads and stores retire in order, and the exchanges order
  // LDS only (lds_barrier), so no barrier waits for a store and the carry reads overlap the
  // forward transform.  In place: each stage row is consumed before the first exchange, each
  // carry entry is read and written by this thread (or its partner value comes through the tile).
  cplx<T> v[E], a1r[E], s0r[E];
#pragma unroll
  for (int e = 0; e < E; ++e)
    v[e] = ld_off<cplx<T>>(sb, (uint32_t)Tile<T, L, E, NCP>::template pos_in<fwd_first<L, V>()>(tc, e) * rowb + cofs);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const uint32_t o = (uint32_t)Tile<T, L, E, NCP>::template pos_in<fwd_last<L, V>()>(tc, e) * rowb + cofs;
    a1r[e] = init ? mk<T>(0, 0) : ld_off<cplx<T>>(cb, o);
    s0r[e] = sums ? ld_off(psin + plane, o) : mk<T>(0, 0);
  }
  FwdFFT<T, L, NCP, V, true>::run(v, S, col, tc, twl);
  const T kyv = p.ky[ky], kzv = p.kz[kz];
  const double wk = herm ? 1.0 : 2.0;
  double ge = 0.0, gq = 0.0;
  cplx<T> a0r[E];
  lds_barrier();  // S receives alpha_n below (self-conjugate planes); the transform's last reads must be done
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int kxi = Tile<T, L, E, NCP>::template pos_in<fwd_last<L, V>()>(tc, e);
    const T kxv = KX[kxi];
    const T k2 = kxv * kxv + kyv * kyv + kzv * kzv;
    const T d = k2 - p.q02;
    const T aLin = p.eps - p.alpha * d * d;
    const T C1 = (T)1 + p.dtd2 * aLin, C2 = (T)1 - p.dtd2 * aLin;
    const T ic2 = recip(C2);
    const cplx<T> a1 = a1r[e];
    const T g = C1 * ic2;
    const cplx<T> a0 = mk(g * a1.x + p.scale * v[e].x, g * a1.y + p.scale * v[e].y);
    if (sums) {  // per-element terms in the working precision, per-thread sums in FP64
      const cplx<T> s0 = s0r[e];
      T re = a1.x * s0.x + a1.y * s0.y;
      if (tail) re += a0.x * s0.x + a0.y * s0.y;
      const T base = re * p.dtd2 * ic2;
      ge += (double)base;
      gq += (double)(base * (k2 - p.q02));
    }
    if (herm) S[kxi * NCP + col] = a0;
    a0r[e] = a0;
    const T m = p.dtd2 * ic2;
    v[e] = mk(m * (cnext * a0.x - a1.x), m * (cnext * a0.y - a1.y));
  }
  if (herm) {  // alpha(k) <- (alpha(k) + conj alpha(-k))/2
    lds_barrier();
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int kxi = Tile<T, L, E, NCP>::template pos_in<fwd_last<L, V>()>(tc, e);
      const cplx<T> a = S[kxi * NCP + col], b = S[((L - kxi) % L) * NCP + pcol];
      a0r[e] = mk((T)0.5 * (a.x + b.x), (T)0.5 * (a.y - b.y));
    }
  }
  double be = 0.0, bq = 0.0;
  if (sums) {  // plane weight (1 on the self-conjugate planes, 2 elsewhere) and dL/dq0 = 4 alpha q0 (k^2 - q0^2)
    be = block_sum(ge, red);
    bq = block_sum(gq, red);
  }
  InvFFT<T, L, NCP, V, true>::run(v, S, col, tc, twl);
#pragma unroll
  for (int e = 0; e < E; ++e)
    st_off(sb, (uint32_t)Tile<T, L, E, NCP>::template pos_in<inv_last<L, V>()>(tc, e) * rowb + cofs, v[e]);
#pragma unroll
  for (int e = 0; e < E; ++e)
    st_off(cb, (uint32_t)Tile<T, L, E, NCP>::template pos_in<fwd_last<L, V>()>(tc, e) * rowb + cofs, a0r[e]);
  if (sums && tid == 0) {  // the block's accumulators, read and written last
    const int b = by * gx + bx;
    acc[2 * b] += wk * be;
    acc[2 * b + 1] += wk * 4.0 * alphad * q0d * bq;
  }
}
template <typename T, int L, int NC, int V = 0>
__global__ __launch_bounds__(NC*(L / PlanV<L, V>::E)) void k_xpass_adj(cplx<T>* stage, cplx<T>* __restrict__ carry,
                                                                   const cplx<T>* __restrict__ psin, PParams<T> p,
                                                                   T cnext, int init, int tail, double q0d,
                                                                   double alphad, double* __restrict__ acc) {
  xpass_adj_tile<T, L, NC, V>(stage, carry, psin, p, cnext, init, tail, q0d, alphad, acc, blockIdx.x, blockIdx.y,
                              gridDim.x);
}

// Per-box arguments of the grouped adjoint x pass (box = blockIdx.z); cnext, init and tail are
// the same for every box of a group at one step.
template <typename T>
struct AdjGrpArgs {
  cplx<T>* stage;
  cplx<T>* carry;
  const cplx<T>* psin;
  double* acc;
  double q0d, alphad;
  PParams<T> p;
};
template <typename T, int L, int NC, int V>
__global__ __launch_bounds__(NC*(L / PlanV<L, V>::E)) void k_xpass_adj_grp(const AdjGrpArgs<T>* __restrict__ g,
                                                                       T cnext, int init, int tail) {
  const AdjGrpArgs<T>& b = g[blockIdx.z];
  xpass_adj_tile<T, L, NC, V>(b.stage, b.carry, b.psin, b.p, cnext, init, tail, b.q0d, b.alphad, b.acc, blockIdx.x,
                              blockIdx.y, gridDim.x);
}

struct PassSched;  // tile schedule of a pass (smx_launch.h)

template <typename T>
struct AdjointOps {
  static hipError_t xpass_adj(cplx<T>* stage, cplx<T>* carry, const cplx<T>* psin, const PParams<T>& p, T cnext,
                              int init, int tail, double q0d, double alphad, double* acc, hipStream_t st);
  // One reverse step of nbox boxes sharing the grid of p and the schedules s: y inverse, z
  // (Z_OP_MUL with each box's recorded N'), y forward on the stage arrays of gz, then the
  // adjoint x pass with gx (both arrays in device memory, one entry per box).
  static hipError_t adj_step_grp(const GrpArgs<T>* gz, const AdjGrpArgs<T>* gx, int nbox, const PParams<T>& p,
                                 int64_t npencils, const PassSched* s, T cnext, int init, int tail, hipStream_t st);
  // the grouped adjoint x pass alone
  static hipError_t xpass_adj_grp(const AdjGrpArgs<T>* gx, int nbox, const PParams<T>& p, T cnext, int init,
                                  int tail, hipStream_t st);
};

}  // namespace smx
