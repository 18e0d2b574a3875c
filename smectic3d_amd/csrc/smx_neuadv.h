// smx_neuadv.h -- kernels of the advected Neumann step: the advected step of pf3d.cpp:399-465
// (force, Stokes projection, N = N(psi) - U.grad psi) on the DCT grid of cosNoAdvMorph.cpp.
// The reference has sine/cosine Stokes machinery for its Neumann box (morphSurfK.cpp:331-348
// wavenumbers, :392-418 plans, :864-882 projection) but never applies it (:896-898) and
// projects every component in one all-sine basis; these kernels use per-component parity.
//
// Parity.  psi, psiSH, mu and N are cosine series along every axis.  d_a psi, the force S_a
// and the velocity U_a are sine series along axis a and cosine series along the other two.
// Sine axes go through the cosine kernels (DPar in smx_neumann.h): a sine-parity real field is
// held folded, (-1)^j x_j along each of its sine axes, and its spectrum in slot order (slot 0 =
// wavenumber pi/d).  Folding is never undone: folded fields only meet as U_a d_a psi (both
// folded along axis a, the signs cancel) or as the input S_a = f d_a psi of the forward sine
// transform along axis a, which takes the folded line.
//
// Projection at a slot triple (sx, sy, sz) with K = pi (sx/Lx, sy/Ly, sz/Lz): the components
// whose own axis slot is nonzero share the wavevector K and are projected together,
//   U_a = s cm (S_a - K_a (K.S)/|K|^2) / |K|^2    (pf3d.cpp:433-447, cm = 1/(rho eta));
// a component whose own axis slot is 0 is alone at the wavevector with pi/d_a on its axis (no
// cosine series has that wavenumber) and keeps its part normal to it,
//   U_a = s cm S_a |K|^2 / (|K|^2 + (pi/d_a)^2)^2.
// That is the periodic step of the box mirrored evenly to [0, 2L) along every axis,
// symmetrised over the mirror images (which removes the one-sided coupling that the periodic
// step's +pi/dx Nyquist convention gives the top mode).
//
// Per step (real arrays [kz][x|kx][y|ky], in place per tile), 15 launches:
//   y inverse:  A -> By (d_y, sine), A, Ax, As         (psi, d_x psi, d_y psi, psiSH at (x,y,kz))
//   k_dctz_force:  A, Ax, By, As -> S0, S1, S2         (z DCT-III, f grad psi, z DCT-II / DST-II)
//   y forward x3 (S1 sine);  k_dctx_stokes: S -> U;  y inverse x3 (U1 sine)
//   k_dctz_advnl:  U0..2, A, Ax, By -> N               (N(psi) - U.grad psi, z DCT-II)
//   y forward;  k_dctx_adv: N, R -> R, A, Ax, As       (DCT-II x, CNAB2, three DCT-III / DST-III x)
#pragma once
#include "smx_advect.h"
#include "smx_neumann.h"

namespace smx {

// z-pass helpers over one tile of NC columns of the (x,y) plane: adjacent columns (2q, 2q+1)
// share one complex FFT and move with one 2-real access per row (as k_dct_zpass).  Real values
// sit at z = mk_src(pos_in<fwd_first>) in every field alike, so products are taken in registers.
template <typename T, int L, int NC>
struct DZ {
  static constexpr int E = Plan<L>::E, TPC = L / E, NPR = NC / 2, NT = NPR * TPC, FP = NPR + 1;
  static constexpr uint32_t CB = sizeof(cplx<T>);

  // One field's rows of the tile, staged through LDS, to the FFT inputs of the inverse
  // transform along z (parity par, uniform): each thread loads its rows k = pos_in<inv_first>
  // once and reads the partner rows (L-k) mod L from the tile.  With DZ2, v2 also receives the
  // inputs of the derivative d/dz (DP_DSIN) of the same (cosine) field.  kq, qw: LDS tables.
  template <bool DZ2 = false>
  __device__ static void stage_form(const T* in, int64_t f0, uint32_t rowb, int q, int tc, int par, const T* kq,
                                    const cplx<T>* qw, cplx<T>* F, cplx<T>* v, cplx<T>* v2 = nullptr) {
    const cplx<T>* src = reinterpret_cast<const cplx<T>*>(in + f0);
    cplx<T> rk[E];
#pragma unroll
    for (int e = 0; e < E; ++e)
      rk[e] = ld_off(src, (uint32_t)Tile<T, L, E, FP>::template pos_in<inv_first<L>()>(tc, e) * rowb +
                              (uint32_t)q * CB);
    __syncthreads();  // the tile's earlier reads are done
#pragma unroll
    for (int e = 0; e < E; ++e) F[Tile<T, L, E, FP>::template pos_in<inv_first<L>()>(tc, e) * FP + q] = rk[e];
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = Tile<T, L, E, FP>::template pos_in<inv_first<L>()>(tc, e);
      const cplx<T> rm = F[((L - k) % L) * FP + q];
      v[e] = (par == DP_COS) ? dct3_par<DP_COS>(rk[e], rm, k, L, kq, qw[k])
                             : (par == DP_SIN ? dct3_par<DP_SIN>(rk[e], rm, k, L, kq, qw[k])
                                              : dct3_par<DP_DSIN>(rk[e], rm, k, L, kq, qw[k]));
      if constexpr (DZ2) v2[e] = dct3_par<DP_DSIN>(rk[e], rm, k, L, kq, qw[k]);
    }
  }
  // forward transform along z (parity par, uniform) of the pairs in v and store of the spectra
  __device__ static void r2c_store(cplx<T>* v, T* out, int par, int64_t f0, uint32_t rowb, int q, int tc, int tid,
                                   const cplx<T>* qw, cplx<T>* F, const cplx<T>* tw) {
    FwdFFT<T, L, FP>::run(v, F, q, tc, tw);
    __syncthreads();
    Tile<T, L, E, FP>::template store<fwd_last<L>()>(v, F, q, tc);
    __syncthreads();
    cplx<T>* dst = reinterpret_cast<cplx<T>*>(out + f0);
    for (int i = tid; i < L * NPR; i += NT) {
      const int k = i / NPR, qq = i - k * NPR;
      const int s = (par == DP_COS) ? k : (L - k) % L;
      st_off(dst, (uint32_t)s * rowb + (uint32_t)qq * CB, dct2_out(F[k * FP + qq], F[((L - k) % L) * FP + qq], qw[k]));
    }
  }
};

// LDS tables of a z kernel: twiddles, quarter-wave twiddles and k_z, filled before any use.
#define SMX_DZ_TABLES(NT_)                                  \
  __shared__ cplx<T> TW[L], QW[L];                         \
  __shared__ T KZ[L];                                      \
  const cplx<T>* twl = tw_lds<L, NT_>(TW, p.twz, tid);      \
  tw_lds<L, NT_>(QW, qtw, tid);                            \
  for (int i = tid; i < L; i += NT_) KZ[i] = p.kz[i];      \
  __syncthreads();

// Force S = f grad psi (pf3d.cpp:417-425), f = psiSH or mu = psiSH - eps psi - N(psi).
// Inputs at (x, y, kz): psi, d_x psi, d_y psi, psiSH; outputs at (x, y, kz'): S0, S1 (cosine
// along z), S2 (sine along z; d_z psi comes from psi_hat with -k_z).  Whole tiles.
template <typename T, int L, int NC>
__global__ __launch_bounds__((NC / 2) * (L / Plan<L>::E)) void k_dctz_force(const T* bpsi, const T* bx, const T* by,
                                                                            const T* bs, T* o0, T* o1, T* o2,
                                                                            AParams<T> ap,
                                                                            const cplx<T>* __restrict__ qtw) {
  using Z = DZ<T, L, NC>;
  constexpr int E = Z::E;
  __shared__ cplx<T> F[L * Z::FP];
  const PParams<T>& p = ap.p;
  const int tid0 = threadIdx.x, q0 = tid0 % Z::NPR, tc0 = tid0 / Z::NPR;
  const int tid = tid0;
  SMX_DZ_TABLES(Z::NT)
  const int64_t f0 = (int64_t)blockIdx.x * NC;
  const uint32_t rowb = (uint32_t)p.nxy * sizeof(T);
  // Fields in turn through one inverse and one forward transform instance (instead of one
  // inlined transform per field):
  //   0 psiSH -> f;  1 psi -> f = mu (F_MU only; also forms d_z psi's inputs);  2 d_z psi -> S2;
  //   3 d_x psi -> S0;  4 d_y psi -> S1.
  const bool mu = (ap.force == F_MU);
  cplx<T> w[E], f[E], g[E];
#pragma unroll 1
  for (int i = 0; i < 5; ++i) {
    if (i == 1 && !mu) continue;
    // per-iteration copies of the lane indices: the address arithmetic of the transforms then
    // stays inside the loop instead of being hoisted into registers live across it
    int q = q0, tc = tc0, tid = tid0;  // shadows the function-scope tid
    asm volatile("" : "+v"(q), "+v"(tc), "+v"(tid));
    if (i == 0) Z::stage_form(bs, f0, rowb, q, tc, DP_COS, KZ, QW, F, w);
    else if (i == 1) Z::template stage_form<true>(bpsi, f0, rowb, q, tc, DP_COS, KZ, QW, F, w, g);
    else if (i == 2 && mu) {
#pragma unroll
      for (int e = 0; e < E; ++e) w[e] = g[e];
    } else Z::stage_form(i == 2 ? bpsi : (i == 3 ? bx : by), f0, rowb, q, tc, i == 2 ? DP_DSIN : DP_COS, KZ, QW, F,
                         w);
    InvFFT<T, L, Z::FP>::run(w, F, q, tc, twl);
    if (i == 0) {
#pragma unroll
      for (int e = 0; e < E; ++e) f[e] = w[e];
    } else if (i == 1) {
#pragma unroll
      for (int e = 0; e < E; ++e)
        f[e] = mk(f[e].x - p.eps * w[e].x - nonlin(p, w[e].x), f[e].y - p.eps * w[e].y - nonlin(p, w[e].y));
    } else {
#pragma unroll
      for (int e = 0; e < E; ++e) w[e] = mk(f[e].x * w[e].x, f[e].y * w[e].y);
      Z::r2c_store(w, i == 2 ? o2 : (i == 3 ? o0 : o1), i == 2 ? DP_SIN : DP_COS, f0, rowb, q, tc, tid, QW, F, twl);
    }
  }
}

// N = N(psi) - U.grad(psi) (pf3d.cpp:455-465), then DCT-II along z; U2 and d_z psi are sine
// series along z.  FP64 tile partials of N for the mixed path's mean mode.
template <typename T, int L, int NC>
__global__ __launch_bounds__((NC / 2) * (L / Plan<L>::E)) void k_dctz_advnl(const T* u0, const T* u1, const T* u2,
                                                                            const T* bpsi, const T* bx, const T* by,
                                                                            T* nout, AParams<T> ap,
                                                                            const cplx<T>* __restrict__ qtw) {
  using Z = DZ<T, L, NC>;
  constexpr int E = Z::E;
  __shared__ cplx<T> F[L * Z::FP];
  __shared__ double red[64];
  const PParams<T>& p = ap.p;
  const int tid0 = threadIdx.x, q0 = tid0 % Z::NPR, tc0 = tid0 / Z::NPR;
  const int tid = tid0;
  SMX_DZ_TABLES(Z::NT)
  const int64_t f0 = (int64_t)blockIdx.x * NC;
  const uint32_t rowb = (uint32_t)p.nxy * sizeof(T);
  // Fields in turn through one inverse transform instance:
  //   0 psi -> N(psi) (also forms d_z psi's inputs);  1 d_z psi;  2 U2;  3 U0;  4 d_x psi;
  //   5 U1;  6 d_y psi,
  // subtracting U_a d_a psi as each pair completes.
  cplx<T> w[E], aux[E], nacc[E];
#pragma unroll 1
  for (int i = 0; i < 7; ++i) {
    int q = q0, tc = tc0;  // per-iteration lane indices (as in k_dctz_force)
    asm volatile("" : "+v"(q), "+v"(tc));
    if (i == 0) Z::template stage_form<true>(bpsi, f0, rowb, q, tc, DP_COS, KZ, QW, F, w, aux);
    else if (i == 1) {
#pragma unroll
      for (int e = 0; e < E; ++e) w[e] = aux[e];
    } else
      Z::stage_form(i == 2 ? u2 : (i == 3 ? u0 : (i == 4 ? bx : (i == 5 ? u1 : by))), f0, rowb, q, tc,
                    i == 2 ? DP_SIN : DP_COS, KZ, QW, F, w);
    InvFFT<T, L, Z::FP>::run(w, F, q, tc, twl);
    if (i == 0) {
#pragma unroll
      for (int e = 0; e < E; ++e) nacc[e] = mk(nonlin(p, w[e].x), nonlin(p, w[e].y));
    } else if (i & 1) {
#pragma unroll
      for (int e = 0; e < E; ++e) aux[e] = w[e];
    } else {
#pragma unroll
      for (int e = 0; e < E; ++e) nacc[e] = mk(nacc[e].x - aux[e].x * w[e].x, nacc[e].y - aux[e].y * w[e].y);
    }
  }
  if (p.partial != nullptr) wave_partial<E, Z::NT>(nacc, red, tid);
  Z::r2c_store(nacc, nout, DP_COS, f0, rowb, q0, tc0, tid, QW, F, twl);  // its barriers order red[]
  if (p.partial != nullptr) store_partial<Z::NT>(red, p.partial + blockIdx.x, tid);
}

#undef SMX_DZ_TABLES

// Stokes x pass: forward x of S0 (sine along x), S1, S2 (cosine), the projection per slot
// triple, inverse x of U0 (sine), U1, U2; in place per tile.  Tile: plane kz, columns
// y0 .. y0+NC-1 (pairs of adjacent ky); V: plan variant.  One LDS tile serves the six transforms: each
// component's spectrum goes from the tile to the registers of the thread that owns its slots
// (k, L-k), where the projection takes all three.  qn: pi/d per axis.
template <typename T, int L, int NC, int V = 0>
__global__ __launch_bounds__((NC / 2) * (L / PlanV<L, V>::E)) void k_dctx_stokes(T* s0, T* s1, T* s2, AParams<T> ap,
                                                                             const cplx<T>* __restrict__ qtw,
                                                                             T qnx, T qny, T qnz) {
  constexpr int E = PlanV<L, V>::E, TPC = L / E, NPR = NC / 2, NT = NPR * TPC, FP = NPR + 1;
  constexpr int NIT = (L / 2 + TPC) / TPC;
  constexpr uint32_t CB = sizeof(cplx<T>);
  __shared__ cplx<T> F[L * FP];
  const PParams<T>& p = ap.p;
  const int tid = threadIdx.x, q = tid % NPR, tc = tid / NPR;
  __shared__ cplx<T> TW[L], QW[L];
  __shared__ T KX[L];
  const cplx<T>* twl = tw_lds<L, NT>(TW, p.twx, tid);
  tw_lds<L, NT>(QW, qtw, tid);
  for (int i = tid; i < L; i += NT) KX[i] = p.kx[i];
  const int kz = blockIdx.y, y0 = blockIdx.x * NC;
  const int64_t plane = (int64_t)kz * p.nxy + y0;
  const uint32_t rowb = (uint32_t)p.ny * sizeof(T);
  cplx<T>* arr[3] = {reinterpret_cast<cplx<T>*>(s0 + plane), reinterpret_cast<cplx<T>*>(s1 + plane),
                     reinterpret_cast<cplx<T>*>(s2 + plane)};
  const T kzv = p.kz[kz], kya = p.ky[y0 + 2 * q], kyb = p.ky[y0 + 2 * q + 1];
  cplx<T>* const a0 = arr[0];
  cplx<T>* const a1 = arr[1];
  cplx<T>* const a2 = arr[2];
  auto comp = [&](int a) { return a == 0 ? a0 : (a == 1 ? a1 : a2); };
  auto load = [&](int a, cplx<T>* w) {
    const cplx<T>* src = comp(a);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int s = mk_src(Tile<T, L, E, FP>::template pos_in<fwd_first<L, V>()>(tc, e), L);
      w[e] = ld_off<cplx<T>>(src, (uint32_t)s * rowb + (uint32_t)q * CB);
    }
  };
  // spectra at slots k = tc + it*TPC and L-k, per component; the components go through one
  // transform instance each way (loops over a, instead of one inlined transform per component)
  cplx<T> v[E];
  cplx<T> fk0[NIT], fm0[NIT], fk1[NIT], fm1[NIT], fk2[NIT], fm2[NIT];
#pragma unroll 1
  for (int a = 0; a < 3; ++a) {
    load(a, v);
    FwdFFT<T, L, FP, V>::run(v, F, q, tc, twl);
    __syncthreads();  // the previous component's reads of the tile are done
    Tile<T, L, E, FP>::template store<fwd_last<L, V>()>(v, F, q, tc);
    __syncthreads();
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
      const int k = tc + it * TPC;
      const int km = (L - k) % L;
      const bool in = k <= L / 2, two = in && k != 0 && 2 * k != L;
      const cplx<T> Zk = F[(in ? k : 0) * FP + q], Zm = F[(in ? km : 0) * FP + q];
      const cplx<T> yk = in ? dct2_out(Zk, Zm, QW[in ? k : 0]) : mk<T>(0, 0);
      const cplx<T> ym = two ? dct2_out(Zm, Zk, QW[km]) : mk<T>(0, 0);
      if (a == 0) {  // sine along x: slot k holds the DCT-II output at (L-k) mod L
        fk0[it] = two ? ym : yk;
        fm0[it] = yk;
      } else if (a == 1) {
        fk1[it] = yk;
        fm1[it] = ym;
      } else {
        fk2[it] = yk;
        fm2[it] = ym;
      }
    }
  }
  const bool ya0 = (y0 + 2 * q == 0), z0 = (kz == 0);  // column b never has ky slot 0
  const T cs = p.scale * ap.cm;
  // U of one column at x slot sx from f = (S0, S1, S2) there; qy / y0s: its k_y, ky-slot-0 flag
  struct U3 { T x, y, z; };
  auto proj1 = [&](T fx, T fy, T fz, int sx, T qy, bool y0s) -> U3 {
    const T qx = KX[sx];
    const T k2 = qx * qx + qy * qy + kzv * kzv;
    const T dot = qx * fx + qy * fy + kzv * fz;
    const T ik2 = (k2 > (T)0) ? recip(k2) : (T)0;
    const T c = cs * ik2;
    const T iax = recip(k2 + qnx * qnx), iay = recip(k2 + qny * qny), iaz = recip(k2 + qnz * qnz);
    U3 u;
    u.x = (sx != 0) ? c * (fx - qx * dot * ik2) : cs * fx * k2 * iax * iax;
    u.y = (!y0s) ? c * (fy - qy * dot * ik2) : cs * fy * k2 * iay * iay;
    u.z = (!z0) ? c * (fz - kzv * dot * ik2) : cs * fz * k2 * iaz * iaz;
    return u;
  };
#pragma unroll
  for (int it = 0; it < NIT; ++it) {
    const int k = tc + it * TPC;
    const int km = (L - k) % L;
    const bool in = k <= L / 2;
    const int kk = in ? k : 0, kmm = in ? km : 0;
    const U3 a = proj1(fk0[it].x, fk1[it].x, fk2[it].x, kk, kya, ya0);
    const U3 b = proj1(fk0[it].y, fk1[it].y, fk2[it].y, kk, kyb, false);
    const U3 c = proj1(fm0[it].x, fm1[it].x, fm2[it].x, kmm, kya, ya0);
    const U3 d = proj1(fm0[it].y, fm1[it].y, fm2[it].y, kmm, kyb, false);
    fk0[it] = mk(a.x, b.x); fk1[it] = mk(a.y, b.y); fk2[it] = mk(a.z, b.z);
    fm0[it] = mk(c.x, d.x); fm1[it] = mk(c.y, d.y); fm2[it] = mk(c.z, d.z);
  }
#pragma unroll 1
  for (int a = 0; a < 3; ++a) {
    __syncthreads();  // the last reads of the tile (items or the previous inverse) are done
    // DCT-III inputs: cosine components from slots k, L-k; the sine one from L-k, k.  One call
    // per component, so that each reads its own arrays with constant indexes.
    auto inputs = [&](const cplx<T>* fk, const cplx<T>* fm, bool sine) {
#pragma unroll
      for (int it = 0; it < NIT; ++it) {
        const int k = tc + it * TPC;
        if (k > L / 2) continue;
        const int km = (L - k) % L;
        const bool two = (k != 0 && 2 * k != L);
        const cplx<T> uk = fk[it], um = fm[it];
        if (!sine || k == 0 || !two) {  // the sine one at k = 0 (slot 0 alone) and L/2 as cosine
          F[k * FP + q] = dct3_in(uk, two ? um : (k == 0 ? mk<T>(0, 0) : uk), QW[k]);
          if (two) F[km * FP + q] = dct3_in(um, uk, QW[km]);
        } else {
          F[k * FP + q] = dct3_in(um, uk, QW[k]);
          F[km * FP + q] = dct3_in(uk, um, QW[km]);
        }
      }
    };
    if (a == 0) inputs(fk0, fm0, true);
    else if (a == 1) inputs(fk1, fm1, false);
    else inputs(fk2, fm2, false);
    __syncthreads();
    Tile<T, L, E, FP>::template load<inv_first<L, V>()>(v, F, q, tc);
    InvFFT<T, L, FP, V>::run(v, F, q, tc, twl);
    cplx<T>* dst = comp(a);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int d = mk_src(Tile<T, L, E, FP>::template pos_in<inv_last<L, V>()>(tc, e), L);
      st_off(dst, (uint32_t)d * rowb + (uint32_t)q * CB, v[e]);
    }
  }
}

// x pass of the advected Neumann step: DCT-II along x of N, the CNAB2 update (MODE X_STEP, R
// carried as C1 psi_hat - dt/2 N_past) or the bootstrap (MODE X_BOOT: psi_hat from psi0,
// R = C1 psi_hat - dt/2 N_hat(psi0)), then three inverse outputs along x: psi (DCT-III),
// d_x psi (DST-III of -k_x psi_hat, folded) and psiSH (DCT-III of opSH psi_hat).
// Tile: plane kz, columns y0 .. y0+NC-1 (pairs of adjacent ky).  in may alias outS.
template <typename T, int L, int NC, int MODE>
__global__ __launch_bounds__((NC / 2) * (L / Plan<L>::E)) void k_dctx_adv(const T* in, T* __restrict__ rbuf,
                                                                          const T* __restrict__ psi0, T* outA,
                                                                          T* outX, T* outS, PParams<T> p,
                                                                          const cplx<T>* __restrict__ qtw) {
  static_assert(MODE == X_STEP || MODE == X_BOOT, "step or bootstrap");
  constexpr int E = Plan<L>::E, TPC = L / E, NPR = NC / 2, NT = NPR * TPC, FP = NPR + 1;
  constexpr uint32_t CB = sizeof(cplx<T>);
  constexpr int NIT = (L / 2 + TPC) / TPC;
  __shared__ cplx<T> F[L * FP];
  __shared__ double red[64];
  const int tid = threadIdx.x, q = tid % NPR, tc = tid / NPR;
  __shared__ cplx<T> TW[L], QW[L];
  __shared__ T KX[L];
  const cplx<T>* twl = tw_lds<L, NT>(TW, p.twx, tid);
  tw_lds<L, NT>(QW, qtw, tid);
  for (int i = tid; i < L; i += NT) KX[i] = p.kx[i];
  const int kz = blockIdx.y, y0 = blockIdx.x * NC;
  const int64_t plane = (int64_t)kz * p.nxy + y0;
  const uint32_t rowb = (uint32_t)p.ny * sizeof(T);
  const cplx<T>* src = reinterpret_cast<const cplx<T>*>(in + plane);
  cplx<T> v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int s = mk_src(Tile<T, L, E, FP>::template pos_in<fwd_first<L>()>(tc, e), L);
    v[e] = ld_off(src, (uint32_t)s * rowb + (uint32_t)q * CB);
  }
  double nsum = 0.0;
  const bool owner0 = (p.partial != nullptr) && kz == 0 && y0 == 0;
  if (owner0) {  // FP64 mean mode of the mixed path: sum of the z pass's partials
    double acc = 0.0;
    for (int i = tid; i < p.npartial; i += NT) acc += p.partial[i];
    nsum = block_sum(acc, red);
  }
  FwdFFT<T, L, FP>::run(v, F, q, tc, twl);
  __syncthreads();
  Tile<T, L, E, FP>::template store<fwd_last<L>()>(v, F, q, tc);
  __syncthreads();
  cplx<T>* rb = reinterpret_cast<cplx<T>*>(rbuf + plane);
  const cplx<T>* cb = (MODE == X_STEP) ? (const cplx<T>*)rb : reinterpret_cast<const cplx<T>*>(psi0 + plane);
  const T kzv = p.kz[kz], kya = p.ky[y0 + 2 * q], kyb = p.ky[y0 + 2 * q + 1];
  // every R (X_STEP) or psi0 (X_BOOT) entry of the thread's items is read before the first
  // store: a wave's loads and stores retire in order, so a read behind a store would wait
  cplx<T> cin[NIT][2], ps[NIT][2];
#pragma unroll
  for (int it = 0; it < NIT; ++it) {
    const int k = tc + it * TPC, km = (L - k) % L;
    const bool inr = k <= L / 2, two = inr && k != 0 && 2 * k != L;
    cin[it][0] = inr ? ld_off(cb, (uint32_t)k * rowb + (uint32_t)q * CB) : mk<T>(0, 0);
    cin[it][1] = two ? ld_off(cb, (uint32_t)km * rowb + (uint32_t)q * CB) : mk<T>(0, 0);
  }
#pragma unroll
  for (int it = 0; it < NIT; ++it) {
    const int k = tc + it * TPC;
    ps[it][0] = ps[it][1] = mk<T>(0, 0);
    if (k > L / 2) continue;
    const int km = (L - k) % L;
    const bool two = (k != 0 && 2 * k != L);
    const cplx<T> Z = F[k * FP + q], Zm = F[km * FP + q];
    cplx<T> y[2];
    y[0] = dct2_out(Z, Zm, QW[k]);
    y[1] = two ? dct2_out(Zm, Z, QW[km]) : mk<T>(0, 0);
    const int kk[2] = {k, km};
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      if (j == 1 && !two) continue;
      const T kxv = KX[kk[j]];
      const T ka = kxv * kxv + kya * kya + kzv * kzv, kb = kxv * kxv + kyb * kyb + kzv * kzv;
      const T La = p.eps - op_sh(p, ka), Lb = p.eps - op_sh(p, kb);  // cosNoAdvMorph.cpp:431-433
      const T C1a = (T)1 + p.dtd2 * La, C2a = (T)1 - p.dtd2 * La;
      const T C1b = (T)1 + p.dtd2 * Lb, C2b = (T)1 - p.dtd2 * Lb;
      const cplx<T> nn = p.scale * y[j];
      cplx<T> s, rn;
      if constexpr (MODE == X_STEP) {  // pf3d.cpp:476-478 with R = C1 psi - dt/2 N_past
        const cplx<T> r = cin[it][j];
        s = mk((r.x + (T)3 * p.dtd2 * nn.x) * recip(C2a), (r.y + (T)3 * p.dtd2 * nn.y) * recip(C2b));
        rn = mk(C1a * s.x - p.dtd2 * nn.x, C1b * s.y - p.dtd2 * nn.y);
        if (owner0 && kk[j] == 0 && q == 0) {
          const double n0 = p.m_scale * nsum;
          const double ps0 = (p.mean_state[0] + 3.0 * p.m_dtd2 * n0) / p.m_c2;
          p.mean_state[0] = p.m_c1 * ps0 - p.m_dtd2 * n0;
          p.mean_state[1] = ps0;
          s.x = (T)ps0;
          rn.x = (T)p.mean_state[0];
        }
      } else {  // the first N_past is N(psi0) without advection (pf3d.cpp:363-374)
        s = cin[it][j];
        rn = mk(C1a * s.x - p.dtd2 * nn.x, C1b * s.y - p.dtd2 * nn.y);
        if (owner0 && kk[j] == 0 && q == 0) {
          const double ps0 = p.mean_state[1];
          p.mean_state[0] = p.m_c1 * ps0 - p.m_dtd2 * p.m_scale * nsum;
          s.x = (T)ps0;
          rn.x = (T)p.mean_state[0];
        }
      }
      st_off(rb, (uint32_t)kk[j] * rowb + (uint32_t)q * CB, rn);
      ps[it][j] = s;
    }
    // psi: DCT-III inputs at k and L-k, in place
    const cplx<T> pk = ps[it][0], pm = two ? ps[it][1] : (k == 0 ? mk<T>(0, 0) : pk);
    F[k * FP + q] = dct3_in(pk, pm, QW[k]);
    if (two) F[km * FP + q] = dct3_in(ps[it][1], pk, QW[km]);
  }
  T* outs[3] = {outA + plane, outX + plane, outS + plane};
#pragma unroll 1
  for (int o = 0; o < 3; ++o) {
    if (o > 0) {
      __syncthreads();  // the previous inverse transform's LDS reads are done
#pragma unroll
      for (int it = 0; it < NIT; ++it) {
        const int k = tc + it * TPC;
        if (k > L / 2) continue;
        const int km = (L - k) % L;
        const bool two = (k != 0 && 2 * k != L);
        if (o == 1) {  // d_x psi: slot s of the sine series = -k_x[s] psi_hat[s] (slot 0: 0)
          const cplx<T> xk = -KX[k] * ps[it][0], xm = -KX[km] * ps[it][1];
          if (k == 0) F[q] = mk<T>(0, 0);
          else if (!two) F[k * FP + q] = dct3_in(xk, xk, QW[k]);
          else {
            F[k * FP + q] = dct3_in(xm, xk, QW[k]);
            F[km * FP + q] = dct3_in(xk, xm, QW[km]);
          }
        } else {  // psiSH = opSH psi_hat (pf3d.cpp:407)
          const T kx2 = KX[k] * KX[k], km2 = KX[km] * KX[km];
          const cplx<T> sk = mk(op_sh(p, kx2 + kya * kya + kzv * kzv) * ps[it][0].x,
                                op_sh(p, kx2 + kyb * kyb + kzv * kzv) * ps[it][0].y);
          const cplx<T> sm = mk(op_sh(p, km2 + kya * kya + kzv * kzv) * ps[it][1].x,
                                op_sh(p, km2 + kyb * kyb + kzv * kzv) * ps[it][1].y);
          const cplx<T> pm = two ? sm : (k == 0 ? mk<T>(0, 0) : sk);
          F[k * FP + q] = dct3_in(sk, pm, QW[k]);
          if (two) F[km * FP + q] = dct3_in(sm, sk, QW[km]);
        }
      }
    }
    __syncthreads();
    Tile<T, L, E, FP>::template load<inv_first<L>()>(v, F, q, tc);
    InvFFT<T, L, FP>::run(v, F, q, tc, twl);
    cplx<T>* dst = reinterpret_cast<cplx<T>*>(outs[o]);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int d = mk_src(Tile<T, L, E, FP>::template pos_in<inv_last<L>()>(tc, e), L);
      st_off(dst, (uint32_t)d * rowb + (uint32_t)q * CB, v[e]);
    }
  }
}

}  // namespace smx
