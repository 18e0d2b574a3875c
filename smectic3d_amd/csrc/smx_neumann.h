// smx_neumann.h -- Neumann (DCT-II / DCT-III) pass kernels of the Swift-Hohenberg step
// (cosFFT/mesabi/cosNoAdvMorph.cpp:483-530).
//
// Every stage array is real, layout [kz][x or kx][y or ky] (y fastest), N entries.
// A step is four in-place launches on the stage buffer A plus the carry R:
//   k_dct_ypass<inv>   (x,ky,kz) -> (x,y,kz)    DCT-III along y
//   k_dct_zpass        (x,y,kz)  -> (x,y,kz)    DCT-III along z, N(psi), DCT-II along z
//   k_dct_ypass<fwd>   (x,y,kz)  -> (x,ky,kz)   DCT-II along y
//   k_dct_xpass        (x,ky,kz) -> (x,ky,kz)   DCT-II along x, CNAB2 update, DCT-III along x
//
// Each 1-D transform of two real lines (a, b) is one length-n complex FFT
// (Makhoul's reordering, IEEE TASSP 28:27-34, 1980):
//   REDFT10: v[m] = z[src(m)], z = a + i b, src(m) = 2m (m < ceil(n/2)) else 2(n-1-m)+1;
//            Z = FFT(v); y_a[k] = 2 Re(e^{-i pi k/2n} (Z[k] + conj Z[n-k])/2),
//            y_b[k] = 2 Re(e^{-i pi k/2n} (Z[k] - conj Z[n-k])/2i).
//   REDFT01: V[k] = e^{i pi k/2n} ((X_k + Y_{n-k}) + i (Y_k - X_{n-k})), X_n = Y_n = 0;
//            z = IFFT(V) (unnormalised); a[src(m)] = Re z[m], b[src(m)] = Im z[m].
// The two lines of a pair are adjacent columns, so a lane moves both with one 2-real
// access (cplx<T> holds the pair: .x column a, .y column b).  The reordering src(m) is
// applied on the HBM side of the first loads and last stores; the z pass needs none
// between its DCT-III and DCT-II, whose FFT positions coincide (forward first-pass inputs
// = inverse last-pass outputs).  In the x pass the DCT-II output at k and n-k, the CNAB2
// update there and the DCT-III input at k and n-k are one work item.
#pragma once
#include "smx_periodic.h"

namespace smx {

__device__ __forceinline__ int mk_src(int m, int n) { return (m < (n + 1) / 2) ? 2 * m : 2 * (n - 1 - m) + 1; }

// Parity of a line along the axis a pass transforms (the advected Neumann step):
//   DP_COS   cosine series, spectrum slot k = wavenumber pi k/L;
//   DP_SIN   sine series (RODFT10/01) through the cosine kernels: the real line is held
//            folded, x_j (-1)^j, and the spectrum in slot order, slot s in 1..n-1 =
//            wavenumber pi s/L, slot 0 = pi n/L.  RODFT10[x]_m = REDFT10[(-1)^j x_j]_{n-1-m}:
//            the DCT-II output at k goes to slot (n-k) mod n; RODFT01[X]_j =
//            (-1)^j REDFT01[X_{n-1-m}]_j: the DCT-III input at k reads slots (n-k) mod n and k;
//   DP_DSIN  inverse only: the derivative of a cosine series, slot s times -kq[s]
//            (d/dx cos(q x) = -q sin(q x); kq[0] = 0), then as DP_SIN.
enum DPar { DP_COS = 0, DP_SIN = 1, DP_DSIN = 2 };

// REDFT10 outputs of a pair at k from Z[k] and Zm = Z[(n-k) mod n]; w = e^{-i pi k/2n}.
template <typename T>
__device__ __forceinline__ cplx<T> dct2_out(cplx<T> Z, cplx<T> Zm, cplx<T> w) {
  return mk(w.x * (Z.x + Zm.x) - w.y * (Z.y - Zm.y), w.x * (Z.y + Zm.y) + w.y * (Z.x - Zm.x));
}
// REDFT01 FFT input at k from the pair at k and its partner pn at n-k (zero for k = 0);
// w = e^{-i pi k/2n} (its conjugate is applied).
template <typename T>
__device__ __forceinline__ cplx<T> dct3_in(cplx<T> p, cplx<T> pn, cplx<T> w) {
  const T ux = p.x + pn.y, uy = p.y - pn.x;
  return mk(w.x * ux + w.y * uy, w.x * uy - w.y * ux);
}

// REDFT01 (or, by PAR, sine) FFT input at k of a line whose spectrum holds rk at slot k and rm
// at slot (L-k) mod L.
template <int PAR, typename T>
__device__ __forceinline__ cplx<T> dct3_par(cplx<T> rk, cplx<T> rm, int k, int L, const T* kq, cplx<T> w) {
  if constexpr (PAR == DP_COS) return dct3_in(rk, (k == 0) ? mk<T>(0, 0) : rm, w);
  cplx<T> P = (k == 0) ? rk : rm, Pn = (k == 0) ? mk<T>(0, 0) : rk;
  if constexpr (PAR == DP_DSIN) {
    P = -kq[(L - k) % L] * P;
    Pn = -kq[k] * Pn;
  }
  return dct3_in(P, Pn, w);
}

// ---------------------------------------------------------------------------
// y pass: contiguous real pencils of length L = ny; pencils g = kz*nx + x, pairs of adjacent
// pencils.  Lanes run along the pencils (tc fastest).  DIR -1: DCT-II, +1: DCT-III.
// The grid covers whole tiles (nx*nz is a multiple of NC; the launcher checks it).
// PAR: parity of the pencils along y (DP_DSIN: DIR +1 only, kq = k_y).  DP_COS with DIR +1 and
// kq = k_y multiplies the cosine coefficients by -k_y^2 (the second derivative).
// TI: the input's type (FP32 rows widened on load into an FP64 transform; surfTrack reads an FP32
// stage this way).
template <typename T, int L, int DIR, int NC, int PAR = DP_COS, typename TI = T>
__global__ __launch_bounds__((NC / 2) * (L / Plan<L>::E)) void k_dct_ypass(const TI* in, T* out,
                                                                           const cplx<T>* __restrict__ tw,
                                                                           const cplx<T>* __restrict__ qtw,
                                                                           const T* __restrict__ kq) {
  static_assert(DIR > 0 || PAR != DP_DSIN, "the derivative is an inverse-transform input");
  constexpr int E = Plan<L>::E, TPC = L / E, NPR = NC / 2, NT = NPR * TPC;
  using AD = ColAddr<L>;
  using TL = TileA<T, L, E, AD>;
  __shared__ cplx<T> S[NPR * AD::LP];
  const int tid = threadIdx.x, tc = tid % TPC, q = tid / TPC;
  __shared__ cplx<T> TW[L];
  const cplx<T>* twl = tw_lds<L, NT>(TW, tw, tid);
  // LDS copy of the quarter-wave twiddles, read after a barrier (the transform's, or the one
  // that stages the DCT-III input rows)
  __shared__ cplx<T> QW[L];
  tw_lds<L, NT>(QW, qtw, tid);
  __shared__ T KQ[PAR == DP_DSIN ? L : 1];  // k_y for the derivative, read after the staging barrier
  if constexpr (PAR == DP_DSIN)
    for (int i = tid; i < L; i += NT) KQ[i] = kq[i];
  const int64_t pa = (int64_t)blockIdx.x * NC + 2 * q;  // pencils pa (column a), pa + 1 (column b)
  const TI* a = in + pa * L;
  const TI* b = a + L;
  cplx<T> v[E];
  if constexpr (DIR < 0) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int s = mk_src(TL::template pos_in<fwd_first<L>()>(tc, e), L);
      v[e] = mk((T)a[s], (T)b[s]);
    }
    FwdFFTA<T, L, AD>::run(v, S, q, tc, twl);
    __syncthreads();
    TL::template store<fwd_last<L>()>(v, S, q, tc);
    __syncthreads();
    T* oa = out + (int64_t)blockIdx.x * NC * L;
    for (int i = tid; i < NPR * L; i += NT) {
      const int qq = i / L, k = i - qq * L;
      const cplx<T> y = dct2_out(S[AD::at(k, qq)], S[AD::at((L - k) % L, qq)], QW[k]);
      const int o = (PAR == DP_COS) ? k : (L - k) % L;
      oa[(2 * qq) * L + o] = y.x;
      oa[(2 * qq + 1) * L + o] = y.y;
    }
  } else {
    // each entry is loaded once: a thread's own entries k stay in registers, the partners L-k
    // come from the tile
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = TL::template pos_in<inv_first<L>()>(tc, e);
      v[e] = mk((T)a[k], (T)b[k]);
    }
    if constexpr (PAR == DP_COS) {
      if (kq != nullptr) {  // -k_y^2, after every load is issued
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const T k = kq[TL::template pos_in<inv_first<L>()>(tc, e)];
          v[e] = (-k * k) * v[e];
        }
      }
    }
    TL::template store<inv_first<L>()>(v, S, q, tc);
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = TL::template pos_in<inv_first<L>()>(tc, e);
      v[e] = dct3_par<PAR>(v[e], S[AD::at((L - k) % L, q)], k, L, KQ, QW[k]);
    }
    InvFFTA<T, L, AD>::run(v, S, q, tc, twl);
    T* oa = out + pa * L;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int d = mk_src(TL::template pos_in<inv_last<L>()>(tc, e), L);
      oa[d] = v[e].x;
      oa[L + d] = v[e].y;
    }
  }
}

// ---------------------------------------------------------------------------
// x pass.  Tile: plane kz = blockIdx.y, columns y0 .. y0+NC-1 (pairs of adjacent ky), rows x.
// Its barriers order LDS only (lds_barrier): R stores issued before the inverse transform drain
// while it runs (each R entry is read and written by one thread; the in-place rows are loaded and
// consumed before the first exchange).
// TI: the input's type, as in k_dct_ypass (an FP32 stage into the FP64 X_FWD_SCALED of surfTrack).
template <typename T, int L, int NC, int MODE, typename TI = T>
__global__ __launch_bounds__((NC / 2) * (L / Plan<L>::E)) void k_dct_xpass(const TI* in, T* out, T* __restrict__ rbuf,
                                                                           const T* __restrict__ psi0, PParams<T> p,
                                                                           const cplx<T>* __restrict__ qtw) {
  constexpr int E = Plan<L>::E, TPC = L / E, NPR = NC / 2, NT = NPR * TPC, FP = NPR + 1;
  constexpr uint32_t CB = sizeof(cplx<T>);
  __shared__ cplx<T> F[L * FP];
  __shared__ double red[64];
  const int tid = threadIdx.x, q = tid % NPR, tc = tid / NPR;
  __shared__ cplx<T> TW[L];
  const cplx<T>* twl = tw_lds<L, NT>(TW, p.twx, tid);
  // LDS copies of qtw and k_x for the item loop (read after the forward transform's barriers)
  __shared__ cplx<T> QW[L];
  __shared__ T KX[L];
  tw_lds<L, NT>(QW, qtw, tid);
  for (int i = tid; i < L; i += NT) KX[i] = p.kx[i];
  const int kz = blockIdx.y, y0 = blockIdx.x * NC;
  const int64_t plane = (int64_t)kz * p.nxy + y0;
  const uint32_t rowb = (uint32_t)p.ny * sizeof(T);
  const cplx<TI>* src = reinterpret_cast<const cplx<TI>*>(in + plane);
  const uint32_t rowbi = (uint32_t)p.ny * sizeof(TI);
  auto ld_in = [&](int row, int col) -> cplx<T> {
    const cplx<TI> w = ld_off(src, (uint32_t)row * rowbi + (uint32_t)col * (uint32_t)sizeof(cplx<TI>));
    return mk((T)w.x, (T)w.y);
  };
  cplx<T>* dst = reinterpret_cast<cplx<T>*>(out + plane);
  cplx<T> v[E];
  double nsum = 0.0;
  const bool owner0 = (p.partial != nullptr) && kz == 0 && y0 == 0;
  if (owner0) {  // FP64 mean mode of the mixed path: sum of the z pass's partials
    double acc = 0.0;
    for (int i = tid; i < p.npartial; i += NT) acc += p.partial[i];
    nsum = block_sum(acc, red);
  }
  if constexpr (MODE == X_INV) {
    // each entry is loaded once: a thread's own slots k stay in registers, the partners L-k
    // come from the tile
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = ld_in(Tile<T, L, E, FP>::template pos_in<inv_first<L>()>(tc, e), q);
    if (p.kop >= 0) {  // the k-space multiplier, columns a and b of the pair (after every load is issued)
      const T kya = p.ky[y0 + 2 * q], kyb = p.ky[y0 + 2 * q + 1], kzv = p.kz[kz];
      lds_barrier();  // the KX fill
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const T qx = KX[Tile<T, L, E, FP>::template pos_in<inv_first<L>()>(tc, e)];
        v[e] = mk(kfactor(p, p.kop, qx, kya, kzv) * v[e].x, kfactor(p, p.kop, qx, kyb, kzv) * v[e].y);
      }
    }
    Tile<T, L, E, FP>::template store<inv_first<L>()>(v, F, q, tc);
    lds_barrier();  // also orders the QW fill before its reads
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = Tile<T, L, E, FP>::template pos_in<inv_first<L>()>(tc, e);
      const cplx<T> pn = (k == 0) ? mk<T>(0, 0) : F[(L - k) * FP + q];
      v[e] = dct3_in(v[e], pn, QW[k]);
    }
  } else {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      v[e] = ld_in(mk_src(Tile<T, L, E, FP>::template pos_in<fwd_first<L>()>(tc, e), L), q);
    }
    FwdFFT<T, L, FP, 0, true>::run(v, F, q, tc, twl);
    lds_barrier();
    Tile<T, L, E, FP>::template store<fwd_last<L>()>(v, F, q, tc);
    lds_barrier();
    if constexpr (MODE == X_MUL2) {
      // Two multiplied inverse transforms of one forward transform: F^-1[m_kop F in] into out,
      // F^-1[m_kop2 F in] into rbuf (p.scale undoes the round trip).  The scaled DCT-II outputs
      // of a thread's items (k and L-k of its own pair, k <= L/2) stay in registers between the
      // two; each inverse takes its DCT-III inputs through the tile, as X_STEP's does.
      constexpr int NIT = (L / 2 + TPC) / TPC;
      const T kzv = p.kz[kz], kya = p.ky[y0 + 2 * q], kyb = p.ky[y0 + 2 * q + 1];
      cplx<T> ys[NIT][2];
#pragma unroll
      for (int it = 0; it < NIT; ++it) {
        const int k = tc + it * TPC, km = (L - k) % L;
        const bool inr = k <= L / 2, two = inr && k != 0 && 2 * k != L;
        ys[it][0] = inr ? p.scale * dct2_out(F[k * FP + q], F[km * FP + q], QW[k]) : mk<T>(0, 0);
        ys[it][1] = two ? p.scale * dct2_out(F[km * FP + q], F[k * FP + q], QW[km]) : mk<T>(0, 0);
      }
      T* outs[2] = {out, rbuf};
#pragma unroll
      for (int pass = 0; pass < 2; ++pass) {
        const int op = pass == 0 ? p.kop : p.kop2;
        lds_barrier();  // every read of the tile (forward outputs, or the first inverse) is done
#pragma unroll
        for (int it = 0; it < NIT; ++it) {
          const int k = tc + it * TPC;
          if (k > L / 2) continue;
          const int km = (L - k) % L;
          const bool two = (k != 0 && 2 * k != L);
          const T kx0 = KX[k], kx1 = KX[km];
          const cplx<T> pk = mk(kfactor(p, op, kx0, kya, kzv) * ys[it][0].x, kfactor(p, op, kx0, kyb, kzv) * ys[it][0].y);
          const cplx<T> pq = mk(kfactor(p, op, kx1, kya, kzv) * ys[it][1].x, kfactor(p, op, kx1, kyb, kzv) * ys[it][1].y);
          // the partner of k = 0 is X_L = Y_L = 0; k = L/2 is its own partner
          const cplx<T> pm = two ? pq : (k == 0 ? mk<T>(0, 0) : pk);
          F[k * FP + q] = dct3_in(pk, pm, QW[k]);
          if (two) F[km * FP + q] = dct3_in(pq, pk, QW[km]);
        }
        lds_barrier();
        Tile<T, L, E, FP>::template load<inv_first<L>()>(v, F, q, tc);
        InvFFT<T, L, FP, 0, true>::run(v, F, q, tc, twl);
        cplx<T>* o = reinterpret_cast<cplx<T>*>(outs[pass] + plane);
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const int d = mk_src(Tile<T, L, E, FP>::template pos_in<inv_last<L>()>(tc, e), L);
          st_off(o, (uint32_t)d * rowb + q * CB, v[e]);
        }
      }
      return;
    }
    // one item: the DCT-II outputs at k and L-k of pair qq, the k-space work there, and the
    // DCT-III inputs at k and L-k
    cplx<T>* rb = rbuf ? reinterpret_cast<cplx<T>*>(rbuf + plane) : nullptr;
    const cplx<T>* p0 = psi0 ? reinterpret_cast<const cplx<T>*>(psi0 + plane) : nullptr;
    const T kzv = p.kz[kz];
    // Thread (q, tc) takes the items k = tc + it*TPC (k <= L/2) of its own pair q.  All of
    // their R (X_STEP) or psi0 (X_BOOT) entries are read before the first store: a wave's
    // loads and stores retire in order, so a read behind a store would wait for it.
    constexpr int NIT = (L / 2 + TPC) / TPC;
    const int qq = q;
    const T kya = p.ky[y0 + 2 * qq], kyb = p.ky[y0 + 2 * qq + 1];
    cplx<T> cin[NIT][2];
    if constexpr (MODE != X_FWD_SCALED) {
      const cplx<T>* cb = (MODE == X_STEP) ? (const cplx<T>*)rb : p0;
#pragma unroll
      for (int it = 0; it < NIT; ++it) {
        const int k = tc + it * TPC, km = (L - k) % L;
        const bool in = k <= L / 2, two = in && k != 0 && 2 * k != L;
        cin[it][0] = in ? ld_off(cb, (uint32_t)k * rowb + qq * CB) : mk<T>(0, 0);
        cin[it][1] = two ? ld_off(cb, (uint32_t)km * rowb + qq * CB) : mk<T>(0, 0);
      }
    }
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
      const int k = tc + it * TPC;
      if (k > L / 2) continue;
      const int km = (L - k) % L;
      const bool two = (k != 0 && 2 * k != L);
      const cplx<T> Z = F[k * FP + qq], Zm = F[km * FP + qq];
      cplx<T> y[2];
      y[0] = dct2_out(Z, Zm, QW[k]);
      y[1] = two ? dct2_out(Zm, Z, QW[km]) : mk<T>(0, 0);
      const int kk[2] = {k, km};
      cplx<T> ps[2];
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        if (j == 1 && !two) {
          ps[1] = mk<T>(0, 0);
          continue;
        }
        const uint32_t o = (uint32_t)kk[j] * rowb + qq * CB;
        if constexpr (MODE == X_FWD_SCALED) {
          cplx<T> r = p.scale * y[j];
          if (owner0 && kk[j] == 0 && qq == 0) {  // k = 0 entry of column a (item 0: thread 0)
            p.mean_state[1] = p.m_scale * nsum;
            r.x = (T)p.mean_state[1];
          }
          st_off(dst, o, r);
        } else {
          const T kxv = KX[kk[j]];
          const T ka = kxv * kxv + kya * kya + kzv * kzv, kb = kxv * kxv + kyb * kyb + kzv * kzv;
          const T da = ka - p.q02, db = kb - p.q02;
          const T La = p.eps - p.alpha * da * da, Lb = p.eps - p.alpha * db * db;  // cosNoAdvMorph.cpp:431-433
          const T C1a = (T)1 + p.dtd2 * La, C2a = (T)1 - p.dtd2 * La;             // :434-435
          const T C1b = (T)1 + p.dtd2 * Lb, C2b = (T)1 - p.dtd2 * Lb;
          const cplx<T> nn = p.scale * y[j];
          cplx<T> s, rn;
          if constexpr (MODE == X_STEP) {  // :524-525 with R = C1 psi - dt/2 N_past
            const cplx<T> r = cin[it][j];
            s = mk((r.x + (T)3 * p.dtd2 * nn.x) * recip(C2a), (r.y + (T)3 * p.dtd2 * nn.y) * recip(C2b));
            rn = mk(C1a * s.x - p.dtd2 * nn.x, C1b * s.y - p.dtd2 * nn.y);
            if (owner0 && kk[j] == 0 && qq == 0) {
              const double n0 = p.m_scale * nsum;
              const double ps0 = (p.mean_state[0] + 3.0 * p.m_dtd2 * n0) / p.m_c2;
              p.mean_state[0] = p.m_c1 * ps0 - p.m_dtd2 * n0;
              p.mean_state[1] = ps0;
              s.x = (T)ps0;
              rn.x = (T)p.mean_state[0];
            }
          } else {  // X_BOOT: N_past = N(psi0), cosNoAdvMorph.cpp:436-442,490
            s = cin[it][j];
            rn = mk(C1a * s.x - p.dtd2 * nn.x, C1b * s.y - p.dtd2 * nn.y);
            if (owner0 && kk[j] == 0 && qq == 0) {
              const double ps0 = p.mean_state[1];
              p.mean_state[0] = p.m_c1 * ps0 - p.m_dtd2 * p.m_scale * nsum;
              s.x = (T)ps0;
              rn.x = (T)p.mean_state[0];
            }
          }
          st_off(rb, o, rn);
          ps[j] = s;
        }
      }
      if constexpr (MODE != X_FWD_SCALED) {
        // DCT-III inputs (the partner of k = 0 is X_L = Y_L = 0; k = L/2 is its own partner)
        const cplx<T> pk = ps[0], pm = two ? ps[1] : (k == 0 ? mk<T>(0, 0) : ps[0]);
        F[k * FP + qq] = dct3_in(pk, pm, QW[k]);
        if (two) F[km * FP + qq] = dct3_in(ps[1], pk, QW[km]);
      }
    }
    if constexpr (MODE == X_FWD_SCALED) return;
    lds_barrier();
    Tile<T, L, E, FP>::template load<inv_first<L>()>(v, F, q, tc);
  }
  InvFFT<T, L, FP, 0, true>::run(v, F, q, tc, twl);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int d = mk_src(Tile<T, L, E, FP>::template pos_in<inv_last<L>()>(tc, e), L);
    st_off(dst, (uint32_t)d * rowb + q * CB, v[e]);
  }
}

// ---------------------------------------------------------------------------
// z pass.  Tile: columns f0 .. f0+NC-1 of the (x,y) plane (pairs of adjacent columns), rows z.
// ZIn / ZOut as in the periodic pass; Z_OUT_REAL writes [x][y][z] (dense).  Whole tiles.
template <typename T, int L, int NC, int IN, int OUT>
__global__ __launch_bounds__((NC / 2) * (L / Plan<L>::E)) void k_dct_zpass(const T* in, T* out,
                                                                           const double* __restrict__ real_in,
                                                                           double* __restrict__ real_out, int op,
                                                                           PParams<T> p,
                                                                           const cplx<T>* __restrict__ qtw) {
  constexpr int E = Plan<L>::E, TPC = L / E, NPR = NC / 2, NT = NPR * TPC, FP = NPR + 1;
  constexpr uint32_t CB = sizeof(cplx<T>);
  __shared__ cplx<T> F[L * FP];
  __shared__ double red[64];
  const int tid = threadIdx.x, q = tid % NPR, tc = tid / NPR;
  __shared__ cplx<T> TW[L];
  const cplx<T>* twl = tw_lds<L, NT>(TW, p.twz, tid);
  // LDS copy of qtw for the DCT-II outputs of the split (read after barriers): a global
  // load there would wait behind the previous row's store in the wave's one counter
  constexpr bool kQW = (OUT & Z_OUT_SPEC) != 0 || IN == Z_IN_SPEC;
  __shared__ cplx<T> QW[kQW ? L : 1];
  if constexpr (kQW) tw_lds<L, NT>(QW, qtw, tid);
  const int64_t f0 = (int64_t)blockIdx.x * NC;
  const uint32_t rowb = (uint32_t)p.nxy * sizeof(T);
  static_assert(fwd_first<L>() == inv_last<L>(), "plan symmetry");
  constexpr int RP = fwd_first<L>();
  cplx<T> v[E];  // (psi_a, psi_b) at z = src(m), m = pos_in<RP>
  if constexpr (IN == Z_IN_SPEC) {
    // each row is loaded once: its own rows k stay in registers, the partner rows L-k come from
    // the tile (one global load per entry instead of two)
    const cplx<T>* src = reinterpret_cast<const cplx<T>*>(in + f0);
#pragma unroll
    for (int e = 0; e < E; ++e)
      v[e] = ld_off(src, (uint32_t)Tile<T, L, E, FP>::template pos_in<inv_first<L>()>(tc, e) * rowb + q * CB);
    Tile<T, L, E, FP>::template store<inv_first<L>()>(v, F, q, tc);
    __syncthreads();  // also orders the QW fill before its reads
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = Tile<T, L, E, FP>::template pos_in<inv_first<L>()>(tc, e);
      const cplx<T> pn = (k == 0) ? mk<T>(0, 0) : F[(L - k) * FP + q];
      v[e] = dct3_in(v[e], pn, QW[k]);
    }
    InvFFT<T, L, FP>::run(v, F, q, tc, twl);
  } else {
    // a lane reads one z of both columns of a pair (two coalesced rows) and writes the pair as
    // one complex: consecutive lanes hit consecutive LDS rows, without bank conflicts
    for (int i = tid; i < NPR * L; i += NT) {
      const int qq = i / L, z = i - qq * L;
      F[z * FP + qq] = mk((T)real_in[(f0 + 2 * qq) * L + z], (T)real_in[(f0 + 2 * qq + 1) * L + z]);
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = F[mk_src(Tile<T, L, E, FP>::template pos_in<RP>(tc, e), L) * FP + q];
  }
  if constexpr ((OUT & Z_OUT_REAL) != 0) {
    // staged through the tile so that each column's L values are written as one contiguous row
    // (lanes of a wave write consecutive z of a column instead of z values L apart)
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) F[mk_src(Tile<T, L, E, FP>::template pos_in<RP>(tc, e), L) * FP + q] = v[e];
    __syncthreads();
    for (int i = tid; i < NPR * L; i += NT) {
      const int qq = i / L, z = i - qq * L;
      const cplx<T> w = F[z * FP + qq];
      real_out[(f0 + 2 * qq) * L + z] = (double)w.x;
      real_out[(f0 + 2 * qq + 1) * L + z] = (double)w.y;
    }
    if constexpr ((OUT & Z_OUT_SPEC) != 0) __syncthreads();  // the forward transform reuses the tile
  }
  if constexpr ((OUT & Z_OUT_SPEC) != 0) {
    if (op == Z_OP_NL) {  // cosNoAdvMorph.cpp:502-503
      if (p.nl_model == 1) {
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = mk(nonlin3(p, v[e].x), nonlin3(p, v[e].y));
      } else {
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = mk(nonlin5(p, v[e].x), nonlin5(p, v[e].y));
      }
    }
    if (p.partial != nullptr) wave_partial<E, NT>(v, red, tid);  // stored after the split
    FwdFFT<T, L, FP>::run(v, F, q, tc, twl);
    __syncthreads();
    Tile<T, L, E, FP>::template store<fwd_last<L>()>(v, F, q, tc);
    __syncthreads();
    cplx<T>* dst = reinterpret_cast<cplx<T>*>(out + f0);
    for (int i = tid; i < L * NPR; i += NT) {
      const int k = i / NPR, qq = i - k * NPR;
      st_off(dst, (uint32_t)k * rowb + qq * CB, dct2_out(F[k * FP + qq], F[((L - k) % L) * FP + qq], QW[k]));
    }
    if (p.partial != nullptr) store_partial<NT>(red, p.partial + blockIdx.x, tid);
  }
}

}  // namespace smx
