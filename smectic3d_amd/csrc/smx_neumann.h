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
#pragma once
#include "smx_periodic.h"

namespace smx {

__device__ __forceinline__ int mk_src(int m, int n) { return (m < (n + 1) / 2) ? 2 * m : 2 * (n - 1 - m) + 1; }

// DCT-II (REDFT10) of the NC columns of a real LDS tile [L][RP] in place.
// Threads: pair q in [0, NC/2), tc in [0, TPC).  F: complex exchange tile [L][FP].
template <typename T, int L, int NC, int RP, int FP>
__device__ __forceinline__ void dct2_tile(T* tile, cplx<T>* F, int q, int tc, int tid, int nt,
                                          const cplx<T>* __restrict__ tw, const cplx<T>* __restrict__ qtw) {
  constexpr int E = Plan<L>::E, NPR = NC / 2;
  cplx<T> v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int m = Tile<T, L, E, FP>::template pos_in<fwd_first<L>()>(tc, e);
    const int s = mk_src(m, L);
    v[e] = mk(tile[s * RP + q], tile[s * RP + q + NPR]);
  }
  FwdFFT<T, L, FP>::run(v, F, q, tc, tw);
  __syncthreads();
  Tile<T, L, E, FP>::template store<fwd_last<L>()>(v, F, q, tc);
  __syncthreads();
  for (int i = tid; i < L * NPR; i += nt) {
    const int k = i / NPR, qq = i - k * NPR;
    const cplx<T> Z = F[k * FP + qq], Zm = F[((L - k) % L) * FP + qq];
    const cplx<T> w = qtw[k];  // e^{-i pi k / 2n}
    tile[k * RP + qq] = w.x * (Z.x + Zm.x) - w.y * (Z.y - Zm.y);
    tile[k * RP + qq + NPR] = w.x * (Z.y + Zm.y) + w.y * (Z.x - Zm.x);
  }
  __syncthreads();
}

// DCT-III (REDFT01) of the NC columns of a real LDS tile [L][RP] in place.
template <typename T, int L, int NC, int RP, int FP>
__device__ __forceinline__ void dct3_tile(T* tile, cplx<T>* F, int q, int tc,
                                          const cplx<T>* __restrict__ tw, const cplx<T>* __restrict__ qtw) {
  constexpr int E = Plan<L>::E, NPR = NC / 2;
  cplx<T> v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int k = Tile<T, L, E, FP>::template pos_in<inv_first<L>()>(tc, e);
    const T X = tile[k * RP + q], Y = tile[k * RP + q + NPR];
    const T Xn = (k == 0) ? (T)0 : tile[(L - k) * RP + q];
    const T Yn = (k == 0) ? (T)0 : tile[(L - k) * RP + q + NPR];
    const cplx<T> w = qtw[k];  // conj gives e^{+i pi k / 2n}
    const T ux = X + Yn, uy = Y - Xn;
    v[e] = mk(w.x * ux + w.y * uy, w.x * uy - w.y * ux);
  }
  InvFFT<T, L, FP>::run(v, F, q, tc, tw);
  __syncthreads();
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int m = Tile<T, L, E, FP>::template pos_in<inv_last<L>()>(tc, e);
    const int d = mk_src(m, L);
    tile[d * RP + q] = v[e].x;
    tile[d * RP + q + NPR] = v[e].y;
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// y pass: contiguous real pencils of length L = ny; pencils g = kz*nx + x.  DIR -1: DCT-II, +1: DCT-III.
template <typename T, int L, int DIR, int NC>
__global__ __launch_bounds__((NC / 2) * (L / Plan<L>::E)) void k_dct_ypass(const T* in, T* out, int64_t npencils,
                                                                           const cplx<T>* __restrict__ tw,
                                                                           const cplx<T>* __restrict__ qtw) {
  constexpr int E = Plan<L>::E, TPC = L / E, NPR = NC / 2, NT = NPR * TPC, RP = NC + 1, FP = NPR + 1;
  __shared__ T tile[L * RP];
  __shared__ cplx<T> F[L * FP];
  const int tid = threadIdx.x, q = tid % NPR, tc = tid / NPR;
  const int64_t p0 = (int64_t)blockIdx.x * NC;
  const int nvalid = (int)((npencils - p0) < NC ? (npencils - p0) : NC);
  for (int i = tid; i < NC * L; i += NT) {
    const int c = i / L, pos = i - c * L;
    tile[pos * RP + c] = (c < nvalid) ? in[p0 * L + i] : (T)0;
  }
  __syncthreads();
  if constexpr (DIR < 0) dct2_tile<T, L, NC, RP, FP>(tile, F, q, tc, tid, NT, tw, qtw);
  else dct3_tile<T, L, NC, RP, FP>(tile, F, q, tc, tw, qtw);
  for (int i = tid; i < NC * L; i += NT) {
    const int c = i / L, pos = i - c * L;
    if (c < nvalid) out[p0 * L + i] = tile[pos * RP + c];
  }
}

// ---------------------------------------------------------------------------
// x pass.  Tile: plane kz = blockIdx.y, columns y0 .. y0+NC-1, rows x.
template <typename T, int L, int NC, int MODE>
__global__ __launch_bounds__((NC / 2) * (L / Plan<L>::E)) void k_dct_xpass(const T* in, T* out, T* __restrict__ rbuf,
                                                                           const T* __restrict__ psi0, PParams<T> p,
                                                                           const cplx<T>* __restrict__ qtw) {
  constexpr int E = Plan<L>::E, TPC = L / E, NPR = NC / 2, NT = NPR * TPC, RP = NC + 1, FP = NPR + 1;
  __shared__ T tile[L * RP];
  __shared__ cplx<T> F[L * FP];
  __shared__ double red[64];
  const int tid = threadIdx.x, q = tid % NPR, tc = tid / NPR;
  const int kz = blockIdx.y, y0 = blockIdx.x * NC;
  const int64_t plane = (int64_t)kz * p.nxy;
  for (int i = tid; i < NC * L; i += NT) {
    const int pos = i / NC, c = i - pos * NC;
    tile[pos * RP + c] = in[plane + (int64_t)pos * p.ny + y0 + c];
  }
  double nsum = 0.0;
  const bool owner0 = (p.partial != nullptr) && kz == 0 && y0 == 0;
  if (owner0) {
    double a = 0.0;
    for (int i = tid; i < p.npartial; i += NT) a += p.partial[i];
    nsum = block_sum(a, red);
  }
  __syncthreads();
  if constexpr (MODE != X_INV) {
    dct2_tile<T, L, NC, RP, FP>(tile, F, q, tc, tid, NT, p.twx, qtw);
    const T kzv = p.kz[kz];
    for (int i = tid; i < NC * L; i += NT) {
      const int kxi = i / NC, c = i - kxi * NC;
      const int64_t idx = plane + (int64_t)kxi * p.ny + y0 + c;
      T v = tile[kxi * RP + c];
      if constexpr (MODE == X_FWD_SCALED) {
        v = p.scale * v;
        if (owner0 && i == 0) {
          p.mean_state[1] = p.m_scale * nsum;
          v = (T)p.mean_state[1];
        }
      } else {
        const T kxv = p.kx[kxi], kyv = p.ky[y0 + c];
        const T k2 = kxv * kxv + kyv * kyv + kzv * kzv;
        const T d = k2 - p.q02;
        const T aLin = p.eps - p.alpha * d * d;  // cosNoAdvMorph.cpp:431-433
        const T C1 = (T)1 + p.dtd2 * aLin, C2 = (T)1 - p.dtd2 * aLin;  // :434-435
        const T nn = p.scale * v;
        T ps, rn;
        if constexpr (MODE == X_STEP) {  // :524-525 with R = C1 psi - dt/2 N_past
          ps = (rbuf[idx] + (T)3 * p.dtd2 * nn) / C2;
          rn = C1 * ps - p.dtd2 * nn;
          if (owner0 && i == 0) {
            const double n0 = p.m_scale * nsum;
            const double ps0 = (p.mean_state[0] + 3.0 * p.m_dtd2 * n0) / p.m_c2;
            p.mean_state[0] = p.m_c1 * ps0 - p.m_dtd2 * n0;
            p.mean_state[1] = ps0;
            ps = (T)ps0;
            rn = (T)p.mean_state[0];
          }
        } else {  // X_BOOT: N_past = N(psi0), cosNoAdvMorph.cpp:436-442,490
          ps = psi0[idx];
          rn = C1 * ps - p.dtd2 * nn;
          if (owner0 && i == 0) {
            const double ps0 = p.mean_state[1];
            p.mean_state[0] = p.m_c1 * ps0 - p.m_dtd2 * p.m_scale * nsum;
            ps = (T)ps0;
            rn = (T)p.mean_state[0];
          }
        }
        rbuf[idx] = rn;
        v = ps;
      }
      tile[kxi * RP + c] = v;
    }
    __syncthreads();
  }
  if constexpr (MODE != X_FWD_SCALED) dct3_tile<T, L, NC, RP, FP>(tile, F, q, tc, p.twx, qtw);
  for (int i = tid; i < NC * L; i += NT) {
    const int pos = i / NC, c = i - pos * NC;
    out[plane + (int64_t)pos * p.ny + y0 + c] = tile[pos * RP + c];
  }
}

// ---------------------------------------------------------------------------
// z pass.  Tile: columns f0 .. f0+NC-1 of the (x,y) plane, rows z.
// ZIn / ZOut as in the periodic pass; Z_OUT_REAL writes [x][y][z] (dense).
template <typename T, int L, int NC, int IN, int OUT>
__global__ __launch_bounds__((NC / 2) * (L / Plan<L>::E)) void k_dct_zpass(const T* in, T* out,
                                                                           const double* __restrict__ real_in,
                                                                           double* __restrict__ real_out, int op,
                                                                           PParams<T> p,
                                                                           const cplx<T>* __restrict__ qtw) {
  constexpr int E = Plan<L>::E, TPC = L / E, NPR = NC / 2, NT = NPR * TPC, RP = NC + 1, FP = NPR + 1;
  __shared__ T tile[L * RP];
  __shared__ cplx<T> F[L * FP];
  __shared__ double red[64];
  const int tid = threadIdx.x, q = tid % NPR, tc = tid / NPR;
  const int64_t f0 = (int64_t)blockIdx.x * NC;
  const int nvalid = (int)((p.nxy - f0) < NC ? (p.nxy - f0) : NC);
  if constexpr (IN == Z_IN_SPEC) {
    for (int i = tid; i < NC * L; i += NT) {
      const int k = i / NC, c = i - k * NC;
      tile[k * RP + c] = (c < nvalid) ? in[(int64_t)k * p.nxy + f0 + c] : (T)0;
    }
    __syncthreads();
    dct3_tile<T, L, NC, RP, FP>(tile, F, q, tc, p.twz, qtw);
  } else {
    for (int i = tid; i < NC * L; i += NT) {
      const int c = i / L, n = i - c * L;
      tile[n * RP + c] = (c < nvalid) ? (T)real_in[(f0 + c) * L + n] : (T)0;
    }
    __syncthreads();
  }
  if constexpr ((OUT & Z_OUT_REAL) != 0) {
    for (int i = tid; i < NC * L; i += NT) {
      const int c = i / L, n = i - c * L;
      if (c < nvalid) real_out[(f0 + c) * L + n] = (double)tile[n * RP + c];
    }
  }
  if constexpr ((OUT & Z_OUT_SPEC) != 0) {
    double acc = 0.0;
    for (int i = tid; i < NC * L; i += NT) {
      const int n = i / NC, c = i - n * NC;
      T v = tile[n * RP + c];
      if (op == Z_OP_NL) v = nonlin(p, v);  // cosNoAdvMorph.cpp:502-503
      tile[n * RP + c] = v;
      acc += (double)v;
    }
    if (p.partial != nullptr) {
      const double s = block_sum(acc, red);
      if (tid == 0) p.partial[blockIdx.x] = s;
    }
    __syncthreads();
    dct2_tile<T, L, NC, RP, FP>(tile, F, q, tc, tid, NT, p.twz, qtw);
    for (int i = tid; i < NC * L; i += NT) {
      const int k = i / NC, c = i - k * NC;
      if (c < nvalid) out[(int64_t)k * p.nxy + f0 + c] = tile[k * RP + c];
    }
  }
}

}  // namespace smx
