// smx_surf.h -- surfTrack's kernels beyond the step's passes (all FP64 arithmetic):
//   k_dct_ypoint / k_ypoint   a field's value at one z index per column, straight from an x stage
//                             (x inverted, y and z spectral): the inverse y transform of every
//                             pencil in LDS, then each output entry's term of the inverse z
//                             transform at its column's index, summed over kz in registers; the
//                             y-inverted field is never written (one read, nxy doubles per split)
//   k_surf_point              the same sum from a y-inverted stage
//   k_dct_zinv_scan / k_zinv_scan   psi's z inverse, the column scan and the k_z derivative at
//                             each column's index in one pass
//   surf_dct_xfwd_f32, surf_dct_yinv_f32, surf_dct_xmul2   Neumann x and y passes reading an
//                             FP32 stage (widened on load) or taking two x multipliers at once
#pragma once
#include <algorithm>

#include "smx_diag.h"
#include "smx_launch.h"

namespace smx {

// kz splits of one evaluation: block (x, s) sums kz units [s U / ns, (s+1) U / ns) (U: pencil
// pairs for the Neumann kernel, spectral planes for the periodic one) and writes its part of
// every column of that x to out[s nxy + c]; the tail kernels add the ns parts in split order.
// Enough splits for about 1024 blocks (more splits repeat each block's set-up for fewer rounds
// of pencils), at most kSurfMaxSplit and at most U.
constexpr int kSurfMaxSplit = 16;
inline int surf_nsplit(bool neu, int nx, int nz) {
  const int units = neu ? (nz + 1) / 2 : nz / 2 + 1;
  return std::max(1, std::min({kSurfMaxSplit, units, (1024 + nx - 1) / nx}));
}

// Trigonometric tables of the z evaluation, one per grid (4 nz doubles): Neumann
// cos(pi j / 2nz), j < 4nz; periodic cos(2 pi j / nz) then sin(2 pi j / nz), j < nz.
__global__ void k_surf_tables(double* __restrict__ t, int nz, int neu) {
  const int n = neu ? 4 * nz : nz;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    if (neu) {
      t[j] = cospi((double)j / (2.0 * nz));
    } else {
      double sv, cv;
      sincospi(2.0 * (double)j / nz, &sv, &cv);
      t[j] = cv;
      t[nz + j] = sv;
    }
  }
}

// Value of a field at z index ke[c] of every column c, from its z spectrum (layout [kz][x][y],
// x and y already inverted): the inverse z transform of one entry per column in place of the
// whole field, the same sum as k_dct_ypoint / k_ypoint below with the y transform already done.
// Block: 64 consecutive columns, one per lane (coalesced rows), and 4 waves that split the kz
// range; grid ceil(nxy / 64).  out: ns parts of nxy (the sum in part 0, zeros in the others).
template <bool NEU>
__global__ __launch_bounds__(256) void k_surf_point(const double* __restrict__ spec, const int* __restrict__ ke,
                                                    int64_t nxy, int nz, const double* __restrict__ kz, int kop,
                                                    const double* __restrict__ trig, int ns,
                                                    double* __restrict__ out) {
  constexpr int kMaxN = 512;
  __shared__ double ct[NEU ? 4 * kMaxN : kMaxN];
  __shared__ double st[NEU ? 1 : kMaxN];
  __shared__ double part[3][64];
  __shared__ double kzs[kMaxN];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int period = NEU ? 4 * nz : nz;
  for (int j = threadIdx.x; j < period; j += 256) {
    ct[j] = trig[j];
    if constexpr (!NEU) st[j] = trig[nz + j];
  }
  if (kop >= 0)
    for (int j = threadIdx.x; j < nz; j += 256) kzs[j] = kz[j];
  __syncthreads();
  const int64_t c = (int64_t)blockIdx.x * 64 + lane;
  const int n = (c < nxy) ? ke[c] : -1;
  const int nk = NEU ? nz : nz / 2 + 1;
  const int k0 = w * nk / 4, k1 = (w + 1) * nk / 4;
  double acc = 0.0;
  if (n >= 0) {
    const int step = NEU ? 2 * n + 1 : n;
    int j = (int)(((uint32_t)k0 * (uint32_t)step) % (uint32_t)period);
    for (int k = k0; k < k1; ++k) {
      if constexpr (NEU) {
        double x = spec[(int64_t)k * nxy + c];
        if (kop == KM_DZZ) x *= -kzs[k] * kzs[k];
        acc += ((k == 0) ? 1.0 : 2.0) * x * ct[j];
      } else {
        const double* e = spec + 2 * ((int64_t)k * nxy + c);
        double re = e[0], im = e[1];
        if (kop == KM_IKZ) {
          const double t = re;
          re = -kzs[k] * im;
          im = kzs[k] * t;
        } else if (kop == KM_DZZ) {
          re *= -kzs[k] * kzs[k];
          im *= -kzs[k] * kzs[k];
        }
        acc += ((k == 0 || 2 * k == nz) ? 1.0 : 2.0) * (re * ct[j] - im * st[j]);
      }
      j += step;
      if (j >= period) j -= period;
    }
  }
  if (w > 0) part[w - 1][lane] = acc;
  __syncthreads();
  if (w == 0 && c < nxy) {
    out[c] = acc + part[0][lane] + part[1][lane] + part[2][lane];
    for (int i = 1; i < ns; ++i) out[i * nxy + c] = 0.0;
  }
}

// Plan variant of the evaluation kernels: E = 8 where the length has that variant (fewer
// registers beside the per-entry accumulators), else the length's plan.
template <int L> constexpr int ypoint_v() { return HasPlanV<L, 1>::value ? 1 : 0; }
// Pencils (periodic) or pencil pairs (Neumann) per FFT round: a 256-thread block, at most 32.
template <int L> constexpr int ypoint_groups() {
  constexpr int TPC = L / PlanV<L, ypoint_v<L>()>::E;
  return (256 / TPC) < 32 ? (256 / TPC) : 32;
}

// Neumann.  Field value at z index ke[c] of every column c = x ny + y, from the field's x stage
// `in` (real, layout [kz][x][ky], FP64 or the FP32 stage itself, widened on load): the DCT-III
// along y of each pencil (kz, x), as k_dct_ypass inverts it, then
//   y_n = w_0 x_0 + 2 sum_{k=1}^{nz-1} w_k x_k cos(pi k (2n+1) / 2nz)
// (REDFT01, cosNoAdvMorphCurv.cpp:750) with w_k the k_z multiplier (zop KM_DZZ: -k_z^2, else 1).
// One FFT carries pencils kz = 2p and 2p + 1 of one x, so both halves of an output entry add
// to the same column's sum.  kq: -k_y^2 on the input (psi_yy) when set.  The cosines come from
// an LDS copy of k_surf_tables' table, indexed by k(2n+1) mod 4nz.  Each round's rows are
// loaded during the previous round's transform (the exchanges order LDS only, lds_barrier).
// Columns with ke < 0 get an unused value.  Grid (nx, ns); dynamic LDS 5 nz doubles.
template <int L, int V, int NPR, typename TI>
__global__ __launch_bounds__(NPR*(L / PlanV<L, V>::E)) void k_dct_ypoint(
    const TI* __restrict__ in, const int* __restrict__ ke, int nx, int nz, const cplx<double>* __restrict__ tw,
    const cplx<double>* __restrict__ qtw, const double* __restrict__ kq, const double* __restrict__ kzt, int zop,
    const double* __restrict__ trig, double* __restrict__ out) {
  using T = double;
  constexpr int E = PlanV<L, V>::E, TPC = L / E, NT = NPR * TPC;
  using AD = ColAddr<L>;
  using TL = TileA<T, L, E, AD>;
  constexpr int RF = inv_first<L, V>(), RL = inv_last<L, V>();
  __shared__ cplx<T> S[NPR * AD::LP];
  __shared__ cplx<T> TW[L];
  __shared__ cplx<T> QW[L];
  __shared__ T KQ[L];
  extern __shared__ double dyn[];
  const int P4 = 4 * nz;
  double* CT = dyn;        // cos(pi j / 2nz), j < 4nz
  double* WZ = dyn + P4;   // weight of each kz: (k == 0 ? 1 : 2) times the k_z multiplier
  const int tid = threadIdx.x, tc0 = tid % TPC, q0 = tid / TPC;
  const cplx<T>* twl = tw_lds<L, NT>(TW, tw, tid);
  tw_lds<L, NT>(QW, qtw, tid);
  if (kq != nullptr)
    for (int i = tid; i < L; i += NT) KQ[i] = kq[i];
  for (int j = tid; j < P4; j += NT) CT[j] = trig[j];
  for (int k = tid; k < nz; k += NT) WZ[k] = ((k == 0) ? 1.0 : 2.0) * ((zop == KM_DZZ) ? -kzt[k] * kzt[k] : 1.0);
  const int x = blockIdx.x, ns = gridDim.y, s = blockIdx.y;
  const int npair = (nz + 1) / 2;
  const int p0 = (int)((int64_t)s * npair / ns), p1 = (int)((int64_t)(s + 1) * npair / ns);
  // each entry's column: its step 2n+1 through the cosine table, the table index of the pair's
  // first pencil, and the index advance per round (2 NPR pencils); all below 2^31 (nz <= 512)
  int stp[E], jj[E], dj[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int n = ke[(int64_t)x * L + mk_src(TL::template pos_in<RL>(tc0, e), L)];
    stp[e] = n >= 0 ? 2 * n + 1 : 0;
    jj[e] = (int)(((uint32_t)(2 * (p0 + q0)) * (uint32_t)stp[e]) % (uint32_t)P4);
    dj[e] = (int)(((uint32_t)(2 * NPR) * (uint32_t)stp[e]) % (uint32_t)P4);
  }
  double acc[E];
#pragma unroll
  for (int e = 0; e < E; ++e) acc[e] = 0.0;
  const int rounds = (p1 - p0 + NPR - 1) / NPR;
  // rows kept in the input's type until they are used, so that a widening conversion does not
  // wait for a prefetch as soon as it is issued
  auto load = [&](cplx<TI>* v, int it, int tc, int q) {
    const int pr = p0 + it * NPR + q;
    const bool va = pr < p1, vb = va && 2 * pr + 1 < nz;
    const TI* a = in + ((int64_t)(2 * pr) * nx + x) * L;
    const TI* b = a + (int64_t)nx * L;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = TL::template pos_in<RF>(tc, e);
      v[e].x = va ? a[k] : (TI)0;
      v[e].y = vb ? b[k] : (TI)0;
    }
  };
  cplx<TI> nv[E];
  load(nv, 0, tc0, q0);
  __syncthreads();  // the LDS tables
  for (int it = 0; it < rounds; ++it) {
    // tc and q laundered per round: index arithmetic recomputed in the loop rather than hoisted
    // out of it into registers held across every round
    int tc = tc0, q = q0;
    asm volatile("" : "+v"(tc), "+v"(q));
    const int pr = p0 + it * NPR + q;
    const bool va = pr < p1, vb = va && 2 * pr + 1 < nz;
    cplx<T> v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = mk((T)nv[e].x, (T)nv[e].y);
    if (it + 1 < rounds) load(nv, it + 1, tc, q);
    if (kq != nullptr) {
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const T k = KQ[TL::template pos_in<RF>(tc, e)];
        v[e] = (-k * k) * v[e];
      }
    }
    lds_barrier();  // the previous round's last tile reads are done
    TL::template store<RF>(v, S, q, tc);
    lds_barrier();
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = TL::template pos_in<RF>(tc, e);
      v[e] = dct3_par<DP_COS>(v[e], S[AD::at((L - k) % L, q)], k, L, (const T*)nullptr, QW[k]);
    }
    InvFFTA<T, L, AD, V, true>::run(v, S, q, tc, twl);
    if (va) {
      const double wa = WZ[2 * pr], wb = vb ? WZ[2 * pr + 1] : 0.0;
#pragma unroll
      for (int e = 0; e < E; ++e) {
        int jb = jj[e] + stp[e];
        if (jb >= P4) jb -= P4;
        acc[e] += wa * v[e].x * CT[jj[e]] + wb * v[e].y * CT[jb];
      }
    }
#pragma unroll
    for (int e = 0; e < E; ++e) {
      jj[e] += dj[e];
      if (jj[e] >= P4) jj[e] -= P4;
    }
  }
  // the NPR pair groups' sums of each entry, added in group order
  lds_barrier();
  double* R = reinterpret_cast<double*>(S);
#pragma unroll
  for (int e = 0; e < E; ++e) R[q0 * L + mk_src(TL::template pos_in<RL>(tc0, e), L)] = acc[e];
  lds_barrier();
  double* o = out + (int64_t)s * nx * L + (int64_t)x * L;
  for (int y = tid; y < L; y += NT) {
    double sum = 0.0;
    for (int g = 0; g < NPR; ++g) sum += R[g * L + y];
    o[y] = sum;
  }
}

// Periodic.  The same from a complex FP64 x stage (layout [kz][x][ky], kz < nz/2 + 1): the
// inverse FFT along y of each pencil (kz, x), as k_ypass inverts it, then the half-spectrum C2R
// term
//   y_n = sum_{k=0}^{nz/2} w_k Re(m_k x_k e^{2 pi i k n / nz}),  w_0 = w_{nz/2} = 1, else 2
// (surfTrack.cpp:611-626; the imaginary parts at k = 0 and nz/2 drop out as in an FFTW c2r) with
// m_k the k_z multiplier (zop KM_IKZ: i k_z, KM_DZZ: -k_z^2, else 1).  kq with kop: the k_y
// multiplier on the input (KM_IKY: i k_y, KM_DYY: -k_y^2), as in k_ypass.  Cosines and sines
// from an LDS copy of k_surf_tables' tables, indexed by kn mod nz; rows prefetched as in
// k_dct_ypoint.  Grid (nx, ns); dynamic LDS 2 nz + 2 (nz/2 + 1) doubles.
template <int L, int V, int NPC>
__global__ __launch_bounds__(NPC*(L / PlanV<L, V>::E)) void k_ypoint(
    const cplx<double>* __restrict__ in, const int* __restrict__ ke, int nx, int nz,
    const cplx<double>* __restrict__ tw, const double* __restrict__ kq, int kop, const double* __restrict__ kzt,
    int zop, const double* __restrict__ trig, double* __restrict__ out) {
  using T = double;
  constexpr int E = PlanV<L, V>::E, TPC = L / E, NT = NPC * TPC;
  using AD = ColAddr<L>;
  using TL = TileA<T, L, E, AD>;
  constexpr int RF = inv_first<L, V>(), RL = inv_last<L, V>();
  __shared__ cplx<T> S[NPC * AD::LP];
  __shared__ cplx<T> TW[L];
  __shared__ T KQ[L];
  extern __shared__ double dyn[];
  const int nzh = nz / 2 + 1;
  double* CT = dyn;        // cos(2 pi j / nz), j < nz
  double* ST = dyn + nz;   // sin(2 pi j / nz)
  double* WR = ST + nz;    // w_k m_k: real part
  double* WI = WR + nzh;   //          imaginary part
  const int tid = threadIdx.x, tc0 = tid % TPC, col0 = tid / TPC;
  const cplx<T>* twl = tw_lds<L, NT>(TW, tw, tid);
  if (kq != nullptr)
    for (int i = tid; i < L; i += NT) KQ[i] = kq[i];
  for (int j = tid; j < 2 * nz; j += NT) CT[j] = trig[j];  // cosines, then sines (ST = CT + nz)
  for (int k = tid; k < nzh; k += NT) {
    const double w = (k == 0 || 2 * k == nz) ? 1.0 : 2.0, kz = kzt[k];
    WR[k] = (zop == KM_DZZ) ? -w * kz * kz : (zop == KM_IKZ) ? 0.0 : w;
    WI[k] = (zop == KM_IKZ) ? w * kz : 0.0;
  }
  const int x = blockIdx.x, ns = gridDim.y, s = blockIdx.y;
  const int k0 = (int)((int64_t)s * nzh / ns), k1 = (int)((int64_t)(s + 1) * nzh / ns);
  int stp[E], jj[E], dj[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int n = ke[(int64_t)x * L + TL::template pos_in<RL>(tc0, e)];
    stp[e] = n >= 0 ? n : 0;
    jj[e] = (int)(((uint32_t)(k0 + col0) * (uint32_t)stp[e]) % (uint32_t)nz);
    dj[e] = (int)(((uint32_t)NPC * (uint32_t)stp[e]) % (uint32_t)nz);
  }
  double acc[E];
#pragma unroll
  for (int e = 0; e < E; ++e) acc[e] = 0.0;
  const int rounds = (k1 - k0 + NPC - 1) / NPC;
  auto load = [&](cplx<T>* v, int it, int tc, int col) {
    const int kz = k0 + it * NPC + col;
    const bool valid = kz < k1;
    const cplx<T>* src = in + ((int64_t)kz * nx + x) * L;
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = valid ? src[TL::template pos_in<RF>(tc, e)] : mk<T>(0, 0);
  };
  cplx<T> nv[E];
  load(nv, 0, tc0, col0);
  __syncthreads();  // the LDS tables
  for (int it = 0; it < rounds; ++it) {
    int tc = tc0, col = col0;  // laundered per round, as in k_dct_ypoint
    asm volatile("" : "+v"(tc), "+v"(col));
    const int kz = k0 + it * NPC + col;
    const bool valid = kz < k1;
    cplx<T> v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = nv[e];
    if (it + 1 < rounds) load(nv, it + 1, tc, col);
    if (kq != nullptr) {
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const T k = KQ[TL::template pos_in<RF>(tc, e)];
        v[e] = (kop == KM_DYY) ? (-k * k) * v[e] : mk(-k * v[e].y, k * v[e].x);
      }
    }
    lds_barrier();  // the previous round's last tile reads are done
    InvFFTA<T, L, AD, V, true>::run(v, S, col, tc, twl);
    if (valid) {
      const double wr = WR[kz], wi = WI[kz];
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const double A = wr * v[e].x - wi * v[e].y, B = wr * v[e].y + wi * v[e].x;
        acc[e] += A * CT[jj[e]] - B * ST[jj[e]];
      }
    }
#pragma unroll
    for (int e = 0; e < E; ++e) {
      jj[e] += dj[e];
      if (jj[e] >= nz) jj[e] -= nz;
    }
  }
  lds_barrier();
  double* R = reinterpret_cast<double*>(S);
#pragma unroll
  for (int e = 0; e < E; ++e) R[col0 * L + TL::template pos_in<RL>(tc0, e)] = acc[e];
  lds_barrier();
  double* o = out + (int64_t)s * nx * L + (int64_t)x * L;
  for (int y = tid; y < L; y += NT) {
    double sum = 0.0;
    for (int g = 0; g < NPC; ++g) sum += R[g * L + y];
    o[y] = sum;
  }
}

template <int L>
static hipError_t surf_ypoint_launch(bool neu, bool f32in, const void* in, const int* ke, int nx, int nz, int ns,
                                     const cplx<double>* tw, const cplx<double>* qtw, const double* kq, int kop,
                                     const double* kzt, int zop, const double* trig, double* out, hipStream_t st) {
  constexpr int V = ypoint_v<L>(), G = ypoint_groups<L>(), NT = G * (L / PlanV<L, V>::E);
  const dim3 grid((unsigned)nx, (unsigned)ns), block(NT);
  if (neu) {
    const size_t dyn = sizeof(double) * 5 * (size_t)nz;
    if (f32in)
      SMX_LAUNCH((k_dct_ypoint<L, V, G, float>), grid, block, dyn, st, (const float*)in, ke, nx, nz, tw, qtw, kq, kzt,
                 zop, trig, out);
    else
      SMX_LAUNCH((k_dct_ypoint<L, V, G, double>), grid, block, dyn, st, (const double*)in, ke, nx, nz, tw, qtw, kq,
                 kzt, zop, trig, out);
  } else {
    if (f32in) return hipErrorInvalidValue;
    const size_t dyn = sizeof(double) * (2 * (size_t)nz + 2 * (size_t)(nz / 2 + 1));
    SMX_LAUNCH((k_ypoint<L, V, G>), grid, block, dyn, st, (const cplx<double>*)in, ke, nx, nz, tw, kq, kop, kzt, zop,
               trig, out);
  }
  return hipGetLastError();
}

// Field values at ke of every column into out[ns][nx*ny] (parts per kz split) from an x stage
// `in` of an nx x ny x nz grid: FP64, or (Neumann, f32in) the FP32 stage.  kq/kop: k_y
// multiplier (Neumann: KM_DYY only); zop: k_z multiplier (Neumann: KM_DZZ only); -1: none.
// tw, qtw: the y-axis twiddles (qtw Neumann); trig: k_surf_tables' table of this nz.
inline hipError_t surf_ypoint(bool neu, bool f32in, const void* in, const int* ke, int nx, int ny, int nz, int ns,
                              const cplx<double>* tw, const cplx<double>* qtw, const double* kq, int kop,
                              const double* kzt, int zop, const double* trig, double* out, hipStream_t st) {
  if (ns < 1 || ns > kSurfMaxSplit) return hipErrorInvalidValue;
  if ((kq == nullptr) != (kop < 0)) return hipErrorInvalidValue;
  if (neu ? (kop >= 0 && kop != KM_DYY) || (zop >= 0 && zop != KM_DZZ)
          : (kop >= 0 && kop != KM_IKY && kop != KM_DYY) || (zop >= 0 && zop != KM_IKZ && zop != KM_DZZ))
    return hipErrorInvalidValue;
#define SMX_CASE(L)                                                                                        \
  if (ny == L)                                                                                             \
    return surf_ypoint_launch<L>(neu, f32in, in, ke, nx, nz, ns, tw, qtw, kq, kop, kzt, zop, trig, out, st);
  SMX_FOR_EACH_LENGTH(SMX_CASE)
#undef SMX_CASE
  return hipErrorInvalidValue;
}

// Neumann: psi, its crossings and psi_zz at them in one pass over the y-inverted stage `in`
// (layout [kz][x][y], FP64).  The z inverse of each column as k_dct_zpass (Z_IN_SPEC ->
// Z_OUT_REAL) writes psi (dense [x][y][z]); the column scan of k_surf_scan<false> (from the top,
// armed above `arm`, first psi < 0) then runs on the block's tile, one wave per column; and psi_zz
// at each column's evaluation index is summed from the spectral rows the thread loaded, kept in
// registers (the sum of k_surf_point<true> with KM_DZZ): each thread adds its own rows, and the
// TPC partial sums of a column pair are added in order.  dzz: ns parts of nxy (the sum in part 0,
// zeros in the others).  The block's threads form whole waves (L >= 96 with 16 columns).
template <int L, int NC>
__global__ __launch_bounds__((NC / 2) * (L / Plan<L>::E)) void k_dct_zinv_scan(
    const double* __restrict__ in, double* __restrict__ psi, int64_t nxy, const cplx<double>* __restrict__ twz,
    const cplx<double>* __restrict__ qtw, const double* __restrict__ kzt, const double* __restrict__ trig,
    double arm, int* __restrict__ kc_out, int* __restrict__ ke_out, int ns, double* __restrict__ dzz) {
  using T = double;
  constexpr int E = Plan<L>::E, TPC = L / E, NPR = NC / 2, NT = NPR * TPC, FP = NPR + 1;
  static_assert(NT % 64 == 0, "the scan takes one whole wave per column");
  static_assert(2 * NT <= 2 * L * FP, "the partial sums fit in the tile");
  constexpr uint32_t CB = sizeof(cplx<T>);
  using TL = Tile<T, L, E, FP>;
  constexpr int RP = fwd_first<L>();
  static_assert(RP == inv_last<L>(), "plan symmetry");
  __shared__ cplx<T> F[L * FP];
  __shared__ cplx<T> TW[L];
  __shared__ cplx<T> QW[L];
  __shared__ int KE[NC];
  const int tid = threadIdx.x, q = tid % NPR, tc = tid / NPR;
  const cplx<T>* twl = tw_lds<L, NT>(TW, twz, tid);
  tw_lds<L, NT>(QW, qtw, tid);
  const int64_t f0 = (int64_t)blockIdx.x * NC;
  const uint32_t rowb = (uint32_t)nxy * sizeof(T);
  const cplx<T>* src = reinterpret_cast<const cplx<T>*>(in + f0);
  cplx<T> v[E], sp[E];  // sp: the loaded spectral rows (pair q at kz = pos_in<inv_first>)
#pragma unroll
  for (int e = 0; e < E; ++e) sp[e] = v[e] = ld_off(src, (uint32_t)TL::template pos_in<inv_first<L>()>(tc, e) * rowb + q * CB);
  TL::template store<inv_first<L>()>(v, F, q, tc);
  __syncthreads();
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int k = TL::template pos_in<inv_first<L>()>(tc, e);
    const cplx<T> pn = (k == 0) ? mk<T>(0, 0) : F[(L - k) * FP + q];
    v[e] = dct3_in(v[e], pn, QW[k]);
  }
  InvFFT<T, L, FP>::run(v, F, q, tc, twl);
  // psi rows through the tile, each column's L values one contiguous row
  __syncthreads();
#pragma unroll
  for (int e = 0; e < E; ++e) F[mk_src(TL::template pos_in<RP>(tc, e), L) * FP + q] = v[e];
  __syncthreads();
  for (int i = tid; i < NPR * L; i += NT) {
    const int qq = i / L, z = i - qq * L;
    const cplx<T> w = F[z * FP + qq];
    psi[(f0 + 2 * qq) * L + z] = w.x;
    psi[(f0 + 2 * qq + 1) * L + z] = w.y;
  }
  // the scan: wave wv takes columns wv, wv + NT/64, ...
  const int lane = tid & 63, wv = tid >> 6;
  for (int c = wv; c < NC; c += NT / 64) {
    const int qq = c >> 1;
    const bool hi = (c & 1) != 0;
    auto P = [&](int z) -> double {
      const cplx<T> w = F[z * FP + qq];
      return hi ? w.y : w.x;
    };
    bool armed = false;
    int kc = -1;
    for (int ch = 0; ch < (L + 63) / 64; ++ch) {
      const int k = L - 1 - ch * 64 - lane;
      const bool inr = k >= 0;
      const double val = inr ? P(k) : 0.0;
      const uint64_t am = __ballot(inr && val > arm);
      uint64_t neg = __ballot(inr && val < 0.0);
      if (kc < 0) {
        if (!armed && am) {
          neg &= lanes_from(__ffsll((unsigned long long)am) - 1);
          armed = true;
        }
        if (armed && neg) kc = L - 1 - ch * 64 - (__ffsll((unsigned long long)neg) - 1);
      }
    }
    if (lane == 0) {
      int ke = kc;
      if (kc >= 0 && kc + 1 < L && fabs(P(kc)) > fabs(P(kc + 1))) ke = kc + 1;
      kc_out[f0 + c] = kc;
      ke_out[f0 + c] = ke;
      KE[c] = ke;
    }
  }
  __syncthreads();  // KE, and every read of the tile is done
  // psi_zz at ke: y_n = sum_k w_k (-k_z^2) x_k cos(pi k (2n+1) / 2L), w_0 = 1, else 2
  const int na = KE[2 * q], nb = KE[2 * q + 1];
  const uint32_t sa = na >= 0 ? 2 * na + 1 : 0, sb = nb >= 0 ? 2 * nb + 1 : 0, P4 = 4 * L;
  double aa = 0.0, ab = 0.0;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int k = TL::template pos_in<inv_first<L>()>(tc, e);
    const double w = ((k == 0) ? 1.0 : 2.0) * (-kzt[k] * kzt[k]);
    aa += w * sp[e].x * trig[((uint32_t)k * sa) % P4];
    ab += w * sp[e].y * trig[((uint32_t)k * sb) % P4];
  }
  double* R = reinterpret_cast<double*>(F);
  R[2 * (tc * NPR + q)] = aa;
  R[2 * (tc * NPR + q) + 1] = ab;
  __syncthreads();
  if (tid < NC) {  // column f0 + tid: pair tid / 2, half tid % 2
    const int qq = tid >> 1, h = tid & 1;
    double sum = 0.0;
    for (int t = 0; t < TPC; ++t) sum += R[2 * (t * NPR + qq) + h];
    dzz[f0 + tid] = sum;
    for (int i = 1; i < ns; ++i) dzz[(int64_t)i * nxy + f0 + tid] = 0.0;
  }
}

// Launch of k_dct_zinv_scan over nxy columns (16 per block); hipErrorNotSupported for a length
// whose block is not whole waves (the caller then runs the z pass, k_surf_scan and k_surf_point).
inline hipError_t surf_dct_zinv_scan(const double* in, double* psi, int64_t nxy, int nz, const cplx<double>* twz,
                                     const cplx<double>* qtw, const double* kzt, const double* trig, double arm,
                                     int* kc, int* ke, int ns, double* dzz, hipStream_t st) {
  constexpr int NC = 16;
  if (nxy % NC != 0) return hipErrorInvalidValue;
#define SMX_CASE(L)                                                                                          \
  if (nz == L) {                                                                                             \
    if constexpr (((NC / 2) * (L / Plan<L>::E)) % 64 == 0) {                                                 \
      SMX_LAUNCH((k_dct_zinv_scan<L, NC>), dim3((unsigned)(nxy / NC)), dim3((NC / 2) * (L / Plan<L>::E)), 0, st, \
                 in, psi, nxy, twz, qtw, kzt, trig, arm, kc, ke, ns, dzz);                                   \
      return hipGetLastError();                                                                              \
    } else {                                                                                                 \
      return hipErrorNotSupported;                                                                           \
    }                                                                                                        \
  }
  SMX_FOR_EACH_LENGTH(SMX_CASE)
#undef SMX_CASE
  return hipErrorInvalidValue;
}

// Periodic: psi, its crossings and psi_z at them in one pass over the y-inverted stage `in`
// (complex, layout [kz][x][y], kz < nz/2 + 1).  The z inverse of each column pair as k_zpass
// (Z_IN_SPEC -> Z_OUT_REAL: the Hermitian extension of the half spectrum, the imaginary parts of
// kz = 0 and nz/2 dropped) writes psi (dense [x][y][z]); the column scan of k_surf_scan<true> (from
// the bottom, armed above `arm`, first psi < 0) runs on the block's tile, one wave per column; and
// psi_z at each column's evaluation index is summed from the raw rows the thread loaded, kept in
// registers: each row kz <= nz/2 is the "low" entry (n = kz) of exactly one thread of the pair,
//   y_n = sum_{k=0}^{nz/2} w_k Re(i k_z X_k e^{2 pi i k n / nz}),  w_0 = w_{nz/2} = 1, else 2
// (the sum of k_surf_point<false> with KM_IKZ).  dz: ns parts of nxy (the sum in part 0, zeros in
// the others).  The block's threads form whole waves.
template <int L, int NC, int V>
__global__ __launch_bounds__((NC / 2) * (L / PlanV<L, V>::E)) void k_zinv_scan(
    const cplx<double>* __restrict__ in, double* __restrict__ psi, int64_t nxy, const cplx<double>* __restrict__ twz,
    const double* __restrict__ kzt, const double* __restrict__ trig, double arm, int* __restrict__ kc_out,
    int* __restrict__ ke_out, int ns, double* __restrict__ dz) {
  using T = double;
  constexpr int E = PlanV<L, V>::E, TPC = L / E, NPR = NC / 2, NT = NPR * TPC, FP = NPR + 1;
  static_assert(NT % 64 == 0, "the scan takes one whole wave per column");
  static_assert(2 * NT <= 2 * L * FP, "the partial sums fit in the tile");
  constexpr uint32_t CB = sizeof(cplx<T>);
  using TL = Tile<T, L, E, FP>;
  constexpr int RP = fwd_first<L, V>();
  static_assert(RP == inv_last<L, V>(), "plan symmetry");
  __shared__ cplx<T> F[L * FP];
  __shared__ cplx<T> TW[L];
  __shared__ int KE[NC];
  const int tid = threadIdx.x, q = tid % NPR, tc = tid / NPR;
  const cplx<T>* twl = tw_lds<L, NT>(TW, twz, tid);
  const int64_t f0 = (int64_t)blockIdx.x * NC;
  const uint32_t rowb = (uint32_t)nxy * CB, qofs = (uint32_t)q * 2 * CB;
  const cplx2<T>* colp = reinterpret_cast<const cplx2<T>*>(in + f0);
  const uint32_t tcr = (uint32_t)tc * rowb + qofs;
  cplx2<T> ab[E];  // raw rows: entry e at n = tc + c_e reads row min(n, L - n)
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int c = TL::template pos_in<inv_first<L, V>()>(0, e);
    const int n = tc + c;
    const bool lo = (c + TPC - 1 <= L / 2) || (c <= L / 2 && n <= L / 2);  // n <= L/2
    const uint32_t off = lo ? tcr + (uint32_t)c * rowb : (uint32_t)(L - c) * rowb - tcr + 2 * qofs;
    ab[e] = ld_off(colp, off);
  }
  cplx<T> v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int n = tc + TL::template pos_in<inv_first<L, V>()>(0, e);
    cplx<T> A = ab[e].a, B = ab[e].b;
    if (n == 0 || n == L / 2) {
      A.y = 0;
      B.y = 0;
    }
    v[e] = (n <= L / 2) ? mk(A.x - B.y, A.y + B.x) : mk(A.x + B.y, B.x - A.y);
  }
  InvFFT<T, L, FP, V>::run(v, F, q, tc, twl);
  __syncthreads();
  TL::template store<RP>(v, F, q, tc);
  __syncthreads();
  for (int i = tid; i < NPR * L; i += NT) {
    const int qq = i / L, n = i - qq * L;
    const cplx<T> w = F[n * FP + qq];
    psi[(f0 + 2 * qq) * L + n] = w.x;
    psi[(f0 + 2 * qq + 1) * L + n] = w.y;
  }
  const int lane = tid & 63, wv = tid >> 6;
  for (int c = wv; c < NC; c += NT / 64) {
    const int qq = c >> 1;
    const bool hi = (c & 1) != 0;
    auto P = [&](int z) -> double {
      const cplx<T> w = F[z * FP + qq];
      return hi ? w.y : w.x;
    };
    bool armed = false;
    int kc = -1;
    for (int ch = 0; ch < (L + 63) / 64; ++ch) {
      const int k = ch * 64 + lane;
      const bool inr = k < L;
      const double val = inr ? P(k) : 0.0;
      const uint64_t am = __ballot(inr && val > arm);
      uint64_t neg = __ballot(inr && val < 0.0);
      if (kc < 0) {
        if (!armed && am) {
          neg &= lanes_from(__ffsll((unsigned long long)am) - 1);
          armed = true;
        }
        if (armed && neg) kc = ch * 64 + (__ffsll((unsigned long long)neg) - 1);
      }
    }
    if (lane == 0) {
      int ke = kc;
      if (kc >= 1 && fabs(P(kc)) > fabs(P(kc - 1))) ke = kc - 1;
      kc_out[f0 + c] = kc;
      ke_out[f0 + c] = ke;
      KE[c] = ke;
    }
  }
  __syncthreads();  // KE, and every read of the tile is done
  const int na = KE[2 * q], nb = KE[2 * q + 1];
  double aa = 0.0, bb = 0.0;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int n = tc + TL::template pos_in<inv_first<L, V>()>(0, e);
    if (n <= L / 2) {
      const double w = (n == 0 || 2 * n == L) ? 1.0 : 2.0, k = kzt[n];
      const cplx<T> A = ab[e].a, B = ab[e].b;
      if (na >= 0) {
        const uint32_t j = ((uint32_t)n * (uint32_t)na) % (uint32_t)L;
        aa += w * ((-k * A.y) * trig[j] - (k * A.x) * trig[L + j]);
      }
      if (nb >= 0) {
        const uint32_t j = ((uint32_t)n * (uint32_t)nb) % (uint32_t)L;
        bb += w * ((-k * B.y) * trig[j] - (k * B.x) * trig[L + j]);
      }
    }
  }
  double* R = reinterpret_cast<double*>(F);
  R[2 * (tc * NPR + q)] = aa;
  R[2 * (tc * NPR + q) + 1] = bb;
  __syncthreads();
  if (tid < NC) {  // column f0 + tid: pair tid / 2, half tid % 2
    const int qq = tid >> 1, h = tid & 1;
    double sum = 0.0;
    for (int t = 0; t < TPC; ++t) sum += R[2 * (t * NPR + qq) + h];
    dz[f0 + tid] = sum;
    for (int i = 1; i < ns; ++i) dz[(int64_t)i * nxy + f0 + tid] = 0.0;
  }
}

// Launch of k_zinv_scan over nxy columns (16 per block, the E = 8 plan where there is one);
// hipErrorNotSupported for a length whose block is not whole waves (the caller then runs the z
// pass, k_surf_scan and k_surf_point).
inline hipError_t surf_zinv_scan(const cplx<double>* in, double* psi, int64_t nxy, int nz, const cplx<double>* twz,
                                 const double* kzt, const double* trig, double arm, int* kc, int* ke, int ns,
                                 double* dz, hipStream_t st) {
  constexpr int NC = 16;
  if (nxy % NC != 0) return hipErrorInvalidValue;
#define SMX_CASE(L)                                                                                       \
  if (nz == L) {                                                                                          \
    constexpr int V = HasPlanV<L, 1>::value ? 1 : 0;                                                      \
    if constexpr (((NC / 2) * (L / PlanV<L, V>::E)) % 64 == 0) {                                          \
      SMX_LAUNCH((k_zinv_scan<L, NC, V>), dim3((unsigned)(nxy / NC)), dim3((NC / 2) * (L / PlanV<L, V>::E)), 0, \
                 st, in, psi, nxy, twz, kzt, trig, arm, kc, ke, ns, dz);                                  \
      return hipGetLastError();                                                                           \
    } else {                                                                                              \
      return hipErrorNotSupported;                                                                        \
    }                                                                                                     \
  }
  SMX_FOR_EACH_LENGTH(SMX_CASE)
#undef SMX_CASE
  return hipErrorInvalidValue;
}

// An FP32 Neumann stage read straight by surfTrack's FP64 passes (rows widened on load), so that
// no FP64 copy of it is made: the scaled forward x transform into the FP64 spectrum (as
// NeumannOps<double>::xpass X_FWD_SCALED) and the inverse y transform into FP64 (as
// NeumannOps<double>::ypass +1), with the FP64 schedule's tile width.
inline hipError_t surf_dct_xfwd_f32(const float* in, double* out, const PParams<double>& p, const cplx<double>* qtw,
                                    hipStream_t st) {
  constexpr int NC = neumann_nc<double>();
  if (p.ny % NC != 0) return hipErrorInvalidValue;
#define SMX_CASE(L)                                                                                            \
  if (p.nx == L) {                                                                                             \
    SMX_LAUNCH((k_dct_xpass<double, L, NC, X_FWD_SCALED, float>), dim3((unsigned)(p.ny / NC), (unsigned)p.nzh), \
               dim3((NC / 2) * (L / Plan<L>::E)), 0, st, in, out, (double*)nullptr, (const double*)nullptr, p,     \
               qtw);                                                                                           \
    return hipGetLastError();                                                                                  \
  }
  SMX_FOR_EACH_LENGTH(SMX_CASE)
#undef SMX_CASE
  return hipErrorInvalidValue;
}
// surfTrack's two x-dependent multipliers of a Neumann x stage (FP64, or the FP32 stage widened on
// load), from one forward transform: F^-1[m_{p.kop} F in] into out1 and F^-1[m_{p.kop2} F in]
// into out2 (k_dct_xpass X_MUL2; p.scale = 1/2nx undoes the DCT-II / DCT-III round trip).
template <typename TI>
inline hipError_t surf_dct_xmul2(const TI* in, double* out1, double* out2, const PParams<double>& p,
                                 const cplx<double>* qtw, hipStream_t st) {
  constexpr int NC = neumann_nc<double>();
  if (p.ny % NC != 0 || p.kop < 0 || p.kop2 < 0) return hipErrorInvalidValue;
#define SMX_CASE(L)                                                                                       \
  if (p.nx == L) {                                                                                        \
    SMX_LAUNCH((k_dct_xpass<double, L, NC, X_MUL2, TI>), dim3((unsigned)(p.ny / NC), (unsigned)p.nzh),      \
               dim3((NC / 2) * (L / Plan<L>::E)), 0, st, in, out1, out2, (const double*)nullptr, p, qtw);   \
    return hipGetLastError();                                                                             \
  }
  SMX_FOR_EACH_LENGTH(SMX_CASE)
#undef SMX_CASE
  return hipErrorInvalidValue;
}
inline hipError_t surf_dct_yinv_f32(const float* in, double* out, int ny, int64_t npencils, const cplx<double>* tw,
                                    const cplx<double>* qtw, hipStream_t st) {
  constexpr int NC = neumann_nc<double>();
  if (npencils % NC != 0) return hipErrorInvalidValue;
#define SMX_CASE(L)                                                                                            \
  if (ny == L) {                                                                                               \
    SMX_LAUNCH((k_dct_ypass<double, L, +1, NC, DP_COS, float>), dim3((unsigned)(npencils / NC)),               \
               dim3((NC / 2) * (L / Plan<L>::E)), 0, st, in, out, tw, qtw, (const double*)nullptr);            \
    return hipGetLastError();                                                                                  \
  }
  SMX_FOR_EACH_LENGTH(SMX_CASE)
#undef SMX_CASE
  return hipErrorInvalidValue;
}

}  // namespace smx
