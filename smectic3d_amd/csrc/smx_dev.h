// smx_dev.h -- device building blocks for libsmectic (gfx950).
//
//  * cplx<T>: 2-component complex in registers / LDS / HBM.
//  * dft<R,SIGN>: in-register radix-R DFT, R in {2,3,4,6,8,12,16}.
//  * BlockFFT: batched length-L Stockham autosort FFT of NC columns held in an
//    LDS tile laid out [pos][col] (col fastest).  Thread (col, tc) owns E =
//    L/TPC elements; pass p with radix R computes E/R butterflies per thread.
//    Inputs of the first pass sit at positions tc + TPC*u + r*(L/R_first);
//    outputs of the last pass at tc + TPC*u + r*(L/R_last), so a forward
//    transform followed by an inverse transform with the radix list reversed
//    needs no data exchange in between.
//
// Stockham pass (Govindaraju et al., SC'08, "High performance discrete Fourier
// transforms on graphics processors"): butterfly j reads X[j + r L/R],
// multiplies by w^{r (j mod Ns)} with w = e^{sign 2 pi i/(Ns R)}, does a
// radix-R DFT and writes Y[(j/Ns) Ns R + (j mod Ns) + r Ns].
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace smx {

// k-space multipliers of the derivative and rate fields (the kmul kernels, and the transform
// passes that apply one to their spectral input as it is loaded)
enum KMode {
  KM_DTPSI = 0,   // out = aLin(k) in + out         (cosNoAdvMorphCurv.cpp:744, surfTrack.cpp:611)
  KM_IKX = 1,     // out = i kx in                 (pf3d.cpp:404)
  KM_IKY = 2,     //                               (:405)
  KM_IKZ = 3,     //                               (:406)
  KM_DXX = 4,     // out = -kx^2 in                (surfTrack.cpp:615, cosNoAdvMorphCurv.cpp:746)
  KM_DYY = 5,     //                               (surfTrack.cpp:616, cosNoAdvMorphCurv.cpp:748)
  KM_DZZ = 6,     //                               (cosNoAdvMorphCurv.cpp:750)
  KM_DXY = 7,     // out = (i ky)(i kx) in         (surfTrack.cpp:617)
  KM_LAP_XY = 8,  // out = -(kx^2 + ky^2) in       (surfTrack.cpp:612-613)
  KM_ALIN = 9,    // out = aLin(k) in              (the linear part of dpsi/dt)
};

template <typename T>
struct __attribute__((aligned(2 * sizeof(T)))) cplx {
  T x, y;
};

// two adjacent complex values moved as one access (16 B in FP32, 32 B in FP64)
template <typename T>
struct __attribute__((aligned(4 * sizeof(T)))) cplx2 {
  cplx<T> a, b;
};

template <typename T> __device__ __forceinline__ cplx<T> mk(T a, T b) { cplx<T> r; r.x = a; r.y = b; return r; }
template <typename T> __device__ __forceinline__ cplx<T> operator+(cplx<T> a, cplx<T> b) { return mk(a.x + b.x, a.y + b.y); }
template <typename T> __device__ __forceinline__ cplx<T> operator-(cplx<T> a, cplx<T> b) { return mk(a.x - b.x, a.y - b.y); }
template <typename T> __device__ __forceinline__ cplx<T> operator*(T s, cplx<T> a) { return mk(s * a.x, s * a.y); }
template <typename T> __device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T> __device__ __forceinline__ cplx<T> conjg(cplx<T> a) { return mk(a.x, -a.y); }
// multiply by SIGN*i
template <int SIGN, typename T> __device__ __forceinline__ cplx<T> mul_si(cplx<T> a) {
  return SIGN > 0 ? mk(-a.y, a.x) : mk(a.y, -a.x);
}
// 1/x: IEEE division in FP64; the hardware reciprocal (1 ulp) in FP32, whose
// passes already round at that level.
__device__ __forceinline__ double recip(double x) { return 1.0 / x; }
__device__ __forceinline__ float recip(float x) { return __builtin_amdgcn_rcpf(x); }

// cos(2*pi*m/48) and sin(2*pi*m/48), m = 0..47, rounded from extended precision
__device__ constexpr double kRootCos48[48] = {1.0, 0.9914448613738104, 0.9659258262890683, 0.9238795325112867, 0.8660254037844386, 0.7933533402912352, 0.7071067811865476, 0.6087614290087207, 0.5, 0.3826834323650898, 0.25881904510252074, 0.1305261922200516, 0.0, -0.1305261922200516, -0.25881904510252074, -0.3826834323650898, -0.5, -0.6087614290087207, -0.7071067811865476, -0.7933533402912352, -0.8660254037844386, -0.9238795325112867, -0.9659258262890683, -0.9914448613738104, -1.0, -0.9914448613738104, -0.9659258262890683, -0.9238795325112867, -0.8660254037844386, -0.7933533402912352, -0.7071067811865476, -0.6087614290087207, -0.5, -0.3826834323650898, -0.25881904510252074, -0.1305261922200516, 0.0, 0.1305261922200516, 0.25881904510252074, 0.3826834323650898, 0.5, 0.6087614290087207, 0.7071067811865476, 0.7933533402912352, 0.8660254037844386, 0.9238795325112867, 0.9659258262890683, 0.9914448613738104};
__device__ constexpr double kRootSin48[48] = {0.0, 0.1305261922200516, 0.25881904510252074, 0.3826834323650898, 0.5, 0.6087614290087207, 0.7071067811865476, 0.7933533402912352, 0.8660254037844386, 0.9238795325112867, 0.9659258262890683, 0.9914448613738104, 1.0, 0.9914448613738104, 0.9659258262890683, 0.9238795325112867, 0.8660254037844386, 0.7933533402912352, 0.7071067811865476, 0.6087614290087207, 0.5, 0.3826834323650898, 0.25881904510252074, 0.1305261922200516, 0.0, -0.1305261922200516, -0.25881904510252074, -0.3826834323650898, -0.5, -0.6087614290087207, -0.7071067811865476, -0.7933533402912352, -0.8660254037844386, -0.9238795325112867, -0.9659258262890683, -0.9914448613738104, -1.0, -0.9914448613738104, -0.9659258262890683, -0.9238795325112867, -0.8660254037844386, -0.7933533402912352, -0.7071067811865476, -0.6087614290087207, -0.5, -0.3826834323650898, -0.25881904510252074, -0.1305261922200516};

// e^{SIGN 2 pi i q / R} times a, for compile-time q, R (R divides 48)
template <int R, int Q, int SIGN, typename T>
__device__ __forceinline__ cplx<T> rot(cplx<T> a) {
  constexpr int m = ((Q % R) * (48 / R)) % 48;
  if constexpr (m == 0) return a;
  else if constexpr (m == 12) return mul_si<SIGN>(a);
  else if constexpr (m == 24) return mk(-a.x, -a.y);
  else if constexpr (m == 36) return mul_si<-SIGN>(a);
  else {
    const T c = (T)kRootCos48[m];
    const T s = (T)(SIGN * kRootSin48[m]);
    return mk(a.x * c - a.y * s, a.x * s + a.y * c);
  }
}

template <int R, int SIGN, typename T> __device__ __forceinline__ void dft(cplx<T>* a);

template <int R, int SIGN, int Q, typename T>
__device__ __forceinline__ void combine_even_odd(cplx<T>* a, const cplx<T>* e, const cplx<T>* o) {
  if constexpr (Q < R / 2) {
    cplx<T> t = rot<R, Q, SIGN>(o[Q]);
    a[Q] = e[Q] + t;
    a[Q + R / 2] = e[Q] - t;
    combine_even_odd<R, SIGN, Q + 1>(a, e, o);
  }
}

template <int R, int SIGN, typename T>
__device__ __forceinline__ void dft(cplx<T>* a) {
  if constexpr (R == 1) {
  } else if constexpr (R == 2) {
    cplx<T> t = a[0];
    a[0] = t + a[1];
    a[1] = t - a[1];
  } else if constexpr (R == 3) {
    const T s3 = (T)(SIGN * 0.8660254037844386);
    cplx<T> t = a[1] + a[2], d = a[1] - a[2];
    cplx<T> m = mk(a[0].x - (T)0.5 * t.x, a[0].y - (T)0.5 * t.y);
    cplx<T> sd = mk(-s3 * d.y, s3 * d.x);
    a[0] = a[0] + t;
    a[1] = m + sd;
    a[2] = m - sd;
  } else if constexpr (R == 4) {
    cplx<T> t0 = a[0] + a[2], t1 = a[0] - a[2], t2 = a[1] + a[3], t3 = mul_si<SIGN>(a[1] - a[3]);
    a[0] = t0 + t2;
    a[2] = t0 - t2;
    a[1] = t1 + t3;
    a[3] = t1 - t3;
  } else {
    static_assert(R % 2 == 0, "radix");
    cplx<T> e[R / 2], o[R / 2];
#pragma unroll
    for (int q = 0; q < R / 2; ++q) {
      e[q] = a[2 * q];
      o[q] = a[2 * q + 1];
    }
    dft<R / 2, SIGN>(e);
    dft<R / 2, SIGN>(o);
    combine_even_odd<R, SIGN, 0>(a, e, o);
  }
}

// ---------------------------------------------------------------------------
// Global access at a 32-bit byte offset from a wave-uniform base pointer: the base
// stays in SGPRs and each lane carries one 32-bit offset (the saddr form of
// global_load/global_store) instead of a 64-bit address pair.  Every array of the
// solver is below 4 GiB (a 512^3 FP64 half spectrum is 1.08 GB).
template <typename X>
__device__ __forceinline__ X ld_off(const X* base, uint32_t off) {
  return *reinterpret_cast<const X*>(reinterpret_cast<const char*>(base) + off);
}
template <typename X>
__device__ __forceinline__ void st_off(X* base, uint32_t off, const X& v) {
  *reinterpret_cast<X*>(reinterpret_cast<char*>(base) + off) = v;
}

// ---------------------------------------------------------------------------
// Radix plans.  Fwd and Inv list the same radices in opposite orders.
template <int... Rs> struct Radices {};
template <int R0, int... Rs> struct First { static constexpr int value = R0; };
template <int R0, int... Rs> struct Last { static constexpr int value = Last<Rs...>::value; };
template <int R0> struct Last<R0> { static constexpr int value = R0; };
template <typename P> struct RadInfo;
template <int... Rs> struct RadInfo<Radices<Rs...>> {
  static constexpr int first = First<Rs...>::value;
  static constexpr int last = Last<Rs...>::value;
};

// L -> (E, forward radices, inverse radices)
template <int L> struct Plan;
#define SMX_PLAN(L_, E_, FWD, INV)   \
  template <> struct Plan<L_> {     \
    static constexpr int E = E_;    \
    using Fwd = Radices<FWD>;       \
    using Inv = Radices<INV>;       \
  };
#define SMX_COMMA ,
SMX_PLAN(8, 8, 8, 8)
SMX_PLAN(12, 12, 4 SMX_COMMA 3, 3 SMX_COMMA 4)
SMX_PLAN(16, 16, 16, 16)
SMX_PLAN(24, 12, 4 SMX_COMMA 6, 6 SMX_COMMA 4)
SMX_PLAN(32, 8, 8 SMX_COMMA 4, 4 SMX_COMMA 8)
SMX_PLAN(48, 12, 4 SMX_COMMA 12, 12 SMX_COMMA 4)
SMX_PLAN(64, 16, 16 SMX_COMMA 4, 4 SMX_COMMA 16)
SMX_PLAN(96, 12, 4 SMX_COMMA 4 SMX_COMMA 6, 6 SMX_COMMA 4 SMX_COMMA 4)
SMX_PLAN(128, 16, 16 SMX_COMMA 8, 8 SMX_COMMA 16)
SMX_PLAN(192, 12, 4 SMX_COMMA 4 SMX_COMMA 12, 12 SMX_COMMA 4 SMX_COMMA 4)
SMX_PLAN(256, 16, 16 SMX_COMMA 16, 16 SMX_COMMA 16)
SMX_PLAN(384, 12, 4 SMX_COMMA 4 SMX_COMMA 4 SMX_COMMA 6, 6 SMX_COMMA 4 SMX_COMMA 4 SMX_COMMA 4)
SMX_PLAN(512, 16, 16 SMX_COMMA 16 SMX_COMMA 2, 2 SMX_COMMA 16 SMX_COMMA 16)
#undef SMX_PLAN

// Alternative schedules of the same length (an autotuner picks per pass kernel).
// Variant 0 is Plan<L>; variant 1 (power-of-two lengths) changes the elements per
// thread: E = 8 for L >= 64 (twice the threads per line, half the registers per
// thread), E = 16 for L = 32.
template <int L, int V> struct PlanV;
template <int L> struct PlanV<L, 0> : Plan<L> {};
#define SMX_PLANV(L_, V_, E_, FWD, INV) \
  template <> struct PlanV<L_, V_> {    \
    static constexpr int E = E_;        \
    using Fwd = Radices<FWD>;           \
    using Inv = Radices<INV>;           \
  };
SMX_PLANV(32, 1, 16, 16 SMX_COMMA 2, 2 SMX_COMMA 16)
SMX_PLANV(64, 1, 8, 8 SMX_COMMA 8, 8 SMX_COMMA 8)
SMX_PLANV(128, 1, 8, 8 SMX_COMMA 4 SMX_COMMA 4, 4 SMX_COMMA 4 SMX_COMMA 8)
SMX_PLANV(256, 1, 8, 8 SMX_COMMA 8 SMX_COMMA 4, 4 SMX_COMMA 8 SMX_COMMA 8)
SMX_PLANV(512, 1, 8, 8 SMX_COMMA 8 SMX_COMMA 8, 8 SMX_COMMA 8 SMX_COMMA 8)
#undef SMX_PLANV
template <int L, int V, typename = void> struct HasPlanV { static constexpr bool value = false; };
template <int L, int V> struct HasPlanV<L, V, decltype((void)PlanV<L, V>::E)> { static constexpr bool value = true; };

// ---------------------------------------------------------------------------
// LDS addressing of an exchange tile of NC columns of length L.
//   RowAddr<NCP>: [pos][col] with row pitch NCP (lanes run across columns).
//   ColAddr<L>:   [col][pos] with one pad element per 16 positions (lanes run along
//                 a column; the stride-R writes of the first Stockham pass stay
//                 conflict-free).
template <int NCP>
struct RowAddr {
  static __device__ __forceinline__ int at(int pos, int col) { return pos * NCP + col; }
};
template <int L>
struct ColAddr {
  static constexpr int LP = L + L / 16 + 1;
  static __device__ __forceinline__ int at(int pos, int col) { return col * LP + pos + (pos >> 4); }
};
template <typename X> struct AddrOf { using type = X; };
template <int L, int NCP> struct AddrSel { using type = RowAddr<NCP>; };

// Block FFT over an LDS tile.  AD: addressing policy (RowAddr / ColAddr).
template <typename T, int L, int E, typename AD>
struct TileA {
  static constexpr int TPC = L / E;
  // element e of a thread for a pass of radix R: butterfly u = e / R, leg r = e % R
  template <int R> __device__ static __forceinline__ int pos_in(int tc, int e) {
    return tc + TPC * (e / R) + (e % R) * (L / R);
  }
  template <int R> __device__ static __forceinline__ void load(cplx<T>* v, const cplx<T>* lds, int col, int tc) {
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = lds[AD::at(pos_in<R>(tc, e), col)];
  }
  template <int R> __device__ static __forceinline__ void store(const cplx<T>* v, cplx<T>* lds, int col, int tc) {
#pragma unroll
    for (int e = 0; e < E; ++e) lds[AD::at(pos_in<R>(tc, e), col)] = v[e];
  }
};
template <typename T, int L, int E, int NCP> using Tile = TileA<T, L, E, RowAddr<NCP>>;

// Block barrier that orders LDS accesses only: global loads and stores in flight are not waited
// for (a workgroup fence restricted to the local address space, as OpenCL's
// barrier(CLK_LOCAL_MEM_FENCE)).  For a kernel whose threads exchange data through LDS alone, so
// that a prefetch of the next tile's rows stays in flight across the transform's exchanges.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}
template <bool LB> __device__ __forceinline__ void tile_barrier() {
  if constexpr (LB) lds_barrier();
  else __syncthreads();
}

template <bool LB, typename T, int L, int E, typename AD, int SIGN, int Ns, int R, int... Rest>
__device__ __forceinline__ void fft_passes(cplx<T>* v, cplx<T>* lds, int col, int tc, const cplx<T>* __restrict__ tw) {
  constexpr int TPC = L / E;
  static_assert(E % R == 0, "E must be a multiple of every radix");
#pragma unroll
  for (int u = 0; u < E / R; ++u) {
    const int b = tc + TPC * u;
    if constexpr (Ns > 1) {
      const int k = b % Ns;
#pragma unroll
      for (int r = 1; r < R; ++r) {
        cplx<T> w = ld_off(tw, (uint32_t)(r * k * (L / (Ns * R))) * (uint32_t)sizeof(cplx<T>));
        if (SIGN > 0) w.y = -w.y;
        v[u * R + r] = cmul(v[u * R + r], w);
      }
    }
    dft<R, SIGN>(&v[u * R]);
  }
  if constexpr (sizeof...(Rest) > 0) {
    tile_barrier<LB>();
#pragma unroll
    for (int u = 0; u < E / R; ++u) {
      const int b = tc + TPC * u;
      const int k = b % Ns;
      const int base = (b - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) lds[AD::at(base + r * Ns, col)] = v[u * R + r];
    }
    tile_barrier<LB>();
    constexpr int Rn = First<Rest...>::value;
    TileA<T, L, E, AD>::template load<Rn>(v, lds, col, tc);
    fft_passes<LB, T, L, E, AD, SIGN, Ns * R, Rest...>(v, lds, col, tc, tw);
  }
}

// LB: exchanges behind lds_barrier (else __syncthreads).
template <typename T, int L, int E, typename AD, int SIGN, typename Rad, bool LB = false> struct BlockFFTA;
template <typename T, int L, int E, typename AD, int SIGN, int... Rs, bool LB>
struct BlockFFTA<T, L, E, AD, SIGN, Radices<Rs...>, LB> {
  // v: first-pass input positions on entry; last-pass output positions on exit.
  // The LDS tile is used for the exchanges; all threads must have finished any
  // earlier reads of it (the first exchange starts with a barrier).
  __device__ static __forceinline__ void run(cplx<T>* v, cplx<T>* lds, int col, int tc, const cplx<T>* __restrict__ tw) {
    fft_passes<LB, T, L, E, AD, SIGN, 1, Rs...>(v, lds, col, tc, tw);
  }
};
// forward (SIGN = -1) with the plan's Fwd radices, inverse (SIGN = +1) with its Inv radices
template <typename T, int L, int NCP, int V = 0, bool LB = false>
using FwdFFT = BlockFFTA<T, L, PlanV<L, V>::E, RowAddr<NCP>, -1, typename PlanV<L, V>::Fwd, LB>;
template <typename T, int L, int NCP, int V = 0, bool LB = false>
using InvFFT = BlockFFTA<T, L, PlanV<L, V>::E, RowAddr<NCP>, +1, typename PlanV<L, V>::Inv, LB>;
template <typename T, int L, typename AD, int V = 0>
using FwdFFTA = BlockFFTA<T, L, PlanV<L, V>::E, AD, -1, typename PlanV<L, V>::Fwd>;
template <typename T, int L, typename AD, int V = 0, bool LB = false>
using InvFFTA = BlockFFTA<T, L, PlanV<L, V>::E, AD, +1, typename PlanV<L, V>::Inv, LB>;
template <int L, int V = 0> constexpr int fwd_first() { return RadInfo<typename PlanV<L, V>::Fwd>::first; }
template <int L, int V = 0> constexpr int fwd_last() { return RadInfo<typename PlanV<L, V>::Fwd>::last; }
template <int L, int V = 0> constexpr int inv_first() { return RadInfo<typename PlanV<L, V>::Inv>::first; }
template <int L, int V = 0> constexpr int inv_last() { return RadInfo<typename PlanV<L, V>::Inv>::last; }

// LDS copy of a length-L twiddle table, filled by the NT threads of the block.  Twiddle
// reads are per-lane gathers repeated in every pass; from LDS they take the short LDS
// path instead of the vector-memory path, where each would also wait behind the tile's
// HBM loads (one in-order counter per wave).  The first pass of every plan reads no
// twiddles and the first exchange starts with a barrier, so the copy is complete before
// its first read without a barrier of its own.
template <int L, int NT, typename T>
__device__ __forceinline__ const cplx<T>* tw_lds(cplx<T>* lds, const cplx<T>* __restrict__ tw, int tid) {
#pragma unroll
  for (int i = tid; i < L; i += NT) lds[i] = tw[i];
  return lds;
}

// ---------------------------------------------------------------------------
// FP64 block reduction helpers (wavefront shuffles, then LDS across waves).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// FP64 partial over a tile of the sum of x + y of the E complex values each thread holds,
// without a barrier of its own: a pairwise tree per thread, a shuffle tree per wave into
// red[] (wave_partial), and thread 0 adds the waves in order (store_partial) after some
// later barrier of the kernel has made red[] visible.  red: >= 64 doubles.
template <int NT> constexpr int partial_slots() { return (NT % 64 == 0) ? NT / 64 : NT; }
template <int E, int NT, typename T>
__device__ __forceinline__ void wave_partial(const cplx<T>* v, double* red, int tid) {
  static_assert(NT % 64 == 0 || NT <= 64, "partial sums: whole waves or one partial wave");
  double t[E];
#pragma unroll
  for (int e = 0; e < E; ++e) t[e] = (double)v[e].x + (double)v[e].y;
#pragma unroll
  for (int w = 1; w < E; w *= 2)
#pragma unroll
    for (int e = 0; e + w < E; e += 2 * w) t[e] += t[e + w];
  if constexpr (NT % 64 == 0) {
    const double ws = wave_sum(t[0]);
    if ((tid & 63) == 0) red[tid >> 6] = ws;
  } else {
    red[tid] = t[0];
  }
}
template <int NT>
__device__ __forceinline__ void store_partial(const double* red, double* out, int tid) {
  if (tid == 0) {
    double s = 0.0;
    for (int w = 0; w < partial_slots<NT>(); ++w) s += red[w];
    *out = s;
  }
}

// Sum over the block; result valid in thread 0.  scratch: >= max(nwaves, 64)
// doubles when blockDim.x is not a multiple of 64, else >= nwaves doubles.
// Shuffles are used only when every wave of the block is full.
__device__ __forceinline__ double block_sum(double v, double* scratch) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  __syncthreads();
  if ((blockDim.x & 63) != 0) {  // small blocks (tiny grids): plain LDS sum, blockDim.x <= 64 here
    scratch[threadIdx.x] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
      for (int i = 0; i < (int)blockDim.x; ++i) s += scratch[i];
    return s;
  }
  v = wave_sum(v);
  if (lane == 0) scratch[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < nw; ++i) s += scratch[i];
  return s;
}

}  // namespace smx
