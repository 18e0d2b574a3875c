// smx_diag.h -- FP64 diagnostics of libsmectic: compensated reductions (L1,
// energy, dominant wavenumber), k-space multipliers for derivative fields, and
// the surfTrack column scans.  Spectra are normalised and laid out
// [kz][kx][ky] (complex, periodic half spectrum) or [kz][kx][ky] (real, DCT).
#pragma once
#include "smx_dev.h"

namespace smx {

struct Kahan {
  double s = 0.0, c = 0.0;
  __device__ __forceinline__ void add(double x) {
    double y = x - c, t = s + y;
    c = (t - s) - y;
    s = t;
  }
};

struct KParams {
  int nx, ny, nz, nplanes;
  int64_t n;  // spectral entries
  int neumann;
  const double* kx;
  const double* ky;
  const double* kz;
  double eps, alpha, q02;
};

// (kz, x, ky) of entry i of a [kz][x][ky] array.  32-bit arithmetic: every supported grid has
// fewer than 2^31 spectral entries (512^3 Neumann: 1.3e8), and 64-bit division costs as much
// issue time as these grid-stride kernels spend on memory.
__device__ __forceinline__ void kidx(const KParams& k, int64_t i, int& iz, int& ix, int& iy) {
  const uint32_t u = (uint32_t)i, nxy = (uint32_t)k.nx * (uint32_t)k.ny;
  const uint32_t z = u / nxy, r = u - z * nxy, x = r / (uint32_t)k.ny;
  iz = (int)z;
  ix = (int)x;
  iy = (int)(r - x * (uint32_t)k.ny);
}

__device__ __forceinline__ double alin(const KParams& k, double k2) {
  const double d = k2 - k.q02;
  return k.eps - k.alpha * d * d;
}

// The k-space multipliers walk the [kz][x][ky] layout row by row: one row (kz, x) per block
// iteration, so qx and qz are uniform over the block and only qy is read per entry.
__global__ __launch_bounds__(256) void k_kmul_cplx(const cplx<double>* __restrict__ in, cplx<double>* __restrict__ out,
                                                   KParams k, int mode) {
  const int nrows = k.nplanes * k.nx;
  for (int row = blockIdx.x; row < nrows; row += gridDim.x) {
    const int iz = row / k.nx, ix = row - iz * k.nx;
    const double qx = k.kx[ix], qz = k.kz[iz];
    const int64_t base = (int64_t)row * k.ny;
    for (int iy = threadIdx.x; iy < k.ny; iy += blockDim.x) {
      const int64_t i = base + iy;
      const double qy = k.ky[iy];
      const cplx<double> a = in[i];
      cplx<double> r;
      switch (mode) {
        case KM_DTPSI: {
          const double m = alin(k, qx * qx + qy * qy + qz * qz);
          const cplx<double> o = out[i];
          r = mk(m * a.x + o.x, m * a.y + o.y);
          break;
        }
        case KM_IKX: r = mk(-qx * a.y, qx * a.x); break;
        case KM_IKY: r = mk(-qy * a.y, qy * a.x); break;
        case KM_IKZ: r = mk(-qz * a.y, qz * a.x); break;
        case KM_DXX: r = (-qx * qx) * a; break;
        case KM_DYY: r = (-qy * qy) * a; break;
        case KM_DZZ: r = (-qz * qz) * a; break;
        case KM_DXY: r = (-qx * qy) * a; break;
        default: r = (-(qx * qx) - qy * qy) * a; break;
      }
      out[i] = r;
    }
  }
}

__global__ __launch_bounds__(256) void k_kmul_real(const double* __restrict__ in, double* __restrict__ out, KParams k,
                                                   int mode) {
  const int nrows = k.nplanes * k.nx;
  for (int row = blockIdx.x; row < nrows; row += gridDim.x) {
    const int iz = row / k.nx, ix = row - iz * k.nx;
    const double qx = k.kx[ix], qz = k.kz[iz];
    const int64_t base = (int64_t)row * k.ny;
    for (int iy = threadIdx.x; iy < k.ny; iy += blockDim.x) {
      const int64_t i = base + iy;
      const double qy = k.ky[iy];
      const double a = in[i];
      double r;
      switch (mode) {
        case KM_DTPSI: r = alin(k, qx * qx + qy * qy + qz * qz) * a + out[i]; break;
        case KM_DXX: r = -qx * qx * a; break;
        case KM_DYY: r = -qy * qy * a; break;
        case KM_DZZ: r = -qz * qz * a; break;
        default: r = (-(qx * qx) - qy * qy) * a; break;
      }
      out[i] = r;
    }
  }
}

// Per-block Kahan partials over the normalised half spectrum:
//   out[3b+0] = sum w (alpha/2 (q0^2 - k^2)^2 - eps/2) |psi_hat|^2
//   out[3b+1] = sum_{k != 0} w |psi_hat|^2,  out[3b+2] = sum_{k != 0} w |k| |psi_hat|^2
// w = 1 on the kz = 0 and kz = nz/2 planes, 2 elsewhere (Hermitian half).
__global__ void k_reduce_spec_cplx(const cplx<double>* __restrict__ a, KParams k, double* __restrict__ out) {
  __shared__ double red[64];
  Kahan s0, s1, s2;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < k.n; i += (int64_t)gridDim.x * blockDim.x) {
    int iz, ix, iy;
    kidx(k, i, iz, ix, iy);
    const double k2 = k.kx[ix] * k.kx[ix] + k.ky[iy] * k.ky[iy] + k.kz[iz] * k.kz[iz];
    const double w = (iz == 0 || 2 * iz == k.nz) ? 1.0 : 2.0;
    const double p2 = w * (a[i].x * a[i].x + a[i].y * a[i].y);
    const double d = k.q02 - k2;
    s0.add((0.5 * k.alpha * d * d - 0.5 * k.eps) * p2);
    if (i != 0) {
      s1.add(p2);
      s2.add(sqrt(k2) * p2);
    }
  }
  const double t0 = block_sum(s0.s, red), t1 = block_sum(s1.s, red), t2 = block_sum(s2.s, red);
  if (threadIdx.x == 0) {
    out[3 * blockIdx.x + 0] = t0;
    out[3 * blockIdx.x + 1] = t1;
    out[3 * blockIdx.x + 2] = t2;
  }
}

// Same over a DCT spectrum; w = prod over axes of (1/2 if k_axis = 0 else 1)
// (discrete Parseval of REDFT10: sum x^2 = (1/2n)(y_0^2/2 + sum_{k>0} y_k^2) per axis).
__global__ void k_reduce_spec_real(const double* __restrict__ a, KParams k, double* __restrict__ out) {
  __shared__ double red[64];
  Kahan s0, s1, s2;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < k.n; i += (int64_t)gridDim.x * blockDim.x) {
    int iz, ix, iy;
    kidx(k, i, iz, ix, iy);
    const double k2 = k.kx[ix] * k.kx[ix] + k.ky[iy] * k.ky[iy] + k.kz[iz] * k.kz[iz];
    const double w = (ix == 0 ? 0.5 : 1.0) * (iy == 0 ? 0.5 : 1.0) * (iz == 0 ? 0.5 : 1.0);
    const double p2 = w * a[i] * a[i];
    const double d = k.q02 - k2;
    s0.add((0.5 * k.alpha * d * d - 0.5 * k.eps) * p2);
    if (i != 0) {
      s1.add(p2);
      s2.add(sqrt(k2) * p2);
    }
  }
  const double t0 = block_sum(s0.s, red), t1 = block_sum(s1.s, red), t2 = block_sum(s2.s, red);
  if (threadIdx.x == 0) {
    out[3 * blockIdx.x + 0] = t0;
    out[3 * blockIdx.x + 1] = t1;
    out[3 * blockIdx.x + 2] = t2;
  }
}

// Per-block partial of sum f_local(psi): energyPsi.m:225-228 (quintic) or the cubic model.
__global__ void k_reduce_local(const double* __restrict__ psi, int64_t n, int nl_model, double beta, double gamma,
                               double* __restrict__ out) {
  __shared__ double red[64];
  Kahan s;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double p = psi[i], p2 = p * p;
    const double f = (nl_model == 1) ? (-beta / 3.0 * p2 * p + gamma / 4.0 * p2 * p2)
                                     : (-beta / 4.0 * p2 * p2 + gamma / 6.0 * p2 * p2 * p2);
    s.add(f);
  }
  const double t = block_sum(s.s, red);
  if (threadIdx.x == 0) out[blockIdx.x] = t;
}

// Per-block partials of sum |a - b| and sum |a| (pf3d.cpp:494-502).
__global__ void k_reduce_l1(const double* __restrict__ a, const double* __restrict__ b, int64_t n,
                            double* __restrict__ out) {
  __shared__ double red[64];
  Kahan s0, s1;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    s0.add(fabs(a[i] - b[i]));
    s1.add(fabs(a[i]));
  }
  const double t0 = block_sum(s0.s, red), t1 = block_sum(s1.s, red);
  if (threadIdx.x == 0) {
    out[2 * blockIdx.x] = t0;
    out[2 * blockIdx.x + 1] = t1;
  }
}

template <typename T>
__global__ void k_count_nonfinite(const T* __restrict__ a, int64_t n, unsigned int* __restrict__ out) {
  unsigned int c = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    c += isfinite((double)a[i]) ? 0u : 1u;
  if (c) atomicAdd(out, c);
}

// Zero fill (a kernel rather than a memset node, so that captured graphs replay it reliably).
__global__ void k_zero_f64(double* __restrict__ a, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    a[i] = 0.0;
}

// Hermitian part of the kz = 0 and kz = nz/2 planes of a periodic x stage ([kz][x][ky],
// complex): G(x, ky) <- (G(x, ky) + conj G(x, -ky)) / 2, the stage of the real field that its
// C2R gives.  An FP32 state carries anti-Hermitian rounding noise on those planes; the C2R drops
// it, but a first derivative at the Nyquist row (k = +pi/d there, pf3d.cpp:162-190) would turn
// it into a real contribution.  One thread per pair (ky, -ky).
__global__ void k_herm_stage(cplx<double>* __restrict__ W, int nx, int ny, int nz) {
  const int half = ny / 2;
  const int64_t per_plane = (int64_t)nx * (half + 1);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < 2 * per_plane;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int pl = (int)(i / per_plane);
    const int64_t r = i - pl * per_plane;
    const int x = (int)(r / (half + 1)), ky = (int)(r - (int64_t)x * (half + 1));
    cplx<double>* row = W + ((int64_t)(pl == 0 ? 0 : nz / 2) * nx + x) * ny;
    const int km = (ny - ky) % ny;
    const cplx<double> a = row[ky], b = row[km];
    const double re = 0.5 * (a.x + b.x), im = 0.5 * (a.y - b.y);
    row[ky] = mk(re, im);
    row[km] = mk(re, -im);
  }
}

// The persistent surfTrack outputs (n doubles) and the crossing count into a mapped pinned host
// buffer (count after them, as a double): written by the kernel in stream order, so the host
// reads them once the stream is idle, without a runtime copy.
__global__ void k_surf_export(const double* __restrict__ a, const unsigned int* __restrict__ count, int64_t n,
                              double* __restrict__ host) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    host[i] = a[i];
  if (blockIdx.x == 0 && threadIdx.x == 0) host[n] = (double)count[0];
}

__global__ void k_f32_to_f64(const float* __restrict__ a, double* __restrict__ b, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (double)a[i];
}

__global__ void k_copy_f4(const float4* __restrict__ a, float4* __restrict__ b, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}

// Fills a with floats in [-1, 1) from a 32-bit integer hash of the word index.  HBM
// throughput depends on the data moved (an all-zero buffer copies about 13% faster
// than one of random words on this hardware), so the copy yardstick moves field-like
// words, not zeros.
__global__ void k_fill_hash_f4(float4* __restrict__ a, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    float w[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      uint32_t h = (uint32_t)(4 * i + c) * 0x9E3779B9u;
      h ^= h >> 16;
      h *= 0x85EBCA6Bu;
      h ^= h >> 13;
      h *= 0xC2B2AE35u;
      h ^= h >> 16;
      w[c] = (float)(int32_t)h * 4.656612873077393e-10f;
    }
    a[i] = make_float4(w[0], w[1], w[2], w[3]);
  }
}

// a[i] = amp * u_i with u_i in [-1, 1) from a 32-bit integer hash of (seed, i): scratch
// fields of realistic bit patterns (tuning runs without a state).
template <typename T>
__global__ void k_fill_hash(T* __restrict__ a, int64_t n, T amp, uint32_t seed) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    uint32_t h = ((uint32_t)i ^ (uint32_t)(i >> 32) * 0x27D4EB2Fu ^ seed) * 0x9E3779B9u;
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    a[i] = amp * (T)((double)(int32_t)h * 4.656612873077393e-10);
  }
}

// Eckhaus scan energies (energyPsi.m:68-234) of up to four perturbed fields in one pass over
// psi0 (dense FP64 [x][y][z], periodic wrap): psi_m = psi0 + amp_m dpsiz cos((1 + q_m)(k + 1) dz)
// with dpsiz the script's 4-point z stencil as written (:100-101), energy density
// -eps/2 psi^2 + alpha/2 (q0^2 psi + lap psi)^2 + f(psi) with second differences (:217-228),
// trapz weights 1/2 at both ends of every axis (:234).  Per-block Kahan partials of the
// weighted sums (times dx dy dz on the host), layout [block][4].
__device__ __forceinline__ double eck_dpz(const double* col, int k, int nz) {
  const int kp = k + 1 == nz ? 0 : k + 1, kpp = (k + 2) % nz, km = k == 0 ? nz - 1 : k - 1, kmm = (k + nz - 2) % nz;
  return -(1.0 / 12) * col[kpp] + (2.0 / 3) * col[kp] - (2.0 / 3) * col[kmm] + (1.0 / 12) * col[km];
}
__global__ void k_eckhaus(const double* __restrict__ psi, int nx, int ny, int nz, double dx, double dy, double dz,
                          double eps, double alpha, double q02, double beta, double gamma, int nl_model, int nm,
                          double q0v, double q1v, double q2v, double q3v, double a0, double a1, double a2, double a3,
                          double* __restrict__ out) {
  __shared__ double red[64];
  const double qv[4] = {q0v, q1v, q2v, q3v}, av[4] = {a0, a1, a2, a3};
  const double idx2 = 1.0 / (dx * dx), idy2 = 1.0 / (dy * dy), idz2 = 1.0 / (dz * dz);
  const int64_t N = (int64_t)nx * ny * nz;
  Kahan acc[4];
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < N; p += (int64_t)gridDim.x * blockDim.x) {
    const int k = (int)(p % nz), j = (int)((p / nz) % ny), i = (int)(p / ((int64_t)ny * nz));
    const int im = i == 0 ? nx - 1 : i - 1, ip = i + 1 == nx ? 0 : i + 1;
    const int jm = j == 0 ? ny - 1 : j - 1, jp = j + 1 == ny ? 0 : j + 1;
    const int km = k == 0 ? nz - 1 : k - 1, kp = k + 1 == nz ? 0 : k + 1;
    const double* c0 = psi + ((int64_t)i * ny + j) * nz;
    const double* cxm = psi + ((int64_t)im * ny + j) * nz;
    const double* cxp = psi + ((int64_t)ip * ny + j) * nz;
    const double* cym = psi + ((int64_t)i * ny + jm) * nz;
    const double* cyp = psi + ((int64_t)i * ny + jp) * nz;
    // psi0 and dpsiz at the 7-point stencil
    const double p0 = c0[k], pkm = c0[km], pkp = c0[kp], pxm = cxm[k], pxp = cxp[k], pym = cym[k], pyp = cyp[k];
    const double d0 = eck_dpz(c0, k, nz), dkm = eck_dpz(c0, km, nz), dkp = eck_dpz(c0, kp, nz);
    const double dxm = eck_dpz(cxm, k, nz), dxp = eck_dpz(cxp, k, nz), dym = eck_dpz(cym, k, nz),
                 dyp = eck_dpz(cyp, k, nz);
    const double w = (i == 0 || i == nx - 1 ? 0.5 : 1.0) * (j == 0 || j == ny - 1 ? 0.5 : 1.0) *
                     (k == 0 || k == nz - 1 ? 0.5 : 1.0);
    for (int m = 0; m < nm; ++m) {
      const double f = (1.0 + qv[m]) * dz;
      const double ck = av[m] * cos(f * (k + 1)), ckm = av[m] * cos(f * (km + 1)), ckp = av[m] * cos(f * (kp + 1));
      const double c = p0 + d0 * ck;
      const double lap = idz2 * ((pkm + dkm * ckm) - 2 * c + (pkp + dkp * ckp)) +
                         idy2 * ((pym + dym * ck) - 2 * c + (pyp + dyp * ck)) +
                         idx2 * ((pxm + dxm * ck) - 2 * c + (pxp + dxp * ck));
      const double t = q02 * c + lap, c2 = c * c;
      const double fl = (nl_model == 1) ? (-beta / 3.0 * c2 * c + gamma / 4.0 * c2 * c2)
                                        : (-beta / 4.0 * c2 * c2 + gamma / 6.0 * c2 * c2 * c2);
      acc[m].add(w * (-0.5 * eps * c2 + 0.5 * alpha * t * t + fl));
    }
  }
  for (int m = 0; m < 4; ++m) {
    const double s = block_sum(acc[m].s, red);
    if (threadIdx.x == 0) out[(int64_t)blockIdx.x * 4 + m] = s;
  }
}

// One axis of the separable psiAvg integral (morphSurfAmp.cpp:1005-1134) on a dense FP64
// [x][y][z] field: out = d (sum of f over the window [c - h, c + h] clipped to [0, n - 1]
// less half its two end values) -- the reference's trapz (:76-86) along that axis, in
// another order of the same terms; f = in^2 when SQUARE.  Axis of length n, stride s.
template <bool SQUARE>
__global__ void k_window_trapz(const double* __restrict__ in, double* __restrict__ out, int64_t N, int n, int64_t s,
                               int h, double d) {
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < N; p += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)((p / s) % n);
    const int a = c - h < 0 ? 0 : c - h, b = c + h > n - 1 ? n - 1 : c + h;
    const double* f = in + (p - (int64_t)(c - a) * s);  // window start, inside the field
    double t = 0.0;
    for (int m = 0; m <= b - a; ++m) {
      const double v = f[(int64_t)m * s];
      t += SQUARE ? v * v : v;
    }
    double fa = f[0], fb = f[(int64_t)(b - a) * s];
    if (SQUARE) {
      fa *= fa;
      fb *= fb;
    }
    out[p] = (b > a) ? d * (t - 0.5 * (fa + fb)) : 0.0;
  }
}

// psiAvg = T / (ref l1 l2 l3 / lambda^3) (morphSurfAmp.cpp:1078-1080, 1118): ref is T at
// (2h, 2h, 2h), the integral over the reference box [h, 3h]^3 (:1009-1049); l the clipped
// window lengths.
__global__ void k_psi_avg_norm(const double* __restrict__ T, double* __restrict__ out, int nx, int ny, int nz, int h,
                               double dx, double dy, double dz, double lam3) {
  const double ref = T[((int64_t)2 * h * ny + 2 * h) * nz + 2 * h];
  const int64_t N = (int64_t)nx * ny * nz;
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < N; p += (int64_t)gridDim.x * blockDim.x) {
    const int k = (int)(p % nz), j = (int)((p / nz) % ny), i = (int)(p / ((int64_t)ny * nz));
    const double l1 = (double)(min(i + h, nx - 1) - max(i - h, 0)) * dx;
    const double l2 = (double)(min(j + h, ny - 1) - max(j - h, 0)) * dy;
    const double l3 = (double)(min(k + h, nz - 1) - max(k - h, 0)) * dz;
    out[p] = T[p] / (ref * l1 * l2 * l3 / lam3);
  }
}

// a_T = psi_T - amp cos(q0 z) in place (FP64 real, dense [x][y][z]); per-block partials of
// J = 1/2 sum a_T^2 and of the explicit target term sum a_T (-dg/dq0) = sum a_T amp z sin(q0 z).
__global__ void k_adj_target(double* __restrict__ a, int64_t n, int nz, double dz, double q0, double amp,
                             double* __restrict__ out) {
  __shared__ double red[64];
  Kahan sj, sq;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double z = (double)(int)(i % nz) * dz;
    const double r = a[i] - amp * cos(q0 * z);
    a[i] = r;
    sj.add(0.5 * r * r);
    sq.add(r * amp * z * sin(q0 * z));
  }
  const double t0 = block_sum(sj.s, red), t1 = block_sum(sq.s, red);
  if (threadIdx.x == 0) {
    out[2 * blockIdx.x] = t0;
    out[2 * blockIdx.x + 1] = t1;
  }
}

// ---------------------------------------------------------------------------
// surfTrack.  Fields are FP64 real, dense [x][y][z]; one thread per (i,j) column.
// out: 4 arrays of nx*ny (surfz, vel, H, K) written only for columns that cross.
struct SurfParams {
  int nx, ny, nz;
  double arm, vsign;
  double tdx, tdy, tdz;  // 2 dx, 2 dy, 2 dz
  // add_nl: the rate field holds F^-1[aLin psi_hat] alone and N(psi) is added at the crossing
  // (F^-1 F N = N without advection); the model, beta, gamma and the mid-plane penalty of N
  int add_nl, nl_model, mid;
  double beta, gamma, pen;
  int ns;  // kz splits of the field values (parts per slot)
};

// N(psi) at z index k of a column (pf3d.cpp:460-461; the cubic model; the mid-plane penalty
// 2 wMid (psi_{k-1} - 2 psi_k + psi_{k+1})/dz^2 of pf3dMorph.cpp:572-581 at k = mid).
__device__ __forceinline__ double surf_nl(const double* psi, const SurfParams& s, int64_t col, int k) {
  const double v = psi[col + k];
  double n;
  if (s.nl_model == 1) {
    n = v * v * (s.beta - s.gamma * v);
  } else {
    const double v2 = v * v;
    n = v2 * v * (s.beta - s.gamma * v2);
  }
  if (k == s.mid && s.mid > 0 && s.mid < s.nz - 1) n += s.pen * (psi[col + k - 1] - 2.0 * v + psi[col + k + 1]);
  return n;
}

// Lanes from `a` on (0 <= a < 64): a crossing may be taken at the arming index itself, as in the
// reference's loop (possible only for a negative threshold, which the host refuses).
__device__ __forceinline__ uint64_t lanes_from(int a) { return ~((1ull << a) - 1ull); }

// Column scan of the reference's sequential loop (arm at psi > arm, then the first psi < 0):
// one wave per column, every 64-value chunk of the column (nz <= 512: at most 8) loaded
// before the tests, which take the chunks in scan order with ballots (lane order = scan
// order).  The crossing index k (-1: none) goes to kc_out[column] and the index the
// derivatives are taken at to ke_out[column]: upward (surfTrack.cpp:779-783) k-1 when
// |psi(k)| > |psi(k-1)|, downward (cosNoAdvMorphCurv.cpp:893-900) k+1 when |psi(k)| >
// |psi(k+1)|, else k.  blockDim = 256: 4 columns per block.
template <bool UP>
__global__ __launch_bounds__(256) void k_surf_scan(const double* __restrict__ psi, int64_t nxy, int nz, double arm,
                                                   int* __restrict__ kc_out, int* __restrict__ ke_out) {
  const int lane = threadIdx.x & 63;
  const int64_t c = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);  // uniform across the wave
  if (c >= nxy) return;
  const double* colp = psi + c * nz;
  const int nch = (nz + 63) >> 6;
  double v[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int k = UP ? q * 64 + lane : nz - 1 - q * 64 - lane;
    v[q] = (q < nch && k >= 0 && k < nz) ? colp[k] : 0.0;
  }
  bool armed = false;
  int kc = -1;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int k = UP ? q * 64 + lane : nz - 1 - q * 64 - lane;
    const bool in = q < nch && k >= 0 && k < nz;
    const uint64_t am = __ballot(in && v[q] > arm);
    uint64_t neg = __ballot(in && v[q] < 0.0);
    if (kc < 0) {
      if (!armed && am) {
        neg &= lanes_from(__ffsll((unsigned long long)am) - 1);
        armed = true;
      }
      if (armed && neg) {
        const int l = __ffsll((unsigned long long)neg) - 1;
        kc = UP ? q * 64 + l : nz - 1 - q * 64 - l;
      }
    }
  }
  if (lane == 0) {
    int ke = kc;
    if (kc >= 0) {
      const int nb = UP ? kc - 1 : kc + 1;
      if (nb >= 0 && nb < nz && fabs(colp[kc]) > fabs(colp[nb])) ke = nb;
    }
    kc_out[c] = kc;
    ke_out[c] = ke;
  }
}

// Field values at each column's evaluation index: slot f holds ns parts of nxy entries (one per
// kz split of the evaluation kernel), added in split order by sval.
enum SurfVal { SV_RATE = 0, SV_DXX = 1, SV_DYY = 2, SV_DZZ = 3, SV_GX = 3, SV_GY = 4, SV_GZ = 5, SV_DXY = 6 };
__device__ __forceinline__ double sval(const double* val, const SurfParams& sp, int f, int64_t nxy, int64_t c) {
  const double* v = val + (int64_t)f * sp.ns * nxy + c;
  double s = v[0];
  for (int i = 1; i < sp.ns; ++i) s += v[i * nxy];
  return s;
}

// surfTrack.cpp:767-806 at the crossing k of k_surf_scan<true> (scan k upward, arm at
// psi > arm, first psi < 0 after arming), with the derivatives at its evaluation index
// (val: rate, psi_xx, psi_yy, psi_x, psi_y, psi_z, psi_xy; the Laplacian numerator psi_xx + psi_yy
// is their sum there); reports k.
// One thread per column.
__global__ __launch_bounds__(256) void k_surf_periodic(const double* __restrict__ psi, const double* __restrict__ val,
                                                       const int* __restrict__ kc_in, const int* __restrict__ ke_in,
                                                       SurfParams sp, double* __restrict__ out,
                                                       unsigned int* __restrict__ nfound) {
  const int64_t nxy = (int64_t)sp.nx * sp.ny;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nxy) return;
  const int kc = kc_in[c];
  if (kc < 0) return;
  const int ke = ke_in[c];
  auto V = [&](int f) { return sval(val, sp, f, nxy, c); };
  const double gx = V(SV_GX), gy = V(SV_GY), gz = V(SV_GZ), dxx = V(SV_DXX), dyy = V(SV_DYY), dxy = V(SV_DXY);
  const double g = sqrt(gx * gx + gy * gy + gz * gz);
  const double rate = V(SV_RATE) + (sp.add_nl ? surf_nl(psi, sp, c * sp.nz, ke) : 0.0);
  out[c] = (double)kc;
  out[nxy + c] = sp.vsign * rate / g;
  out[2 * nxy + c] = (dxx + dyy) / g;  // psi_xx + psi_yy (surfTrack.cpp:614)
  out[3 * nxy + c] = (dxx * dyy - dxy * dxy) / (g * g);
  atomicAdd(nfound, 1u);
}

// Central differences with one-sided (x2) stencils at the walls,
// cosNoAdvMorphCurv.cpp:811-873.
__device__ __forceinline__ double fd_x(const double* p, const SurfParams& s, int i, int j, int k) {
  auto P = [&](int a) { return p[((int64_t)a * s.ny + j) * s.nz + k]; };
  if (i > 0 && i < s.nx - 1) return (P(i + 1) - P(i - 1)) / s.tdx;
  if (i == 0) return 2.0 * (P(1) - P(0)) / s.tdx;
  return 2.0 * (P(i) - P(i - 1)) / s.tdx;
}
__device__ __forceinline__ double fd_y(const double* p, const SurfParams& s, int i, int j, int k) {
  auto P = [&](int b) { return p[((int64_t)i * s.ny + b) * s.nz + k]; };
  if (j > 0 && j < s.ny - 1) return (P(j + 1) - P(j - 1)) / s.tdy;
  if (j == 0) return 2.0 * (P(1) - P(0)) / s.tdy;
  return 2.0 * (P(j) - P(j - 1)) / s.tdy;
}
__device__ __forceinline__ double fd_z(const double* p, const SurfParams& s, int i, int j, int k) {
  auto P = [&](int c) { return p[((int64_t)i * s.ny + j) * s.nz + c]; };
  if (k > 0 && k < s.nz - 1) return (P(k + 1) - P(k - 1)) / s.tdz;
  if (k == 0) return 2.0 * (P(1) - P(0)) / s.tdz;
  return 2.0 * (P(k) - P(k - 1)) / s.tdz;
}

// cosNoAdvMorphCurv.cpp:877-966 at the crossing k of k_surf_scan<false> (scan k downward,
// arm at psi > arm, first psi < 0), evaluated at its index k2 (k+1 when |psi(k)| >
// |psi(k+1)|): Megrabov H = div(grad psi/|grad psi|) and Gaussian curvature K from the FD
// gradient and the spectral second derivatives (val: rate, psi_xx, psi_yy, psi_zz at k2).
// One thread per column.
__global__ __launch_bounds__(256) void k_surf_neumann(const double* __restrict__ psi, const double* __restrict__ val,
                                                      const int* __restrict__ kc_in, const int* __restrict__ ke_in,
                                                      SurfParams sp, double* __restrict__ out,
                                                      unsigned int* __restrict__ nfound) {
  const int64_t nxy = (int64_t)sp.nx * sp.ny;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nxy) return;
  const int kc = kc_in[c];
  if (kc < 0) return;
  const int i = (int)(c / sp.ny), j = (int)(c - (int64_t)i * sp.ny);
  const int64_t col = c * sp.nz;
  const int k = kc, k2 = ke_in[c];
  const double px = fd_x(psi, sp, i, j, k2), py = fd_y(psi, sp, i, j, k2), pz = fd_z(psi, sp, i, j, k2);
  const double g = sqrt(px * px + py * py + pz * pz);
  double d_xy;
  if (j > 0 && j < sp.ny - 1) d_xy = (fd_x(psi, sp, i, j + 1, k2) - fd_x(psi, sp, i, j - 1, k2)) / sp.tdy;
  else if (j == 0) d_xy = 2.0 * (fd_x(psi, sp, i, 1, k2) - fd_x(psi, sp, i, 0, k2)) / sp.tdy;
  else d_xy = 2.0 * (fd_x(psi, sp, i, j, k2) - fd_x(psi, sp, i, j - 1, k2)) / sp.tdy;
  // the reference picks the z stencil by k while centring it on k2 (:926-940); a
  // centred stencil that would leave the column falls back to one-sided
  double d_xz, d_yz;
  if (k > 0 && k < sp.nz - 1 && k2 - 1 >= 0 && k2 + 1 <= sp.nz - 1) {
    d_xz = (fd_x(psi, sp, i, j, k2 + 1) - fd_x(psi, sp, i, j, k2 - 1)) / sp.tdz;
    d_yz = (fd_y(psi, sp, i, j, k2 + 1) - fd_y(psi, sp, i, j, k2 - 1)) / sp.tdz;
  } else if (k == 0 || k2 - 1 < 0) {
    d_xz = 2.0 * (fd_x(psi, sp, i, j, k2 + 1) - fd_x(psi, sp, i, j, k2)) / sp.tdz;
    d_yz = 2.0 * (fd_y(psi, sp, i, j, k2 + 1) - fd_y(psi, sp, i, j, k2)) / sp.tdz;
  } else {
    d_xz = 2.0 * (fd_x(psi, sp, i, j, k2) - fd_x(psi, sp, i, j, k2 - 1)) / sp.tdz;
    d_yz = 2.0 * (fd_y(psi, sp, i, j, k2) - fd_y(psi, sp, i, j, k2 - 1)) / sp.tdz;
  }
  const double fxx = sval(val, sp, SV_DXX, nxy, c), fyy = sval(val, sp, SV_DYY, nxy, c),
               fzz = sval(val, sp, SV_DZZ, nxy, c);
  const double Hc = ((py * py + pz * pz) * fxx + (px * px + pz * pz) * fyy + (px * px + py * py) * fzz -
                     2.0 * (px * py * d_xy + px * pz * d_xz + py * pz * d_yz)) /
                    (g * g * g);
  const double Kc = (pz * pz * (fxx * fyy - d_xy * d_xy) + px * px * (fyy * fzz - d_yz * d_yz) +
                     py * py * (fxx * fzz - d_xz * d_xz) +
                     2.0 * (py * d_xy * (pz * d_xz - px * fzz) + px * d_xz * (py * d_yz - pz * fyy) +
                            pz * d_yz * (px * d_xy - py * fxx))) /
                    (g * g * g * g);
  const double rate = sval(val, sp, SV_RATE, nxy, c) + (sp.add_nl ? surf_nl(psi, sp, col, k2) : 0.0);
  out[c] = (double)k2;
  out[nxy + c] = sp.vsign * rate / g;
  out[2 * nxy + c] = Hc;
  out[3 * nxy + c] = Kc;
  atomicAdd(nfound, 1u);
}

}  // namespace smx
