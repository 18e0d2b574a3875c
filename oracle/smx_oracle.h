/*
 * smx_oracle.h -- FP64 CPU oracle for the Swift-Hohenberg time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * loops of evitral/smectic3D with its own mixed-radix FP64 FFT (FFTW is not
 * available to this build).  Only the tests, the smoke check and the CPU
 * baseline leg of the benchmark may load it; the product library never links
 * or calls it.
 *
 * Layouts (same as the reference):
 *   real      [x][y][z], z contiguous, dense (row length nz)
 *   periodic  spectra [x][y][nz/2+1] complex (pf3d.cpp:403)
 *   Neumann   spectra [x][y][z] real (cosNoAdvMorph.cpp:296)
 */
#ifndef SMX_ORACLE_H
#define SMX_ORACLE_H

#include <stdint.h>
#include "../include/smectic.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- transforms (FFTW conventions, unnormalised) ---- */
void or_fft1d(int n, double *re_im, int sign);                      /* in place, interleaved complex */
int or_rfft3(int nx, int ny, int nz, const double *in, double *out); /* r2c 3d, out = interleaved complex */
int or_irfft3(int nx, int ny, int nz, const double *in, double *out);/* c2r 3d, input preserved */
int or_redft10_3d(int nx, int ny, int nz, const double *in, double *out); /* DCT-II all axes */
int or_redft01_3d(int nx, int ny, int nz, const double *in, double *out); /* DCT-III all axes */
void or_redft10_1d(int n, const double *in, double *out);
void or_redft01_1d(int n, const double *in, double *out);
void or_rodft10_1d(int n, const double *in, double *out); /* DST-II */
void or_rodft01_1d(int n, const double *in, double *out); /* DST-III */

/* ---- solver following the reference dataflow ---- */
typedef struct or_state or_state;
or_state *or_create(const smx_config *cfg);
void or_destroy(or_state *s);
int or_set_psi(or_state *s, const double *psi);    /* dense [x][y][z]; bootstrap NqPast */
int or_get_psi(const or_state *s, double *psi);
int or_step(or_state *s, int64_t nsteps);
int or_velocity(or_state *s, double *u);        /* Neumann Stokes velocity, 3 dense fields */
double or_l1(const or_state *s);                   /* over the last step */
double or_energy(const or_state *s);               /* spectral quadrature */
double or_energy_fd_trapz(const or_state *s);      /* energyPsi.m:217-234 */
double or_dominant_wavenumber(const or_state *s);
int or_dtpsi(or_state *s, double *out);
int or_surftrack(or_state *s, double arm, double vel_sign, double *surfz, double *vel, double *H,
                 double *K, int64_t *found);
int64_t or_step_count(const or_state *s);
/* Raw access for tests: periodic NqPast as interleaved complex, or Neumann NqPast real. */
int or_get_nqpast(const or_state *s, double *out);

/* ---- initial conditions (host) ---- */
double or_splitmix_uniform(uint64_t seed, uint64_t box, uint64_t idx);
void or_ic_noise(int nx, int ny, int nz, uint64_t seed, uint64_t box, double amp, double *psi);
void or_ic_hyperboloid(int nx, int ny, int nz, double dx, double dy, double dz, double q0, double a,
                       double noise, uint64_t seed, double *psi);

/* pf3dVal.cpp:214-329: growth rate at the box centre after 3 steps for
 * psi0 = 1e-4 cos(mq*k*dz), for nq values mq = mq0 + i*dmq. */
int or_growth_sweep(const smx_config *cfg, double mq0, double dmq, int nq, double *growth);

/* energyPsi.m:68-234 (Eckhaus scan): with dpsiz the script's 4-point z stencil of psi0
 * (:100-101, coefficients as written), energy[m] = FD + trapz energy (or_energy_fd_trapz)
 * of psi0 + amp[m] dpsiz cos((1 + q[m]) (k + 1) dz), k the 0-based z index (:118). */
int or_eckhaus_scan(const smx_config *cfg, const double *psi0, int nq, const double *q, const double *amp,
                    double *energy);

/* morphSurfAmp.cpp:1005-1134: local amplitude psiAvg of a dense [x][y][z] field --
 * trapezoid integral of psi^2 over the window of +-nw/2 points about each point (clipped
 * to the grid), over the same integral on the reference box [nw/2, 3nw/2]^3 times
 * l1 l2 l3 / lambda^3, lambda = 2 pi/q0.  Literal O(N^3 nw^3) restatement.  Returns -1
 * when the reference box does not fit the grid. */
int or_psi_avg(int nx, int ny, int nz, double dx, double dy, double dz, int nw, double q0, const double *psi,
               double *out);

/* Discrete adjoint of the periodic CNAB2 step (no advection/penalty):
 * J = 1/2 sum (psi_T - amp cos(q0 z))^2 and dJ/dpsi0, dJ/deps, dJ/dq0. */
int or_adjoint_grad(const smx_config *cfg, const double *psi0, int64_t T, double amp, double *loss, double *dpsi0,
                    double *deps, double *dq0);

#ifdef __cplusplus
}
#endif
#endif
