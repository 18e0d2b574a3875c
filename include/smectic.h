/*
 * smectic.h -- C ABI of libsmectic, the MI355X (gfx950) implementation of the
 * semi-implicit pseudo-spectral Swift-Hohenberg time step of evitral/smectic3D.
 *
 * The reference has no library API: every driver is a standalone main() whose
 * time loop calls FFTW-MPI plans on caller-owned arrays.  This ABI replaces the
 * loop body of those drivers (the boundary sits at the loop, not at
 * fftw_execute, so that pointwise and k-space work can be fused into the
 * transform passes).  Each entry point names the reference code it replaces.
 *
 * Conventions
 *  - Plain C types only.  Host pointers are borrowed for the duration of a call.
 *  - Return 0 on success, a negative SMX_ERR_* code on failure; the message is
 *    available from smx_last_error() (thread-local).
 *  - A handle owns its device buffers and one HIP stream.  A handle is not
 *    thread-safe; distinct handles may be used from distinct threads.
 *  - Real fields cross the boundary in the reference layout: x slowest, z
 *    contiguous, element (i,j,k) at (i*ny + j)*row_stride + k.  row_stride is
 *    nz (dense, as in cosFFT/mesabi/cosNoAdvMorph.cpp:296) or 2*(nz/2+1)
 *    (FFTW's padded r2c layout, periodicFFT/pf3d.cpp:312).
 */
#ifndef SMECTIC_H
#define SMECTIC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SMX_ABI_VERSION 1

/* boundary conditions */
#define SMX_BC_PERIODIC 0 /* R2C/C2R, periodicFFT/pf3d.cpp:226-291 */
#define SMX_BC_NEUMANN 1  /* REDFT10/REDFT01, cosFFT/mesabi/cosNoAdvMorph.cpp:249-262 */

/* arithmetic of the transforms and pointwise work */
#define SMX_FP64 0       /* everything in FP64 (the reference's arithmetic) */
#define SMX_FP32_MIXED 1 /* FP32 passes, FP64 twiddle tables, FP64 mean mode and reductions */

/* nonlinearity N(psi) */
#define SMX_NL_QUINTIC 0 /* beta*psi^3 - gamma*psi^5, periodicFFT/pf3d.cpp:460-461 */
#define SMX_NL_CUBIC 1   /* beta*psi^2 - gamma*psi^3 */

/* advective force fed to the Stokes solve */
#define SMX_FORCE_SH 0 /* S = psiSH * grad(psi), periodicFFT/pf3d.cpp:422-424 */
#define SMX_FORCE_MU 1 /* S = mu * grad(psi), mu = dF/dpsi */

/* error codes */
#define SMX_OK 0
#define SMX_ERR_INVALID (-1)     /* bad argument or configuration */
#define SMX_ERR_UNSUPPORTED (-2) /* valid but not implemented (e.g. a grid length without a radix plan) */
#define SMX_ERR_HIP (-3)         /* HIP runtime failure */
#define SMX_ERR_NONFINITE (-4)   /* state contains NaN/Inf */
#define SMX_ERR_STATE (-5)       /* call out of order (e.g. step before set_psi) */

typedef struct smx_handle smx_handle;

/*
 * Step operator and grid.  Replaces the compile-time constants of the drivers:
 * periodicFFT/pf3d.cpp:35-75 (grid, eps, alpha, beta, gamma, q0, eta, rho, dt,
 * box L = N*2*pi/(q0*Nw)), periodicFFT/pf3dMorph.cpp:94-102 (mid-plane
 * penalty wMid), cosFFT/mesabi/cosNoAdvMorph.cpp:90-112 (Neumann variant).
 */
typedef struct smx_config {
  int abi_version; /* must be SMX_ABI_VERSION */
  int bc;          /* SMX_BC_* */
  int precision;   /* SMX_FP64 or SMX_FP32_MIXED */
  int nx, ny, nz;  /* grid; each a product of 2s and 3s with a radix plan */
  double lx, ly, lz; /* box lengths; dx = lx/nx etc. */
  double eps, q0, alpha, beta, gamma, dt; /* L(k) = eps - alpha*(|k|^2 - q0^2)^2 */
  int nl_model;    /* SMX_NL_* */
  int advect;      /* 1: spectral Stokes advection, pf3d.cpp:399-465; with SMX_BC_NEUMANN the
                      sine/cosine projection of per-component parity (morphSurfK.cpp:331-348,
                      864-882 has the machinery, never applied, :896-898) */
  double eta, rho; /* Stokes viscosity and density, pf3d.cpp:54-55,210 */
  int force_form;  /* SMX_FORCE_* */
  double w_mid;    /* mid-plane penalty weight, pf3dMorph.cpp:102 (0 = off) */
  int mid_plane;   /* z index of the penalty plane (reference: nz/2); <0 = off */
  int device;      /* HIP device ordinal */
} smx_config;

/* Fill *cfg with the reference's pf3d.cpp constants at grid n^3 (Nw = 16). */
void smx_config_default(smx_config *cfg, int n);

/* Create a solver.  Replaces the set-up of pf3d.cpp:84-291 (MPI/FFTW init,
 * wavenumbers :160-190, operator tables :195-222, 13 plans :226-291). */
int smx_create(const smx_config *cfg, smx_handle **out);
void smx_destroy(smx_handle *h);
const char *smx_last_error(void);
/* Config the handle was created with. */
int smx_get_config(const smx_handle *h, smx_config *cfg);

/* Load psi (host, reference layout) and bootstrap the Adams-Bashforth history
 * from N(psi) without advection.  Replaces the IC scatter/load and the "First
 * NqPast" block: pf3d.cpp:363-374, pf3dMorph.cpp:416-472,
 * cosFFT/mesabi/cosNoAdvMorph.cpp:423-442. */
int smx_set_psi(smx_handle *h, const double *psi, int64_t row_stride);

/* Copy the current psi to the host in the reference layout.  Replaces the
 * Gather of psi to rank 0 before a dump, pf3d.cpp:526-527,554-555. */
int smx_get_psi(smx_handle *h, double *psi, int64_t row_stride);

/* Advance nsteps CNAB2 steps.  Replaces the body of the time loop:
 * pf3d.cpp:390-485,549 (periodic, with advection when cfg.advect),
 * pf3dMorph.cpp:501-610 (mid-plane penalty),
 * cosFFT/mesabi/cosNoAdvMorph.cpp:483-530,619 (Neumann; with cfg.advect the
 * same advected step on the cosine grid, velocity components as sine series
 * along their own axis).
 * Asynchronous on the handle's stream.  Every finite_every-th step (default 100, the
 * reference's L1 cadence pf3d.cpp:489) the carried state is checked for NaN/Inf without
 * waiting for the device; a check that found one makes the next smx_step and smx_sync
 * return SMX_ERR_NONFINITE (the reference's only guard is its L1 test, which a NaN state
 * turns into a NaN without stopping the loop).  With L1 tracking on (the default) the
 * call keeps one device copy of the state from before its last step for smx_l1. */
int smx_step(smx_handle *h, int64_t nsteps);

/* smx_step for callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  No counterpart in the reference.  The same
 * steps, kernels and results as smx_step, with two differences: the FP64 partial sums of the
 * z passes (FP32-mixed mean mode) go to a static device array of 512*512/8 + 64 doubles
 * (262,656 bytes, one per device, held until the process ends) instead of the handle's buffer,
 * and no L1 snapshot is kept (smx_l1 then returns SMX_ERR_STATE).  Allocates nothing and calls
 * no allocating routine.  SMX_ERR_INVALID when a grid plane needs more partial sums than the
 * array's 32,832 entries (nx*ny/8 + 64 of them; every supported plane fits).  Every call on one
 * device shares that device's array
 * and leaves its contents behind: calls that overlap on one device (handles on different
 * streams stepping at once) race. */
int smx_step_noalloc(smx_handle *h, int64_t nsteps);

/* L1 = sum|psi^{n+1}-psi^n| / (dt * sum|psi^{n+1}|) over the last step taken
 * by the most recent smx_step call, FP64 compensated.  Replaces the L1 block
 * pf3d.cpp:489-519 (Reduce x2 + Bcast).  SMX_ERR_STATE when L1 tracking is off or no
 * step was taken since the state was set; SMX_ERR_NONFINITE for a NaN/Inf state. */
int smx_l1(smx_handle *h, double *l1);

/* L1 tracking (on = 1, the default): smx_step snapshots the state before its last step
 * (one device copy of the spectral stage array per call) so that smx_l1 can compare the
 * two.  Off: no snapshot, smx_l1 returns SMX_ERR_STATE.  No counterpart in the reference
 * (it keeps psi and psiNew, pf3d.cpp:494-509). */
int smx_set_l1_tracking(smx_handle *h, int on);
/* Steps between the non-finite checks of smx_step (0 turns them off; default 100). */
int smx_set_finite_check(smx_handle *h, int64_t every);

/* Lyapunov energy F of the current state, FP64 compensated, spectral
 * (Parseval) quadrature of -eps/2 psi^2 + alpha/2((lap+q0^2)psi)^2 plus the
 * model's local terms (-beta/4 psi^4 + gamma/6 psi^6 for the quintic model,
 * energyPsi.m:217-228; -beta/3 psi^3 + gamma/4 psi^4 for the cubic one).
 * For SMX_BC_NEUMANN the quadrature uses the cosine basis. */
int smx_energy(smx_handle *h, double *F);

/* Power-weighted mean |k| of the current state (k = 0 excluded); the layer
 * wavenumber diagnostic used to compare runs (cf. periodicFFT/pf3dVal.cpp:214). */
int smx_dominant_wavenumber(smx_handle *h, double *q);

/* dpsi/dt = F^-1[L(k) psi_hat + N_hat] of the current state (host, reference
 * layout); with advection on, N_hat includes -U.grad psi of the current state
 * (surfTrack.cpp:585-606).  Replaces the dTpsi transform of
 * cosNoAdvMorphCurv.cpp:744,764 and surfTrack.cpp:606,622. */
int smx_dtpsi(smx_handle *h, double *dtpsi, int64_t row_stride);

/* Eckhaus scan of energyPsi.m:68-234 on the current state psi0: energy[m] is the
 * finite-difference + trapz energy (periodic second differences, trapezoid weights 1/2
 * at both ends of every axis, energyPsi.m:217-234) of psi0 + amp[m] dpsiz
 * cos((1 + q[m]) (k + 1) dz), with dpsiz the script's 4-point z stencil of psi0 exactly
 * as written (:100-101) and k the 0-based z index (:118); amp[m] = 0 gives psi0's own
 * energy, the scan's first entry.  All nq energies come from ceil(nq/4) FP64 passes over
 * the GPU-resident field (the script re-evaluates a full energy per Q on the host). */
int smx_eckhaus_scan(smx_handle *h, int nq, const double *q, const double *amp, double *energy);

/* Local amplitude psiAvg of the current state (host, reference layout): the trapezoid
 * integral of psi^2 over the window of +-nw/2 points about each point (clipped to the
 * grid), divided by the same integral over the reference box [nw/2, 3nw/2]^3 times
 * l1 l2 l3 / lambda^3 (l: clipped window lengths, lambda = 2 pi/q0).  Replaces the
 * O(N^3 nw^3) loops of cosFFT/comet/morphSurfAmp.cpp:1005-1134 (rank 0, lambda^3 = 8 pi^3
 * for q0 = 1) by three separable window passes on the GPU in FP64.  SMX_ERR_INVALID unless
 * nw >= 2 and 3(nw/2) <= n - 1 on every axis. */
int smx_psi_avg(smx_handle *h, int nw, double *out, int64_t row_stride);

/* surfTrack column scan at the psi = 0 crossing of every (i,j) column.
 * Periodic: surfTrack.cpp:767-806 (scan k upward, arm at psi > arm, spectral
 * derivatives, H ~ (psi_xx+psi_yy)/|grad psi|, K = (psi_xx psi_yy - psi_xy^2)/|grad psi|^2).
 * Neumann: cosNoAdvMorphCurv.cpp:811-966 (scan k downward, arm at psi > arm,
 * central finite-difference gradients, Megrabov mean curvature
 * H = div(grad psi/|grad psi|) and Gaussian curvature K).
 * Outputs are nx*ny arrays in (i*ny + j) order.  Like the reference's
 * persistent surf arrays, a column without a crossing keeps the value the
 * handle last stored for it (zero-initialised).  *found = number of columns
 * that crossed on this call.  The derivative fields are evaluated only at each
 * column's crossing (x and y transforms of the whole field, the z transform at one
 * point); the first call allocates the scratch (real psi, an FP64 copy of a periodic
 * FP32 stage, up to 112 values and 2 indices per column, a 4*nz table) and a mapped
 * pinned host buffer of 4*nx*ny+1 doubles, both kept by the handle.  Synchronous. */
typedef struct smx_surf_opts {
  double arm;      /* arming threshold: 0.1 (surfTrack.cpp:775) or 0.7 (cosNoAdvMorphCurv.cpp:889) */
  double vel_sign; /* +1: v = +dtpsi/|grad psi| as in the reference; -1: v = -dtpsi/|grad psi| */
} smx_surf_opts;
int smx_surftrack(smx_handle *h, const smx_surf_opts *opts, double *surfz, double *vel,
                  double *H, double *K, int64_t *found);

/*
 * Gradient of J = 1/2 sum_x (psi_T(x) - amp cos(q0 z))^2, z = k dz, after nsteps periodic
 * CNAB2 steps from psi0 (host, reference layout), with respect to psi0 (host array,
 * same layout), eps and q0 (the grid spacing does not move with q0).  Hand-written
 * discrete adjoint of the step (the transpose of pf3d.cpp:471-481 with the bootstrap
 * :363-374); the forward trajectory is checkpointed every ckpt_every steps and each
 * segment is recomputed once.  No counterpart in the reference (it has no adjoint);
 * the periodic, non-advected, penalty-free step only.  Leaves the handle at psi0.
 */
int smx_adjoint_grad(smx_handle *h, const double *psi0, int64_t row_stride, int64_t nsteps, int64_t ckpt_every,
                     double target_amp, double *loss, double *dpsi0, double *deps, double *dq0);
/* Device memory the adjoint may spend on keeping recorded segments resident (N'(psi_n) and
 * psi_hat_n of whole checkpoint segments, recorded during the forward sweep and never
 * recomputed).  bytes < 0 (default): up to 80% of the free device memory, for the duration
 * of one smx_adjoint_grad call -- the workspace is released when the call returns;
 * bytes >= 0: at most that many bytes, and the workspace stays allocated in the handle so
 * that repeated calls of the same length reuse it (0: checkpoints only, every segment
 * recomputed once).  The gradient does not depend on it.  smx_destroy and the next call of
 * this function release a kept workspace.  No counterpart in the reference. */
int smx_set_adjoint_budget(smx_handle *h, int64_t bytes);
/* Bookkeeping of the last smx_adjoint_grad: steps recorded in the forward sweep, steps
 * recomputed from checkpoints, device bytes allocated for the sweep. */
int smx_adjoint_info(const smx_handle *h, int64_t *resident_steps, int64_t *recomputed_steps, int64_t *device_bytes);

/*
 * Ensembles of independent boxes of uneven sizes (parameter sweeps).  Each box is
 * a full solver (its own config and buffers, on its own stream); smx_batch_run
 * advances every box by nsteps through one captured hipGraph whose parallel branches
 * (fork/join on the batch stream) are the step chains of single boxes and of groups:
 * periodic, non-advected boxes of one precision, grid and tile schedule advance with one
 * grouped launch per pass (equal to their single-box runs to rounding: a few multiply-adds
 * are fused differently).  No counterpart in the reference (it runs one box per MPI job).
 */
typedef struct smx_batch smx_batch;
int smx_batch_create(int nbox, const smx_config *cfgs, smx_batch **out);
void smx_batch_destroy(smx_batch *b);
/* Per-box handle for set/get/diagnostics; owned by the batch. */
int smx_batch_handle(smx_batch *b, int box, smx_handle **h);
/* Advance every box by nsteps: graph launches of `graph_steps` steps each
 * (captured once per batch) plus a direct remainder. */
int smx_batch_run(smx_batch *b, int64_t nsteps, int64_t graph_steps);
/* Number of kernel nodes in the captured graph (0 before the first run). */
int smx_batch_graph_nodes(smx_batch *b, int64_t *nodes);

/*
 * Gradients of every box of a batch (periodic, non-advected, penalty-free boxes): for box i,
 * J_i = 1/2 sum_x (psi_T(x) - target_amp[i] cos(q0_i z))^2 after nsteps CNAB2 steps from
 * psi0[i] (host, reference layout, row stride >= that box's nz), and dJ_i/dpsi0 (into
 * dpsi0[i]), dJ_i/deps, dJ_i/dq0 -- smx_adjoint_grad of each box, with checkpoints every
 * ckpt_every steps, but the adjoint sweeps of all boxes are the parallel branches of one
 * captured hipGraph, launched once (boxes of one grid, schedule and record layout sweep
 * together through grouped launches, as in smx_batch_run).  Record budgets: a box with an explicit
 * smx_set_adjoint_budget uses it and keeps its workspace (the batch then keeps the captured
 * graph for the next call of the same shape when every box does); the other boxes share
 * 80% of the free device memory in proportion to the records they need and release their
 * workspaces on return.  Each box is left at its psi0.  No counterpart in the reference
 * (it has no adjoint and runs one box per MPI job; the step is pf3d.cpp:471-481). */
int smx_batch_adjoint_grad(smx_batch *b, const double *const *psi0, int64_t row_stride, int64_t nsteps,
                           int64_t ckpt_every, const double *target_amp, double *loss, double *const *dpsi0,
                           double *deps, double *dq0);
/* Kernel and copy nodes of the last captured adjoint batch graph (0 before the first call). */
int smx_batch_adjoint_nodes(smx_batch *b, int64_t *nodes);

/*
 * Tile schedules of the four kernels of the step, in launch order (slot 0: inverse y pass, 1: z
 * pass with N(psi), 2: forward y pass, 3: x pass with the CNAB2 update).  Periodic handles: the
 * columns per tile (8, 16 or 32) and the radix-plan variant (0: default; 1: the alternative
 * plan of the power-of-two lengths); slot 3 also takes variants 2 to 7: plans 0 and 1 with the x
 * pass's R stores issued after its inverse transform (2, 3), with all of R read before any of it
 * is rewritten (4, 5), or with the new R staged in LDS and written after the update loop (6, 7),
 * instead of R read and rewritten element by element (bit-identical results).  The default is
 * nc = 16, variant = 0 in every slot.  The advected step uses the schedules of slots 0 and 2 for
 * its eleven y passes.  Neumann handles: the tile width only (variant 0): 32 (default) or 16
 * columns in FP32, 16 (default) or 8 in FP64; the advected Neumann step keeps the defaults.
 * The step's result depends on the schedule only through rounding.
 */
typedef struct smx_schedule {
  int nc[4];
  int variant[4];
} smx_schedule;
int smx_get_schedule(const smx_handle *h, smx_schedule *s);
/* Install a schedule; SMX_ERR_UNSUPPORTED if some slot has no kernel for this grid (too
 * many threads or too much LDS per tile, or ny / nx*ny not a multiple of nc), or if a
 * non-default schedule is given to an advected Neumann handle. */
int smx_set_schedule(smx_handle *h, const smx_schedule *s);
/* Autotuner: slot by slot, time every valid candidate (reps whole step chains each -- up to
 * 16x more on grids smaller than 256^3, whose kernels are short --, HIP
 * events on the handle's stream) on a scratch copy of the state -- hashed noise when none
 * is set; the state itself is not touched -- install the fastest and return it in *chosen
 * (may be NULL); with SMX_TUNE_VERBOSE=1 in the environment each candidate's median is
 * printed to stderr.  Results are cached per (device,
 * precision, grid) for the life of the process.  Replaces the FFTW_MEASURE planning of
 * pf3d.cpp:226-291 (which times candidate FFT algorithms at set-up).  An advected Neumann
 * handle keeps its fixed tiles (the call returns them); not callable during stream capture. */
int smx_autotune(smx_handle *h, int reps, smx_schedule *chosen);
/* smx_autotune on every box (boxes of one shape are tuned once; advected Neumann boxes keep
 * their fixed tiles); the next smx_batch_run captures its graph again with the new schedules. */
int smx_batch_autotune(smx_batch *b, int reps);

/* Block until the handle's stream is idle.  SMX_ERR_NONFINITE if a step check of an
 * earlier smx_step found NaN/Inf in the state. */
int smx_sync(smx_handle *h);

/*
 * Exact checkpoint of the solver state -- the carried CNAB2 arrays (stage buffer and
 * R = C1 psi_hat - dt/2 N_hat_past; with advection also the two derivative stages), the FP64
 * mean mode and the step count -- so that a run resumed from it continues bit for bit.
 * Replaces the reference's restart from a 6-digit text dump of psi, which re-bootstraps
 * N_hat_past from N(psi) without advection (pf3dMorph.cpp:416-472) and is therefore not
 * continuation-exact.  Binary file: header (magic "SMXCKPT1", ABI version, smx_config, step
 * count, array sizes), the arrays, the mean mode and an FNV-1a 64-bit checksum.  Loading
 * fails with SMX_ERR_INVALID for a file of another grid, precision or operator (the configs
 * must agree in every field but device) or with a bad checksum.
 */
int smx_save_checkpoint(smx_handle *h, const char *path);
int smx_load_checkpoint(smx_handle *h, const char *path);

/* Number of steps taken since the last smx_set_psi. */
int smx_step_count(const smx_handle *h, int64_t *n);

/*
 * Measurement.  Times every kernel launch of nsteps steps with HIP events on
 * the handle's stream and reports, per kernel slot of the step chain, the
 * average launch duration (ms) and the algorithmic bytes per launch.  names
 * receives '\n'-separated kernel slot names (at most names_cap bytes).
 * No counterpart in the reference (its only timers are the wall/CPU clocks of
 * cosFFT/comet/cosNoAdvMorph.cpp:562-571,765-776).
 */
int smx_profile_step(smx_handle *h, int64_t nsteps, int max_slots, int *nslots, double *avg_ms,
                     double *alg_bytes, char *names, int names_cap);

/* Plain device copy kernel over nbytes (grid-stride float4 copy of hashed,
 * non-zero words), used to calibrate hardware byte counters against a kernel of
 * known traffic and as the achievable-bandwidth yardstick: avg_ms is the
 * launch time of the fastest of three launch shapes (2048 and 8192 blocks of
 * 256 threads, and one float4 per thread). */
int smx_calibration_copy(smx_handle *h, int64_t nbytes, int reps, double *avg_ms);

#ifdef __cplusplus
}
#endif

#endif /* SMECTIC_H */
