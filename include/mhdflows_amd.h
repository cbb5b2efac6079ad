/*
 * MI355X-native MHDFlows hot path: C ABI of libmhdflows_amd.so.
 *
 * The library implements the 3D periodic incompressible HD/MHD right-hand side
 * and its RK4 / LSRK54 time step of MHDFlows.jl v0.2.2 on one AMD Instinct
 * MI355X (gfx950).  Every entry point takes plain pointers and sizes; no C++ or
 * torch types cross this boundary.  All functions return 0 on success or a
 * negative MHD_ERR_* code; mhd_last_error() gives the message of the calling
 * thread's last failure.
 *
 * Array layout at the boundary is byte-identical to the reference's Julia
 * arrays: the state `sol` is Complex{T}[kr, l, m, f] column-major, i.e. C order
 * [f][z][y][kx] with nkr = nx/2+1 and complex values interleaved (re, im); real
 * fields are T[x, y, z] column-major, i.e. C order [z][y][x].  Field order
 * f = ux, uy, uz, bx, by, bz (src/Structure/datastructure.jl:90).
 *
 * Ownership: a handle owns all of its device memory.  Host pointers passed in are
 * copied before the call returns and never retained.  Device pointers passed in
 * (mhd_calcN and friends) must hold the documented number of fields.
 *
 * Threading: one handle is used by one host thread at a time; distinct handles are
 * independent.  Work is ordered on the handle's HIP stream.  Calls that return
 * host data synchronize that stream.
 */
#ifndef MHDFLOWS_AMD_H
#define MHDFLOWS_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MHD_ABI_VERSION 1

enum mhd_status {
  MHD_OK = 0,
  MHD_ERR_INVALID = -1,      /* invalid argument (e.g. odd grid size, bad field index) */
  MHD_ERR_UNSUPPORTED = -2,  /* grid length without a compiled transform */
  MHD_ERR_BUDGET = -3,       /* device memory needed exceeds the configured HBM budget */
  MHD_ERR_HIP = -4,          /* HIP runtime error (out of memory, launch failure, no device) */
  MHD_ERR_NAN = -5,          /* NaN detected in the fields */
};

enum mhd_precision { MHD_F32 = 0, MHD_F64 = 1 };
enum mhd_physics { MHD_PHYS_HD = 0, MHD_PHYS_MHD = 1 };
enum mhd_stepper { MHD_RK4 = 0, MHD_LSRK54 = 1 };

typedef struct mhd_handle mhd_handle;

/* Problem(dev; nx, ny, nz, Lx, Ly, Lz, ν, nν, η, nη, B_field, stepper, T, dt)
 * -- src/pgen.jl:64-127.  As in the reference, nη is not forwarded to the
 * parameters (src/pgen.jl:114-116), so the resistive hyper term stays off. */
typedef struct mhd_config {
  int32_t nx, ny, nz;        /* even grid sizes */
  double Lx, Ly, Lz;         /* box lengths (2π by default in the reference) */
  double nu, eta;            /* viscosity / resistivity, kept in Float64 (MHDParams.jl:45-47) */
  int32_t n_nu, n_eta;       /* hyper-viscosity order (active when > 1) */
  int32_t precision;         /* mhd_precision: T = Float32 or Float64 */
  int32_t physics;           /* mhd_physics: B_field = false / true */
  int32_t stepper;           /* mhd_stepper */
  int32_t device;            /* HIP device ordinal */
  void* stream;              /* hipStream_t to order work on, or NULL for a library-owned stream */
  uint64_t hbm_budget;       /* bytes of device memory the handle may use; 0 = no limit */
  double dt;                 /* initial clock.dt */
} mhd_config;

/* Diagnostics in Float64 (ProbDiagnostic, src/utils/UserInterface.jl:65-86, without
 * the 3-significant-digit display rounding; helicities as src/utils/MHDAnalysis.jl:94-174). */
typedef struct mhd_diag {
  double E_kin;      /* Σ|u|² dV (no 1/2, as the reference) */
  double E_mag;      /* Σ|b|² dV */
  double H_kin;      /* Σ u·(∇×u) dV */
  double H_mag;      /* Σ A·b dV, Coulomb gauge: the volume integral; the reference's h_m
                        (MHDAnalysis.jl:113-117) returns the density A·b without dV, so
                        sum(h_m) = H_mag / dV, while sum(h_k) = H_kin (h_k includes dV) */
  double H_cross;    /* Σ u·b dV */
  double max_div_b;  /* max |∇·b| (Div, src/utils/VectorCalculus.jl:72-94) */
  double max_div_u;  /* max |∇·u| */
  int32_t has_nan;   /* 1 if E_kin or E_mag is NaN */
  int32_t pad;
} mhd_diag;

/* Per-kernel timing collected with HIP events on the handle's stream. */
typedef struct mhd_ktime {
  char name[32];
  int64_t launches;
  double total_ms;
  double bytes;      /* algorithmic bytes moved by all timed launches of this kernel */
} mhd_ktime;

/* ---- library ---------------------------------------------------------------- */
int32_t mhd_abi_version(void);
const char* mhd_last_error(void);
/* 1 if a transform of length n is compiled in (n even, 8..1024, factors 2 and 3). */
int32_t mhd_supported_length(int32_t n);

/* ---- problem lifetime: Problem(dev; ...) src/pgen.jl:64-127 ------------------ */
/* Device bytes a handle with this configuration allocates. */
int32_t mhd_query_bytes(const mhd_config* cfg, uint64_t* bytes);
int32_t mhd_create(const mhd_config* cfg, mhd_handle** out);
int32_t mhd_destroy(mhd_handle* h);

/* ---- state boundary ----------------------------------------------------------- */
/* SetUpProblemIC!(prob; ux..bz) src/utils/IC.jl:41-109: copy real host fields
 * (nf = 3 HD / 6 MHD entries, each nx*ny*nz values of T; NULL entries are skipped)
 * into vars and R2C them into sol. */
int32_t mhd_set_ic_real(mhd_handle* h, const void* const* fields);
/* Copy prob.sol (nf*nz*ny*nkr complex T) from / to host memory. */
int32_t mhd_get_sol(mhd_handle* h, void* host_out);
int32_t mhd_set_sol(mhd_handle* h, const void* host_in);
/* Zero prob.sol and write nmodes sparse half-spectrum entries: idx holds (kx, l, m)
 * array indices per entry, values holds nf complex T per entry (field-fastest).
 * vars are refreshed from the dealiased state (a spectral SetUpProblemIC!). */
int32_t mhd_set_sol_modes(mhd_handle* h, int64_t nmodes, const int32_t* idx, const void* values);
/* mhd_set_sol_modes that allocates nothing.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (tight serving loops admitting requests,
 * code that forbids hipMalloc).  The staging of idx and values lives in a library-owned
 * 4 MiB device array that each GPU holds until the process ends, so nmodes is limited to
 * what fits (about 38800 Float64 MHD entries, 69900 Float32 MHD entries); larger calls
 * fail with MHD_ERR_INVALID.  All calls on one GPU share that array: calls that overlap
 * on one GPU (from several threads or handles) race.  sol is zeroed by a fill kernel, not
 * a memset.  Same result as mhd_set_sol_modes. */
int32_t mhd_set_sol_modes_noalloc(mhd_handle* h, int64_t nmodes, const int32_t* idx, const void* values);
/* Read nmodes entries of prob.sol: idx holds (kx, l, m) array indices per entry, values_out
 * receives nf complex T per entry (field-fastest).  A partial read of the state (no
 * counterpart in the reference, which copies whole arrays), e.g. the modes a coarser grid
 * keeps, without moving the whole half spectrum over PCIe. */
int32_t mhd_get_sol_modes(mhd_handle* h, int64_t nmodes, const int32_t* idx, void* values_out);
/* Copy prob.vars real field f (0..nf-1) to host memory. */
int32_t mhd_get_vars(mhd_handle* h, int32_t field, void* host_out);
/* Device pointers of the handle's sol (nf fields) and vars (nf real fields). */
int32_t mhd_device_ptrs(mhd_handle* h, void** sol, void** vars);
/* Clock (src/Problems.jl:53-60): t and dt as stored in T, step counter. */
int32_t mhd_get_clock(mhd_handle* h, double* t, double* dt, int64_t* step);
int32_t mhd_set_clock(mhd_handle* h, double t, double dt, int64_t step);

/* ---- operator boundary (device pointers, reference layout) --------------------- */
/* MHDcalcN!/HDcalcN!(N, sol, t, ...) src/pgen.jl:153-181: dealias sol in place,
 * refresh vars, write the right-hand side of all nf fields into N (forcing
 * added for MHD after the advection; for HD it is added first and then
 * discarded, as in the reference). */
int32_t mhd_calcN(mhd_handle* h, void* N_dev, void* sol_dev, double t);
/* MHDcalcN_advection! / HDcalcN_advection! src/Solver/MHDSolver.jl:330-351,
 * src/Solver/HDSolver.jl:95-108: no dealiasing, no forcing. */
int32_t mhd_advection(mhd_handle* h, void* N_dev, void* sol_dev, double t);
/* UᵢUpdate!(N, ...; direction) src/Solver/MHDSolver.jl:27-103 (HD: HDSolver.jl:25-93)
 * and BᵢUpdate! MHDSolver.jl:106-177: read the handle's vars only and write
 * component dir (0=x, 1=y, 2=z) of N (field dir, resp. 3+dir). */
int32_t mhd_u_update(mhd_handle* h, void* N_dev, int32_t dir);
int32_t mhd_b_update(mhd_handle* h, void* N_dev, int32_t dir);

/* ---- time stepping ------------------------------------------------------------ */
/* stepforward! src/timestepper/timestepper.jl:4-6 with the configured stepper,
 * nsteps times with the current clock.dt (no host synchronisation). */
int32_t mhd_step(mhd_handle* h, int32_t nsteps);
/* getCFL! src/integrator.jl:158-198: set clock.dt from the current vars. */
int32_t mhd_cfl_dt(mhd_handle* h, double cfl_coef, double* dt_out);
/* TimeIntegrator!(prob, t0, N0; usr_dt, CFL_Coef) hot loop src/integrator.jl:31-156:
 * clock.step = 0, then while N0 >= step && t0 >= t: [getCFL! unless usr_dt != 0],
 * stepforward!.  dt is computed on the device; with t0 = +Inf no host
 * synchronisation happens inside the loop.  Returns the number of steps taken in
 * *steps_out (may be NULL). */
int32_t mhd_integrate(mhd_handle* h, double t0, int64_t N0, double usr_dt, double cfl_coef,
                      int64_t* steps_out);
/* Capture one step into a HIP graph and replay it for subsequent mhd_step calls
 * (enable = 1) or launch kernels one by one (enable = 0). */
int32_t mhd_use_graph(mhd_handle* h, int32_t enable);
/* Dealiased steps (default enable = 1): a stage input is dealiased before use, so the
 * right-hand side at aliased modes never reaches a kept mode; it only sets the aliased
 * modes of sol at the end of a step, which the next step discards.  Every step of an
 * mhd_step / mhd_integrate (t0 = Inf) call except the last therefore transforms, stores
 * and updates the kept modes only; the last step is computed in full, so the state seen
 * after the call equals the one every-step-full stepping gives (enable = 0).  Has no
 * counterpart in the reference, which computes every step in full. */
int32_t mhd_set_step_pruning(mhd_handle* h, int32_t enable);

/* ---- diagnostics and forcing -------------------------------------------------- */
/* FP64 reductions over the current vars (the reference's diagnostics read vars). */
int32_t mhd_diagnostics(mhd_handle* h, mhd_diag* out);
/* Refresh vars from sol (C2R of the dealiased state) without stepping. */
int32_t mhd_refresh_vars(mhd_handle* h);
/* calcF! hook (src/pgen.jl:231-234).  A problem has one calcF!: each call below replaces the
 * forcing set before; forcing acts on the MHD right-hand side after the advection and is
 * discarded for HD, as in the reference (src/pgen.jl:173-181 adds it before
 * HDcalcN_advection! zeroes N).  Module buffers are allocated when a module is set.
 *
 * Seeded time-independent solenoidal forcing on the shell round(|k|) == kf with amplitude
 * amp (the large-scale forcing of the 1024^3 precision run; amp = 0 turns every forcing off). */
int32_t mhd_set_forcing(mhd_handle* h, uint64_t seed, double kf, double amp);
/* N97ForceDriving! (src/pgen/TaylorGreenDynamo.jl:12-16) with SetUpN97!(prob; F0, kf) (:18-33),
 * or any static forcing field: f3 holds 3 real fields (nx*ny*nz values of T each, NULL = zero;
 * for N97 the caller evaluates F0 sin(kf x) cos(kf y) cos(kf z) and -F0 cos(kf x) sin(kf y)
 * cos(kf z) on the grid in T, as SetUpN97! does), transformed on the device and added to the
 * velocity right-hand side, N_u += R2C(f).  Allocates 3 complex fields. */
int32_t mhd_set_forcing_field(mhd_handle* h, const void* const* f3);
/* NDForceDriving! (src/pgen/NegativeDamping.jl:23-44) with SetUpND!(prob, P, fx, fy, fz)
 * (:16-22): every right-hand side adds A R2C(f_i u_i) to N_u_i, A = P / (sum|u_i^2 f_i| dV)
 * with u the real fields of the stage input; the sum is accumulated in Float64 (the
 * reference sums in T).  f3: 3 real fields (NULL = zero).  Single-run problems only. */
int32_t mhd_set_forcing_nd(mhd_handle* h, double P, const void* const* f3);
/* Volume penalisation (Problem(...; VP_method = true), src/pgen.jl:84; MHDParams_VP /
 * HDParams_VP src/Structure/MHDParams.jl:1-40): every right-hand side adds VP_UᵢUpdate! and,
 * for MHD, VP_BᵢUpdate! (src/Solver/VPSolver.jl:21-56): N_a += -sum_j (delta_aj - k_j k_a
 * k^-2) R2C(chi/eta (u_j - U0_j)) (b, B0 likewise) with eta = clock.dt * 13 / 7, in T.
 * chi (params.χ), U0[3] (params.U₀x/y/z, set by SetUpProblemIC!'s U₀x..U₀z), B0[3] (MHD):
 * real fields of nx*ny*nz T; a NULL pointer keeps that field (zero when first enabled).
 * enable = 0 turns it off.  Single-run problems; not combined with negative damping; the
 * per-direction updates (mhd_u_update, mhd_b_update) refuse VP problems.  TimeIntegrator!'s
 * DivVCorrection!/DivBCorrection! around every step (src/integrator.jl:84-87,118-121) belong
 * to the host loop (mhd_div_correction between mhd_step calls); mhd_step and mhd_integrate
 * do not apply them. */
int32_t mhd_set_vp(mhd_handle* h, int32_t enable, const void* chi, const void* const* U0, const void* const* B0);
/* A99ForceDriving! (src/pgen/A99ForceDriving.jl:33-60) with SetUpFk(prob; kf, P, σ²)
 * (:93-127), usr_vars.b = b and usr_vars.A = A.  The reference draws unseeded rand() per
 * mode and call; here the four draws of a mode (phase 1, the complex Φ, phase 2) are
 * counter-based: splitmix64(seed·0x9E3779B97F4A7C15 + (counter << 32) + 1 + mode index),
 * counter = 4·c + draw, c = 8·clock.step + stage inside a step and 2^40 + n for the n-th
 * mhd_calcN call of the handle; mode index = (m·ny + l)·nkr + kx. */
int32_t mhd_set_forcing_a99(mhd_handle* h, uint64_t seed, double kf, double P, double sigma2, double b, double A);

/* ---- analysis of the current vars on the device ------------------------------- */
/* spectralline(A) src/utils/MHDAnalysis.jl:236-255 for every vars field: *nbins = krmax =
 * round(max|k| + 1); out (when not NULL, nbins_cap >= *nbins) receives nf + 1 rows of nbins:
 * row f = sum of |R2C(vars f)|^2 over the half-spectrum modes with round(|k|) = r (round half
 * to even, |k| in T), row nf = the number of those modes (the reference's kr marks the shells
 * that have any).  Accumulated in Float64 (the reference sums in T). */
int32_t mhd_spectra(mhd_handle* h, double* out, int32_t nbins_cap, int32_t* nbins);
/* ScaleDecomposition(B1, B2, B3; kf = [k1, k2]) src/utils/MHDAnalysis.jl:55-82 of vars fields
 * field0 .. field0 + ncomp - 1 (1 to 3 components of u or of b): modes with k1 <= |k| <= k2
 * kept, back to real space; ncomp real fields of nx*ny*nz T to host_out. */
int32_t mhd_scale_decomposition(mhd_handle* h, int32_t field0, int32_t ncomp, double k1, double k2, void* host_out);
/* Vector calculus of host real fields (nx*ny*nz T each, the problem's grid) on the device, in T
 * as the reference evaluates them: op 3/4/5 = ∂i(B1, "x"/"y"/"z") (src/utils/VectorCalculus.jl:
 * 101-123, 1 in, 1 out), 6 = LaplaceSolver(B) (:155-168: Φ = C2R(R2C(B) k^-2), the reference's
 * sign, 1 in, 1 out), 7 = Curl(B1, B2, B3) (:27-50, 3 in, 3 out), 8 = Div(B1, B2, B3) (:70-92,
 * 3 in, 1 out), 9 = VectorPotential(B1, B2, B3) (src/utils/MHDAnalysis.jl:156-180, Coulomb
 * gauge, 3 in, 3 out).  A NULL input is a zero field.  Uses the problem's work space (not
 * during a call in flight on the same handle). */
int32_t mhd_field_op(mhd_handle* h, int32_t op, const void* const* in, void* const* out);
/* DivVCorrection!(prob) (field = 0) / DivBCorrection!(prob) (field = 1, MHD)
 * src/Solver/VPSolver.jl:60-139 (exported by src/MHDFlows.jl:62-63): remove the gradient part
 * of u (b) in sol, phi = -i (k . F) k^-2, F -= i k phi (T arithmetic, k^-2 = 0 at k = 0), then
 * refresh the three vars fields from all modes (ldiv! of the corrected spectrum). */
int32_t mhd_div_correction(mhd_handle* h, int32_t field);
/* Velocity structure functions src/utils/TurbStatTool.jl: SFV(R) = sum_c SFC(v_c)(R) with
 * SFC(V) = 2 (mean(V) - CF(V)), CF(V) = fftshift(real(ifft(|fft(V)|^2))) (:64-71), binned over
 * the lags R of the fftshifted cube with the reference's index arithmetic (idx = i - nx/2 at
 * the 1-based fftshifted position i, so idx = lag + 1).  mode 0: SF2 1D (:88-117), bin
 * round(|(idx, jdx, kdx)|) = 1 .. nbins, nbins = 2 ceil(|(nx/2, ny/2, nz/2)|).  mode 1:
 * SFr_Global (:119-160, MHD, nonzero mean field): bin (round(|R cos theta|), round(|R sin
 * theta|)) with theta the angle to (mean(bx), mean(by), mean(bz)) in T, an nrow x nrow matrix
 * (nrow = ceil(|(nx/2, ny/2, nz/2)|) + 1) stored column-major as Julia's SFVr[rpar, rperp].
 * out[0, nbins) receives sum / count per bin (NaN for an empty bin, as the reference's 0/0),
 * out[nbins, 2 nbins) the counts; out == NULL only reports *nbins and *nrow.  Sums in Float64
 * (the reference sums SFV of T values in Float64 too). */
int32_t mhd_structure_function(mhd_handle* h, int32_t mode, double* out, int32_t cap, int32_t* nbins,
                               int32_t* nrow);

/* ---- checkpoint / restart ------------------------------------------------------ */
/* savefile(prob, file_number; file_path_and_name) src/integrator.jl:259-288: write the real
 * fields prob.vars as datasets i/j/k_velocity and i/j/k_mag_field (MHD) and clock.t as time.
 * libhdf5 is not available, so the file is a small self-describing container (layout in
 * mhd_ckpt.hip): named datasets with Julia-order dimensions, readable without the library.
 * exact = 1 adds the spectral state sol and the clock (step, dt and the maxima getCFL! uses
 * next), so that mhd_restart(.., exact = 1) continues bit for bit like an uninterrupted run
 * (the reference's restart, from real fields, is not bit exact). */
int32_t mhd_savefile(mhd_handle* h, const char* path, int32_t exact);
/* Restart!(prob, file_path_and_name) src/integrator.jl:208-257: exact = 0 reads the real fields
 * (converted to T), sets vars, R2Cs them into sol (Move_Data_to_Prob!, src/utils/utils.jl:172-176)
 * and sets clock.t; exact = 1 restores sol and the whole clock saved with exact = 1. */
int32_t mhd_restart(mhd_handle* h, const char* path, int32_t exact);

/* ---- measurement -------------------------------------------------------------- */
/* Autotuner: time every compiled variant of the hot launches (column-pass strip
 * widths, row-transform plans) with reps launches each, keep the fastest for this
 * grid and precision (for the process) and rebuild the captured step.  A text
 * report of the timings is written to report (may be NULL). */
int32_t mhd_autotune(mhd_handle* h, int32_t reps, char* report, int32_t report_len);
/* Slab pipeline of each right-hand side: the y-inverse pass, the row kernel and the
 * y-forward pass run over `slabs` slabs of z-planes as independent chains on two streams, so
 * one slab's row kernel (arithmetic-bound) overlaps another slab's y passes (memory-bound);
 * 1 turns it off.  The results are bitwise the same.  No counterpart in the reference (a
 * launch schedule of this library). */
int32_t mhd_set_pipeline(mhd_handle* h, int32_t slabs);
/* Set (v5 != NULL; negative entries keep the current value) and/or read (out5 != NULL)
 * the launch variants {zpass_inv, ypass_inv, ypass_fwd, zpass_fwd: strip width
 * 0 = 128 B, 1 = 64 B, 2 = 256 B, 3 = 512 B of kx per row; xrow_fused: bit 0 plan (Float32
 * powers of two: 0 = eight, 1 = four elements per thread), bit 1 XOR-swizzled LDS index
 * (0..3)}.  Results do not depend on the variant beyond floating-point rounding of the
 * transform order. */
int32_t mhd_set_variants(mhd_handle* h, const int32_t* v5, int32_t* out5);
/* Time nsteps steps with HIP events on the handle's stream (whole-step wall time). */
int32_t mhd_time_steps(mhd_handle* h, int32_t nsteps, double* ms_out);
/* Run nsteps steps with an event pair around every kernel launch and report per
 * kernel class launches, time and algorithmic bytes (at most max entries). */
int32_t mhd_profile_steps(mhd_handle* h, int32_t nsteps, mhd_ktime* out, int32_t max, int32_t* n_out);
/* Reference streaming copy of `bytes` bytes (one 16 B element per lane), timed over reps launches:
 * average ms per launch.  Used to calibrate HBM byte counters and as the practical
 * bandwidth ceiling next to the 8 TB/s peak. */
int32_t mhd_bench_copy(uint64_t bytes, int32_t reps, double* ms_out);
/* Algorithmic bytes of one step (the 3-pass model: MHD RK4 = 438 V_c + 120 V_r). */
double mhd_step_bytes_model(const mhd_config* cfg);

/* ---- batched ensembles: many independent runs stepped together ----------------- */
typedef struct mhd_batch mhd_batch;
/* Runs are grouped by (grid size, precision); each group steps with one launch per
 * kernel class covering all its runs, one captured HIP graph per group step, and the groups
 * step concurrently on their own streams.  With mhd_batch_multi, groups of the same
 * precision, physics and stepper whose lengths are all in {32, 48, 64, 96, 128} (and whose
 * runs have no forcing module or volume penalisation) step together instead: each column
 * pass, row kernel and spectral update of their step is one multi-size launch covering every
 * such group, and their whole step is one captured HIP graph. */
int32_t mhd_batch_create(const mhd_config* cfgs, int32_t nruns, mhd_batch** out);
int32_t mhd_batch_destroy(mhd_batch* b);
int32_t mhd_batch_set_ic_real(mhd_batch* b, int32_t run, const void* const* fields);
int32_t mhd_batch_set_sol(mhd_batch* b, int32_t run, const void* host_in);
int32_t mhd_batch_set_sol_modes(mhd_batch* b, int32_t run, int64_t nmodes, const int32_t* idx, const void* values);
int32_t mhd_batch_get_sol_modes(mhd_batch* b, int32_t run, int64_t nmodes, const int32_t* idx, void* values_out);
int32_t mhd_batch_get_sol(mhd_batch* b, int32_t run, void* host_out);
int32_t mhd_batch_get_clock(mhd_batch* b, int32_t run, double* t, double* dt, int64_t* step);
/* Set per-run CFL control (coef > 0: dt from getCFL! before every step). */
int32_t mhd_batch_set_cfl(mhd_batch* b, int32_t run, double cfl_coef);
/* Advance every run by nsteps steps. */
int32_t mhd_batch_step(mhd_batch* b, int32_t nsteps);
int32_t mhd_batch_use_graph(mhd_batch* b, int32_t enable);
/* Multi-size launches: enable = 1 (the default) steps every eligible group by them, 2 the
 * same except that the group with the most grid points of each set (of three or more groups)
 * steps on its own stream beside them, 0 (the default) steps every group on its own stream;
 * enable < 0 keeps the setting.  ngroups (may be NULL) receives the number of groups stepped by
 * multi-size launches.  No counterpart in the reference (its ensembles are separate
 * problems); for ensemble and serving callers. */
int32_t mhd_batch_multi(mhd_batch* b, int32_t enable, int32_t* ngroups);
/* Per kernel class of the multi-size sets: column passes (col) and the row kernel (row) as
 * one multi-size launch (1) or as each group's own launch on the same stream (0); negative
 * keeps the choice.  mhd_batch_autotune times both forms and keeps the faster; the spectral
 * update is always one multi-size launch.  out2 (may be NULL) receives the choice of the
 * first multi-size set, -1 if there is none. */
int32_t mhd_batch_multi_classes(mhd_batch* b, int32_t col, int32_t row, int32_t* out2);
int32_t mhd_batch_time_steps(mhd_batch* b, int32_t nsteps, double* ms_out);
/* Serving-pool control.  A run of a batch is a slot: set its physical parameters,
 * then start it with a fresh clock (t = 0, step = 0), a dt (or CFL coefficient > 0 for
 * getCFL! before every step) and a step target; step launches skip a run once its
 * step count reaches the target, so slots join and retire at step boundaries while
 * the captured ensemble step stays the same.  New batch runs start with no target. */
int32_t mhd_batch_set_params(mhd_batch* b, int32_t run, double nu, double eta, int32_t n_nu);
int32_t mhd_batch_start_run(mhd_batch* b, int32_t run, double dt, double cfl_coef, int64_t target_steps);
/* Enqueue nsteps steps of every run without waiting; query reports 1 while work is
 * pending; sync waits for it. */
int32_t mhd_batch_launch(mhd_batch* b, int32_t nsteps);
/* mhd_set_step_pruning for every run of the batch (per run, the last step of a launch or
 * the step reaching its target is the full one). */
int32_t mhd_batch_set_step_pruning(mhd_batch* b, int32_t enable);
int32_t mhd_batch_query(mhd_batch* b, int32_t* busy);
int32_t mhd_batch_sync(mhd_batch* b);
/* Tune every group's launch variants (mhd_autotune), then the multi-size sets' per-class
 * choice (mhd_batch_multi_classes). */
int32_t mhd_batch_autotune(mhd_batch* b, int32_t reps);

#ifdef __cplusplus
}
#endif
#endif /* MHDFLOWS_AMD_H */
