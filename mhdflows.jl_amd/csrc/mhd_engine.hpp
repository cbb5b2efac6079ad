// Host-side engine: buffers, twiddle tables and launch sequences of one group of
// same-size runs (one run for a plain problem, many for an ensemble group).
#pragma once
#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "mhd_launch.hpp"

namespace mhd {

extern thread_local std::string g_err;
int fail(int code, const std::string& msg);

// error codes (mirror include/mhdflows_amd.h)
constexpr int E_OK = 0, E_INVALID = -1, E_UNSUPPORTED = -2, E_BUDGET = -3, E_HIP = -4, E_NAN = -5;

#define MHD_HIP(x)                                                                                   \
  do {                                                                                               \
    hipError_t e_ = (x);                                                                             \
    if (e_ != hipSuccess) return ::mhd::fail(::mhd::E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)
#define MHD_TRY(x)             \
  do {                         \
    int r_ = (x);              \
    if (r_ != 0) return r_;    \
  } while (0)

// Per-run physical parameters.
struct RunParams {
  double nu = 0, eta = 0;
  int nnu = 0, neta = 0;
  double dt = 0;
};

struct KTime {
  std::string name;
  int64_t launches = 0;
  double ms = 0;
  double bytes = 0;
};

class EngineBase {
 public:
  virtual ~EngineBase() {}
  virtual int nruns() const = 0;
  virtual int set_ic_real(int run, const void* const* fields) = 0;
  virtual int get_sol(int run, void* host) = 0;
  virtual int set_sol(int run, const void* host) = 0;
  virtual int set_sol_modes(int run, int64_t nmodes, const int32_t* idx, const void* vals) = 0;
  // set_sol_modes with the device staging of idx and vals in a caller's buffer of cap bytes
  // (at least mode_stage_bytes(nmodes)); allocates nothing
  virtual int set_sol_modes_staged(int run, int64_t nmodes, const int32_t* idx, const void* vals, void* stage,
                                   size_t cap) = 0;
  virtual size_t mode_stage_bytes(int64_t nmodes) const = 0;
  // read nmodes half-spectrum entries of sol (field-fastest, nf complex T per entry)
  virtual int get_sol_modes(int run, int64_t nmodes, const int32_t* idx, void* vals) = 0;
  virtual int get_vars(int run, int f, void* host) = 0;
  virtual int device_ptrs(int run, void** sol, void** vars) = 0;
  virtual int get_clock(int run, double* t, double* dt, int64_t* step) = 0;
  // the whole device clock of a run (write = 1 replaces it), for bit-exact checkpoints
  virtual int clock_io(int run, RunClock* c, int write) = 0;
  virtual int precision_bytes() const = 0;  // sizeof(T)
  virtual const Geom& geom() const = 0;
  virtual int nfields() const = 0;
  virtual int set_clock(int run, double t, double dt, int64_t step) = 0;
  virtual int set_cfl(int run, int use, double coef) = 0;
  virtual int calcN(void* N, void* sol, int advection_only) = 0;
  virtual int component_update(void* N, int field) = 0;
  virtual int step(int nsteps) = 0;
  virtual int set_pruning(int enable) = 0;
  virtual int cfl_now(double coef, double* dt) = 0;
  virtual int integrate(double t0, int64_t N0, double usr_dt, double coef, int64_t* taken) = 0;
  virtual int set_graph(int enable) = 0;
  virtual int diagnostics(double* out8) = 0;
  virtual int refresh_vars() = 0;
  // analysis of vars (src/utils/MHDAnalysis.jl): shell spectra and band filter
  virtual int spectrum_bins() const = 0;
  virtual int spectra(double* host_out, int nbins) = 0;
  virtual int scale_decomposition(int f0, int ncomp, double k1, double k2, void* host_out) = 0;
  virtual int div_correction(int which) = 0;
  virtual int field_op(int op, const void* const* in, void* const* out) = 0;
  virtual int sf_shape(int mode, int* nb, int* nbx) const = 0;
  virtual int structure_function(int mode, double* host_out) = 0;
  virtual int set_forcing(int run, uint64_t seed, double kf, double amp) = 0;
  virtual int set_forcing_field(int run, const void* const* f3) = 0;
  virtual int set_forcing_nd(int run, double P, const void* const* f3) = 0;
  virtual int set_vp(int run, int enable, const void* chi, const void* const* U0, const void* const* B0) = 0;
  virtual int set_forcing_a99(int run, uint64_t seed, double kf, double P, double s2, double b, double A) = 0;
  virtual int time_steps(int nsteps, double* ms) = 0;
  virtual int profile_steps(int nsteps, std::vector<KTime>* out) = 0;
  virtual int tune(int reps, char* report, int len) = 0;
  // serving-pool slot control: physical parameters, and a fresh clock with a step target
  virtual int set_run_params(int run, double nu, double eta, int nnu) = 0;
  virtual int start_run(int run, double dt, double cfl_coef, int64_t target) = 0;
  virtual int query() = 0;  // 1 while the stream has work, 0 when idle
  virtual int set_variants(const int32_t* v5, int32_t* out5) = 0;
  virtual int sync() = 0;
  virtual hipStream_t stream() const = 0;
  // multi-size stepping (make_fused): eligibility, the key groups must share (precision,
  // physics, stepper) and a counter that changes whenever the group's step launches change
  // slab pipeline of the right-hand side (1 = off; see Engine::rhs)
  virtual int set_pipeline(int slabs) = 0;
  virtual bool multi_ok() const = 0;
  virtual int multi_key() const = 0;
  virtual uint64_t graph_gen() const = 0;
};

// Several groups of an ensemble stepped together by multi-size launches.
class FusedBase {
 public:
  virtual ~FusedBase() {}
  // queue nsteps steps of every group (asynchronous; ordered after, and before, the work on
  // the groups' own streams)
  virtual int step(int nsteps) = 0;
  virtual int set_graph(int enable) = 0;
  // choose per kernel class (column passes, row kernel) between one multi-size launch and
  // each group's own launch by timing both; choice[0..1] receives 1 for multi-size
  virtual int tune(int reps, int* choice) = 0;
  // set the choice directly (negative: keep); get_choice reads it
  virtual int set_choice(int col, int row) = 0;
  virtual void get_choice(int* c2) const = 0;
};

// groups: same multi_key(), all multi_ok(); returns null (with *err set) on failure
FusedBase* make_fused(const std::vector<EngineBase*>& groups, int* err);

// Device bytes of one run.
int64_t run_bytes(const Geom& g, int phys, int prec);

// savefile / Restart! (src/integrator.jl:208-288): see mhd_ckpt.hip for the file layout
int save_checkpoint(EngineBase* e, int run, const char* path, int exact);
int restart_checkpoint(EngineBase* e, int run, const char* path, int exact);

EngineBase* make_engine(int prec, const Geom& g, const double* L, int phys, int stepper,
                        const std::vector<RunParams>& runs, int device, hipStream_t stream, int* err);

}  // namespace mhd
