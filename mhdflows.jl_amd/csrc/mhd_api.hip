// extern "C" entry points of libmhdflows_amd.so (declared in include/mhdflows_amd.h).
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>

#include "../../include/mhdflows_amd.h"
#include "mhd_engine.hpp"

using namespace mhd;

// plain streaming copy (16 B per lane), the reference kernel for HBM byte counters
__global__ void __launch_bounds__(256)
copy16_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t n16) {
  // one 16 B element per lane and no loop: the most waves in flight, the fastest plain
  // copy shape measured on MI355X
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n16) dst[i] = src[i];
}

struct mhd_handle {
  mhd_config cfg;
  EngineBase* e;
};

// Device staging of mhd_set_sol_modes_noalloc: the (kx, l, m) triples and the values of
// one call.  Every GPU the library runs on holds its own copy until the process ends; all
// calls on one GPU share it, so such calls must not overlap on that GPU.
constexpr size_t kScratchBytes = size_t(4) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
// Device address of `scratch` per device ordinal.  The symbol's host-side name is not a
// device address; this one, taken once per device with that device current, is.
std::mutex g_scratch_mu;
std::map<int, void*> g_scratch_ptr;

int scratch_ptr(int device, void** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_ptr.find(device);
  if (it != g_scratch_ptr.end()) {
    *out = it->second;
    return 0;
  }
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_ptr[device] = p;
  *out = p;
  return 0;
}
}  // namespace

struct mhd_batch {
  struct Group {
    EngineBase* e;
    std::vector<int> runs;  // global run ids in this group
  };
  std::vector<Group> groups;
  std::vector<std::pair<int, int>> where;  // global run -> (group, local run)
  std::vector<mhd_config> cfgs;
  // Stepping plan: groups that can share multi-size launches (same precision, physics and
  // stepper, small lengths, no modules) step together, the others on their own streams.
  // 0: none (default: on a 64-run 32^3..96^3 ensemble the groups' own streams overlap one
  // group's row kernel with another's passes, 11.0k against 9.6k (1) and 10.1k (2) run-steps/s),
  // 1: every eligible group, 2: the largest group of a set on its own stream
  int multi = 0;
  bool use_graph = true;
  bool planned = false;
  std::vector<bool> plan_ok;                    // multi_ok() of each group when planned
  std::vector<std::unique_ptr<FusedBase>> fused;
  std::vector<int> solo;                        // groups stepped alone
  int fused_groups = 0;
};

namespace {

int validate(const mhd_config* c) {
  if (!c) return fail(E_INVALID, "config is NULL");
  const int n[3] = {c->nx, c->ny, c->nz};
  for (int v : n) {
    if (v <= 0 || v % 2) return fail(E_INVALID, "grid sizes must be positive and even (Int(nx/2+1), utils.jl:20)");
    if (!length_supported(v))
      return fail(E_UNSUPPORTED, "grid length " + std::to_string(v) + " has no compiled transform");
  }
  // The x-row kernel of nx = 24 for Float64 MHD (six fields of a 12-point half row per
  // thread pair) once returned wrong values and then faulted the GPU; no cause has been
  // shown, so that instance is not launched.  Float32, HD, and 24 points along y or z are
  // unaffected (other kernel instances).
  if (c->nx == 24 && c->precision == MHD_F64 && c->physics == MHD_PHYS_MHD)
    return fail(E_UNSUPPORTED, "nx = 24 is not supported for Float64 MHD (its row kernel is disabled)");
  if (!(c->Lx > 0 && c->Ly > 0 && c->Lz > 0)) return fail(E_INVALID, "box lengths must be positive");
  if (c->precision != MHD_F32 && c->precision != MHD_F64) return fail(E_INVALID, "precision must be F32 or F64");
  if (c->physics != MHD_PHYS_HD && c->physics != MHD_PHYS_MHD) return fail(E_INVALID, "physics must be HD or MHD");
  if (c->stepper != MHD_RK4 && c->stepper != MHD_LSRK54) return fail(E_INVALID, "stepper must be RK4 or LSRK54");
  if (c->nu < 0 || c->eta < 0) return fail(E_INVALID, "viscosity and resistivity must be >= 0");
  return 0;
}

Geom geom_of(const mhd_config* c) {
  return make_geom(c->nx, c->ny, c->nz, c->Lx, c->Ly, c->Lz, c->precision == MHD_F64 ? 16 : 8);
}

RunParams params_of(const mhd_config* c) {
  RunParams p;
  p.nu = c->nu;
  p.eta = c->physics == MHD_PHYS_MHD ? c->eta : 0.0;
  p.nnu = c->n_nu;
  p.neta = 0;  // nη is not forwarded to SetParams (src/pgen.jl:114-116)
  p.dt = c->dt;
  return p;
}

// (Re)build the batch's stepping plan when multi-size launches are switched or a group's
// eligibility changed.
int batch_plan(mhd_batch* b) {
  std::vector<bool> ok;
  for (auto& g : b->groups) ok.push_back(b->multi > 0 && g.e->multi_ok());
  if (b->planned && ok == b->plan_ok) return 0;
  for (auto& g : b->groups) MHD_TRY(g.e->sync());  // multi-size steps run on group streams
  b->fused.clear();
  b->solo.clear();
  b->fused_groups = 0;
  std::map<int, std::vector<int>> sets;
  for (size_t i = 0; i < b->groups.size(); ++i) {
    if (ok[i]) sets[b->groups[i].e->multi_key()].push_back((int)i);
    else b->solo.push_back((int)i);
  }
  for (auto& kv : sets) {
    if (b->multi == 2 && kv.second.size() >= 3) {
      // the group with the most grid points steps on its own stream, overlapping the others
      size_t big = 0;
      double most = -1;
      for (size_t j = 0; j < kv.second.size(); ++j) {
        const auto& gr = b->groups[kv.second[j]];
        const Geom& G = gr.e->geom();
        const double pts = (double)gr.runs.size() * G.nx * G.ny * G.nz;
        if (pts > most) { most = pts; big = j; }
      }
      b->solo.push_back(kv.second[big]);
      kv.second.erase(kv.second.begin() + (long)big);
    }
    if (kv.second.size() < 2) {  // a single group keeps its own tuned launches
      b->solo.push_back(kv.second[0]);
      continue;
    }
    std::vector<EngineBase*> es;
    for (int i : kv.second) es.push_back(b->groups[i].e);
    int err = 0;
    FusedBase* f = make_fused(es, &err);
    if (!f) return err ? err : fail(E_HIP, "multi-size step setup failed");
    MHD_TRY(f->set_graph(b->use_graph ? 1 : 0));
    b->fused.emplace_back(f);
    b->fused_groups += (int)kv.second.size();
  }
  b->plan_ok = ok;
  b->planned = true;
  return 0;
}

// queue nsteps steps of every run (asynchronous)
int batch_issue(mhd_batch* b, int nsteps) {
  MHD_TRY(batch_plan(b));
  for (auto& f : b->fused) MHD_TRY(f->step(nsteps));
  for (int i : b->solo) MHD_TRY(b->groups[i].e->step(nsteps));
  return 0;
}

int check_handle(mhd_handle* h) {
  if (!h || !h->e) return fail(E_INVALID, "handle is NULL");
  const hipError_t e = hipSetDevice(h->cfg.device);
  if (e != hipSuccess) return fail(E_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
  return 0;
}

}  // namespace

#define BATCH_RUN(b, run)                                                          \
  if (!(b)) return fail(E_INVALID, "batch is NULL");                              \
  if ((run) < 0 || (run) >= (int)(b)->where.size()) return fail(E_INVALID, "run out of range"); \
  auto& G_ = (b)->groups[(b)->where[(run)].first];                                 \
  const int L_ = (b)->where[(run)].second;

extern "C" {

int32_t mhd_abi_version(void) { return MHD_ABI_VERSION; }
const char* mhd_last_error(void) { return g_err.c_str(); }
int32_t mhd_supported_length(int32_t n) { return length_supported(n) ? 1 : 0; }

int32_t mhd_query_bytes(const mhd_config* cfg, uint64_t* bytes) {
  MHD_TRY(validate(cfg));
  if (!bytes) return fail(E_INVALID, "bytes is NULL");
  *bytes = (uint64_t)run_bytes(geom_of(cfg), cfg->physics, cfg->precision);
  return 0;
}

int32_t mhd_create(const mhd_config* cfg, mhd_handle** out) {
  MHD_TRY(validate(cfg));
  if (!out) return fail(E_INVALID, "out is NULL");
  *out = nullptr;
  const Geom g = geom_of(cfg);
  const uint64_t need = (uint64_t)run_bytes(g, cfg->physics, cfg->precision);
  if (cfg->hbm_budget && need > cfg->hbm_budget)
    return fail(E_BUDGET, "problem needs " + std::to_string(need) + " bytes, budget is " +
                              std::to_string(cfg->hbm_budget));
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(E_HIP, "no HIP device available");
  if (cfg->device < 0 || cfg->device >= ndev) return fail(E_INVALID, "device ordinal out of range");
  const double L[3] = {cfg->Lx, cfg->Ly, cfg->Lz};
  int err = 0;
  EngineBase* e = make_engine(cfg->precision, g, L, cfg->physics, cfg->stepper, {params_of(cfg)}, cfg->device,
                              (hipStream_t)cfg->stream, &err);
  if (!e) return err ? err : fail(E_HIP, "engine creation failed");
  *out = new mhd_handle{*cfg, e};
  return 0;
}

int32_t mhd_destroy(mhd_handle* h) {
  if (!h) return 0;
  if (h->e) {
    (void)hipSetDevice(h->cfg.device);
    delete h->e;
  }
  delete h;
  return 0;
}

int32_t mhd_set_ic_real(mhd_handle* h, const void* const* fields) {
  MHD_TRY(check_handle(h));
  if (!fields) return fail(E_INVALID, "fields is NULL");
  return h->e->set_ic_real(0, fields);
}
int32_t mhd_get_sol(mhd_handle* h, void* host_out) {
  MHD_TRY(check_handle(h));
  if (!host_out) return fail(E_INVALID, "output pointer is NULL");
  return h->e->get_sol(0, host_out);
}
int32_t mhd_set_sol(mhd_handle* h, const void* host_in) {
  MHD_TRY(check_handle(h));
  if (!host_in) return fail(E_INVALID, "input pointer is NULL");
  return h->e->set_sol(0, host_in);
}
int32_t mhd_set_sol_modes(mhd_handle* h, int64_t nmodes, const int32_t* idx, const void* values) {
  MHD_TRY(check_handle(h));
  if (nmodes > 0 && (!idx || !values)) return fail(E_INVALID, "idx and values are required");
  return h->e->set_sol_modes(0, nmodes, idx, values);
}
int32_t mhd_batch_set_sol_modes(mhd_batch* b, int32_t run, int64_t nmodes, const int32_t* idx, const void* values) {
  BATCH_RUN(b, run);
  if (nmodes > 0 && (!idx || !values)) return fail(E_INVALID, "idx and values are required");
  return G_.e->set_sol_modes(L_, nmodes, idx, values);
}
int32_t mhd_set_sol_modes_noalloc(mhd_handle* h, int64_t nmodes, const int32_t* idx, const void* values) {
  MHD_TRY(check_handle(h));  // makes the handle's device current
  if (nmodes < 0) return fail(E_INVALID, "nmodes must be >= 0");
  if (nmodes > 0 && (!idx || !values)) return fail(E_INVALID, "idx and values are required");
  const size_t need = h->e->mode_stage_bytes(nmodes);
  if (need > kScratchBytes) {
    int64_t most = 0;
    while (h->e->mode_stage_bytes(most + 1024) <= kScratchBytes) most += 1024;
    while (h->e->mode_stage_bytes(most + 1) <= kScratchBytes) ++most;
    return fail(E_INVALID, std::to_string(nmodes) + " modes need " + std::to_string(need) +
                               " bytes of staging, more than the 4 MiB scratch (at most " + std::to_string(most) +
                               " modes for this precision and physics); use mhd_set_sol_modes");
  }
  void* p = nullptr;
  MHD_TRY(scratch_ptr(h->cfg.device, &p));
  return h->e->set_sol_modes_staged(0, nmodes, idx, values, p, kScratchBytes);
}
int32_t mhd_get_sol_modes(mhd_handle* h, int64_t nmodes, const int32_t* idx, void* values_out) {
  MHD_TRY(check_handle(h));
  if (nmodes > 0 && (!idx || !values_out)) return fail(E_INVALID, "idx and values are required");
  return h->e->get_sol_modes(0, nmodes, idx, values_out);
}
int32_t mhd_batch_get_sol_modes(mhd_batch* b, int32_t run, int64_t nmodes, const int32_t* idx, void* values_out) {
  BATCH_RUN(b, run);
  if (nmodes > 0 && (!idx || !values_out)) return fail(E_INVALID, "idx and values are required");
  return G_.e->get_sol_modes(L_, nmodes, idx, values_out);
}
int32_t mhd_get_vars(mhd_handle* h, int32_t field, void* host_out) {
  MHD_TRY(check_handle(h));
  if (!host_out) return fail(E_INVALID, "output pointer is NULL");
  return h->e->get_vars(0, field, host_out);
}
int32_t mhd_device_ptrs(mhd_handle* h, void** sol, void** vars) {
  MHD_TRY(check_handle(h));
  return h->e->device_ptrs(0, sol, vars);
}
int32_t mhd_get_clock(mhd_handle* h, double* t, double* dt, int64_t* step) {
  MHD_TRY(check_handle(h));
  return h->e->get_clock(0, t, dt, step);
}
int32_t mhd_set_clock(mhd_handle* h, double t, double dt, int64_t step) {
  MHD_TRY(check_handle(h));
  return h->e->set_clock(0, t, dt, step);
}

int32_t mhd_calcN(mhd_handle* h, void* N_dev, void* sol_dev, double /*t*/) {
  MHD_TRY(check_handle(h));
  if (!N_dev || !sol_dev) return fail(E_INVALID, "N and sol must be device pointers");
  return h->e->calcN(N_dev, sol_dev, 0);
}
int32_t mhd_advection(mhd_handle* h, void* N_dev, void* sol_dev, double /*t*/) {
  MHD_TRY(check_handle(h));
  if (!N_dev || !sol_dev) return fail(E_INVALID, "N and sol must be device pointers");
  return h->e->calcN(N_dev, sol_dev, 1);
}
int32_t mhd_u_update(mhd_handle* h, void* N_dev, int32_t dir) {
  MHD_TRY(check_handle(h));
  if (!N_dev) return fail(E_INVALID, "N must be a device pointer");
  if (dir < 0 || dir > 2) return fail(E_INVALID, "Warning : Unknown direction is declerad");  // MHDSolver.jl:52
  return h->e->component_update(N_dev, dir);
}
int32_t mhd_b_update(mhd_handle* h, void* N_dev, int32_t dir) {
  MHD_TRY(check_handle(h));
  if (!N_dev) return fail(E_INVALID, "N must be a device pointer");
  if (h->cfg.physics != MHD_PHYS_MHD) return fail(E_INVALID, "BᵢUpdate! needs B_field = true");
  if (dir < 0 || dir > 2) return fail(E_INVALID, "Warning : Unknown direction is declerad");
  return h->e->component_update(N_dev, 3 + dir);
}

int32_t mhd_step(mhd_handle* h, int32_t nsteps) {
  MHD_TRY(check_handle(h));
  return h->e->step(nsteps);
}
int32_t mhd_cfl_dt(mhd_handle* h, double cfl_coef, double* dt_out) {
  MHD_TRY(check_handle(h));
  return h->e->cfl_now(cfl_coef, dt_out);
}
int32_t mhd_integrate(mhd_handle* h, double t0, int64_t N0, double usr_dt, double cfl_coef, int64_t* steps_out) {
  MHD_TRY(check_handle(h));
  return h->e->integrate(t0, N0, usr_dt, cfl_coef, steps_out);
}
int32_t mhd_use_graph(mhd_handle* h, int32_t enable) {
  MHD_TRY(check_handle(h));
  return h->e->set_graph(enable);
}

int32_t mhd_diagnostics(mhd_handle* h, mhd_diag* out) {
  MHD_TRY(check_handle(h));
  if (!out) return fail(E_INVALID, "out is NULL");
  double v[8];
  MHD_TRY(h->e->diagnostics(v));
  out->E_kin = v[0];
  out->E_mag = v[1];
  out->H_kin = v[2];
  out->H_mag = v[3];
  out->H_cross = v[4];
  out->max_div_b = v[5];
  out->max_div_u = v[6];
  out->has_nan = v[7] != 0.0;
  out->pad = 0;
  return 0;
}
int32_t mhd_refresh_vars(mhd_handle* h) {
  MHD_TRY(check_handle(h));
  return h->e->refresh_vars();
}
int32_t mhd_set_forcing(mhd_handle* h, uint64_t seed, double kf, double amp) {
  MHD_TRY(check_handle(h));
  if (!(kf >= 0)) return fail(E_INVALID, "kf must be >= 0");
  return h->e->set_forcing(0, seed, kf, amp);
}

int32_t mhd_set_forcing_field(mhd_handle* h, const void* const* f3) {
  MHD_TRY(check_handle(h));
  return h->e->set_forcing_field(0, f3);
}
int32_t mhd_set_forcing_nd(mhd_handle* h, double P, const void* const* f3) {
  MHD_TRY(check_handle(h));
  if (!f3) return fail(E_INVALID, "f3 is NULL");
  return h->e->set_forcing_nd(0, P, f3);
}
int32_t mhd_set_vp(mhd_handle* h, int32_t enable, const void* chi, const void* const* U0, const void* const* B0) {
  MHD_TRY(check_handle(h));
  return h->e->set_vp(0, enable, chi, U0, B0);
}
int32_t mhd_set_forcing_a99(mhd_handle* h, uint64_t seed, double kf, double P, double sigma2, double b, double A) {
  MHD_TRY(check_handle(h));
  return h->e->set_forcing_a99(0, seed, kf, P, sigma2, b, A);
}

int32_t mhd_spectra(mhd_handle* h, double* out, int32_t nbins_cap, int32_t* nbins) {
  MHD_TRY(check_handle(h));
  const int need = h->e->spectrum_bins();
  if (nbins) *nbins = need;
  if (!out) return 0;
  if (nbins_cap < need) return fail(E_INVALID, "out holds fewer than nbins bins per row");
  return h->e->spectra(out, need);
}
int32_t mhd_scale_decomposition(mhd_handle* h, int32_t field0, int32_t ncomp, double k1, double k2, void* host_out) {
  MHD_TRY(check_handle(h));
  if (!host_out) return fail(E_INVALID, "host_out is NULL");
  return h->e->scale_decomposition(field0, ncomp, k1, k2, host_out);
}

int32_t mhd_field_op(mhd_handle* h, int32_t op, const void* const* in, void* const* out) {
  MHD_TRY(check_handle(h));
  return h->e->field_op(op, in, out);
}
int32_t mhd_div_correction(mhd_handle* h, int32_t field) {
  MHD_TRY(check_handle(h));
  return h->e->div_correction(field);
}

int32_t mhd_structure_function(mhd_handle* h, int32_t mode, double* out, int32_t cap, int32_t* nbins,
                               int32_t* nrow) {
  MHD_TRY(check_handle(h));
  if (mode != 0 && mode != 1) return fail(E_INVALID, "mode must be 0 (SF2 1D) or 1 (SFr_Global)");
  int nb = 0, nbx = 0;
  h->e->sf_shape(mode, &nb, &nbx);
  if (nbins) *nbins = nb;
  if (nrow) *nrow = nbx;
  if (!out) return 0;
  if (cap < 2 * nb) return fail(E_INVALID, "out holds fewer than 2 * nbins values");
  return h->e->structure_function(mode, out);
}

int32_t mhd_savefile(mhd_handle* h, const char* path, int32_t exact) {
  MHD_TRY(check_handle(h));
  return save_checkpoint(h->e, 0, path, exact);
}
int32_t mhd_restart(mhd_handle* h, const char* path, int32_t exact) {
  MHD_TRY(check_handle(h));
  return restart_checkpoint(h->e, 0, path, exact);
}

int32_t mhd_autotune(mhd_handle* h, int32_t reps, char* report, int32_t report_len) {
  MHD_TRY(check_handle(h));
  return h->e->tune(reps, report, report_len);
}
int32_t mhd_set_pipeline(mhd_handle* h, int32_t slabs) {
  MHD_TRY(check_handle(h));
  return h->e->set_pipeline(slabs);
}
int32_t mhd_set_variants(mhd_handle* h, const int32_t* v5, int32_t* out5) {
  MHD_TRY(check_handle(h));
  return h->e->set_variants(v5, out5);
}
int32_t mhd_batch_autotune(mhd_batch* b, int32_t reps) {
  if (!b) return fail(E_INVALID, "batch is NULL");
  if (reps <= 0) return fail(E_INVALID, "reps must be > 0");
  for (auto& g : b->groups) MHD_TRY(g.e->tune(reps, nullptr, 0));
  MHD_TRY(batch_plan(b));
  for (auto& f : b->fused) MHD_TRY(f->tune(reps, nullptr));
  return 0;
}
int32_t mhd_bench_copy(uint64_t bytes, int32_t reps, double* ms_out) {
  if (bytes < 16 || reps <= 0 || !ms_out) return fail(E_INVALID, "bytes >= 16 and reps > 0 required");
  const int64_t n16 = (int64_t)(bytes / 16);
  if ((n16 + 255) / 256 > 0x7fffffffLL) return fail(E_INVALID, "copy too large for one launch");
  // buffers and events are released on every return path
  struct Res {
    void *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~Res() {
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
      if (a) (void)hipFree(a);
      if (b) (void)hipFree(b);
    }
  } r;
  MHD_HIP(hipMalloc(&r.a, n16 * 16));
  MHD_HIP(hipMalloc(&r.b, n16 * 16));
  MHD_HIP(hipMemset(r.a, 1, n16 * 16));
  MHD_HIP(hipEventCreate(&r.e0));
  MHD_HIP(hipEventCreate(&r.e1));
  const unsigned nb = (unsigned)((n16 + 255) / 256);
  hipLaunchKernelGGL(copy16_kernel, dim3(nb), dim3(256), 0, 0, (const uint4*)r.a, (uint4*)r.b, n16);
  MHD_HIP(hipGetLastError());
  MHD_HIP(hipEventRecord(r.e0, 0));
  for (int k = 0; k < reps; ++k)
    hipLaunchKernelGGL(copy16_kernel, dim3(nb), dim3(256), 0, 0, (const uint4*)r.a, (uint4*)r.b, n16);
  MHD_HIP(hipGetLastError());
  MHD_HIP(hipEventRecord(r.e1, 0));
  MHD_HIP(hipEventSynchronize(r.e1));
  float ms = 0;
  MHD_HIP(hipEventElapsedTime(&ms, r.e0, r.e1));
  *ms_out = ms / reps;
  return 0;
}
int32_t mhd_set_step_pruning(mhd_handle* h, int32_t enable) {
  if (!h) return fail(E_INVALID, "handle is NULL");
  return h->e->set_pruning(enable);
}
int32_t mhd_time_steps(mhd_handle* h, int32_t nsteps, double* ms_out) {
  MHD_TRY(check_handle(h));
  if (!ms_out) return fail(E_INVALID, "ms_out is NULL");
  return h->e->time_steps(nsteps, ms_out);
}
int32_t mhd_profile_steps(mhd_handle* h, int32_t nsteps, mhd_ktime* out, int32_t max, int32_t* n_out) {
  MHD_TRY(check_handle(h));
  std::vector<KTime> k;
  MHD_TRY(h->e->profile_steps(nsteps, &k));
  const int n = std::min<int>((int)k.size(), max);
  for (int i = 0; i < n; ++i) {
    std::memset(out[i].name, 0, sizeof(out[i].name));
    std::strncpy(out[i].name, k[i].name.c_str(), sizeof(out[i].name) - 1);
    out[i].launches = k[i].launches;
    out[i].total_ms = k[i].ms;
    out[i].bytes = k[i].bytes;
  }
  if (n_out) *n_out = n;
  return 0;
}

double mhd_step_bytes_model(const mhd_config* cfg) {
  if (!cfg) return 0.0;
  const double s = cfg->precision == MHD_F32 ? 4.0 : 8.0;
  const double n3 = (double)cfg->nx * cfg->ny * cfg->nz;
  const double Vr = s * n3;
  const double Vc = 2.0 * s * (cfg->nx / 2 + 1) * (double)cfg->ny * cfg->nz;
  if (cfg->physics == MHD_PHYS_MHD)
    return cfg->stepper == MHD_RK4 ? 438 * Vc + 120 * Vr : 534 * Vc + 150 * Vr;
  return cfg->stepper == MHD_RK4 ? 255 * Vc + 72 * Vr : 0.0;
}

// ---- batched ensembles -----------------------------------------------------------
int32_t mhd_batch_create(const mhd_config* cfgs, int32_t nruns, mhd_batch** out) {
  if (!cfgs || nruns <= 0 || !out) return fail(E_INVALID, "need at least one config");
  *out = nullptr;
  uint64_t need = 0, budget = 0;
  for (int i = 0; i < nruns; ++i) {
    MHD_TRY(validate(&cfgs[i]));
    need += (uint64_t)run_bytes(geom_of(&cfgs[i]), cfgs[i].physics, cfgs[i].precision);
    if (cfgs[i].hbm_budget) budget = cfgs[i].hbm_budget;
  }
  if (budget && need > budget) return fail(E_BUDGET, "ensemble exceeds the HBM budget");
  // group by everything that fixes the kernel instances and launch shapes
  using Key = std::tuple<int, int, int, double, double, double, int, int, int, int>;
  std::map<Key, int> gid;
  auto* b = new mhd_batch();
  b->cfgs.assign(cfgs, cfgs + nruns);
  std::vector<std::vector<int>> members;
  for (int i = 0; i < nruns; ++i) {
    const mhd_config& c = cfgs[i];
    Key k{c.nx, c.ny, c.nz, c.Lx, c.Ly, c.Lz, c.precision, c.physics, c.stepper, c.device};
    if (!gid.count(k)) {
      gid[k] = (int)members.size();
      members.emplace_back();
    }
    members[gid[k]].push_back(i);
  }
  b->where.resize(nruns);
  for (size_t gi = 0; gi < members.size(); ++gi) {
    const mhd_config& c0 = cfgs[members[gi][0]];
    std::vector<RunParams> rp;
    for (size_t j = 0; j < members[gi].size(); ++j) {
      rp.push_back(params_of(&cfgs[members[gi][j]]));
      b->where[members[gi][j]] = {(int)gi, (int)j};
    }
    const double L[3] = {c0.Lx, c0.Ly, c0.Lz};
    int err = 0;
    EngineBase* e = make_engine(c0.precision, geom_of(&c0), L, c0.physics, c0.stepper, rp, c0.device, nullptr, &err);
    if (!e) {
      mhd_batch_destroy(b);
      return err ? err : fail(E_HIP, "engine creation failed");
    }
    b->groups.push_back({e, members[gi]});
  }
  *out = b;
  return 0;
}

int32_t mhd_batch_destroy(mhd_batch* b) {
  if (!b) return 0;
  b->fused.clear();  // before the engines whose streams and buffers they use
  for (auto& g : b->groups) delete g.e;
  delete b;
  return 0;
}


int32_t mhd_batch_set_ic_real(mhd_batch* b, int32_t run, const void* const* fields) {
  BATCH_RUN(b, run);
  return G_.e->set_ic_real(L_, fields);
}
int32_t mhd_batch_set_sol(mhd_batch* b, int32_t run, const void* host_in) {
  BATCH_RUN(b, run);
  return G_.e->set_sol(L_, host_in);
}
int32_t mhd_batch_get_sol(mhd_batch* b, int32_t run, void* host_out) {
  BATCH_RUN(b, run);
  return G_.e->get_sol(L_, host_out);
}
int32_t mhd_batch_get_clock(mhd_batch* b, int32_t run, double* t, double* dt, int64_t* step) {
  BATCH_RUN(b, run);
  return G_.e->get_clock(L_, t, dt, step);
}
int32_t mhd_batch_set_cfl(mhd_batch* b, int32_t run, double cfl_coef) {
  BATCH_RUN(b, run);
  return G_.e->set_cfl(L_, cfl_coef > 0 ? 1 : 0, cfl_coef);
}
int32_t mhd_batch_step(mhd_batch* b, int32_t nsteps) {
  if (!b) return fail(E_INVALID, "batch is NULL");
  if (nsteps < 0) return fail(E_INVALID, "nsteps must be >= 0");
  // multi-size sets on their first group's stream, the other groups on their own streams;
  // they overlap on the device
  MHD_TRY(batch_issue(b, nsteps));
  for (auto& g : b->groups) MHD_TRY(g.e->sync());
  return 0;
}
int32_t mhd_batch_set_params(mhd_batch* b, int32_t run, double nu, double eta, int32_t n_nu) {
  BATCH_RUN(b, run);
  if (nu < 0 || eta < 0) return fail(E_INVALID, "viscosity and resistivity must be >= 0");
  return G_.e->set_run_params(L_, nu, eta, n_nu);
}
int32_t mhd_batch_start_run(mhd_batch* b, int32_t run, double dt, double cfl_coef, int64_t target_steps) {
  BATCH_RUN(b, run);
  if (target_steps < 0) return fail(E_INVALID, "target_steps must be >= 0");
  return G_.e->start_run(L_, dt, cfl_coef, target_steps);
}
int32_t mhd_batch_set_step_pruning(mhd_batch* b, int32_t enable) {
  if (!b) return fail(E_INVALID, "batch is NULL");
  for (auto& g : b->groups) MHD_TRY(g.e->set_pruning(enable));
  return 0;
}
int32_t mhd_batch_launch(mhd_batch* b, int32_t nsteps) {
  if (!b) return fail(E_INVALID, "batch is NULL");
  if (nsteps < 0) return fail(E_INVALID, "nsteps must be >= 0");
  return batch_issue(b, nsteps);
}
int32_t mhd_batch_query(mhd_batch* b, int32_t* busy) {
  if (!b || !busy) return fail(E_INVALID, "batch or busy is NULL");
  *busy = 0;
  for (auto& g : b->groups) {
    const int q = g.e->query();
    if (q < 0) return q;
    *busy |= q;
  }
  return 0;
}
int32_t mhd_batch_sync(mhd_batch* b) {
  if (!b) return fail(E_INVALID, "batch is NULL");
  for (auto& g : b->groups) MHD_TRY(g.e->sync());
  return 0;
}
int32_t mhd_batch_use_graph(mhd_batch* b, int32_t enable) {
  if (!b) return fail(E_INVALID, "batch is NULL");
  b->use_graph = enable != 0;
  for (auto& g : b->groups) MHD_TRY(g.e->set_graph(enable));
  for (auto& f : b->fused) MHD_TRY(f->set_graph(enable));
  return 0;
}
int32_t mhd_batch_multi(mhd_batch* b, int32_t enable, int32_t* ngroups) {
  if (!b) return fail(E_INVALID, "batch is NULL");
  if (enable > 2) return fail(E_INVALID, "enable must be 0, 1 or 2 (negative keeps the setting)");
  if (enable >= 0 && enable != b->multi) {
    b->multi = enable;
    b->planned = false;
  }
  MHD_TRY(batch_plan(b));
  if (ngroups) *ngroups = b->fused_groups;
  return 0;
}
int32_t mhd_batch_multi_classes(mhd_batch* b, int32_t col, int32_t row, int32_t* out2) {
  if (!b) return fail(E_INVALID, "batch is NULL");
  MHD_TRY(batch_plan(b));
  for (auto& f : b->fused) MHD_TRY(f->set_choice(col, row));
  if (out2) {
    int c[2] = {-1, -1};  // no multi-size set
    if (!b->fused.empty()) b->fused[0]->get_choice(c);
    out2[0] = c[0];
    out2[1] = c[1];
  }
  return 0;
}
int32_t mhd_batch_time_steps(mhd_batch* b, int32_t nsteps, double* ms_out) {
  if (!b || !ms_out) return fail(E_INVALID, "batch or ms_out is NULL");
  // warm the graphs, then time all groups' streams together
  MHD_TRY(batch_issue(b, 0));
  for (auto& g : b->groups) MHD_TRY(g.e->sync());
  // every event created here is released on every return path
  struct Events {
    std::vector<hipEvent_t> v;
    ~Events() {
      for (hipEvent_t x : v)
        if (x) (void)hipEventDestroy(x);
    }
    hipError_t make(hipEvent_t* out) {
      hipError_t r = hipEventCreate(out);
      if (r == hipSuccess) v.push_back(*out);
      return r;
    }
  } evs;
  hipEvent_t a = nullptr, e = nullptr;
  MHD_HIP(evs.make(&a));
  MHD_HIP(evs.make(&e));
  hipStream_t s0 = b->groups[0].e->stream();
  MHD_HIP(hipEventRecord(a, s0));
  for (size_t i = 1; i < b->groups.size(); ++i) MHD_HIP(hipStreamWaitEvent(b->groups[i].e->stream(), a, 0));
  MHD_TRY(batch_issue(b, nsteps));
  for (size_t i = 1; i < b->groups.size(); ++i) {
    hipEvent_t d = nullptr;
    MHD_HIP(evs.make(&d));
    MHD_HIP(hipEventRecord(d, b->groups[i].e->stream()));
    MHD_HIP(hipStreamWaitEvent(s0, d, 0));
  }
  MHD_HIP(hipEventRecord(e, s0));
  MHD_HIP(hipEventSynchronize(e));
  float ms = 0;
  MHD_HIP(hipEventElapsedTime(&ms, a, e));
  *ms_out = ms;
  return 0;
}

}  // extern "C"
