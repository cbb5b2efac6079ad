// Engine<T>: launch sequences of the MHD/HD right-hand side and time steps for a
// group of runs sharing one grid size and precision.
#include "mhd_engine.hpp"
#include "mhd_multi.hpp"

#include <algorithm>
#include <functional>
#include <cmath>
#include <cstring>
#include <limits>

namespace mhd {

thread_local std::string g_err;
int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

// ---------------------------------------------------------------------------
// twiddle tables, generated in extended precision and rounded to T
// ---------------------------------------------------------------------------
namespace {
std::mutex g_tab_mu;
std::map<std::tuple<int, int, int, int>, void*> g_tabs;  // (device, sizeof(T), n, kind) -> device table

template <class T>
const cx<T>* twiddles(int device, int n, int kind, int* err) {
  // kind 0: tw[m] = exp(-2 pi i m / n), m < n ; kind 1: exp(-2 pi i k / n), k <= n/2
  std::lock_guard<std::mutex> lk(g_tab_mu);
  auto key = std::make_tuple(device, (int)sizeof(T), n, kind);
  auto it = g_tabs.find(key);
  if (it != g_tabs.end()) return (const cx<T>*)it->second;
  const int cnt = kind == 0 ? n : n / 2 + 1;
  std::vector<cx<T>> h(cnt);
  const long double pi = 3.14159265358979323846264338327950288L;
  for (int m = 0; m < cnt; ++m) {
    // reduce to the first octant so the long-double sin/cos see small arguments
    long double ang = 2.0L * pi * (long double)m / (long double)n;
    h[m].re = (T)cosl(ang);
    h[m].im = (T)(-sinl(ang));
    // exact values at the quarter points
    if (4 * (int64_t)m % n == 0) {
      const int q = (int)(4 * (int64_t)m / n) % 4;
      const T c[4] = {T(1), T(0), T(-1), T(0)}, s[4] = {T(0), T(-1), T(0), T(1)};
      h[m].re = c[q];
      h[m].im = s[q];
    }
  }
  void* d = nullptr;
  if (hipMalloc(&d, cnt * sizeof(cx<T>)) != hipSuccess) {
    *err = fail(E_HIP, "hipMalloc of twiddle table failed");
    return nullptr;
  }
  if (hipMemcpy(d, h.data(), cnt * sizeof(cx<T>), hipMemcpyHostToDevice) != hipSuccess) {
    *err = fail(E_HIP, "hipMemcpy of twiddle table failed");
    return nullptr;
  }
  g_tabs[key] = d;
  return (const cx<T>*)d;
}

// splitmix64-based uniform in [0,1), as the forcing definition in the oracle
uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
double uniform01(uint64_t seed, uint64_t counter) {
  const uint64_t x = seed * 0x9E3779B97F4A7C15ull + counter + 1;
  return (double)(splitmix64(x) >> 11) * (1.0 / 9007199254740992.0);
}
}  // namespace

int64_t run_bytes(const Geom& g, int phys, int prec) {
  const int s = prec == 0 ? 4 : 8;
  const int nf = phys == PHYS_MHD ? 6 : 3;
  const int np = phys == PHYS_MHD ? 9 : 6;
  const int64_t rows = (int64_t)g.ny * g.nz;
  const int64_t blocks = prec == 0 ? row_blocks_max<float>(g.nx, rows) : row_blocks_max<double>(g.nx, rows);
  return (int64_t)(3 * nf) * g.nc * 2 * s + (int64_t)np * g.wnc * 2 * s + (int64_t)nf * g.nr * s + 2 * blocks * s +
         (int64_t)nf * g.ny * g.nz * 2 * s + 4096;
}

template <class T>
class Engine : public EngineBase {
 public:
  Geom g{};
  int phys = 0, nf = 0, np = 0, stepper = 0, nr_ = 0, dev = 0;
  hipStream_t st = nullptr;
  // Second stream and events of the slab pipeline (see rhs), and the stream the hot launches
  // go to (st, or st2 for odd slabs).
  hipStream_t st2 = nullptr, cs_ = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  bool own_stream = false;
  std::vector<RunBuf<T>> rbh;
  RunBuf<T>* rbd = nullptr;
  std::vector<void*> allocs;
  std::vector<RunClock> clk_init;
  const cx<T>*twy = nullptr, *twz = nullptr, *twM = nullptr, *twh = nullptr;
  int64_t row_blocks = 0;
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  bool use_graph = true;
  bool profiling = false;
  struct Ev {
    const char* name;
    hipEvent_t a, b;
    double bytes;
  };
  std::vector<Ev> evs;
  std::vector<void*> force_tabs;
  // calcF! module buffers per run: static forcing spectrum, negative-damping f fields and sums
  std::vector<void*> fdense_bufs, ndf_bufs, ndpart_bufs;
  bool nd_any = false;     // a run has negative damping: MHD row launches form 12 products
  std::vector<void*> vp_bufs;  // volume penalisation fields per run (chi, U0, B0)
  bool vp_any = false;     // a run has volume penalisation: row launches form nf more products
  int np_alloc = 0;        // product fields each run's work buffer holds
  int64_t calls_ = 0;      // right-hand sides evaluated by calcN (draw counter of stochastic forcing)
  double* dpart = nullptr;     // FP64 diagnostic partials
  int dpart_blocks = 0;

  ~Engine() override {
    if (st) hipStreamSynchronize(st);
    if (gexec) hipGraphExecDestroy(gexec);
    if (graph) hipGraphDestroy(graph);
    for (void* p : allocs) hipFree(p);
    for (void* p : force_tabs)
      if (p) hipFree(p);
    for (auto* v : {&fdense_bufs, &ndf_bufs, &ndpart_bufs, &vp_bufs})
      for (void* p : *v)
        if (p) hipFree(p);
    if (dpart) hipFree(dpart);
    if (stage_) hipFree(stage_);
    if (spec_) hipFree(spec_);
    if (rbd) hipFree(rbd);
    if (st2) (void)hipStreamSynchronize(st2);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (st2) (void)hipStreamDestroy(st2);
    if (own_stream && st) hipStreamDestroy(st);
  }

  int nruns() const override { return nr_; }
  hipStream_t stream() const override { return st; }
  int sync() override {
    MHD_HIP(hipStreamSynchronize(st));
    return 0;
  }

  int init(const Geom& geom, const double* L, int ph, int stp, const std::vector<RunParams>& rp, int device,
           hipStream_t s) {
    g = geom;
    lens[0] = L[0];
    lens[1] = L[1];
    lens[2] = L[2];
    phys = ph;
    nf = ph == PHYS_MHD ? 6 : 3;
    np = ph == PHYS_MHD ? 9 : 6;
    stepper = stp;
    nr_ = (int)rp.size();
    dev = device;
    MHD_HIP(hipSetDevice(device));
    {
      int lds = 0;
      if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && lds > 0)
        lds_per_block_limit() = lds;
    }
    if (s) {
      st = s;
    } else {
      MHD_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
      own_stream = true;
    }
    cs_ = st;
    {
      const int64_t rows = (int64_t)g.ny * g.nz;
      pipe_ = (rows >= (int64_t)256 * 256 && rows <= (int64_t)512 * 512) ? 2 : 1;
    }
    MHD_HIP(hipStreamCreateWithFlags(&st2, hipStreamNonBlocking));  // default priority
    MHD_HIP(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
    MHD_HIP(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
    int err = 0;
    twy = twiddles<T>(device, g.ny, 0, &err);
    twz = twiddles<T>(device, g.nz, 0, &err);
    twM = twiddles<T>(device, g.nx / 2, 0, &err);
    twh = twiddles<T>(device, g.nx, 1, &err);
    if (err) return err;
    row_blocks = row_blocks_max<T>(g.nx, (int64_t)g.ny * g.nz);  // partial capacity for any row variant
    apply_cached_tuning();
    rbh.resize(nr_);
    force_tabs.assign(nr_, nullptr);
    fdense_bufs.assign(nr_, nullptr);
    ndf_bufs.assign(nr_, nullptr);
    ndpart_bufs.assign(nr_, nullptr);
    vp_bufs.assign(nr_, nullptr);
    np_alloc = np;
    for (int r = 0; r < nr_; ++r) {
      RunBuf<T>& b = rbh[r];
      std::memset(&b, 0, sizeof(b));
      auto alloc = [&](size_t bytes, void** p) -> int {
        if (hipMalloc(p, bytes) != hipSuccess) return fail(E_HIP, "hipMalloc failed (out of device memory?)");
        allocs.push_back(*p);
        return hipMemsetAsync(*p, 0, bytes, st) == hipSuccess ? 0 : fail(E_HIP, "hipMemsetAsync failed");
      };
      void* p;
      const size_t cb = sizeof(cx<T>) * (size_t)g.nc;
      MHD_TRY(alloc(cb * nf, &p));
      b.buf[B_SOL] = (cx<T>*)p;
      MHD_TRY(alloc(cb * nf, &p));
      b.buf[B_STIN] = (cx<T>*)p;
      b.rscr = (T*)p;
      MHD_TRY(alloc(cb * nf, &p));
      b.buf[B_ACC] = (cx<T>*)p;
      MHD_TRY(alloc(sizeof(cx<T>) * (size_t)g.wnc * np, &p));
      b.buf[B_WORK] = (cx<T>*)p;
      MHD_TRY(alloc(sizeof(T) * (size_t)g.nr * nf, &p));
      b.vars = (T*)p;
      MHD_TRY(alloc(sizeof(T) * 2 * (size_t)row_blocks, &p));
      b.partial = (T*)p;
      MHD_TRY(alloc(sizeof(cx<T>) * (size_t)nf * g.ny * g.nz, &p));
      b.plane = (cx<T>*)p;
      MHD_TRY(alloc(sizeof(RunClock), &p));
      b.clk = (RunClock*)p;
      b.nu = rp[r].nu;
      b.eta = rp[r].eta;
      b.nnu = rp[r].nnu;
      b.neta = rp[r].neta;
      RunClock c;
      std::memset(&c, 0, sizeof(c));
      c.dt = (double)(T)rp[r].dt;
      c.dl = (double)std::min({clk_dx(0), clk_dx(1), clk_dx(2)});
      c.t_diff = std::numeric_limits<double>::infinity();
      c.target = std::numeric_limits<int64_t>::max();
      clk_init.push_back(c);
      MHD_HIP(hipMemcpyAsync(b.clk, &c, sizeof(c), hipMemcpyHostToDevice, st));
    }
    MHD_HIP(hipMalloc(&rbd, sizeof(RunBuf<T>) * nr_));
    MHD_TRY(upload_runs());
    dpart_blocks = 1024;
    MHD_HIP(hipMalloc(&dpart, sizeof(double) * 3 * dpart_blocks));
    MHD_HIP(hipStreamSynchronize(st));
    return 0;
  }

  int upload_runs() {
    MHD_HIP(hipMemcpyAsync(rbd, rbh.data(), sizeof(RunBuf<T>) * nr_, hipMemcpyHostToDevice, st));
    return 0;
  }

  // ---- launch helpers (optionally timed per kernel class) -------------------
  template <class F>
  int timed(const char* name, double bytes, F&& f) {
    hipEvent_t a = nullptr, b = nullptr;
    if (profiling) {
      MHD_HIP(hipEventCreate(&a));
      MHD_HIP(hipEventCreate(&b));
      MHD_HIP(hipEventRecord(a, st));
    }
    hipError_t e = f();
    if (e != hipSuccess) return fail(E_HIP, std::string("launch of ") + name + " failed: " + hipGetErrorString(e));
    if (profiling) {
      MHD_HIP(hipEventRecord(b, st));
      evs.push_back({name, a, b, bytes});
    }
    return 0;
  }

  // tuned variants: col_var[c] for the hot column passes c = zpass_inv, ypass_inv,
  // ypass_fwd, zpass_fwd (strip width), row_var for the fused row kernel (plan)
  int col_var[4] = {0, 0, 0, 0};
  int row_var = 0;

  static int col_class(int axis, int dir) { return dir > 0 ? (axis == 2 ? 0 : 1) : (axis == 1 ? 2 : 3); }

  int gate_ = 0;  // set while a time step is being launched
  // Slab pipeline of a right-hand side: the y-inverse pass, the row kernel and the y-forward
  // pass act plane by plane, so over slabs of z-planes they run as pipe_ independent chains,
  // alternately on st and st2; the row kernel (VALU-bound) of one slab then overlaps the y passes
  // (memory-bound) of another.  slab_z0_/slab_z1_ bound the current slab (z1 < 0: none).
  // Default: 2 slabs for planes of 256^2 to 512^2 rows (256^3 Float32 RK4: 321.5 -> 324.7
  // steps/s; 4 slabs 322.9, 8 slabs 313.8); off below (launch-bound grids) and above (1024^3:
  // 4.33 against 4.32 steps/s).
  int pipe_ = 1;
  int slab_z0_ = 0, slab_z1_ = -1;
  // slab boundaries of the pipeline for row variant rvar (each slab starts on a whole row
  // block); returns the number of slabs, 1 when the pipeline is off
  int pipe_slabs(int rvar, std::vector<int>* zb) const {
    if (pipe_ <= 1 || rec_ || profiling || nd_any || vp_any) return 1;
    const int rpb = row_rows_per_block<T>(g.nx, rvar);
    int q = 1;  // z-planes per row-block boundary: q*ny is a multiple of rpb
    while ((int64_t)q * g.ny % rpb) ++q;
    zb->assign(1, 0);
    for (int k = 1; k <= pipe_; ++k) {
      int z = (int)((int64_t)g.nz * k / pipe_);
      z = std::min(g.nz, (z + q - 1) / q * q);
      if (z <= zb->back()) return 1;
      zb->push_back(z);
    }
    return zb->back() == g.nz ? pipe_ : 1;
  }
  int set_pipeline(int slabs) override {
    if (slabs < 1 || slabs > 16) return fail(E_INVALID, "pipeline slabs must be 1..16");
    pipe_ = slabs;
    drop_graph();
    return 0;
  }
  bool prune_enabled = true;  // steps before the last of a call skip the aliased modes (step_pruned)
  bool host_pruned_ = false;  // byte models of profiled launches: the step being profiled is pruned

  // Launch scope: helper launches cover runs [run0_, run0_ + cur_n()) of the group
  // (all runs unless a RunScope narrows them to one, e.g. to set up a single slot).
  int run0_ = 0, nrun_ = -1;
  int cur_n() const { return nrun_ < 0 ? nr_ : nrun_; }
  RunBuf<T>* cur_rb() const { return rbd + run0_; }
  struct RunScope {
    Engine* e;
    RunScope(Engine* en, int r) : e(en) {
      e->run0_ = r;
      e->nrun_ = 1;
    }
    ~RunScope() {
      e->run0_ = 0;
      e->nrun_ = -1;
    }
  };

  int colpass(const char* name, int axis, int dir, int src, int dst, int src_f0, int dst_f0, int nfields,
              bool prune, bool hot = false, bool plane_out = false, bool prune_out = false) {
    const int n = axis == 1 ? g.ny : g.nz;
    const int S = (rec_ && rec_multi_) ? multi_col_strip(n) : col_strip<T>(n, hot ? col_var[col_class(axis, dir)] : 0);
    ColArgs a{};
    a.g = g;
    a.axis = axis;
    a.nfields = nfields;
    a.src = src;
    a.dst = dst;
    a.src_f0 = src_f0;
    a.dst_f0 = dst_f0;
    a.ncols = prune ? g.kx_keep : g.nkr;
    // strips cover the stored columns: whole lines of the work buffer (see colpass_kernel)
    const int scols = dst != B_WORK ? a.ncols : (a.ncols >= g.nkr ? g.wrs : g.kx_keep_w);
    a.nstrips = (std::max(a.ncols, scols) + S - 1) / S;
    if (axis == 2 && prune) {  // only pencils with kept l
      a.nother = nkept_y(g);
      a.other_lo_end = g.ly_lo_end;
      a.other_gap = g.ly_hi_begin - g.ly_lo_end;
    } else if (axis == 1 && slab_z1_ > 0) {  // the z-planes of one pipeline slab
      a.nother = slab_z1_ - slab_z0_;
      a.other_lo_end = a.nother;
      a.other_gap = 0;
      a.other0 = slab_z0_;
    } else {
      a.nother = axis == 1 ? g.nz : g.ny;
      a.other_lo_end = a.nother;
      a.other_gap = 0;
    }
    if (prune) {
      a.mask_lo = axis == 1 ? g.ly_lo_end : g.mz_lo_end;
      a.mask_hi = axis == 1 ? g.ly_hi_begin : g.mz_hi_begin;
    } else {
      a.mask_lo = a.mask_hi = 0;
    }
    a.nruns = cur_n();
    a.nitems = (int64_t)cur_n() * nfields * a.nother * a.nstrips;
    a.gate = gate_;
    a.plane_out = plane_out ? 1 : 0;
    a.prune_out = prune_out ? 1 : 0;
    const double cxb = sizeof(cx<T>);
    double bytes;
    if (prune_out && host_pruned_) {  // pruned step: kept columns (and pencils) only, kept outputs
      const double ocols = axis == 2 ? (double)nkept_y(g) : (double)a.nother;
      const double cols = (double)cur_n() * nfields * ocols * g.kx_keep;
      bytes = cols * n * cxb + cols * (axis == 1 ? nkept_y(g) : nkept_z(g)) * cxb;
    } else {
      const double cols = (double)cur_n() * nfields * a.nother * a.ncols;
      bytes = cols * (n - (a.mask_hi - a.mask_lo)) * cxb + cols * n * cxb;
    }
    const cx<T>* tw = axis == 1 ? twy : twz;
    if (rec_) {
      Rec r;
      r.type = Rec::COL;
      r.n = n;
      r.dir = dir;
      r.S = S;
      r.ca = a;
      r.runs = cur_rb();
      r.tw = tw;
      rec_->push_back(r);
      return 0;
    }
    return timed(name, bytes, [&] { return launch_colpass<T>(a, n, dir, S, cur_rb(), tw, cs_); });
  }

  // row launches: variant 0 everywhere except the fused products kernel (row_var)
  RowArgs row_args(int var = 0) const {
    RowArgs r{};
    r.g = g;
    r.nruns = cur_n();
    r.rows = (int64_t)g.ny * g.nz;
    const int64_t rpb = row_rows_per_block<T>(g.nx, var);
    r.blocks_per_run = (r.rows + rpb - 1) / rpb;
    r.in_ncols = g.nkr;
    r.gate = gate_;
    return r;
  }

  int xrow(const char* name, const RowArgs& a, int kind, int nfk, double bytes, int var = 0) {
    if (rec_) {
      Rec r;
      r.type = Rec::ROW;
      r.n = g.nx;
      r.kind = kind;
      r.var = var;
      r.nfk = nfk;
      r.ra = a;
      r.runs = cur_rb();
      rec_->push_back(r);
      return 0;
    }
    return timed(name, bytes, [&] { return launch_xrow<T>(a, kind, nfk, var, cur_rb(), twM, twh, cs_); });
  }

  // A launch that is not one of the hot kinds: run now, or recorded as a closure over the
  // stream it is later issued on.
  template <class F>
  int other(const char* name, F&& launch) {
    if (rec_) {
      Rec r;
      r.type = Rec::OTHER;
      r.fn = launch;
      rec_->push_back(r);
      return 0;
    }
    return timed(name, 0, [&] { return launch(st); });
  }

  // ---- recording of a time step for multi-size ensemble launches -----------------
  // While rec_ is set, step_launches_gated() records the hot launches with their arguments
  // instead of launching them, and the other launches as closures.  With rec_multi_ the
  // shapes are the multi-size kernels' (column strips of 128 threads, the row plan without
  // the LDS swizzle); otherwise the group's own tuned shapes.
  struct Rec {
    enum { OTHER = 0, COL = 1, ROW = 2, SPEC = 3 };
    int type = OTHER;
    int n = 0, dir = 0, kind = 0, var = 0, S = 0, nfk = 0;
    ColArgs ca{};
    RowArgs ra{};
    SpecArgs sa{};
    const RunBuf<T>* runs = nullptr;
    const cx<T>* tw = nullptr;
    std::function<hipError_t(hipStream_t)> fn;
  };
  std::vector<Rec>* rec_ = nullptr;
  bool rec_multi_ = false;
  uint64_t graph_gen_ = 0;  // bumped whenever the captured step is dropped (its launches changed)

  int record_step(std::vector<Rec>* out, bool multi_shapes) {
    out->clear();
    rec_ = out;
    rec_multi_ = multi_shapes;
    gate_ = 1;
    const int rc = step_launches_gated();
    gate_ = 0;
    rec_ = nullptr;
    return rc;
  }
  // launch one recorded launch of this group alone (its single-size kernel)
  hipError_t launch_rec(const Rec& r, hipStream_t s) {
    switch (r.type) {
      case Rec::COL: return launch_colpass<T>(r.ca, r.n, r.dir, r.S, r.runs, r.tw, s);
      case Rec::ROW: return launch_xrow<T>(r.ra, r.kind, r.nfk, r.var, r.runs, twM, twh, s);
      case Rec::SPEC: return launch_spectral<T>(r.sa, phys, r.runs, s);
      default: return r.fn(s);
    }
  }
  void reset_prune(hipStream_t s) { (void)launch_small<T>(SK_PRUNE, rbd, nr_, phys, 0, 0, s); }
  // the group can join multi-size launches: every axis has a multi-size transform and no run
  // has a calcF! module or volume penalisation (their launches have no multi-size form)
  bool multi_ok() const override {
    return multi_length(g.nx) && multi_length(g.ny) && multi_length(g.nz) && !any_module() && !nd_any && !vp_any;
  }
  int multi_key() const override { return ((int)sizeof(T) * 4 + phys) * 4 + stepper; }
  uint64_t graph_gen() const override { return graph_gen_; }
  // a call's prune marks: the last of its nsteps steps (per run: or its target) is full
  int mark_prune(int nsteps, hipStream_t s) {
    if (nsteps <= 0) return 0;
    const hipError_t e = launch_small<T>(SK_PRUNE, rbd, nr_, phys, nsteps, prune_enabled ? 1 : 0, s);
    if (e != hipSuccess) return fail(E_HIP, std::string("launch of prune_mark failed: ") + hipGetErrorString(e));
    return 0;
  }

  // ---- autotuning of the hot launches ------------------------------------------
  static std::map<std::string, std::vector<int>>& tune_cache() {
    static std::map<std::string, std::vector<int>> c;
    return c;
  }
  std::string tune_key() const {
    return std::to_string(dev) + ":" + std::to_string(sizeof(T)) + ":" + std::to_string(g.nx) + "x" +
           std::to_string(g.ny) + "x" + std::to_string(g.nz) + ":" + std::to_string(phys);
  }
  void apply_cached_tuning() {
    std::lock_guard<std::mutex> lk(g_tab_mu);
    auto it = tune_cache().find(tune_key());
    if (it == tune_cache().end()) return;
    for (int c = 0; c < 4; ++c) col_var[c] = it->second[c];
    row_var = it->second[4];
  }

  // time `reps` launches of one hot kernel with the given variant (work/stin only)
  int time_variant(int cls, int var, int reps, float* ms) {
    // The variant under test is set for this call only; the previous choice is restored and
    // the events are released on every return path (error returns included).
    struct Guard {
      Engine* e;
      int cls, col, row;
      hipEvent_t a = nullptr, b = nullptr;
      ~Guard() {
        if (cls < 4) e->col_var[cls] = col;
        else e->row_var = row;
        if (a) hipEventDestroy(a);
        if (b) hipEventDestroy(b);
      }
    } gd{this, cls, cls < 4 ? col_var[cls] : 0, row_var};
    if (cls < 4) col_var[cls] = var;
    else row_var = var;
    MHD_HIP(hipEventCreate(&gd.a));
    MHD_HIP(hipEventCreate(&gd.b));
    const hipEvent_t a = gd.a, b = gd.b;
    int rc = 0;
    if (cls == 4) {
      // The row kernel is timed where a step runs it: right after the inverse passes
      // that produce its input, so the input is real data and as warm in the caches as
      // in a step.  Only the row launch lies between the events.
      float sum = 0;
      for (int k = 0; k <= reps && rc == 0; ++k) {
        rc = colpass("tune", 2, 1, B_SOL, B_WORK, 0, 0, nf, true, true);
        if (rc == 0) rc = colpass("tune", 1, 1, B_WORK, B_WORK, 0, 0, nf, true, true);
        if (rc) break;
        RowArgs ra = row_args(row_var);
        ra.prune_out = gate_;
        ra.in_sel = B_WORK;
        ra.in_ncols = g.kx_keep;
        ra.out_sel = B_WORK;
        MHD_HIP(hipEventRecord(a, st));
        rc = xrow("tune", ra, phys == PHYS_MHD ? ROW_MHD : ROW_HD, nf, 0, row_var);
        MHD_HIP(hipEventRecord(b, st));
        MHD_HIP(hipEventSynchronize(b));
        float e = 0;
        MHD_HIP(hipEventElapsedTime(&e, a, b));
        if (k > 0) sum += e;  // the first launch warms up
      }
      *ms = sum;
      return rc;
    }
    for (int k = 0; k <= reps && rc == 0; ++k) {
      if (k == 1) MHD_HIP(hipEventRecord(a, st));  // first launch warms up
      switch (cls) {
        case 0: rc = colpass("tune", 2, 1, B_SOL, B_WORK, 0, 0, nf, true, true); break;
        case 1: rc = colpass("tune", 1, 1, B_WORK, B_WORK, 0, 0, nf, true, true); break;
        case 2: rc = colpass("tune", 1, -1, B_WORK, B_WORK, 0, 0, np, false, true, false, gate_ != 0); break;
        case 3: rc = colpass("tune", 2, -1, B_WORK, B_WORK, 0, 0, np, false, true, false, gate_ != 0); break;
        default: {
          RowArgs ra = row_args(row_var);
          ra.prune_out = gate_;
          ra.in_sel = B_WORK;
          ra.in_ncols = g.kx_keep;
          ra.out_sel = B_WORK;
          rc = xrow("tune", ra, phys == PHYS_MHD ? ROW_MHD : ROW_HD, nf, 0, row_var);
        }
      }
    }
    MHD_HIP(hipEventRecord(b, st));
    MHD_HIP(hipEventSynchronize(b));
    MHD_HIP(hipEventElapsedTime(ms, a, b));
    return rc;
  }

  // Pick the fastest variant of each hot launch; the captured step is rebuilt.
  // Uses work and stin as scratch (their contents are dead between steps).
  int autotune(int reps, std::string* report) {
    if (reps <= 0) reps = 5;
    MHD_HIP(hipStreamSynchronize(st));
    const char* names[5] = {"zpass_inv", "ypass_inv", "ypass_fwd", "zpass_fwd", "xrow_fused"};
    std::vector<int> best(5, 0);
    // The forward passes and the row kernel are timed as dealiased steps run them (the
    // common case): a step launch on a clock marked for a pruned step.
    if (prune_enabled) MHD_HIP(launch_small<T>(SK_PRUNE, rbd, nr_, phys, 2, 1, st));
    for (int cls = 0; cls < 5; ++cls) {
      gate_ = (prune_enabled && cls >= 2) ? 1 : 0;
      float bms = 1e30f;
      const int nvar = 4;
      std::string line = std::string(names[cls]) + ":";
      for (int v = 0; v < nvar; ++v) {
        if (cls < 4) {
          const int n = (cls == 0 || cls == 3) ? g.nz : g.ny;
          bool dup = false;  // a strip width equal to a lower variant's (or invalid) is not timed again
          for (int u = 0; u < v; ++u)
            if (col_strip<T>(n, u) == col_strip<T>(n, v)) dup = true;
          if (dup) continue;
        } else {
          // row variants: bit 0 plan, bit 1 swizzled LDS; skip a plan equal to a lower variant's
          bool dup = false;
          for (int u = 0; u < v; ++u)
            if ((u >> 1) == (v >> 1) && row_rows_per_block<T>(g.nx, u) == row_rows_per_block<T>(g.nx, v)) dup = true;
          if (dup) continue;
        }
        float ms = 0;
        const int rc = time_variant(cls, v, reps, &ms);
        if (rc) {
          gate_ = 0;
          return rc;
        }
        line += " v" + std::to_string(v) + "=" + std::to_string(ms / reps) + "ms";
        if (ms < bms) {
          bms = ms;
          best[cls] = v;
        }
      }
      if (report) *report += line + " -> v" + std::to_string(best[cls]) + "\n";
    }
    gate_ = 0;
    MHD_HIP(launch_small<T>(SK_PRUNE, rbd, nr_, phys, 0, 0, st));
    for (int c = 0; c < 4; ++c) col_var[c] = best[c];
    row_var = best[4];
    {
      std::lock_guard<std::mutex> lk(g_tab_mu);
      tune_cache()[tune_key()] = best;
    }
    drop_graph();
    return 0;
  }

  // v5 = {zpass_inv, ypass_inv, ypass_fwd, zpass_fwd, xrow_fused}; entries < 0 keep the current choice
  int set_variants(const int32_t* v5, int32_t* out5) override {
    if (v5) {
      for (int c = 0; c < 4; ++c)
        if (v5[c] >= 0) {
          if (v5[c] > 3) return fail(E_INVALID, "column-pass variant must be 0..3");
          col_var[c] = v5[c];
        }
      if (v5[4] >= 0) {
        if (v5[4] > 3) return fail(E_INVALID, "row variant must be 0..3");
        row_var = v5[4];
      }
      drop_graph();
    }
    if (out5) {
      for (int c = 0; c < 4; ++c) out5[c] = col_var[c];
      out5[4] = row_var;
    }
    return 0;
  }

  int tune(int reps, char* report, int len) override {
    std::string r;
    MHD_TRY(autotune(reps, &r));
    if (report && len > 0) {
      std::strncpy(report, r.c_str(), (size_t)len - 1);
      report[len - 1] = 0;
    }
    return 0;
  }

  // ---- one right-hand side evaluation ----------------------------------------
  // in_sel: stage input; prune: the stage input is dealiased; last: write vars + maxima.
  int rhs(int in_sel, bool prune, SpecArgs sp, bool last) {
    // a masked stage input has its kx = 0 plane copied aside by the z-pass, so the
    // spectral launch covers every mode at once (kx = 0 mirrors read from the copy)
    const bool fused_planes = prune && sp.mask_input && sp.self_sel < 0;
    MHD_TRY(colpass("zpass_inv", 2, 1, in_sel, B_WORK, 0, 0, nf, prune, true, fused_planes));
    // negative damping adds three products (f_i u_i) and its sum to the MHD row launch;
    // volume penalisation adds nf products chi/eta (v - V0) (u, and b for MHD) to every one
    const bool nd = nd_any && phys == PHYS_MHD && sp.add_force;
    const bool vp = vp_any;
    // a recorded step keeps the row plan (variant bit 0) and drops the LDS swizzle (bit 1),
    // which the multi-size row kernel does not have and which changes no value
    const int rvar = (nd || vp) ? 0 : ((rec_ && rec_multi_) ? (row_var & 1) : row_var);
    const int npf = nd ? 12 : (vp ? np + nf : np);
    RowArgs ra = row_args(rvar);
    ra.in_sel = B_WORK;
    ra.in_ncols = prune ? g.kx_keep : g.nkr;
    ra.out_sel = B_WORK;
    ra.write_vars = last ? 1 : 0;
    ra.want_max = last ? 1 : 0;
    const double rows = (double)nr_ * g.ny * g.nz;
    ra.prune_out = gate_;
    const int kout = host_pruned_ ? g.kx_keep : g.nkr;
    // the real fields are stored at the last stage of a full step only
    double bytes = rows * (nf * ra.in_ncols + npf * kout) * sizeof(cx<T>) +
                   (last && !host_pruned_ ? rows * nf * g.nx * sizeof(T) : 0) + (nd ? rows * 3 * g.nx * sizeof(T) : 0) +
                   (vp ? rows * (1 + nf) * g.nx * sizeof(T) : 0);
    const int kind = phys == PHYS_MHD ? (nd ? ROW_MHD_ND : (vp ? ROW_MHD_VP : ROW_MHD)) : (vp ? ROW_HD_VP : ROW_HD);
    std::vector<int> zb;
    const int nslab = pipe_slabs(rvar, &zb);
    if (nslab <= 1) {
      MHD_TRY(colpass("ypass_inv", 1, 1, B_WORK, B_WORK, 0, 0, nf, prune, true));
      MHD_TRY(xrow("xrow_fused", ra, kind, nf, bytes, rvar));
      MHD_TRY(colpass("ypass_fwd", 1, -1, B_WORK, B_WORK, 0, 0, npf, false, true, false, gate_ != 0));
    } else {
      struct SlabReset {  // back to the whole grid on st on every return path
        Engine* e;
        ~SlabReset() {
          e->cs_ = e->st;
          e->slab_z0_ = 0;
          e->slab_z1_ = -1;
        }
      } reset{this};
      const int64_t rpb = row_rows_per_block<T>(g.nx, rvar);
      MHD_HIP(hipEventRecord(ev_fork, st));
      MHD_HIP(hipStreamWaitEvent(st2, ev_fork, 0));
      for (int k = 0; k < nslab; ++k) {
        cs_ = (k & 1) ? st2 : st;
        slab_z0_ = zb[k];
        slab_z1_ = zb[k + 1];
        MHD_TRY(colpass("ypass_inv", 1, 1, B_WORK, B_WORK, 0, 0, nf, prune, true));
        RowArgs rs = ra;
        rs.blk0 = (int64_t)slab_z0_ * g.ny / rpb;
        rs.rows = (int64_t)slab_z1_ * g.ny;
        rs.blocks_per_run = (rs.rows - rs.blk0 * rpb + rpb - 1) / rpb;
        MHD_TRY(xrow("xrow_fused", rs, kind, nf, bytes * (slab_z1_ - slab_z0_) / g.nz, rvar));
        MHD_TRY(colpass("ypass_fwd", 1, -1, B_WORK, B_WORK, 0, 0, npf, false, true, false, gate_ != 0));
      }
      MHD_HIP(hipEventRecord(ev_join, st2));
      MHD_HIP(hipStreamWaitEvent(st, ev_join, 0));
    }
    MHD_TRY(colpass("zpass_fwd", 2, -1, B_WORK, B_WORK, 0, 0, npf, false, true, false, gate_ != 0));
    if (nd) {
      const double dV = (double)clk_dx(0) * (double)clk_dx(1) * (double)clk_dx(2);
      MHD_TRY(other("nd_amp", [this, nb = ra.blocks_per_run, dV, gt = gate_](hipStream_t s) {
        return launch_nd_amp<T>(rbd, nr_, nb, dV, gt, s);
      }));
    }
    const int64_t used_blocks = ra.blocks_per_run;
    sp.g = g;
    sp.nruns = nr_;
    sp.nitems = (int64_t)nr_ * g.nc;
    sp.prod_sel = B_WORK;
    sp.gate = gate_;
    sp.planes = fused_planes ? 2 : 0;
    if (rec_) {
      Rec r;
      r.type = Rec::SPEC;
      r.sa = sp;
      r.runs = rbd;
      rec_->push_back(r);
    } else {
      MHD_TRY(timed("spectral", spectral_bytes(sp), [&] { return launch_spectral<T>(sp, phys, rbd, st); }));
    }
    if (last) MHD_TRY(other("reduce", [this, used_blocks, adv = sp.mode == SP_RHS ? 0 : 1](hipStream_t s) {
      return launch_small<T>(SK_REDUCE, rbd, nr_, phys, used_blocks, adv, s);
    }));
    return 0;
  }

  double spectral_bytes(const SpecArgs& sp) const {
    const double c = sizeof(cx<T>) * (double)nr_;
    const double kept = (double)g.kx_keep * nkept_y(g) * nkept_z(g);
    const double all = (sp.gate && host_pruned_) ? kept : (double)g.nc;  // pruned step: kept modes only
    double rd = (np + forcing_fields()) * all, wr = 0;
    if (sp.mode == SP_RK4) {
      rd += nf * kept;                                  // stage input
      if (sp.stage > 0) rd += nf * kept + nf * all;     // base sol + acc
      wr += sp.stage < 3 ? nf * all + nf * kept : nf * all;
    } else if (sp.mode == SP_LSRK54) {
      rd += nf * kept + (sp.stage > 0 ? nf * all : 0);
      wr += 2 * nf * all;
    } else {
      rd += nf * (sp.self_sel >= 0 ? all : kept);
      wr += (sp.f_hi - sp.f_lo) * all;
    }
    return (rd + wr) * c;
  }

  // extra complex fields the spectral launch reads per mode for forcing (dense spectrum or
  // negative-damping products)
  int forcing_fields() const {
    int n = 0;
    for (int r = 0; r < nr_; ++r)
      if (rbh[r].fdense || rbh[r].ndf) n = 3;
    return n + (vp_any ? nf : 0);
  }

  bool any_module() const {
    for (int r = 0; r < nr_; ++r)
      if (rbh[r].fdense || rbh[r].ndf || rbh[r].a99 || rbh[r].vp) return true;
    return false;
  }

  SpecArgs spec_base() const {
    SpecArgs s{};
    s.g = g;
    s.modules = any_module() ? 1 : 0;
    s.mask_input = 1;
    s.add_force = 1;
    s.self_sel = -1;
    s.f_lo = 0;
    s.f_hi = nf;
    return s;
  }

  // ---- one time step -----------------------------------------------------------
  int step_launches() {
    gate_ = 1;
    int rc = step_launches_gated();
    gate_ = 0;
    return rc;
  }

  int step_launches_gated() {
    MHD_TRY(other("cfl", [this](hipStream_t s) { return launch_small<T>(SK_CFL, rbd, nr_, phys, 0, 0, s); }));
    if (stepper == STEP_RK4) {
      for (int s = 0; s < 4; ++s) {
        SpecArgs sp = spec_base();
        sp.mode = SP_RK4;
        sp.stage = s;
        sp.in_sel = s == 0 ? B_SOL : B_STIN;
        MHD_TRY(rhs(sp.in_sel, true, sp, s == 3));
      }
    } else {
      static const double A[5] = {0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
                                  -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0};
      static const double B[5] = {1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                                  1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
                                  2277821191437.0 / 14882151754819.0};
      for (int s = 0; s < 5; ++s) {
        SpecArgs sp = spec_base();
        sp.mode = SP_LSRK54;
        sp.stage = s;
        sp.in_sel = B_SOL;
        sp.ca = A[s];
        sp.cb = B[s];
        MHD_TRY(rhs(B_SOL, true, sp, s == 4));
      }
    }
    return 0;
  }

  int ensure_graph() {
    if (gexec) return 0;
    MHD_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int rc = step_launches();
    hipGraph_t gr = nullptr;
    hipError_t e = hipStreamEndCapture(st, &gr);
    if (rc) {
      if (gr) hipGraphDestroy(gr);
      return rc;
    }
    if (e != hipSuccess) return fail(E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    graph = gr;
    MHD_HIP(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
    return 0;
  }

  int step(int nsteps) override {
    if (nsteps < 0) return fail(E_INVALID, "nsteps must be >= 0");
    MHD_TRY(mark_prune(nsteps, st));  // the last of these steps (per run: or its target) is computed in full
    if (use_graph && !profiling) {
      MHD_TRY(ensure_graph());
      for (int i = 0; i < nsteps; ++i) MHD_HIP(hipGraphLaunch(gexec, st));
      return 0;
    }
    for (int i = 0; i < nsteps; ++i) MHD_TRY(step_launches());
    return 0;
  }

  int set_pruning(int enable) override {
    prune_enabled = enable != 0;
    return 0;
  }

  int set_graph(int enable) override {
    use_graph = enable != 0;
    return 0;
  }

  // ---- state boundary -------------------------------------------------------------
  int check_run(int run) const {
    if (run < 0 || run >= nr_) return fail(E_INVALID, "run index out of range");
    return 0;
  }

  int set_ic_real(int run, const void* const* fields) override {
    MHD_TRY(check_run(run));
    RunBuf<T>& b = rbh[run];
    bool any = false;
    for (int f = 0; f < nf; ++f) {
      if (!fields[f]) continue;
      any = true;
      MHD_HIP(hipMemcpyAsync(b.vars + (int64_t)f * g.nr, fields[f], sizeof(T) * g.nr, hipMemcpyHostToDevice, st));
    }
    if (!any) return 0;
    RunScope scope(this, run);  // only this run's buffers are transformed
    MHD_TRY(r2c_vars_to_work(run));
    for (int f = 0; f < nf; ++f) {
      if (!fields[f]) continue;
      // work rows and planes are padded: copy into the unpadded state layout
      MHD_TRY(timed("unpad", 0, [&] {
        return launch_unpad<T>(g, b.buf[B_WORK] + (int64_t)f * g.wnc, b.buf[B_SOL] + (int64_t)f * g.nc, st);
      }));
    }
    MHD_TRY(vars_maxima());
    MHD_HIP(hipStreamSynchronize(st));
    return 0;
  }

  // R2C of all vars fields into work[0..nf) (every run of the group)
  int r2c_vars_to_work(int /*run*/) {
    for (int f0 = 0; f0 < nf; f0 += 3) {
      RowArgs ra = row_args();
      ra.in_f0 = f0;
      ra.real_sel = 0;
      ra.out2_sel = B_WORK;
      ra.out2_f0 = f0;
      MHD_TRY(xrow("xrow_r2c", ra, ROW_R2C, 3, 0));
    }
    MHD_TRY(colpass("ypass_fwd", 1, -1, B_WORK, B_WORK, 0, 0, nf, false));
    MHD_TRY(colpass("zpass_fwd", 2, -1, B_WORK, B_WORK, 0, 0, nf, false));
    return 0;
  }

  // max u^2 / b^2 over vars into the clocks (getCFL! input after the fields are set)
  int vars_maxima() {
    const int64_t nb = std::min<int64_t>(row_blocks, 1024);
    MHD_TRY(timed("varsmax", 0, [&] { return launch_small<T>(SK_VARSMAX, cur_rb(), cur_n(), nf, nb, (int)g.nr, st); }));
    MHD_TRY(timed("reduce", 0, [&] { return launch_small<T>(SK_REDUCE, cur_rb(), cur_n(), phys, nb, 0, st); }));
    return 0;
  }

  int get_sol(int run, void* host) override {
    MHD_TRY(check_run(run));
    MHD_HIP(hipMemcpyAsync(host, rbh[run].buf[B_SOL], sizeof(cx<T>) * g.nc * nf, hipMemcpyDeviceToHost, st));
    MHD_HIP(hipStreamSynchronize(st));
    return 0;
  }
  int set_sol(int run, const void* host) override {
    MHD_TRY(check_run(run));
    MHD_HIP(hipMemcpyAsync(rbh[run].buf[B_SOL], host, sizeof(cx<T>) * g.nc * nf, hipMemcpyHostToDevice, st));
    MHD_HIP(hipStreamSynchronize(st));
    return 0;
  }
  // ---- sparse half-spectrum entries ----------------------------------------------
  // Device staging of (kx, l, m) index triples and values: idx first (padded to 256 B),
  // then nf complex T per entry.  The engine keeps one buffer that grows on demand.
  void* stage_ = nullptr;
  size_t stage_cap_ = 0;
  static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
  size_t mode_stage_bytes(int64_t nmodes) const override {
    return align256(sizeof(int32_t) * 3 * (size_t)nmodes) + sizeof(cx<T>) * (size_t)nf * (size_t)nmodes;
  }
  int grow_stage(size_t bytes) {
    if (bytes <= stage_cap_) return 0;
    MHD_HIP(hipStreamSynchronize(st));
    if (stage_) hipFree(stage_);
    if (spec_) hipFree(spec_);
    stage_ = nullptr;
    stage_cap_ = 0;
    MHD_HIP(hipMalloc(&stage_, bytes));
    stage_cap_ = bytes;
    return 0;
  }
  int check_modes(int64_t nmodes, const int32_t* idx) const {
    if (nmodes < 0) return fail(E_INVALID, "nmodes must be >= 0");
    for (int64_t i = 0; i < nmodes; ++i) {
      if (idx[3 * i] < 0 || idx[3 * i] >= g.nkr || idx[3 * i + 1] < 0 || idx[3 * i + 1] >= g.ny ||
          idx[3 * i + 2] < 0 || idx[3 * i + 2] >= g.nz)
        return fail(E_INVALID, "mode index outside the half spectrum");
    }
    return 0;
  }

  // zero sol, scatter sparse modes, then refresh vars (and the CFL maxima) from it
  int set_sol_modes(int run, int64_t nmodes, const int32_t* idx, const void* vals) override {
    MHD_TRY(check_run(run));
    MHD_TRY(check_modes(nmodes, idx));
    if (nmodes > 0) MHD_TRY(grow_stage(mode_stage_bytes(nmodes)));
    return set_sol_modes_staged(run, nmodes, idx, vals, stage_, stage_cap_);
  }

  int set_sol_modes_staged(int run, int64_t nmodes, const int32_t* idx, const void* vals, void* stage,
                           size_t cap) override {
    MHD_TRY(check_run(run));
    MHD_TRY(check_modes(nmodes, idx));
    if (nmodes > 0 && (!stage || mode_stage_bytes(nmodes) > cap))
      return fail(E_INVALID, "staging buffer too small for " + std::to_string(nmodes) + " modes");
    RunBuf<T>& b = rbh[run];
    MHD_TRY(timed("zero", 0, [&] { return launch_zero<T>(b.buf[B_SOL], g.nc * nf, st); }));
    if (nmodes > 0) {
      int32_t* di = (int32_t*)stage;
      cx<T>* dv = (cx<T>*)((char*)stage + align256(sizeof(int32_t) * 3 * (size_t)nmodes));
      MHD_HIP(hipMemcpyAsync(di, idx, sizeof(int32_t) * 3 * nmodes, hipMemcpyHostToDevice, st));
      MHD_HIP(hipMemcpyAsync(dv, vals, sizeof(cx<T>) * nf * nmodes, hipMemcpyHostToDevice, st));
      MHD_TRY(timed("scatter", 0, [&] { return launch_scatter<T>(g, b.buf[B_SOL], nf, nmodes, di, dv, st); }));
    }
    return refresh_vars_run(run);  // ends with a stream synchronisation: the staging is free again
  }

  int get_sol_modes(int run, int64_t nmodes, const int32_t* idx, void* vals) override {
    MHD_TRY(check_run(run));
    MHD_TRY(check_modes(nmodes, idx));
    if (nmodes == 0) return 0;
    MHD_TRY(grow_stage(mode_stage_bytes(nmodes)));
    int32_t* di = (int32_t*)stage_;
    cx<T>* dv = (cx<T>*)((char*)stage_ + align256(sizeof(int32_t) * 3 * (size_t)nmodes));
    MHD_HIP(hipMemcpyAsync(di, idx, sizeof(int32_t) * 3 * nmodes, hipMemcpyHostToDevice, st));
    MHD_TRY(timed("gather", 0, [&] { return launch_gather<T>(g, rbh[run].buf[B_SOL], nf, nmodes, di, dv, st); }));
    MHD_HIP(hipMemcpyAsync(vals, dv, sizeof(cx<T>) * nf * nmodes, hipMemcpyDeviceToHost, st));
    MHD_HIP(hipStreamSynchronize(st));
    return 0;
  }

  int get_vars(int run, int f, void* host) override {
    MHD_TRY(check_run(run));
    if (f < 0 || f >= nf) return fail(E_INVALID, "field index out of range");
    MHD_HIP(hipMemcpyAsync(host, rbh[run].vars + (int64_t)f * g.nr, sizeof(T) * g.nr, hipMemcpyDeviceToHost, st));
    MHD_HIP(hipStreamSynchronize(st));
    return 0;
  }
  int device_ptrs(int run, void** sol, void** vars) override {
    MHD_TRY(check_run(run));
    if (sol) *sol = rbh[run].buf[B_SOL];
    if (vars) *vars = rbh[run].vars;
    return 0;
  }
  int get_clock(int run, double* t, double* dt, int64_t* step) override {
    MHD_TRY(check_run(run));
    RunClock c;
    MHD_HIP(hipMemcpyAsync(&c, rbh[run].clk, sizeof(c), hipMemcpyDeviceToHost, st));
    MHD_HIP(hipStreamSynchronize(st));
    if (t) *t = c.t;
    if (dt) *dt = c.dt;
    if (step) *step = c.step;
    return 0;
  }
  int clock_io(int run, RunClock* c, int write) override {
    MHD_TRY(check_run(run));
    return write ? write_clock(run, *c) : read_clock(run, c);
  }
  int precision_bytes() const override { return (int)sizeof(T); }
  const Geom& geom() const override { return g; }
  int nfields() const override { return nf; }

  int read_clock(int run, RunClock* c) {
    MHD_HIP(hipMemcpyAsync(c, rbh[run].clk, sizeof(RunClock), hipMemcpyDeviceToHost, st));
    MHD_HIP(hipStreamSynchronize(st));
    return 0;
  }
  int write_clock(int run, const RunClock& c) {
    MHD_HIP(hipMemcpyAsync(rbh[run].clk, &c, sizeof(RunClock), hipMemcpyHostToDevice, st));
    MHD_HIP(hipStreamSynchronize(st));
    return 0;
  }
  int set_clock(int run, double t, double dt, int64_t step) override {
    MHD_TRY(check_run(run));
    RunClock c;
    MHD_TRY(read_clock(run, &c));
    c.t = (double)(T)t;
    c.dt = (double)(T)dt;
    c.step = step;
    return write_clock(run, c);
  }
  int set_run_params(int run, double nu, double eta, int nnu) override {
    MHD_TRY(check_run(run));
    rbh[run].nu = nu;
    rbh[run].eta = phys == PHYS_MHD ? eta : 0.0;
    rbh[run].nnu = nnu;
    MHD_TRY(upload_runs());
    MHD_HIP(hipStreamSynchronize(st));
    return 0;
  }
  int start_run(int run, double dt, double cfl_coef, int64_t target) override {
    MHD_TRY(check_run(run));
    RunClock c;
    MHD_TRY(read_clock(run, &c));
    c.t = 0.0;
    c.step = 0;
    c.dt = (double)(T)dt;
    c.use_cfl = cfl_coef > 0 ? 1 : 0;
    c.cfl_coef = cfl_coef;
    c.t_diff = t_diff(run, cfl_coef > 0 ? cfl_coef : 1.0);
    c.target = target;
    return write_clock(run, c);
  }
  int query() override {
    hipError_t e = hipStreamQuery(st);
    if (e == hipSuccess) return 0;
    if (e == hipErrorNotReady) return 1;
    return fail(E_HIP, std::string("hipStreamQuery: ") + hipGetErrorString(e));
  }
  int set_cfl(int run, int use, double coef) override {
    MHD_TRY(check_run(run));
    RunClock c;
    MHD_TRY(read_clock(run, &c));
    c.use_cfl = use;
    c.cfl_coef = coef;
    c.t_diff = t_diff(run, coef);
    return write_clock(run, c);
  }

  // diffusive bound t_diff (integrator.jl:60-72); nη is never forwarded, so it is 0
  double t_diff(int run, double coef) const {
    const RunBuf<T>& b = rbh[run];
    double vi, nv;
    if (phys == PHYS_MHD) {
      vi = std::max(b.nu, b.eta);
      nv = std::max(b.nnu, b.neta);
    } else {
      vi = b.nu;
      nv = b.nnu;
    }
    // dl^nv is evaluated in T before the Float64 coefficient multiplies it
    const T dl = (T)clk_init[run].dl;
    if (vi == 0) return std::numeric_limits<double>::infinity();
    T p = dl;
    for (int i = 1; i < (nv > 1 ? (int)nv : 2); ++i) p = p * dl;
    return coef * (double)p / vi;
  }

  // ---- operator boundary -------------------------------------------------------
  int calcN(void* N, void* sol, int advection_only) override {
    if (nr_ != 1) return fail(E_INVALID, "calcN acts on single-run problems");
    rbh[0].buf[B_EXT_SOL] = (cx<T>*)sol;
    rbh[0].buf[B_EXT_N] = (cx<T>*)N;
    MHD_TRY(upload_runs());
    SpecArgs sp = spec_base();
    sp.mode = SP_RHS;
    sp.in_sel = B_EXT_SOL;
    sp.out_sel = B_EXT_N;
    sp.out_f0 = 0;
    sp.mask_input = advection_only ? 0 : 1;
    sp.zero_aliased = advection_only ? 0 : 1;
    sp.add_force = advection_only ? 0 : 1;
    if (!advection_only) sp.draw_ctr = (int64_t(1) << 40) + calls_++;
    MHD_TRY(rhs(B_EXT_SOL, !advection_only, sp, true));
    MHD_HIP(hipStreamSynchronize(st));
    return 0;
  }

  int component_update(void* N, int field) override {
    if (nr_ != 1) return fail(E_INVALID, "component updates act on single-run problems");
    if (vp_any)
      return fail(E_UNSUPPORTED, "the per-direction updates do not include volume penalisation; use calcN");
    if (field < 0 || field >= nf) return fail(E_INVALID, "direction out of range");
    rbh[0].buf[B_EXT_N] = (cx<T>*)N;
    MHD_TRY(upload_runs());
    RowArgs ra = row_args();
    ra.in_f0 = 0;
    ra.real_sel = 0;
    ra.out_sel = B_WORK;
    ra.out_f0 = 0;
    ra.out2_sel = B_STIN;
    ra.out2_f0 = 0;
    MHD_TRY(xrow("xrow_vars", ra, phys == PHYS_MHD ? ROW_VARS_MHD : ROW_VARS_HD, nf, 0));
    MHD_TRY(colpass("ypass_fwd", 1, -1, B_WORK, B_WORK, 0, 0, np, false));
    MHD_TRY(colpass("zpass_fwd", 2, -1, B_WORK, B_WORK, 0, 0, np, false));
    MHD_TRY(colpass("ypass_fwd", 1, -1, B_STIN, B_STIN, 0, 0, nf, false));
    MHD_TRY(colpass("zpass_fwd", 2, -1, B_STIN, B_STIN, 0, 0, nf, false));
    SpecArgs sp = spec_base();
    sp.nruns = 1;
    sp.nitems = g.nc;
    sp.mode = SP_RHS;
    sp.in_sel = B_STIN;
    sp.mask_input = 0;
    sp.add_force = 0;
    sp.prod_sel = B_WORK;
    sp.self_sel = B_STIN;
    sp.self_f0 = 0;
    sp.out_sel = B_EXT_N;
    sp.out_f0 = field;
    sp.f_lo = field;
    sp.f_hi = field + 1;
    MHD_TRY(timed("spectral", 0, [&] { return launch_spectral<T>(sp, phys, rbd, st); }));
    MHD_HIP(hipStreamSynchronize(st));
    return 0;
  }

  int cfl_now(double coef, double* dt) override {
    for (int r = 0; r < nr_; ++r) MHD_TRY(set_cfl(r, 1, coef));
    MHD_TRY(timed("cfl", 0, [&] { return launch_small<T>(SK_CFL, rbd, nr_, phys, 0, 0, st); }));
    RunClock c;
    MHD_TRY(read_clock(0, &c));
    for (int r = 0; r < nr_; ++r) {
      RunClock cr;
      MHD_TRY(read_clock(r, &cr));
      cr.use_cfl = 0;
      MHD_TRY(write_clock(r, cr));
    }
    if (dt) *dt = c.dt;
    return 0;
  }

  int integrate(double t0, int64_t N0, double usr_dt, double coef, int64_t* taken) override {
    if (nr_ != 1) return fail(E_INVALID, "integrate acts on single-run problems");
    RunClock c;
    MHD_TRY(read_clock(0, &c));
    c.step = 0;  // integrator.jl:76
    if (usr_dt != 0.0) {
      c.dt = (double)(T)usr_dt;
      c.use_cfl = 0;
    } else {
      c.use_cfl = 1;
      c.cfl_coef = coef;
      c.t_diff = t_diff(0, coef);
    }
    MHD_TRY(write_clock(0, c));
    int64_t n = 0;
    if (std::isinf(t0) && t0 > 0) {
      // while N0 >= step: exactly N0+1 steps, no host round trips
      if (N0 >= 0) {
        int64_t left = N0 + 1;
        while (left > 0) {
          const int chunk = (int)std::min<int64_t>(left, 1 << 20);
          MHD_TRY(step(chunk));
          left -= chunk;
        }
        n = N0 + 1;
      }
    } else {
      while (true) {
        MHD_TRY(read_clock(0, &c));
        if (!(N0 >= c.step && t0 >= c.t)) break;
        MHD_TRY(step(1));
        ++n;
      }
    }
    MHD_TRY(read_clock(0, &c));
    c.use_cfl = 0;
    MHD_TRY(write_clock(0, c));
    if (taken) *taken = n;
    return 0;
  }

  // ---- vars refresh and diagnostics -------------------------------------------
  int refresh_vars_run(int run) {
    RunScope scope(this, run);
    return refresh_vars();
  }

  int refresh_vars() override {
    MHD_TRY(colpass("zpass_inv", 2, 1, B_SOL, B_WORK, 0, 0, nf, true));
    MHD_TRY(colpass("ypass_inv", 1, 1, B_WORK, B_WORK, 0, 0, nf, true));
    for (int f0 = 0; f0 < nf; f0 += 3) {
      RowArgs ra = row_args();
      ra.in_sel = B_WORK;
      ra.in_f0 = f0;
      ra.in_ncols = g.kx_keep;
      ra.real_sel = 0;
      ra.real_out_f0 = f0;
      ra.write_vars = 1;
      MHD_TRY(xrow("xrow_c2r", ra, ROW_C2R, 3, 0));
    }
    MHD_TRY(vars_maxima());
    MHD_HIP(hipStreamSynchronize(st));
    return 0;
  }

  // spectral op on work[0..3) -> work[3..), back to real space into rscr[0..ncomp)
  int diag_field(int op, int ncomp) {
    for (int r = 0; r < nr_; ++r)
      MHD_TRY(timed("diag", 0, [&] {
        return launch_diag_spectral<T>(g, rbh[r].buf[B_WORK], rbh[r].buf[B_WORK] + 3 * g.wnc, op, st);
      }));
    MHD_TRY(colpass("zpass_inv", 2, 1, B_WORK, B_WORK, 3, 3, ncomp, false));
    MHD_TRY(colpass("ypass_inv", 1, 1, B_WORK, B_WORK, 3, 3, ncomp, false));
    RowArgs ra = row_args();
    ra.in_sel = B_WORK;
    ra.in_f0 = 3;
    ra.in_ncols = g.nkr;
    ra.real_sel = 1;
    ra.real_out_f0 = 0;
    ra.write_vars = 1;
    MHD_TRY(xrow("xrow_c2r", ra, ROW_C2R, ncomp, 0));
    return 0;
  }

  int reduce_dpart(int nblk, int width, std::vector<double>* sums) {
    std::vector<double> h((size_t)nblk * width);
    MHD_HIP(hipMemcpyAsync(h.data(), dpart, sizeof(double) * nblk * width, hipMemcpyDeviceToHost, st));
    MHD_HIP(hipStreamSynchronize(st));
    sums->assign(width, 0.0);
    for (int i = 0; i < nblk; ++i)
      for (int w = 0; w < width; ++w) {
        if (w == 1 && width == 2) {
          const double v = h[(size_t)i * width + w];
          (*sums)[w] = (v > (*sums)[w] || v != v) ? v : (*sums)[w];
        } else {
          (*sums)[w] += h[(size_t)i * width + w];
        }
      }
    return 0;
  }

  int diagnostics(double* out) override {
    if (nr_ != 1) return fail(E_INVALID, "diagnostics act on single-run problems");
    const RunBuf<T>& b = rbh[0];
    const double dV = (double)(T)((T)(clk_dx(0))) * (double)(T)clk_dx(1) * (double)(T)clk_dx(2);
    const int nblk = dpart_blocks;
    std::vector<double> s;
    MHD_TRY(timed("diag", 0, [&] { return launch_field_sums<T>(b.vars, g.nr, nf, dpart, nblk, st); }));
    MHD_TRY(reduce_dpart(nblk, 3, &s));
    out[0] = s[0] * dV;
    out[1] = s[1] * dV;
    out[4] = s[2] * dV;
    // kinetic helicity and max |div u|
    MHD_TRY(r2c_fields(0));
    MHD_TRY(diag_field(OP_CURL, 3));
    MHD_TRY(timed("diag", 0, [&] { return launch_dot_max<T>(b.rscr, b.vars, g.nr, 3, dpart, nblk, st); }));
    MHD_TRY(reduce_dpart(nblk, 2, &s));
    out[2] = s[0] * dV;
    MHD_TRY(diag_field(OP_DIV, 1));
    MHD_TRY(timed("diag", 0, [&] { return launch_dot_max<T>(b.rscr, nullptr, g.nr, 1, dpart, nblk, st); }));
    MHD_TRY(reduce_dpart(nblk, 2, &s));
    out[6] = s[1];
    if (phys == PHYS_MHD) {
      MHD_TRY(r2c_fields(3));
      MHD_TRY(diag_field(OP_VECPOT, 3));
      MHD_TRY(timed("diag", 0, [&] {
        return launch_dot_max<T>(b.rscr, b.vars + 3 * g.nr, g.nr, 3, dpart, nblk, st);
      }));
      MHD_TRY(reduce_dpart(nblk, 2, &s));
      out[3] = s[0] * dV;
      MHD_TRY(diag_field(OP_DIV, 1));
      MHD_TRY(timed("diag", 0, [&] { return launch_dot_max<T>(b.rscr, nullptr, g.nr, 1, dpart, nblk, st); }));
      MHD_TRY(reduce_dpart(nblk, 2, &s));
      out[5] = s[1];
    } else {
      out[3] = out[5] = 0.0;
    }
    out[7] = (out[0] != out[0] || out[1] != out[1]) ? 1.0 : 0.0;
    return 0;
  }
  double lens[3] = {2 * M_PI, 2 * M_PI, 2 * M_PI};
  T clk_dx(int axis) const {
    const int n = axis == 0 ? g.nx : axis == 1 ? g.ny : g.nz;
    return (T)((T)lens[axis] / (T)n);
  }

  // R2C of vars fields f0..f0+2 into work[0..3)
  int r2c_fields(int f0) {
    RowArgs ra = row_args();
    ra.in_f0 = f0;
    ra.real_sel = 0;
    ra.out2_sel = B_WORK;
    ra.out2_f0 = 0;
    MHD_TRY(xrow("xrow_r2c", ra, ROW_R2C, 3, 0));
    MHD_TRY(colpass("ypass_fwd", 1, -1, B_WORK, B_WORK, 0, 0, 3, false));
    MHD_TRY(colpass("zpass_fwd", 2, -1, B_WORK, B_WORK, 0, 0, 3, false));
    return 0;
  }

  // ---- analysis (src/utils/MHDAnalysis.jl) ----------------------------------------------
  // krmax = round(Int, maximum(sqrt(Krsq)) + 1) (MHDAnalysis.jl:244), |k| in T
  int spectrum_bins() const override {
    const T qx = (T)((double)(g.nx / 2) * g.multx), qy = (T)((double)(g.ny / 2) * g.multy);
    const T qz = (T)((double)(g.nz / 2) * g.multz);
    const T kmax = std::sqrt((qx * qx + qy * qy) + qz * qz);
    return (int)std::nearbyint((double)(T)(kmax + T(1)));
  }

  double* spec_ = nullptr;  // shell sums of one analysis call, grown on demand
  size_t spec_cap_ = 0;

  // spectralline of every vars field: out[f * nbins + r] = sum |R2C(vars f)|^2 over the modes with
  // round(|k|) = r, out[nf * nbins + r] = their number
  int spectra(double* host_out, int nbins) override {
    if (nr_ != 1) return fail(E_INVALID, "spectra act on single-run problems");
    const int need = spectrum_bins();
    if (nbins < need) return fail(E_INVALID, "spectra need " + std::to_string(need) + " bins");
    const size_t per = (size_t)4 * nbins;  // 3 fields + counts per transformed triple
    const size_t total = per * (size_t)((nf + 2) / 3);
    if (total > spec_cap_) {
      MHD_HIP(hipStreamSynchronize(st));
      if (spec_) (void)hipFree(spec_);
      spec_ = nullptr;
      spec_cap_ = 0;
      MHD_HIP(hipMalloc(&spec_, sizeof(double) * total));
      spec_cap_ = total;
    }
    MHD_HIP(hipMemsetAsync(spec_, 0, sizeof(double) * total, st));
    for (int f0 = 0; f0 < nf; f0 += 3) {
      MHD_TRY(r2c_fields(f0));
      MHD_TRY(timed("shell", 0, [&] { return launch_shell<T>(g, rbh[0].buf[B_WORK], 3, spec_ + per * (f0 / 3), nbins, st); }));
    }
    std::vector<double> h(total);
    MHD_HIP(hipMemcpyAsync(h.data(), spec_, sizeof(double) * total, hipMemcpyDeviceToHost, st));
    MHD_HIP(hipStreamSynchronize(st));
    for (int f = 0; f < nf; ++f)
      std::memcpy(host_out + (size_t)f * nbins, h.data() + per * (f / 3) + (size_t)(f % 3) * nbins, sizeof(double) * nbins);
    std::memcpy(host_out + (size_t)nf * nbins, h.data() + 3 * (size_t)nbins, sizeof(double) * nbins);
    return 0;
  }

  // ScaleDecomposition of vars fields f0 .. f0 + ncomp - 1 (within one transformed triple):
  // keep k1 <= |k| <= k2, back to real space, ncomp real fields to host_out
  int scale_decomposition(int f0, int ncomp, double k1, double k2, void* host_out) override {
    if (nr_ != 1) return fail(E_INVALID, "ScaleDecomposition acts on single-run problems");
    if (ncomp < 1 || ncomp > 3 || f0 < 0 || f0 + ncomp > nf || (f0 / 3) != ((f0 + ncomp - 1) / 3))
      return fail(E_INVALID, "fields must be 1 to 3 components of u or of b");
    const int base = f0 - f0 % 3;
    MHD_TRY(r2c_fields(base));
    MHD_TRY(timed("band", 0, [&] { return launch_band<T>(g, rbh[0].buf[B_WORK], 3, std::min(k1, k2), std::max(k1, k2), st); }));
    MHD_TRY(colpass("zpass_inv", 2, 1, B_WORK, B_WORK, 0, 0, 3, false));
    MHD_TRY(colpass("ypass_inv", 1, 1, B_WORK, B_WORK, 0, 0, 3, false));
    RowArgs ra = row_args();
    ra.in_sel = B_WORK;
    ra.in_f0 = 0;
    ra.in_ncols = g.nkr;
    ra.real_sel = 1;
    ra.real_out_f0 = 0;
    ra.write_vars = 1;
    MHD_TRY(xrow("xrow_c2r", ra, ROW_C2R, 3, 0));
    MHD_HIP(hipMemcpyAsync(host_out, rbh[0].rscr + (int64_t)(f0 - base) * g.nr, sizeof(T) * g.nr * ncomp,
                           hipMemcpyDeviceToHost, st));
    MHD_HIP(hipStreamSynchronize(st));
    return 0;
  }

  // Vector calculus of host real fields on the device (VectorCalculus.jl, MHDAnalysis.jl):
  // upload nin fields, R2C, the spectral op in T, C2R of nout fields, download
  int field_op(int op, const void* const* in, void* const* out) override {
    if (nr_ != 1) return fail(E_INVALID, "field operations act on single-run problems");
    int nin = 0, nout = 0;
    switch (op) {
      case OP_CURL_T: case OP_VECPOT_T: nin = 3; nout = 3; break;
      case OP_DIV_T: nin = 3; nout = 1; break;
      case OP_DX: case OP_DY: case OP_DZ: case OP_INVLAP: nin = 1; nout = 1; break;
      default: return fail(E_INVALID, "unknown field operation");
    }
    if (!in || !out) return fail(E_INVALID, "NULL field arrays");
    for (int c = 0; c < nout; ++c)
      if (!out[c]) return fail(E_INVALID, "NULL output field");
    RunBuf<T>& b = rbh[0];
    const size_t fb = sizeof(T) * (size_t)g.nr;
    for (int c = 0; c < 3; ++c) {
      if (c < nin && in[c]) MHD_HIP(hipMemcpyAsync(b.rscr + (size_t)c * g.nr, in[c], fb, hipMemcpyHostToDevice, st));
      else MHD_HIP(hipMemsetAsync(b.rscr + (size_t)c * g.nr, 0, fb, st));
    }
    RowArgs ra = row_args();
    ra.in_f0 = 0;
    ra.real_sel = 1;
    ra.out2_sel = B_WORK;
    ra.out2_f0 = 0;
    MHD_TRY(xrow("xrow_r2c", ra, ROW_R2C, 3, 0));
    MHD_TRY(colpass("ypass_fwd", 1, -1, B_WORK, B_WORK, 0, 0, 3, false));
    MHD_TRY(colpass("zpass_fwd", 2, -1, B_WORK, B_WORK, 0, 0, 3, false));
    MHD_TRY(diag_field(op, nout == 1 ? 1 : 3));
    for (int c = 0; c < nout; ++c)
      MHD_HIP(hipMemcpyAsync(out[c], b.rscr + (size_t)c * g.nr, fb, hipMemcpyDeviceToHost, st));
    MHD_HIP(hipStreamSynchronize(st));
    return 0;
  }

  // DivVCorrection! (which = 0) / DivBCorrection! (which = 1), src/Solver/VPSolver.jl:60-139:
  // project the field triple of sol on k . F = 0, then the triple's vars = C2R of all its modes
  int div_correction(int which) override {
    if (nr_ != 1) return fail(E_INVALID, "the divergence corrections act on single-run problems");
    if (which != 0 && which != 1) return fail(E_INVALID, "field must be 0 (velocity) or 1 (magnetic)");
    if (which == 1 && phys != PHYS_MHD) return fail(E_INVALID, "DivBCorrection! needs the magnetic field");
    const int f0 = 3 * which;
    cx<T>* sol = (cx<T>*)rbh[0].buf[B_SOL] + (int64_t)f0 * g.nc;
    MHD_TRY(timed("div_correction", 0, [&] { return launch_div_correction<T>(g, sol, st); }));
    MHD_TRY(colpass("zpass_inv", 2, 1, B_SOL, B_WORK, f0, 0, 3, false));
    MHD_TRY(colpass("ypass_inv", 1, 1, B_WORK, B_WORK, 0, 0, 3, false));
    RowArgs ra = row_args();
    ra.in_sel = B_WORK;
    ra.in_f0 = 0;
    ra.in_ncols = g.nkr;
    ra.real_sel = 0;
    ra.real_out_f0 = f0;
    ra.write_vars = 1;
    MHD_TRY(xrow("xrow_c2r", ra, ROW_C2R, 3, 0));
    MHD_TRY(vars_maxima());
    MHD_HIP(hipStreamSynchronize(st));
    return 0;
  }

  // Structure functions of the velocity (src/utils/TurbStatTool.jl): mode 0 SF2 1D
  // (:88-117, bins of round(|R|)), mode 1 SFr_Global (:119-160, bins of |R| parallel and
  // perpendicular to the mean magnetic field).  Output: normalized bin values (sum / count,
  // NaN for empty bins as 0/0), Julia column-major for mode 1, then the counts.
  int sf_shape(int mode, int* nb, int* nbx) const override {
    const double R = std::sqrt((double)(g.nx / 2) * (g.nx / 2) + (double)(g.ny / 2) * (g.ny / 2) +
                               (double)(g.nz / 2) * (g.nz / 2));
    const int Rc = (int)std::ceil(R);
    if (mode == 0) {
      *nb = 2 * Rc;
      *nbx = *nb;
    } else {
      *nbx = Rc + 1;
      *nb = (Rc + 1) * (Rc + 1);
    }
    return 0;
  }

  int structure_function(int mode, double* host_out) override {
    if (nr_ != 1) return fail(E_INVALID, "structure functions act on single-run problems");
    if (mode != 0 && mode != 1) return fail(E_INVALID, "mode must be 0 (SF2 1D) or 1 (SFr_Global)");
    if (mode == 1 && phys != PHYS_MHD) return fail(E_INVALID, "SFr_Global bins against the mean magnetic field");
    int nb = 0, nbx = 0;
    sf_shape(mode, &nb, &nbx);
    const size_t total = 2 * (size_t)nb + 3;
    if (total > spec_cap_) {
      MHD_HIP(hipStreamSynchronize(st));
      if (spec_) (void)hipFree(spec_);
      spec_ = nullptr;
      spec_cap_ = 0;
      MHD_HIP(hipMalloc(&spec_, sizeof(double) * total));
      spec_cap_ = total;
    }
    double* hist = spec_;
    double* means = spec_ + 2 * (size_t)nb;
    double mb[3] = {0, 0, 0};
    if (mode == 1) {  // mean(bx), mean(by), mean(bz) in T
      MHD_TRY(r2c_fields(3));
      MHD_TRY(timed("sf_power", 0, [&] { return launch_sf_power<T>(g, rbh[0].buf[B_WORK], means, st); }));
      MHD_HIP(hipMemcpyAsync(mb, means, sizeof(mb), hipMemcpyDeviceToHost, st));
      MHD_HIP(hipStreamSynchronize(st));
      if (mb[0] == 0.0 && mb[1] == 0.0 && mb[2] == 0.0)
        return fail(E_INVALID, "SFr_Global needs a nonzero mean magnetic field");
    }
    MHD_TRY(r2c_fields(0));
    MHD_TRY(timed("sf_power", 0, [&] { return launch_sf_power<T>(g, rbh[0].buf[B_WORK], means, st); }));
    double mu[3];
    MHD_HIP(hipMemcpyAsync(mu, means, sizeof(mu), hipMemcpyDeviceToHost, st));
    // CF = C2R of the summed power (work field 0; fields 1 and 2 ride along unused)
    MHD_TRY(colpass("zpass_inv", 2, 1, B_WORK, B_WORK, 0, 0, 3, false));
    MHD_TRY(colpass("ypass_inv", 1, 1, B_WORK, B_WORK, 0, 0, 3, false));
    RowArgs ra = row_args();
    ra.in_sel = B_WORK;
    ra.in_f0 = 0;
    ra.in_ncols = g.nkr;
    ra.real_sel = 1;
    ra.real_out_f0 = 0;
    ra.write_vars = 1;
    MHD_TRY(xrow("xrow_c2r", ra, ROW_C2R, 3, 0));
    MHD_HIP(hipStreamSynchronize(st));
    const double summean2 = 2.0 * ((double)(T)mu[0] + (double)(T)mu[1] + (double)(T)mu[2]);
    MHD_HIP(hipMemsetAsync(hist, 0, sizeof(double) * 2 * nb, st));
    MHD_TRY(timed("sf_bin", 0, [&] {
      return launch_sf_bin<T>(g, rbh[0].rscr, summean2, mb, hist, nb, nbx, mode == 1, st);
    }));
    std::vector<double> h(2 * (size_t)nb);
    MHD_HIP(hipMemcpyAsync(h.data(), hist, sizeof(double) * 2 * nb, hipMemcpyDeviceToHost, st));
    MHD_HIP(hipStreamSynchronize(st));
    for (int i = 0; i < nb; ++i) {
      host_out[i] = h[i] / h[nb + i];
      host_out[nb + i] = h[nb + i];
    }
    return 0;
  }

  // ---- forcing -------------------------------------------------------------------
  // Drop every calcF! module of a run (the reference holds one calcF! per problem, so setting
  // a module replaces the previous one).  The captured step is rebuilt on the next step.
  int clear_forcing(int run) {
    MHD_HIP(hipStreamSynchronize(st));
    for (auto* v : {&force_tabs, &fdense_bufs, &ndf_bufs, &ndpart_bufs}) {
      if ((*v)[run]) (void)hipFree((*v)[run]);
      (*v)[run] = nullptr;
    }
    RunBuf<T>& b = rbh[run];
    b.force = nullptr;
    b.force_K = 0;
    b.fdense = nullptr;
    b.ndf = nullptr;
    b.ndpart = nullptr;
    b.nd_P = 0;
    b.a99 = 0;
    nd_any = false;
    for (int r = 0; r < nr_; ++r) nd_any = nd_any || rbh[r].ndf != nullptr;
    drop_graph();
    return 0;
  }
  void drop_graph() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    gexec = nullptr;
    graph = nullptr;
    ++graph_gen_;
  }
  int forcing_done() {
    MHD_TRY(upload_runs());
    MHD_HIP(hipStreamSynchronize(st));
    return 0;
  }
  int device_alloc(size_t bytes, void** p) {
    *p = nullptr;
    if (hipMalloc(p, bytes) != hipSuccess) {
      *p = nullptr;
      return fail(E_HIP, "hipMalloc failed (out of device memory?)");
    }
    return 0;
  }

  // N97ForceDriving! / a static forcing field: R2C of the real fields f3 (NULL = zero) on the
  // device (SetUpN97!, src/pgen/TaylorGreenDynamo.jl:29-32), added to N_u of every MHD call
  int set_forcing_field(int run, const void* const* f3) override {
    MHD_TRY(check_run(run));
    MHD_TRY(clear_forcing(run));
    RunBuf<T>& b = rbh[run];
    for (int c = 0; c < 3; ++c) {
      if (f3 && f3[c]) MHD_HIP(hipMemcpyAsync(b.rscr + (int64_t)c * g.nr, f3[c], sizeof(T) * g.nr, hipMemcpyHostToDevice, st));
      else MHD_HIP(hipMemsetAsync(b.rscr + (int64_t)c * g.nr, 0, sizeof(T) * g.nr, st));
    }
    {
      RunScope scope(this, run);
      RowArgs ra = row_args();
      ra.in_f0 = 0;
      ra.real_sel = 1;
      ra.out2_sel = B_WORK;
      ra.out2_f0 = 0;
      MHD_TRY(xrow("xrow_r2c", ra, ROW_R2C, 3, 0));
      MHD_TRY(colpass("ypass_fwd", 1, -1, B_WORK, B_WORK, 0, 0, 3, false));
      MHD_TRY(colpass("zpass_fwd", 2, -1, B_WORK, B_WORK, 0, 0, 3, false));
    }
    void* d = nullptr;
    MHD_TRY(device_alloc(sizeof(cx<T>) * 3 * (size_t)g.nc, &d));
    fdense_bufs[run] = d;
    for (int c = 0; c < 3; ++c)
      MHD_TRY(timed("unpad", 0, [&] {
        return launch_unpad<T>(g, b.buf[B_WORK] + (int64_t)c * g.wnc, (cx<T>*)d + (int64_t)c * g.nc, st);
      }));
    b.fdense = (const cx<T>*)d;
    return forcing_done();
  }

  // NDForceDriving! with SetUpND!(prob, P, fx, fy, fz) (src/pgen/NegativeDamping.jl:16-22)
  int set_forcing_nd(int run, double P, const void* const* f3) override {
    MHD_TRY(check_run(run));
    if (nr_ != 1) return fail(E_INVALID, "negative damping acts on single-run problems");
    if (rbh[run].vp) return fail(E_UNSUPPORTED, "negative damping with volume penalisation is not supported");
    MHD_TRY(clear_forcing(run));
    RunBuf<T>& b = rbh[run];
    void* d = nullptr;
    MHD_TRY(device_alloc(sizeof(T) * 3 * (size_t)g.nr, &d));
    ndf_bufs[run] = d;
    for (int c = 0; c < 3; ++c) {
      T* dst = (T*)d + (int64_t)c * g.nr;
      if (f3 && f3[c]) MHD_HIP(hipMemcpyAsync(dst, f3[c], sizeof(T) * g.nr, hipMemcpyHostToDevice, st));
      else MHD_HIP(hipMemsetAsync(dst, 0, sizeof(T) * g.nr, st));
    }
    MHD_TRY(device_alloc(sizeof(double) * (size_t)row_blocks, &d));
    ndpart_bufs[run] = d;
    if (phys == PHYS_MHD) MHD_TRY(grow_work(run, 12));  // room for the three f_i u_i products
    b.ndf = (const T*)ndf_bufs[run];
    b.ndpart = (double*)ndpart_bufs[run];
    b.nd_P = P;
    nd_any = true;
    return forcing_done();
  }

  // the work buffer of a run holds at least nfields product fields
  int grow_work(int run, int nfields) {
    if (np_alloc >= nfields) return 0;
    RunBuf<T>& b = rbh[run];
    MHD_HIP(hipStreamSynchronize(st));
    void* w = nullptr;
    MHD_TRY(device_alloc(sizeof(cx<T>) * (size_t)g.wnc * nfields, &w));
    for (auto& p : allocs)
      if (p == (void*)b.buf[B_WORK]) p = w;
    (void)hipFree(b.buf[B_WORK]);
    b.buf[B_WORK] = (cx<T>*)w;
    np_alloc = nfields;
    return 0;
  }

  // Volume penalisation (VP_method, MHDParams_VP/HDParams_VP): chi, U0 (and B0 for MHD) as real
  // fields; a NULL pointer leaves that field as it was (zero when first enabled).  enable = 0
  // turns it off and frees the fields.
  int set_vp(int run, int enable, const void* chi, const void* const* U0, const void* const* B0) override {
    MHD_TRY(check_run(run));
    if (nr_ != 1) return fail(E_INVALID, "volume penalisation acts on single-run problems");
    RunBuf<T>& b = rbh[run];
    MHD_HIP(hipStreamSynchronize(st));
    if (!enable) {
      if (vp_bufs[run]) (void)hipFree(vp_bufs[run]);
      vp_bufs[run] = nullptr;
      b.vp = nullptr;
      vp_any = false;
      drop_graph();
      return forcing_done();
    }
    if (b.ndf) return fail(E_UNSUPPORTED, "volume penalisation with negative damping is not supported");
    const int nfield = 1 + nf;  // chi, then U0 (and B0)
    if (!vp_bufs[run]) {
      void* d = nullptr;
      MHD_TRY(device_alloc(sizeof(T) * nfield * (size_t)g.nr, &d));
      MHD_HIP(hipMemsetAsync(d, 0, sizeof(T) * nfield * (size_t)g.nr, st));
      vp_bufs[run] = d;
    }
    T* base = (T*)vp_bufs[run];
    const size_t fb = sizeof(T) * (size_t)g.nr;
    if (chi) MHD_HIP(hipMemcpyAsync(base, chi, fb, hipMemcpyHostToDevice, st));
    for (int c = 0; c < 3; ++c) {
      if (U0 && U0[c]) MHD_HIP(hipMemcpyAsync(base + (1 + c) * (size_t)g.nr, U0[c], fb, hipMemcpyHostToDevice, st));
      if (phys == PHYS_MHD && B0 && B0[c])
        MHD_HIP(hipMemcpyAsync(base + (4 + c) * (size_t)g.nr, B0[c], fb, hipMemcpyHostToDevice, st));
    }
    MHD_TRY(grow_work(run, np + nf));
    b.vp = base;
    vp_any = true;
    drop_graph();
    return forcing_done();
  }

  // A99ForceDriving! with SetUpFk(prob; kf, P, sigma2) (src/pgen/A99ForceDriving.jl:93-127),
  // usr_vars.b = b and usr_vars.A = A; evaluated in T as the reference's grid arrays are
  int set_forcing_a99(int run, uint64_t seed, double kf, double P, double s2, double bb, double A) override {
    MHD_TRY(check_run(run));
    if (!(s2 > 0)) return fail(E_INVALID, "sigma^2 must be > 0");
    MHD_TRY(clear_forcing(run));
    const T kfT = (T)kf, s2T = (T)s2;
    double integral = 0.0;  // sum(exp(-(k-kf)^2/s2) / (k+1)^2) over the half spectrum
    for (int m = 0; m < g.nz; ++m)
      for (int l = 0; l < g.ny; ++l)
        for (int kx = 0; kx < g.nkr; ++kx) {
          const T qx = (T)((double)kx * g.multx), qy = (T)((double)signed_index(l, g.ny) * g.multy);
          const T qz = (T)((double)signed_index(m, g.nz) * g.multz);
          const T K = (qx * qx + qy * qy) + qz * qz;
          const T k = std::sqrt(K), d = k - kfT;
          const T e = std::exp(-(d * d) / s2T);
          const T k1 = k + T(1);
          integral += (double)(T)(e * (T(1) / (k1 * k1)));
        }
    const double dx = (double)clk_dx(0), dy = (double)clk_dx(1), dz = (double)clk_dx(2);
    const double amp =
        std::sqrt(P * 3 * (lens[0] / dx) * (lens[1] / dy) * (lens[2] / dz) / integral * (1 / dx / dy / dz));
    RunBuf<T>& b = rbh[run];
    b.a99 = 1;
    b.a99_seed = seed;
    b.a99_kf = (double)kfT;
    b.a99_s2 = (double)s2T;
    b.a99_b = (double)(T)bb;
    b.a99_A = (double)(T)A;
    b.a99_amp = (double)(T)amp;
    const T bT = (T)bb;
    b.a99_den = (double)(T)std::tanh((T)((T)(bT * (T)3.14159265358979323846) / (T)2));
    return forcing_done();
  }

  int set_forcing(int run, uint64_t seed, double kf, double amp) override {
    MHD_TRY(check_run(run));
    MHD_TRY(clear_forcing(run));
    if (amp != 0.0 && kf > 0) {
      const int K = (int)std::ceil(kf + 0.5);
      const int64_t per = (int64_t)(K + 1) * (2 * K + 1) * (2 * K + 1);
      std::vector<cx<T>> tab(3 * per, cx<T>{T(0), T(0)});
      const double lo2 = (kf - 0.5) * (kf - 0.5), hi2 = (kf + 0.5) * (kf + 0.5);
      const double Ntot = (double)g.nx * g.ny * g.nz;
      int64_t i = 0;
      for (int r = -K; r <= K; ++r)
        for (int q = -K; q <= K; ++q)
          for (int p = -K; p <= K; ++p) {
            const double s2 = (double)(p * p + q * q + r * r);
            if (!(lo2 <= s2 && s2 < hi2)) continue;
            if (!(p > 0 || (p == 0 && q > 0) || (p == 0 && q == 0 && r > 0))) continue;
            double fr[3], fi[3];
            for (int c = 0; c < 3; ++c) {
              const double ph = 2 * M_PI * uniform01(seed, (uint64_t)(3 * i + c));
              fr[c] = amp * std::cos(ph);
              fi[c] = amp * std::sin(ph);
            }
            const double k[3] = {p * 2 * M_PI / lens[0], q * 2 * M_PI / lens[1], r * 2 * M_PI / lens[2]};
            const double kk = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
            const double dr = (k[0] * fr[0] + k[1] * fr[1] + k[2] * fr[2]) / kk;
            const double di = (k[0] * fi[0] + k[1] * fi[1] + k[2] * fi[2]) / kk;
            for (int c = 0; c < 3; ++c) {
              fr[c] -= k[c] * dr;
              fi[c] -= k[c] * di;
            }
            for (int sg = 1; sg >= -1; sg -= 2) {
              const int kx = sg * p, ky = sg * q, kz = sg * r;
              if (kx < 0) continue;
              const int64_t fidx = ((int64_t)(kz + K) * (2 * K + 1) + (ky + K)) * (K + 1) + kx;
              for (int c = 0; c < 3; ++c)
                tab[c * per + fidx] = cx<T>{(T)(Ntot * fr[c]), (T)(Ntot * (sg > 0 ? fi[c] : -fi[c]))};
            }
            ++i;
          }
      void* d = nullptr;
      MHD_HIP(hipMalloc(&d, sizeof(cx<T>) * tab.size()));
      MHD_HIP(hipMemcpy(d, tab.data(), sizeof(cx<T>) * tab.size(), hipMemcpyHostToDevice));
      force_tabs[run] = d;
      rbh[run].force = (const cx<T>*)d;
      rbh[run].force_K = K;
    }
    MHD_TRY(upload_runs());
    MHD_HIP(hipStreamSynchronize(st));
    // the captured step reads the descriptor array, which stays at the same address
    return 0;
  }

  // ---- measurement -------------------------------------------------------------
  int time_steps(int nsteps, double* ms) override {
    struct Pair {  // released on every return path
      hipEvent_t a = nullptr, b = nullptr;
      ~Pair() {
        if (a) hipEventDestroy(a);
        if (b) hipEventDestroy(b);
      }
    } ev;
    MHD_HIP(hipEventCreate(&ev.a));
    MHD_HIP(hipEventCreate(&ev.b));
    if (use_graph) MHD_TRY(ensure_graph());
    MHD_HIP(hipStreamSynchronize(st));
    MHD_HIP(hipEventRecord(ev.a, st));
    MHD_TRY(step(nsteps));
    MHD_HIP(hipEventRecord(ev.b, st));
    MHD_HIP(hipEventSynchronize(ev.b));
    float f = 0;
    MHD_HIP(hipEventElapsedTime(&f, ev.a, ev.b));
    *ms = f;
    return 0;
  }

  int profile_steps(int nsteps, std::vector<KTime>* out) override {
    MHD_HIP(hipStreamSynchronize(st));
    if (nsteps > 0) {
      const hipError_t e = launch_small<T>(SK_PRUNE, rbd, nr_, phys, nsteps, prune_enabled ? 1 : 0, st);
      if (e != hipSuccess) return fail(E_HIP, std::string("launch of prune_mark failed: ") + hipGetErrorString(e));
    }
    profiling = true;
    evs.clear();
    int rc = 0;
    for (int i = 0; i < nsteps && rc == 0; ++i) {
      host_pruned_ = prune_enabled && i + 1 < nsteps;  // byte models follow the device's choice
      rc = step_launches();
    }
    host_pruned_ = false;
    profiling = false;
    if (rc) {
      for (auto& e : evs) {
        hipEventDestroy(e.a);
        hipEventDestroy(e.b);
      }
      evs.clear();
      return rc;
    }
    MHD_HIP(hipStreamSynchronize(st));
    std::map<std::string, KTime> agg;
    std::vector<std::string> order;
    for (auto& e : evs) {
      float ms = 0;
      MHD_HIP(hipEventElapsedTime(&ms, e.a, e.b));
      if (!agg.count(e.name)) order.push_back(e.name);
      KTime& k = agg[e.name];
      k.name = e.name;
      k.launches += 1;
      k.ms += ms;
      k.bytes += e.bytes;
      hipEventDestroy(e.a);
      hipEventDestroy(e.b);
    }
    evs.clear();
    out->clear();
    for (auto& n : order) out->push_back(agg[n]);
    return 0;
  }
};

// ---------------------------------------------------------------------------
// Multi-size stepping of several groups of an ensemble (same precision, physics and
// stepper): each group's step is recorded once, and the recordings are replayed position by
// position -- one multi-size launch per hot launch (column passes, row kernel, spectral
// update) for up to MULTI_MAX groups (row launches: groups of one row plan), the small
// launches (CFL, maxima) per group -- on the first group's stream, captured as one HIP
// graph per step.
// ---------------------------------------------------------------------------
template <class T>
class FusedStep : public FusedBase {
 public:
  using Rec = typename Engine<T>::Rec;
  std::vector<Engine<T>*> es;
  // each group's step recorded in the multi-size shapes (recm) and in its own shapes (recs)
  std::vector<std::vector<Rec>> recm, recs;
  std::vector<uint64_t> gens;
  std::vector<hipEvent_t> evs;  // per group: its stream's pending work, then (index 0) our steps
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  bool use_graph = true;
  // per kernel class: one multi-size launch for all groups, or each group's own launch
  // (set by tune(); the spectral update is always multi-size)
  bool multi_col = true, multi_row = true;

  ~FusedStep() override {
    if (!es.empty() && es[0]->st) (void)hipStreamSynchronize(es[0]->st);
    drop();
    for (hipEvent_t e : evs)
      if (e) (void)hipEventDestroy(e);
  }
  hipStream_t stream() const { return es[0]->st; }

  void drop() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    gexec = nullptr;
    graph = nullptr;
  }

  int init(const std::vector<EngineBase*>& groups) {
    for (EngineBase* b : groups) es.push_back(static_cast<Engine<T>*>(b));
    evs.assign(es.size(), nullptr);
    for (auto& e : evs) MHD_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return record();
  }

  int record() {
    drop();
    recm.assign(es.size(), {});
    recs.assign(es.size(), {});
    gens.clear();
    for (size_t k = 0; k < es.size(); ++k) {
      MHD_TRY(es[k]->record_step(&recm[k], true));
      MHD_TRY(es[k]->record_step(&recs[k], false));
      gens.push_back(es[k]->graph_gen());
    }
    // every group runs the same launch sequence (same physics and stepper, no modules)
    for (size_t k = 0; k < es.size(); ++k) {
      if (recm[k].size() != recm[0].size() || recs[k].size() != recm[0].size())
        return fail(E_INVALID, "multi-size step: launch sequences differ");
      for (size_t i = 0; i < recm[0].size(); ++i) {
        const Rec &a = recm[0][i], &b = recm[k][i], &c = recs[k][i];
        if (a.type != b.type || a.dir != b.dir || a.kind != b.kind || c.type != a.type)
          return fail(E_INVALID, "multi-size step: launch sequences differ");
        if (b.type == Rec::SPEC && (b.sa.planes != 2 || b.sa.modules != 0))
          return fail(E_INVALID, "multi-size step: spectral launch has no multi-size form");
      }
    }
    return 0;
  }

  bool stale() const {
    for (size_t k = 0; k < es.size(); ++k)
      if (es[k]->graph_gen() != gens[k]) return true;
    return false;
  }

  // position i of the step: one multi-size launch per chunk of groups, or every group's own
  int launch_at(size_t i, hipStream_t s) {
    const int type = recm[0][i].type;
    const int ng_all = (int)es.size();
    const bool multi = type == Rec::SPEC || (type == Rec::COL && multi_col) || (type == Rec::ROW && multi_row);
    if (!multi) {
      // the small launches (CFL, maxima) follow the row launches' shapes (block counts)
      const auto& R = (type == Rec::OTHER && multi_row) ? recm : recs;
      for (int k = 0; k < ng_all; ++k) {
        const hipError_t e = es[k]->launch_rec(R[k][i], s);
        if (e != hipSuccess) return fail(E_HIP, std::string("multi-size step launch: ") + hipGetErrorString(e));
      }
      return 0;
    }
    // groups that can share a launch (row launches: the same row plan), in chunks of MULTI_MAX
    for (int pv = 0; pv < 2; ++pv) {
      std::vector<int> sel;
      for (int k = 0; k < ng_all; ++k)
        if (type != Rec::ROW || (recm[k][i].var & 1) == pv) sel.push_back(k);
      for (size_t c0 = 0; c0 < sel.size(); c0 += MULTI_MAX) {
        const int ng = (int)std::min<size_t>(MULTI_MAX, sel.size() - c0);
        hipError_t e = hipSuccess;
        if (type == Rec::COL) {
          MultiCol<T> m{};
          m.ngroups = ng;
          for (int j = 0; j < ng; ++j) {
            const Rec& r = recm[sel[c0 + j]][i];
            m.n[j] = r.n;
            m.a[j] = r.ca;
            m.runs[j] = r.runs;
            m.tw[j] = r.tw;
            m.start[j + 1] = m.start[j] + r.ca.nitems;
          }
          e = launch_colpass_multi<T>(m, recm[sel[c0]][i].dir, s);
        } else if (type == Rec::ROW) {
          MultiRow<T> m{};
          m.ngroups = ng;
          for (int j = 0; j < ng; ++j) {
            const int k = sel[c0 + j];
            const Rec& r = recm[k][i];
            m.nx[j] = r.n;
            m.a[j] = r.ra;
            m.runs[j] = r.runs;
            m.twM[j] = es[k]->twM;
            m.twh[j] = es[k]->twh;
            m.start[j + 1] = m.start[j] + r.ra.blocks_per_run * r.ra.nruns;
          }
          e = launch_xrow_multi<T>(m, recm[sel[c0]][i].kind, pv, s);
        } else {
          MultiSpec<T> m{};
          m.ngroups = ng;
          for (int j = 0; j < ng; ++j) {
            const Rec& r = recm[sel[c0 + j]][i];
            const Geom& G = r.sa.g;
            m.bpx[j] = (int)(((int64_t)G.ny * G.nkr + 255) / 256);
            m.a[j] = r.sa;
            m.runs[j] = r.runs;
            m.start[j + 1] = m.start[j] + (int64_t)m.bpx[j] * G.nz * r.sa.nruns;
          }
          e = launch_spectral_multi<T>(m, es[sel[c0]]->phys, s);
        }
        if (e != hipSuccess) return fail(E_HIP, std::string("multi-size launch: ") + hipGetErrorString(e));
      }
      if (type != Rec::ROW) break;  // one pass over all groups
    }
    return 0;
  }

  int launch_step(hipStream_t s) {
    for (size_t i = 0; i < recm[0].size(); ++i) MHD_TRY(launch_at(i, s));
    return 0;
  }

  int ensure_graph() {
    if (stale()) MHD_TRY(record());
    if (gexec) return 0;
    hipStream_t s = stream();
    MHD_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int rc = launch_step(s);
    hipGraph_t gr = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &gr);
    if (rc) {
      if (gr) hipGraphDestroy(gr);
      return rc;
    }
    if (e != hipSuccess) return fail(E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    graph = gr;
    MHD_HIP(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
    return 0;
  }

  // order the fused stream after the other groups' queued work
  int join_in(hipStream_t s) {
    for (size_t k = 1; k < es.size(); ++k) {
      MHD_HIP(hipEventRecord(evs[k], es[k]->st));
      MHD_HIP(hipStreamWaitEvent(s, evs[k], 0));
    }
    return 0;
  }
  // later work queued on the other groups' streams waits for the fused stream's work
  int join_out(hipStream_t s) {
    MHD_HIP(hipEventRecord(evs[0], s));
    for (size_t k = 1; k < es.size(); ++k) MHD_HIP(hipStreamWaitEvent(es[k]->st, evs[0], 0));
    return 0;
  }

  int step(int nsteps) override {
    if (nsteps < 0) return fail(E_INVALID, "nsteps must be >= 0");
    hipStream_t s = stream();
    MHD_TRY(join_in(s));
    for (Engine<T>* e : es) MHD_TRY(e->mark_prune(nsteps, s));
    if (use_graph) {
      MHD_TRY(ensure_graph());
      for (int i = 0; i < nsteps; ++i) MHD_HIP(hipGraphLaunch(gexec, s));
    } else {
      if (stale()) MHD_TRY(record());
      for (int i = 0; i < nsteps; ++i) MHD_TRY(launch_step(s));
    }
    return join_out(s);
  }

  int set_graph(int enable) override {
    use_graph = enable != 0;
    return 0;
  }

  // Time the column passes and the row launch of the first right-hand side of a step (they
  // write only the work buffers; the first stage stores no real fields or maxima) as
  // multi-size launches and as each group's own launches, with every run's prune marks set
  // as in the dealiased steps of a call, and keep the faster form per class.
  int tune(int reps, int* choice) override {
    if (stale()) MHD_TRY(record());
    hipStream_t s = stream();
    MHD_TRY(join_in(s));
    for (Engine<T>* e : es) {
      const hipError_t r = launch_small<T>(SK_PRUNE, e->rbd, e->nr_, e->phys, 1 << 20, 1, s);
      if (r != hipSuccess) return fail(E_HIP, std::string("prune marks: ") + hipGetErrorString(r));
    }
    size_t first = 0, end = 0;  // [first, end): the first right-hand side's column and row launches
    while (first < recm[0].size() && recm[0][first].type == Rec::OTHER) ++first;
    end = first;
    while (end < recm[0].size() && recm[0][end].type != Rec::SPEC) ++end;
    struct Ev {
      hipEvent_t a = nullptr, b = nullptr;
      ~Ev() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
      }
    } ev;
    MHD_HIP(hipEventCreate(&ev.a));
    MHD_HIP(hipEventCreate(&ev.b));
    const bool keep_col = multi_col, keep_row = multi_row;
    auto time_class = [&](int type, bool multi, float* ms) -> int {
      (type == Rec::COL ? multi_col : multi_row) = multi;
      float sum = 0;
      for (int k = 0; k <= reps; ++k) {
        MHD_HIP(hipEventRecord(ev.a, s));
        for (size_t i = first; i < end; ++i)
          if (recm[0][i].type == type) MHD_TRY(launch_at(i, s));
        MHD_HIP(hipEventRecord(ev.b, s));
        MHD_HIP(hipEventSynchronize(ev.b));
        float e = 0;
        MHD_HIP(hipEventElapsedTime(&e, ev.a, ev.b));
        if (k > 0) sum += e;  // the first round warms up
      }
      *ms = sum;
      return 0;
    };
    float cm = 0, cs = 0, rm = 0, rs = 0;
    int rc = time_class(Rec::COL, true, &cm);
    if (!rc) rc = time_class(Rec::COL, false, &cs);
    if (!rc) rc = time_class(Rec::ROW, true, &rm);
    if (!rc) rc = time_class(Rec::ROW, false, &rs);
    for (Engine<T>* e : es) e->reset_prune(s);
    if (rc) {
      multi_col = keep_col;
      multi_row = keep_row;
      return rc;
    }
    multi_col = cm <= cs;
    multi_row = rm <= rs;
    drop();
    if (choice) {
      choice[0] = multi_col ? 1 : 0;
      choice[1] = multi_row ? 1 : 0;
    }
    return join_out(s);
  }

  int set_choice(int col, int row) override {
    if (col >= 0) multi_col = col != 0;
    if (row >= 0) multi_row = row != 0;
    drop();
    return 0;
  }
  void get_choice(int* c2) const override {
    c2[0] = multi_col ? 1 : 0;
    c2[1] = multi_row ? 1 : 0;
  }
};

FusedBase* make_fused(const std::vector<EngineBase*>& groups, int* err) {
  *err = 0;
  if (groups.empty()) return nullptr;
  FusedBase* f = nullptr;
  if (groups[0]->precision_bytes() == 4) {
    auto* x = new FusedStep<float>();
    *err = x->init(groups);
    f = x;
  } else {
    auto* x = new FusedStep<double>();
    *err = x->init(groups);
    f = x;
  }
  if (*err) {
    delete f;
    return nullptr;
  }
  return f;
}

EngineBase* make_engine(int prec, const Geom& g, const double* L, int phys, int stepper,
                        const std::vector<RunParams>& runs, int device, hipStream_t stream, int* err) {
  if (prec == 0) {
    auto* e = new Engine<float>();
    *err = e->init(g, L, phys, stepper, runs, device, stream);
    if (*err) {
      delete e;
      return nullptr;
    }
    return e;
  }
  auto* e = new Engine<double>();
  *err = e->init(g, L, phys, stepper, runs, device, stream);
  if (*err) {
    delete e;
    return nullptr;
  }
  return e;
}

}  // namespace mhd
