// Shared definitions of the MI355X MHD library: grid geometry, per-run buffer
// descriptors and the FourierFlows 1/3 box-mask arithmetic.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>

#include "mhd_fft.hpp"

namespace mhd {

enum Phys { PHYS_HD = 0, PHYS_MHD = 1 };
enum Stepper { STEP_RK4 = 0, STEP_LSRK54 = 1 };

// Buffer selectors inside a RunBuf.
enum BufSel { B_SOL = 0, B_STIN = 1, B_ACC = 2, B_WORK = 3, B_EXT_SOL = 4, B_EXT_N = 5, B_SCR = 6, NBUF = 7 };

// Geometry of one grid size.  Arrays are [f][z][y][kx] (complex, nkr = nx/2+1)
// and [f][z][y][x] (real), i.e. Julia's column-major (kr,l,m,f) / (x,y,z).
struct Geom {
  int nx, ny, nz, nkr;
  // kept index sets of the 1/3 box mask: x keeps kx < kx_keep; y keeps
  // l < ly_lo_end or l >= ly_hi_begin; z likewise.
  int kx_keep, ly_lo_end, ly_hi_begin, mz_lo_end, mz_hi_begin;
  int64_t nc;      // complex elements per field  = nz*ny*nkr
  int64_t nr;      // real elements per field     = nz*ny*nx
  double multx, multy, multz;  // wavenumber multipliers ((2pi/L)*n)/n
  double inv_n;    // 1/(nx*ny*nz) in Float64 (rounded to T by the kernels)
  uint32_t fd_nkr_m, fd_nkr_l;  // division by nkr: q = (umulhi(x, m) + x) >> l for x < 2^31
  uint32_t fd_nin_m, fd_nin_l;  // division by nkr - 2 (interior kx count)
  // The work buffer's z-planes are padded: planes a multiple of 8 KB apart put the nz
  // elements of a z column on few DRAM channels, and the z passes, which read and write
  // such columns, then run well below the copy rate.
  // The work buffer's rows start on WORK_ROW_ALIGN-byte boundaries: with nkr-long rows
  // (e.g. 129 complex) the strips of columns that the y and z passes move straddle an extra
  // cache line in most rows, and the column access patterns run 30-50 % slower.
  int wrs;         // work buffer row stride (complex elements) >= nkr
  // Row stores of the x-row kernel and column-pass stores into the work buffer cover whole
  // 128-byte lines: writing part of a line is slow (a probe: row stores of 86 of 144 slots ran
  // at 3.2 TB/s, of 96 at 5.2 TB/s).  A dealiased store of kx < kx_keep extends to kx_keep_w
  // (kx_keep rounded up to a line, values no reader uses), a full store of kx < nkr to wrs
  // (zeros past nkr).
  int kx_keep_w;
  int64_t wps;     // work buffer plane stride (complex elements) = ny*wrs + WORK_PLANE_PAD
  int64_t wnc;     // work buffer field stride = nz*wps
};

constexpr int WORK_PLANE_PAD = 32;     // complex elements added to every z-plane of the work buffer
// Measured on MI355X at 256^3 Float32: 129-complex rows 288.6 steps/s, 64 B aligned rows
// 306.0, 128 B aligned rows 318.6 (column passes 12-21 % faster).
constexpr int WORK_ROW_ALIGN = 128;  // bytes

// Row, plane and field strides of a complex buffer (only the work buffer is padded) and the
// offset of row (z, y) = row index z*ny + y in it.
__host__ __device__ inline int64_t row_stride(const Geom& g, int sel) { return sel == B_WORK ? g.wrs : g.nkr; }
__host__ __device__ inline int64_t plane_stride(const Geom& g, int sel) {
  return sel == B_WORK ? g.wps : (int64_t)g.ny * g.nkr;
}
__host__ __device__ inline int64_t field_stride(const Geom& g, int sel) { return sel == B_WORK ? g.wnc : g.nc; }
__host__ __device__ inline int64_t row_offset(const Geom& g, int sel, int64_t row) {
  if (sel != B_WORK) return row * g.nkr;
  const int r = (int)row;  // rows of one run: ny*nz < 2^31
  const int z = r / g.ny;
  return (int64_t)z * g.wps + (int64_t)(r - z * g.ny) * g.wrs;
}

// Division by an invariant integer d via multiply and shift (l = ceil(log2 d),
// m = ceil(2^(32+l)/d) - 2^32); exact for 0 <= x < 2^31.
inline void fastdiv_make(uint32_t d, uint32_t* m, uint32_t* l) {
  uint32_t s = 0;
  while ((1ull << s) < d) ++s;
  *l = s;
  *m = (uint32_t)((((1ull << (32 + s)) + d - 1) / d) - (1ull << 32));
}
__device__ __forceinline__ uint32_t fastdiv(uint32_t x, uint32_t m, uint32_t l) {
  return (uint32_t)(((uint64_t)__umulhi(x, m) + x) >> l);
}

// FourierFlows getaliasedwavenumbers with aliased fraction 1/3 (as called at
// src/utils/utils.jl:82-84): L=(1-a)/2, R=(1+a)/2, iL=floor(L n)+1, iR=ceil(R n)
// (1-based), aliased = iL:iR (full axes) and iL:nkr (half axis).  Returns the
// 0-based [lo, hi) aliased block.
inline void aliased_block(int n, int* lo, int* hi) {
  const double a = 1.0 / 3.0;
  const double L = (1.0 - a) / 2.0, R = (1.0 + a) / 2.0;
  const int iL = (int)std::floor(L * n) + 1;
  const int iR = (int)std::ceil(R * n);
  *lo = iL - 1;
  *hi = iR;
}

// cx_bytes: bytes of one complex value of the run's precision (8 or 16)
inline Geom make_geom(int nx, int ny, int nz, double Lx, double Ly, double Lz, int cx_bytes) {
  Geom g;
  g.nx = nx; g.ny = ny; g.nz = nz; g.nkr = nx / 2 + 1;
  int lo, hi;
  aliased_block(nx, &lo, &hi);
  g.kx_keep = lo;
  aliased_block(ny, &lo, &hi);
  g.ly_lo_end = lo; g.ly_hi_begin = hi;
  aliased_block(nz, &lo, &hi);
  g.mz_lo_end = lo; g.mz_hi_begin = hi;
  g.nc = (int64_t)nz * ny * g.nkr;
  const int ra = WORK_ROW_ALIGN / cx_bytes > 1 ? WORK_ROW_ALIGN / cx_bytes : 1;
  g.wrs = (g.nkr + ra - 1) / ra * ra;
  g.kx_keep_w = (g.kx_keep + ra - 1) / ra * ra;
  if (g.kx_keep_w > g.wrs) g.kx_keep_w = g.wrs;
  g.wps = (int64_t)ny * g.wrs + WORK_PLANE_PAD;
  g.wnc = (int64_t)nz * g.wps;
  g.nr = (int64_t)nz * ny * nx;
  const double twopi = 6.283185307179586;  // Julia's 2π as Float64
  g.multx = ((twopi / Lx) * nx) / nx;
  g.multy = ((twopi / Ly) * ny) / ny;
  g.multz = ((twopi / Lz) * nz) / nz;
  g.inv_n = 1.0 / ((double)nx * (double)ny * (double)nz);
  fastdiv_make((uint32_t)g.nkr, &g.fd_nkr_m, &g.fd_nkr_l);
  fastdiv_make((uint32_t)(g.nkr - 2), &g.fd_nin_m, &g.fd_nin_l);
  return g;
}

__host__ __device__ inline bool kept_y(const Geom& g, int l) { return l < g.ly_lo_end || l >= g.ly_hi_begin; }
__host__ __device__ inline bool kept_z(const Geom& g, int m) { return m < g.mz_lo_end || m >= g.mz_hi_begin; }
__host__ __device__ inline int nkept_y(const Geom& g) { return g.ly_lo_end + g.ny - g.ly_hi_begin; }
__host__ __device__ inline int nkept_z(const Geom& g) { return g.mz_lo_end + g.nz - g.mz_hi_begin; }
__host__ __device__ inline int signed_index(int i, int n) { return i < (n + 1) / 2 ? i : i - n; }

// Device-side clock of one run.  dt is chosen on the device (CFL) so a whole step
// can be replayed from a captured graph without a host round trip.
struct RunClock {
  double t;          // clock.t (holds a value representable in T)
  double dt;         // clock.dt (representable in T)
  int64_t step;      // clock.step
  double maxu2;      // max_a max_x u_a^2 of the current vars (T value)
  double maxb2;      // max_a max_x b_a^2 of the current vars (T value)
  double cfl_coef;   // CFL coefficient
  double t_diff;     // diffusive bound (Float64)
  double dl;         // min(dx,dy,dz) in T
  int32_t use_cfl;   // 1: dt from getCFL! before each step
  int32_t nan_flag;  // set when the maxima are NaN
  int64_t target;    // step launches skip this run once step >= target (pool slots)
  int64_t prune_until;  // steps ending before this step count skip the aliased modes (see step_pruned)
  double nd_amp;     // negative damping: A = P / (sum |u_i^2 f_i| dV) of the current right-hand side
};

// Step launches of a run whose clock reached its target do nothing (finished or
// empty slots of a serving pool); the step counter only moves at the end of a step.
template <class RB>
__device__ __forceinline__ bool gated_off(int gate, const RB& rb) {
  return gate && rb.clk->step >= rb.clk->target;
}

// A stage input is dealiased before use, so the right-hand side at aliased modes never
// reaches a kept mode: it only sets the aliased modes of sol at the end of the step, and the
// next step discards them.  Steps that are not the last of the current call (the step count
// reaching prune_until) therefore store and update kept modes only; the last step is full.
template <class RB>
__device__ __forceinline__ bool step_pruned(const RB& rb) {
  return rb.clk->step + 1 < rb.clk->prune_until;
}

// Per-run descriptor read by every kernel (one per run of a batched launch).
template <class T>
struct RunBuf {
  cx<T>* buf[NBUF];   // complex buffers by BufSel (nf or np fields each)
  T* vars;            // nf real fields
  T* rscr;            // real scratch (>= 3 real fields), aliases the stage-input buffer
  T* partial;         // per-workgroup maxima (2 per row block), CFL input
  cx<T>* plane;       // kx = 0 plane of the masked stage input, nf x nz x ny (zero off the kept set),
                      // copied by the inverse z-pass for the spectral kernel's Hermitian projection
  RunClock* clk;
  const cx<T>* force; // forcing table (3 comps x (K+1) x (2K+1)^2) or null
  int force_K;
  int nnu, neta;
  double nu, eta;
  // calcF! modules of the reference (at most one is set):
  const cx<T>* fdense;  // static forcing spectrum added to N_u (3 fields of nc; N97, user fields) or null
  const T* ndf;         // negative damping: the real fields f_x, f_y, f_z (nr each) or null
  double* ndpart;       // negative damping: per row-block FP64 sums of |u_i^2 f_i| (row kernel output)
  double nd_P;          // negative damping: injected power P
  int a99;              // A99 stochastic forcing on
  uint64_t a99_seed;
  double a99_kf, a99_s2, a99_b, a99_A, a99_amp;  // kf, sigma^2, b, usr A, setup amplitude (all rounded to T)
  double a99_den;       // tanh(b pi / 2) evaluated in T
  // volume penalisation (MHDParams_VP / HDParams_VP): real fields chi, U0x, U0y, U0z and, for
  // MHD, B0x, B0y, B0z (nr each), or null
  const T* vp;
};

}  // namespace mhd
