// Multi-size kernels: one launch covers the runs of several size groups of an ensemble.
//
// An ensemble of runs of uneven grid size is stepped as groups of same-size runs.  For the
// small grids (lengths in kMultiLengths) each hot launch of a time step -- the four column
// passes, the fused x-row kernel and the spectral update -- covers every group at once: the
// work items of all groups form one index range (prefix sums per group), and each workgroup
// takes one item and runs the transform of its group's length.  Workgroups are 128 threads
// for every length (column passes: S = 128 / (N/E) columns per strip), so one launch shape
// serves all groups; the per-item code is the same device function the single-size kernels
// run, so every run's result is bitwise the one of its group's own launches.  LDS is sized
// per launch for the longest transform in it, and the column kernel is compiled in two
// length classes (up to 96, up to 128) so that shorter sets do not carry the registers of
// the 128-point body (84 against 59 VGPRs).  (A grid-stride loop over the items, tried as a
// persistent form, raised the column kernel from 84 to 166 VGPRs.)
#pragma once
#include "mhd_kernels.hpp"

namespace mhd {

constexpr int MULTI_MAX = 4;        // size groups per multi-size launch
constexpr int MULTI_THREADS = 128;  // threads per workgroup of every multi-size kernel
constexpr int kMultiLengths[] = {32, 48, 64, 96, 128};
inline bool multi_length(int n) {
  for (int v : kMultiLengths)
    if (v == n) return true;
  return false;
}

// strip width (columns per work item) of a column pass of length N in a 128-thread workgroup
template <int N> struct MultiColShape {
  static constexpr int E = Plan<N>::E;
  static constexpr int TH = N / E;
  static constexpr int S = MULTI_THREADS / TH;
  static_assert(S * TH == MULTI_THREADS, "length has no 128-thread column shape");
};
// dynamic LDS of a multi-size launch, reinterpreted per kernel
extern __shared__ __align__(16) unsigned char multi_lds[];

inline int multi_col_strip(int n) {
  switch (n) {
    case 32: return MultiColShape<32>::S;
    case 48: return MultiColShape<48>::S;
    case 64: return MultiColShape<64>::S;
    case 96: return MultiColShape<96>::S;
    case 128: return MultiColShape<128>::S;
    default: return 0;
  }
}

// Kernel arguments (by value): per group the single-size launch's arguments, its run
// descriptors and twiddles, and the first index of its items in the combined range.
template <class T>
struct MultiCol {
  int ngroups;
  int n[MULTI_MAX];              // transform length along the pass axis
  int64_t start[MULTI_MAX + 1];  // first work item of each group; start[ngroups] = total
  ColArgs a[MULTI_MAX];
  const RunBuf<T>* runs[MULTI_MAX];
  const cx<T>* tw[MULTI_MAX];
};

template <class T>
struct MultiRow {
  int ngroups;
  int nx[MULTI_MAX];
  int64_t start[MULTI_MAX + 1];
  RowArgs a[MULTI_MAX];
  const RunBuf<T>* runs[MULTI_MAX];
  const cx<T>* twM[MULTI_MAX];
  const cx<T>* twh[MULTI_MAX];
};

template <class T>
struct MultiSpec {
  int ngroups;
  int bpx[MULTI_MAX];            // 256-thread blocks per m slab (ny * nkr modes)
  int64_t start[MULTI_MAX + 1];  // first block of each group
  SpecArgs a[MULTI_MAX];
  const RunBuf<T>* runs[MULTI_MAX];
};

__device__ __forceinline__ int multi_group(const int64_t* start, int ngroups, int64_t it) {
  int g = 0;
  while (g + 1 < ngroups && it >= start[g + 1]) ++g;
  return g;
}

// LMAX: the longest transform compiled in (96 or 128).  Dynamic LDS: N*S exchange slots and
// N twiddles for the longest N of the launch.
template <class T, int DIR, int LMAX>
__global__ void __launch_bounds__(MULTI_THREADS) colpass_multi_kernel(MultiCol<T> m) {
  cx<T>* const lds = reinterpret_cast<cx<T>*>(multi_lds);
  const int64_t it = blockIdx.x;
  if (it >= m.start[m.ngroups]) return;
  {
    const int g = multi_group(m.start, m.ngroups, it);
    const int64_t item = it - m.start[g];
    const ColArgs& a = m.a[g];
    switch (m.n[g]) {
#define MHD_MULTI_COL(N_)                                                                            \
  case N_:                                                                                           \
    if constexpr (N_ <= LMAX)                                                                        \
      colpass_item<T, N_, MultiColShape<N_>::S, DIR>(a, item, m.runs[g], m.tw[g], lds,               \
                                                     lds + N_ * MultiColShape<N_>::S);               \
    break;
      MHD_MULTI_COL(32)
      MHD_MULTI_COL(48)
      MHD_MULTI_COL(64)
      MHD_MULTI_COL(96)
      MHD_MULTI_COL(128)
#undef MHD_MULTI_COL
      default: break;
    }
  }
}

// LDS bytes of a multi-size launch: column passes N*S + N complex values, rows the
// RowShape exchange slots (plain index: M+1 per row) of the longest length in it
template <class T> inline size_t multi_col_lds(int n) {
  return (size_t)(n * multi_col_strip(n) + n) * sizeof(cx<T>);
}
template <class T> inline size_t multi_row_lds(int nx, int pv) {
  const int m = nx / 2;
  const int tx = m / row_plan_E(m, sizeof(T) == 8, pv);
  return (size_t)(128 / tx) * (m + 1) * sizeof(cx<T>);
}

// PV: the row plan (bit 0 of the row variant) of every group of the launch, so each run
// computes with the plan its group's own launches use; the LDS index is the plain one
// (the swizzled index, bit 1, changes no value).
template <class T, int KIND, int NF, int PV>
__global__ void __launch_bounds__(MULTI_THREADS) xrow_multi_kernel(MultiRow<T> m) {
  cx<T>* const xb = reinterpret_cast<cx<T>*>(multi_lds);  // dynamic: exchange slots of the longest row
  __shared__ T red[2 * 128];
  __shared__ cx<T> stw[2 * 64 + 1];  // 2M+1 twiddles, M <= 64
  const int64_t it = blockIdx.x;
  if (it >= m.start[m.ngroups]) return;
  {
    const int g = multi_group(m.start, m.ngroups, it);
    const int64_t item = it - m.start[g];
    const RowArgs& a = m.a[g];
    switch (m.nx[g]) {
#define MHD_MULTI_ROW(N_)                                                                        \
  case N_:                                                                                       \
    xrow_item<T, N_, KIND, NF, PV>(a, item, m.runs[g], m.twM[g], m.twh[g], xb, red, stw);        \
    break;
      MHD_MULTI_ROW(32)
      MHD_MULTI_ROW(48)
      MHD_MULTI_ROW(64)
      MHD_MULTI_ROW(96)
      MHD_MULTI_ROW(128)
#undef MHD_MULTI_ROW
      default: break;
    }
  }
}

// Spectral update of a time step (masked stage input, every mode in one pass, no modules).
template <class T, int PHYS>
__global__ void __launch_bounds__(256) spectral_multi_kernel(MultiSpec<T> m) {
  const int64_t b = blockIdx.x;
  if (b >= m.start[m.ngroups]) return;
  {
    const int g = multi_group(m.start, m.ngroups, b);
    const int64_t bx = b - m.start[g];
    const int bpx = m.bpx[g];
    const int64_t per_run = (int64_t)bpx * m.a[g].g.nz;
    const int run = (int)(bx / per_run);
    const int r = (int)(bx - (int64_t)run * per_run);
    const int mz = r / bpx;
    const uint32_t i = (uint32_t)(r - mz * bpx) * 256u + threadIdx.x;
    spectral_item<T, PHYS, 2, 0>(m.a[g], m.runs[g], i, mz, run);
  }
}

// Launchers (k_multi.inc): one workgroup per item, dynamic LDS for the longest length of the
// launch.  Items are taken in plain order: consecutive workgroups go to the 8 XCDs in turn,
// so every XCD gets its share of each group's items (the XCD-aware order of the single-size
// kernels gives each XCD a contiguous range, i.e. whole groups of unequal cost per item:
// 8200 against 9180 run-steps/s on a 64-run 32^3..96^3 ensemble).
template <class T>
hipError_t launch_colpass_multi(const MultiCol<T>& m, int dir, hipStream_t s);
template <class T>
hipError_t launch_xrow_multi(const MultiRow<T>& m, int kind, int plan_var, hipStream_t s);
template <class T>
hipError_t launch_spectral_multi(const MultiSpec<T>& m, int phys, hipStream_t s);

}  // namespace mhd
