// Host-side launchers: map runtime grid lengths onto the compiled kernel
// instances.  Implemented per precision in k_launch_f32.hip / k_launch_f64.hip.
#pragma once
#include <algorithm>

#include "mhd_kernels.hpp"

namespace mhd {

// lengths with compiled column and row transforms
constexpr int kLengths[] = {8, 16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512, 768, 1024};
inline bool length_supported(int n) {
  for (int v : kLengths)
    if (v == n) return true;
  return false;
}

template <class T> constexpr int strip_width() { return 128 / (int)sizeof(cx<T>); }
template <class T> inline int col_threads(int n) {
  // S columns x TH threads per column (TH = N / E)
  int e = 16;
  switch (n) {
    case 8: case 16: case 32: case 64: case 512: e = 8; break;
    case 24: case 48: case 96: case 192: case 384: case 768: e = 12; break;
    default: e = 16;
  }
  return strip_width<T>() * (n / e);
}
template <class T> inline int row_rows_per_block(int nx, int var = 0) {
  const int m = nx / 2;
  const int tx = m / row_plan_E(m, sizeof(T) == 8, var & 1);
  return 128 / tx;
}
// most row blocks any variant needs (sizes the per-block partials)
template <class T> inline int64_t row_blocks_max(int nx, int64_t rows) {
  int64_t b = 0;
  for (int v = 0; v < 2; ++v) {
    const int64_t rpb = row_rows_per_block<T>(nx, v);
    b = std::max<int64_t>(b, (rows + rpb - 1) / rpb);
  }
  return b;
}

// LDS bytes one workgroup may allocate on the device in use (hipDeviceAttributeMaxSharedMemoryPerBlock,
// read when an engine is created); wider column strips are used only within it.
inline int64_t& lds_per_block_limit() {
  static int64_t v = 64 * 1024;
  return v;
}

// column-pass strip widths: variant 0 = one 128 B segment, 1 = half, 2 = double, 3 = four
template <class T> inline int col_strip(int n, int var) {
  const int sw = strip_width<T>();
  const int s = var == 1 ? sw / 2 : var == 2 ? sw * 2 : var == 3 ? sw * 4 : sw;
  const int e = (n == 8 || n == 16 || n == 32 || n == 64 || n == 512) ? 8 : (n % 3 == 0) ? 12 : 16;
  const int threads = s * (n / e);
  const int64_t lds = (int64_t)n * (s + 1) * (int64_t)sizeof(cx<T>);  // tile + twiddle table
  if (threads > 1024 || lds > lds_per_block_limit()) return sw;
  return s;
}

// S = col_strip<T>(n, var) columns per tile (ColArgs.nstrips must use the same S)
template <class T>
hipError_t launch_colpass(const ColArgs& a, int n, int dir, int S, const RunBuf<T>* runs, const cx<T>* tw,
                          hipStream_t s);
template <class T>
hipError_t launch_xrow(const RowArgs& a, int kind, int nf, int var, const RunBuf<T>* runs, const cx<T>* twM,
                       const cx<T>* twh, hipStream_t s);
template <class T>
hipError_t launch_spectral(const SpecArgs& a, int phys, const RunBuf<T>* runs, hipStream_t s);
template <class T>
hipError_t launch_small(int which, const RunBuf<T>* runs, int nruns, int phys, int64_t nblocks, int advance,
                        hipStream_t s);
template <class T>
hipError_t launch_field_sums(const T* vars, int64_t nr, int nf, double* part, int nblocks, hipStream_t s);
template <class T>
hipError_t launch_dot_max(const T* a, const T* b, int64_t nr, int ncomp, double* part, int nblocks, hipStream_t s);
template <class T>
hipError_t launch_diag_spectral(const Geom& g, const cx<T>* src, cx<T>* dst, int op, hipStream_t s);
template <class T>
hipError_t launch_dealias(const Geom& g, cx<T>* x, int nf, hipStream_t s);
template <class T>
hipError_t launch_scatter(const Geom& g, cx<T>* x, int nf, int64_t nmodes, const int32_t* idx, const cx<T>* vals,
                          hipStream_t s);
template <class T>
hipError_t launch_gather(const Geom& g, const cx<T>* x, int nf, int64_t nmodes, const int32_t* idx, cx<T>* vals,
                         hipStream_t s);
template <class T>
hipError_t launch_zero(cx<T>* x, int64_t n, hipStream_t s);
template <class T>
hipError_t launch_unpad(const Geom& g, const cx<T>* work_field, cx<T>* dst, hipStream_t s);
template <class T>
hipError_t launch_shell(const Geom& g, const cx<T>* work, int ncomp, double* out, int nbins, hipStream_t s);
template <class T>
hipError_t launch_band(const Geom& g, cx<T>* work, int ncomp, double k1, double k2, hipStream_t s);
template <class T>
hipError_t launch_div_correction(const Geom& g, cx<T>* f, hipStream_t s);
template <class T>
hipError_t launch_sf_power(const Geom& g, cx<T>* work, double* means, hipStream_t s);
template <class T>
hipError_t launch_sf_bin(const Geom& g, const T* cf, double summean2, const double* mb, double* hist, int nb,
                         int nbx, bool global, hipStream_t s);
template <class T>
hipError_t launch_nd_amp(const RunBuf<T>* runs, int nruns, int64_t nblocks, double dV, int gate, hipStream_t s);

enum SmallKernel { SK_CFL = 0, SK_REDUCE = 1, SK_ADVANCE = 2, SK_VARSMAX = 3, SK_PRUNE = 4 };

}  // namespace mhd
