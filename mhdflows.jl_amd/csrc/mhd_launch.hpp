// Host-side launchers: map runtime grid lengths onto the compiled kernel
// instances.  Implemented per precision in k_launch_f32.hip / k_launch_f64.hip.
#pragma once
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
template <class T> inline int row_rows_per_block(int nx) {
  // must match RowPlanT<T, nx/2>::E (mhd_fft.hpp)
  const int m = nx / 2;
  int e = 8;
  const bool three = (m % 3 == 0);
  if (m == 4) e = 4;
  else if (three) e = 6;
  else if (sizeof(T) == 8) e = 4;
  else e = 8;
  const int tx = m / e;
  return 128 / tx;
}

template <class T>
hipError_t launch_colpass(const ColArgs& a, int n, int dir, const RunBuf<T>* runs, const cx<T>* tw,
                          hipStream_t s);
template <class T>
hipError_t launch_xrow(const RowArgs& a, int kind, int nf, const RunBuf<T>* runs, const cx<T>* twM,
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

enum SmallKernel { SK_CFL = 0, SK_REDUCE = 1, SK_ADVANCE = 2, SK_VARSMAX = 3 };

}  // namespace mhd
