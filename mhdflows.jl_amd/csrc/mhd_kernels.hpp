// Kernel templates of the MI355X MHD right-hand side and time step.
//
// One right-hand-side evaluation (MHDcalcN!, src/pgen.jl:153-162) is six launches:
//   colpass z (inverse)  stage input  -> work[0..nf)   only kept (kx, l) pencils, kept m rows read
//   colpass y (inverse)  work -> work                  only kx < kx_keep columns, kept l rows read
//   xrow (fused)         work[0..nf) -> work[0..np)    c2r, products u(x)u-b(x)b and u x b, r2c
//   colpass y (forward)  work -> work                  all columns
//   colpass z (forward)  work -> work                  all columns
//   spectral             products + stage input -> RK update (or N)
// Masked (aliased) modes of the stage input are exactly zero, so the inverse passes
// never read them; every forward output mode is produced, as the reference
// computes the right-hand side on the whole half spectrum.
#pragma once
#include "mhd_common.hpp"

namespace mhd {

// XCD-aware block order: hardware hands consecutive workgroups to the 8 XCDs in
// turn; remapping gives each XCD a contiguous range of items, so neighbouring
// column strips (which share cache lines when rows are not 128 B aligned) land
// in the same L2.
__device__ __forceinline__ int64_t xcd_item(int64_t b, int64_t nb) {
  const int64_t per = (nb + 7) / 8;
  return (b % 8) * per + b / 8;
}

// ---------------------------------------------------------------------------
// column pass: FFT along y or z of S adjacent kx columns
// ---------------------------------------------------------------------------
struct ColArgs {
  Geom g;
  int axis;            // 1: y, 2: z
  int nfields;         // fields per run
  int src, dst;        // BufSel
  int src_f0, dst_f0;  // first field index inside src / dst
  int ncols;           // kx columns processed [0, ncols)
  int nstrips;         // ceil(ncols / S)
  int nother;          // other-axis indices processed
  int other_lo_end;    // pruned other set: o < lo_end -> o, else o + other_gap
  int other_gap;
  int other0;          // first other index (a y pass over a slab of z-planes)
  int mask_lo, mask_hi;  // input positions in [mask_lo, mask_hi) are zero (not read)
  int nruns;
  int64_t nitems;
  int gate;            // 1 inside a time step: skip runs whose clock reached its target
  int plane_out;       // inverse z-pass of a masked stage input: copy the kx = 0 column into rb.plane
  int prune_out;       // forward pass of a time step: in a pruned step only kept outputs are produced
};

// One work item (a strip of S columns of one pencil line) of a column pass, run by the
// S*(N/E) threads of a workgroup.  lds holds N*S exchange slots and stw N twiddles.  Every
// early return is taken by the whole workgroup.
template <class T, int N, int S, int DIR>
__device__ __forceinline__ void colpass_item(const ColArgs& a, int64_t item, const RunBuf<T>* __restrict__ runs,
                                             const cx<T>* __restrict__ tw, cx<T>* lds, cx<T>* stw) {
  constexpr int E = Plan<N>::E;
  constexpr int TH = N / E;
  using Rad = typename Plan<N>::R;
  constexpr int R0 = RFirst<Rad>::value;
  constexpr int RL = RLast<Rad>::value;

  const int t = threadIdx.x;
  const int c = t % S;
  const int g = t / S;
  const Geom& G = a.g;
  // element (along the axis) and other-index strides of the source and the destination
  const int64_t sps = plane_stride(G, a.src), dps = plane_stride(G, a.dst);
  const int64_t srs = row_stride(G, a.src), drs = row_stride(G, a.dst);
  const int64_t s_estride = a.axis == 1 ? srs : sps;
  const int64_t s_ostride = a.axis == 1 ? sps : srs;
  const int64_t d_estride = a.axis == 1 ? drs : dps;
  const int64_t d_ostride = a.axis == 1 ? dps : drs;
  auto ldsf = [&](int pos) -> cx<T>& { return lds[pos * S + c]; };

  int64_t it = item;
  const int strip = (int)(it % a.nstrips); it /= a.nstrips;
  const int oi = (int)(it % a.nother); it /= a.nother;
  const int f = (int)(it % a.nfields);
  const int run = (int)(it / a.nfields);
  const int o = (oi < a.other_lo_end ? oi : oi + a.other_gap) + a.other0;
  const RunBuf<T>& rb = runs[run];
  if (gated_off(a.gate, rb)) return;
  const bool pruned = a.prune_out && step_pruned(rb);
  // stored columns: the computed ones, rounded up to whole 128-byte lines of the work buffer
  // (the extra columns load nothing and store zeros).  Measured at 256^3 (A/B on one box): the
  // forward y pass 168 -> 157 us, the forward z pass 125 -> 122 us, the step 319 -> 323 steps/s.
  const bool dwork = a.dst == B_WORK;
  const int ccols = pruned ? (a.ncols < G.kx_keep ? a.ncols : G.kx_keep) : a.ncols;
  const int scols = !dwork ? ccols : (ccols >= G.nkr ? G.wrs : G.kx_keep_w);
  if (pruned && (strip * S >= scols || (a.axis == 2 && !kept_y(G, o)))) return;  // no kept output
  const int col = strip * S + c;
  const bool col_ok = col < a.ncols && !(pruned && col >= G.kx_keep);
  const bool col_st = col < scols;
  const gptr<const cx<T>> __restrict__ src = gmem<const cx<T>>(rb.buf[a.src]) +
                                             (int64_t)(a.src_f0 + f) * field_stride(G, a.src) + o * s_ostride + col;
  const gptr<cx<T>> __restrict__ dst =
      gmem(rb.buf[a.dst]) + (int64_t)(a.dst_f0 + f) * field_stride(G, a.dst) + o * d_ostride + col;

  cx<T> x[E];
#pragma unroll
  for (int i = 0; i < E / R0; ++i) {
#pragma unroll
    for (int r = 0; r < R0; ++r) {
      const int p = g + i * TH + r * (N / R0);
      const bool ok = col_ok && !(p >= a.mask_lo && p < a.mask_hi);
      x[i * R0 + r] = ok ? ldst<false, T>(src + p * s_estride) : cx<T>{T(0), T(0)};
    }
  }
  // twiddles into LDS while the column loads are in flight: the passes then read them
  // from LDS instead of waiting on the cache hierarchy in their dependency chains
  for (int i = t; i < N; i += S * TH) stw[i] = tw[i];
  __syncthreads();
  if (a.plane_out && col == 0) {  // axis 2: o = l, positions p = m
    const gptr<cx<T>> pl = gmem(rb.plane) + ((int64_t)f * G.nz) * G.ny + o;
#pragma unroll
    for (int i = 0; i < E / R0; ++i)
#pragma unroll
      for (int r = 0; r < R0; ++r) pl[(int64_t)(g + i * TH + r * (N / R0)) * G.ny] = x[i * R0 + r];
  }
  fft_regs<T, N, E, DIR, Rad>(x, g, ldsf, stw);
  if (col_st) {
#pragma unroll
    for (int i = 0; i < E / RL; ++i) {
#pragma unroll
      for (int r = 0; r < RL; ++r) {
        const int p = g + i * TH + r * (N / RL);
        if (pruned && !(a.axis == 1 ? kept_y(G, p) : kept_z(G, p))) continue;
        stst<false, T>(dst + p * d_estride, x[i * RL + r]);
      }
    }
  }
}

template <class T, int N, int S, int DIR>
__global__ void __launch_bounds__(S*(N / Plan<N>::E))
colpass_kernel(ColArgs a, const RunBuf<T>* __restrict__ runs, const cx<T>* __restrict__ tw) {
  __shared__ cx<T> lds[N * S];
  __shared__ cx<T> stw[N];  // the twiddle table, staged once per workgroup
  const int64_t item = xcd_item(blockIdx.x, (int64_t)gridDim.x);
  if (item >= a.nitems) return;
  colpass_item<T, N, S, DIR>(a, item, runs, tw, lds, stw);
}

// ---------------------------------------------------------------------------
// row kernel: transforms along x of whole rows, with the real-space work fused
// ---------------------------------------------------------------------------
enum RowKind {
  ROW_MHD = 0,        // 6 complex rows -> c2r -> 9 products -> r2c
  ROW_HD = 1,         // 3 complex rows -> c2r -> 6 products -> r2c
  ROW_C2R = 2,        // NF complex rows -> NF real rows
  ROW_R2C = 3,        // NF real rows -> NF complex rows
  ROW_VARS_MHD = 4,   // 6 real rows -> 9 products and the 6 inputs -> r2c (15 rows)
  ROW_VARS_HD = 5,    // 3 real rows -> 6 products and the 3 inputs -> r2c (9 rows)
  ROW_MHD_ND = 6,     // ROW_MHD plus negative damping: 3 more products f_i u_i and sum |u_i^2 f_i|
  ROW_MHD_VP = 7,     // ROW_MHD plus volume penalisation: 6 more products chi/eta (u_j - U0_j), chi/eta (b_j - B0_j)
  ROW_HD_VP = 8,      // ROW_HD plus volume penalisation: 3 more products chi/eta (u_j - U0_j)
};

struct RowArgs {
  Geom g;
  int nruns;
  int in_sel;         // BufSel of the complex input (c2r kinds)
  int in_f0;          // first input field (complex input, or real input rows)
  int real_sel;       // real rows: 0 = vars, 1 = real scratch (rscr)
  int in_ncols;       // kx entries read per input row (others are zero)
  int out_sel;        // BufSel of the product spectra
  int out_f0;         // first product field
  int out2_sel;       // BufSel of the r2c of the inputs themselves (ROW_R2C / ROW_VARS_*)
  int out2_f0;        // first field of those
  int real_out_f0;    // c2r kinds: first real field written
  int write_vars;     // c2r kinds: store the real rows into vars
  int want_max;       // c2r kinds: per-block max of u^2 and b^2 into partial
  int64_t rows;       // end of the rows covered (ny*nz)
  int64_t blocks_per_run;
  int64_t blk0;       // first row block (a slab of z-planes starts at row blk0 * rows-per-block)
  int gate;           // 1 inside a time step: skip runs whose clock reached its target
  int prune_out;      // products of a time step: a pruned step stores the kept kx only
};

// inputs, products and self-transforms (r2c of the inputs) of each row kind
template <int KIND, int NF> struct RowKindInfo;
template <int NF> struct RowKindInfo<ROW_MHD, NF> { static constexpr int NIN = 6, NPROD = 9, NSELF = 0; };
template <int NF> struct RowKindInfo<ROW_HD, NF> { static constexpr int NIN = 3, NPROD = 6, NSELF = 0; };
template <int NF> struct RowKindInfo<ROW_C2R, NF> { static constexpr int NIN = NF, NPROD = 0, NSELF = 0; };
template <int NF> struct RowKindInfo<ROW_R2C, NF> { static constexpr int NIN = NF, NPROD = 0, NSELF = NF; };
template <int NF> struct RowKindInfo<ROW_VARS_MHD, NF> { static constexpr int NIN = 6, NPROD = 9, NSELF = 6; };
template <int NF> struct RowKindInfo<ROW_VARS_HD, NF> { static constexpr int NIN = 3, NPROD = 6, NSELF = 3; };
template <int NF> struct RowKindInfo<ROW_MHD_ND, NF> { static constexpr int NIN = 6, NPROD = 12, NSELF = 0; };
template <int NF> struct RowKindInfo<ROW_MHD_VP, NF> { static constexpr int NIN = 6, NPROD = 15, NSELF = 0; };
template <int NF> struct RowKindInfo<ROW_HD_VP, NF> { static constexpr int NIN = 3, NPROD = 9, NSELF = 0; };

// Real-space products of one point; u = v[0..2], b = v[3..5].
// MHD: T_ij = b_i b_j - u_i u_j (i<=j: 00,01,02,11,12,22), M_aj = u_a b_j - b_a u_j (01,02,12).
// HD:  U_ij = u_i u_j.
template <class T, int KIND>
__device__ __forceinline__ T product(int p, const T* v) {
  if constexpr (KIND == ROW_MHD || KIND == ROW_VARS_MHD || KIND == ROW_MHD_ND || KIND == ROW_MHD_VP) {
    switch (p) {
      case 0: return v[3] * v[3] - v[0] * v[0];
      case 1: return v[3] * v[4] - v[0] * v[1];
      case 2: return v[3] * v[5] - v[0] * v[2];
      case 3: return v[4] * v[4] - v[1] * v[1];
      case 4: return v[4] * v[5] - v[1] * v[2];
      case 5: return v[5] * v[5] - v[2] * v[2];
      case 6: return v[0] * v[4] - v[3] * v[1];
      case 7: return v[0] * v[5] - v[3] * v[2];
      default: return v[1] * v[5] - v[4] * v[2];
    }
  } else {
    switch (p) {
      case 0: return v[0] * v[0];
      case 1: return v[0] * v[1];
      case 2: return v[0] * v[2];
      case 3: return v[1] * v[1];
      case 4: return v[1] * v[2];
      default: return v[2] * v[2];
    }
  }
}

template <class T>
__device__ __forceinline__ T nanmax(T a, T b) { return (a > b || a != a) ? a : b; }

// Row LDS index.  With PAD the slots of each 16-entry line (128 B of Float32 pairs, all
// 32 banks) are permuted by an XOR with 5 * line: the strided exchange writes of the
// Stockham passes (position strides 2 and 4, and blocks of 4 or 8 at stride 16) then
// spread over the banks, while unit-stride accesses stay conflict-free.
template <bool PAD>
__device__ __forceinline__ int rx(int pos) {
  return PAD ? pos ^ ((5 * (pos >> 4)) & 15) : pos;
}
template <int M, bool PAD> constexpr int row_slots() { return PAD ? ((M >> 4) + 1) * 16 : M + 1; }

// Forward real transform of the values z (one complex per pair of real samples,
// in the pass-0 layout of Rfwd) and store the nkr half-spectrum entries.
template <class T, int M, int E, int TX, class Rfwd, bool WS, bool PAD, class Lds>
__device__ __forceinline__ void row_r2c_store(cx<T>* z, int tr, Lds& ldsf, cx<T>* xrow,
                                              const cx<T>* __restrict__ twM, const cx<T>* __restrict__ twh,
                                              gptr<cx<T>> __restrict__ dst, bool row_ok, int kend) {
  // kend: entries [0, kend) are stored; entries past M (the padding of a work-buffer row,
  // written so the row's last line is whole) are zeros
  constexpr int RfL = RLast<Rfwd>::value;
  fft_regs<T, M, E, -1, Rfwd, WS>(z, tr, ldsf, twM);
  xsync<WS>();
#pragma unroll
  for (int i = 0; i < E / RfL; ++i)
#pragma unroll
    for (int r = 0; r < RfL; ++r) xrow[rx<PAD>(tr + i * TX + r * (M / RfL))] = z[i * RfL + r];
  xsync<WS>();
  // entries k < min(M, kend); the Nyquist entry M shares Z_0 with k = 0 and is finished there.
  // The thread's own outputs are the positions k = tr + j*TX (j < E), held in z: Float32
  // rows take Z_k from there and only the mirror Z_{M-k} through LDS.
  constexpr int NI = E / RfL;
  constexpr bool OWN = sizeof(T) == 4;
  const int kl = kend < M ? kend : M;
#pragma unroll
  for (int j = 0; j < E; ++j) {
    const int k = tr + j * TX;
    if (k < kl) {
      const cx<T> Zk = OWN ? z[(j % NI) * RfL + j / NI] : xrow[rx<PAD>(k)];
      const cx<T> Zm = conj(xrow[rx<PAD>(k == 0 ? 0 : M - k)]);
      const cx<T> Ev = (Zk + Zm) * T(0.5);
      const cx<T> Od = mul_si<-1>(Zk - Zm) * T(0.5);  // (Zk - Zm) / (2i)
      if (row_ok) {
        stst<false, T>(dst + k, Ev + cmul(twh[k], Od));
        if (k == 0 && M < kend) stst<false, T>(dst + M, Ev + cmul(twh[M], Od));
      }
    }
  }
  if (row_ok)
    for (int k = M + 1 + tr; k < kend; k += TX) stst<false, T>(dst + k, cx<T>{T(0), T(0)});
}

// Launch shape of the x-row kernel for a row length and variant: VAR bit 0 selects the
// plan (elements per thread), bit 1 the swizzled LDS index; RPW rows per 128-thread
// workgroup, each with SLOTS LDS exchange slots.
template <class T, int NX, int VAR> struct RowShape {
  static constexpr int M = NX / 2;
  using RP = RowPlanV<T, M, (VAR & 1)>;
  static constexpr bool PAD = (VAR & 2) != 0;
  static constexpr int E = RP::E;
  static constexpr int TX = M / E;
  static constexpr int RPW = 128 / TX;
  static constexpr int SLOTS = row_slots<M, PAD>();
  static constexpr int LDS_CX = RPW * SLOTS;  // exchange slots of the workgroup
};

// One work item (RPW rows of one run) of the x-row kernel, run by a 128-thread workgroup.
// xbuf holds RowShape::LDS_CX exchange slots, red 2*128 values and stw 2*M+1 twiddles.
// Every early return is taken by the whole workgroup.
template <class T, int NX, int KIND, int NF, int VAR>
__device__ __forceinline__ void xrow_item(const RowArgs& a, int64_t item, const RunBuf<T>* __restrict__ runs,
                                          const cx<T>* __restrict__ twM, const cx<T>* __restrict__ twh,
                                          cx<T>* xbuf, T* red, cx<T>* stw) {
  using SH = RowShape<T, NX, VAR>;
  constexpr int M = SH::M;
  using RP = typename SH::RP;
  constexpr bool PAD = SH::PAD;
  constexpr int E = SH::E;
  constexpr int TX = SH::TX;
  static_assert(TX <= 128, "row transform too long for one workgroup");
  constexpr int RPW = SH::RPW;
  // a row's TX threads are lanes of one wavefront when TX <= 64 (TX is a power of
  // two): its LDS exchanges then need only wavefront-scope ordering
  static_assert((TX & (TX - 1)) == 0, "threads per row must be a power of two");
  constexpr bool WS = TX <= 64;
  using Rinv = typename RP::R;
  using Rfwd = typename RReverse<Rinv>::type;
  constexpr int R0 = RFirst<Rinv>::value;   // first inverse radix == last forward radix
  constexpr int RL = RLast<Rinv>::value;    // last inverse radix == first forward radix
  constexpr bool C2R_IN = (KIND == ROW_MHD || KIND == ROW_HD || KIND == ROW_C2R || KIND == ROW_MHD_ND ||
                           KIND == ROW_MHD_VP || KIND == ROW_HD_VP);
  constexpr bool ND = KIND == ROW_MHD_ND;
  // volume penalisation: the advective products come first, then one product per field
  constexpr bool VP = KIND == ROW_MHD_VP || KIND == ROW_HD_VP;
  constexpr int NADV = KIND == ROW_MHD_VP ? 9 : (KIND == ROW_HD_VP ? 6 : 0);
  constexpr int NIN = RowKindInfo<KIND, NF>::NIN;
  constexpr int NPROD = RowKindInfo<KIND, NF>::NPROD;

  const int t = threadIdx.x;
  const int rw = t / TX;
  const int tr = t % TX;
  const Geom& G = a.g;
  const int run = (int)(item / a.blocks_per_run);
  const int64_t blk = a.blk0 + item % a.blocks_per_run;  // row block within the run
  const int64_t row = blk * RPW + rw;
  const bool row_ok = row < a.rows;
  const RunBuf<T>& rb = runs[run];
  if (gated_off(a.gate, rb)) return;
  // output entries stored: kept kx of a dealiased step or all, rounded up to whole lines in
  // the work buffer (compact destinations take the exact range)
  const bool row_pruned = a.prune_out && step_pruned(rb);
  const int kend_c = row_pruned ? G.kx_keep : M + 1;
  const int kend_w = row_pruned ? G.kx_keep_w : G.wrs;
  // The passes read their twiddles from LDS: a table load from the cache hierarchy would
  // sit in every pass's dependency chain.
  for (int i = t; i < 2 * M + 1; i += 128) stw[i] = i < M ? twM[i] : twh[i - M];
  __syncthreads();
  const cx<T>* const sM = stw;
  const cx<T>* const sh = stw + M;
  const gptr<T> rbase = gmem(a.real_sel ? rb.rscr : rb.vars);
  cx<T>* xrow = xbuf + rw * SH::SLOTS;
  auto ldsf = [&](int pos) -> cx<T>& { return xrow[rx<PAD>(pos)]; };

  // v[f][e]: the real samples 2j and 2j+1 of field f at the thread's positions
  // j = tr + i*TX + r*(M/RL) (last inverse pass layout == first forward pass layout).
  cx<T> v[NIN][E];
  cx<T> fd[ND ? 3 : 1][E];  // negative damping: f_x, f_y, f_z at the same samples

  if constexpr (C2R_IN) {
    const T scale = T(G.inv_n);
    // Every field's row is requested before the first transform starts, so the loads'
    // latencies overlap instead of adding up: entries k = tr + i*TX (i < E) land in
    // v[f][i]; with at least NIN threads per row, lane tr = f holds the Nyquist entry
    // k = M of field f, otherwise lane 0 loads it when its field is staged.
    const int64_t ifs = field_stride(G, a.in_sel);
    const gptr<const cx<T>> __restrict__ src0 =
        gmem<const cx<T>>(rb.buf[a.in_sel]) + (int64_t)a.in_f0 * ifs + row_offset(G, a.in_sel, row);
#pragma unroll
    for (int f = 0; f < NIN; ++f)
#pragma unroll
      for (int i = 0; i < E; ++i) {
        const int k = tr + i * TX;
        v[f][i] = (row_ok && k < a.in_ncols) ? ldst<true, T>(src0 + (int64_t)f * ifs + k) : cx<T>{T(0), T(0)};
      }
    constexpr bool NYQ_LANES = TX >= NIN;
    cx<T> nyq = cx<T>{T(0), T(0)};
    if constexpr (NYQ_LANES)
      if (row_ok && tr < NIN && M < a.in_ncols) nyq = ldst<true, T>(src0 + (int64_t)tr * ifs + M);
#pragma unroll
    for (int f = 0; f < NIN; ++f) {
      xsync<WS>();
#pragma unroll
      for (int i = 0; i < E; ++i) xrow[rx<PAD>(tr + i * TX)] = v[f][i];
      if constexpr (NYQ_LANES) {
        if (tr == f) xrow[rx<PAD>(M)] = nyq;
      } else if (tr == 0) {
        xrow[rx<PAD>(M)] = (row_ok && M < a.in_ncols) ? ldst<true, T>(src0 + (int64_t)f * ifs + M) : cx<T>{T(0), T(0)};
      }
      xsync<WS>();
      // Position p = tr + i*TX + r*(M/R0) = tr + (i + r*E/R0)*TX was loaded by this thread
      // into raw[i + r*E/R0]: Float32 rows take X_p from there and only the mirror X_{M-p}
      // through LDS (Float64 rows, short of registers, read both from LDS).
      constexpr bool OWN = sizeof(T) == 4;
      cx<T> raw[E];
#pragma unroll
      for (int e = 0; e < E; ++e) raw[e] = v[f][e];
#pragma unroll
      for (int i = 0; i < E / R0; ++i) {
#pragma unroll
        for (int r = 0; r < R0; ++r) {
          const int p = tr + i * TX + r * (M / R0);
          cx<T> Xk = OWN ? raw[i + r * (E / R0)] : xrow[rx<PAD>(p)];
          cx<T> Xm = xrow[rx<PAD>(M - p)];
          if (p == 0) { Xk.im = T(0); Xm.im = T(0); }  // c2r ignores Im of the DC and Nyquist bins
          const cx<T> A = Xk + conj(Xm);
          const cx<T> B = Xk - conj(Xm);
          v[f][i * R0 + r] = A + mul_si<1>(cmulc(B, sh[p]));
        }
      }
      fft_regs<T, M, E, 1, Rinv, WS>(v[f], tr, ldsf, sM);
#pragma unroll
      for (int e = 0; e < E; ++e) v[f][e] = v[f][e] * scale;
    }
    // A dealiased step's real fields would be overwritten by a later step of the same call
    // before anything reads them (the CFL takes the maxima below, not vars), so only the
    // call's last step, which is full, stores them.
    if (a.write_vars && row_ok && !(a.prune_out && step_pruned(rb))) {
#pragma unroll
      for (int f = 0; f < NIN; ++f) {
        const gptr<T> __restrict__ vr = rbase + (int64_t)(a.real_out_f0 + f) * G.nr + row * NX;
#pragma unroll
        for (int i = 0; i < E / RL; ++i)
#pragma unroll
          for (int r = 0; r < RL; ++r) {
            const int j = tr + i * TX + r * (M / RL);
            *(gptr<cx<T>>)(vr + 2 * j) = v[f][i * RL + r];
          }
      }
    }
    if (a.want_max) {
      T mu = T(0), mbm = T(0);
      if (row_ok) {
#pragma unroll
        for (int f = 0; f < NIN; ++f)
#pragma unroll
          for (int e = 0; e < E; ++e) {
            const T s1 = v[f][e].re * v[f][e].re, s2 = v[f][e].im * v[f][e].im;
            if (f < 3) mu = nanmax(nanmax(mu, s1), s2);
            else mbm = nanmax(nanmax(mbm, s1), s2);
          }
      }
      T* const red1 = red + 128;
      red[t] = mu;
      red1[t] = mbm;
      __syncthreads();
      for (int s = 64; s > 0; s >>= 1) {
        if (t < s) {
          red[t] = nanmax(red[t], red[t + s]);
          red1[t] = nanmax(red1[t], red1[t + s]);
        }
        __syncthreads();
      }
      if (t == 0) {
        const gptr<T> part = gmem(rb.partial);
        part[2 * blk] = red[0];
        part[2 * blk + 1] = red1[0];
      }
    }
    if constexpr (ND) {
      // NDForceDriving! (src/pgen/NegativeDamping.jl:23-44): f_i at the thread's samples and the
      // block's share of sum |u_i^2 f_i|, accumulated in Float64 (one partial per row block)
      __shared__ double nds[128];
      double sum = 0.0;
      const gptr<const T> fb = gmem<const T>(rb.ndf);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const gptr<const T> fr = fb + (int64_t)c * G.nr + row * NX;
#pragma unroll
        for (int i = 0; i < E / RL; ++i)
#pragma unroll
          for (int r = 0; r < RL; ++r) {
            const int j = tr + i * TX + r * (M / RL);
            const int e = i * RL + r;
            fd[c][e] = row_ok ? ldg((gptr<const cx<T>>)(fr + 2 * j)) : cx<T>{T(0), T(0)};
            const double ur = v[c][e].re, ui = v[c][e].im;
            sum += fabs(ur * ur * (double)fd[c][e].re) + fabs(ui * ui * (double)fd[c][e].im);
          }
      }
      nds[t] = sum;
      __syncthreads();
      for (int st = 64; st > 0; st >>= 1) {
        if (t < st) nds[t] += nds[t + st];
        __syncthreads();
      }
      if (t == 0) gmem(rb.ndpart)[blk] = nds[0];
    }
  } else {
    // real input rows (vars), loaded straight into the forward pass-0 layout
#pragma unroll
    for (int f = 0; f < NIN; ++f) {
      const gptr<const T> __restrict__ vr = (gptr<const T>)rbase + (int64_t)(a.in_f0 + f) * G.nr + row * NX;
#pragma unroll
      for (int i = 0; i < E / RL; ++i)
#pragma unroll
        for (int r = 0; r < RL; ++r) {
          const int j = tr + i * TX + r * (M / RL);
          v[f][i * RL + r] = row_ok ? ldg((gptr<const cx<T>>)(vr + 2 * j)) : cx<T>{T(0), T(0)};
        }
    }
  }

  if constexpr (KIND == ROW_C2R) {
    return;
  } else {
    // forward transforms: products (if any) then the inputs themselves (ROW_R2C / ROW_VARS_*)
    constexpr int NOUT = NPROD + RowKindInfo<KIND, NF>::NSELF;
    // Float32 rows unroll this loop (each product's formula is then resolved at compile
    // time); Float64 rows keep it rolled, as their six resident fields leave no room for
    // the longer live ranges of the unrolled form.
    constexpr int UNR = sizeof(T) == 4 ? NOUT : 1;
    // VP_UᵢUpdate! / VP_BᵢUpdate! (src/Solver/VPSolver.jl:21-56): chi / eta with
    // eta = clock.dt * 13 / 7, all in T as the reference's broadcasts evaluate it
    const T eta = VP ? T(rb.clk->dt) * T(13) / T(7) : T(1);
#pragma unroll UNR
    for (int p = 0; p < NOUT; ++p) {
      cx<T> z[E];
      if (VP && p >= NADV) {
        // chi/eta * (v_c - W_c) with W = U0 (c < 3) or B0 (c >= 3): rb.vp fields 0 (chi) and
        // 1 + c, read at the thread's samples for each product (no registers held across them)
        const gptr<const T> cr = gmem<const T>(rb.vp) + row * NX;
#pragma unroll
        for (int c = 0; c < NIN; ++c)
          if (c == p - NADV) {
            const gptr<const T> wr = cr + (int64_t)(1 + c) * G.nr;
#pragma unroll
            for (int i = 0; i < E / RL; ++i)
#pragma unroll
              for (int r = 0; r < RL; ++r) {
                const int j = tr + i * TX + r * (M / RL);
                const int e = i * RL + r;
                const cx<T> x = row_ok ? ldg((gptr<const cx<T>>)(cr + 2 * j)) : cx<T>{T(0), T(0)};
                const cx<T> w = row_ok ? ldg((gptr<const cx<T>>)(wr + 2 * j)) : cx<T>{T(0), T(0)};
                z[e] = cx<T>{(x.re / eta) * (v[c][e].re - w.re), (x.im / eta) * (v[c][e].im - w.im)};
              }
          }
      } else if (ND && p >= 9) {
        // F_i = f_i u_i (NegativeDamping.jl:39); c is matched against p so that the register
        // arrays are indexed by constants only
#pragma unroll
        for (int c = 0; c < 3; ++c)
          if (c == p - 9)
#pragma unroll
            for (int e = 0; e < E; ++e) z[e] = cx<T>{v[c][e].re * fd[c][e].re, v[c][e].im * fd[c][e].im};
      } else if (p < NPROD) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
          T pr[NIN > 0 ? NIN : 1], pi[NIN > 0 ? NIN : 1];
#pragma unroll
          for (int f = 0; f < NIN; ++f) { pr[f] = v[f][e].re; pi[f] = v[f][e].im; }
          z[e] = cx<T>{product<T, KIND>(p, pr), product<T, KIND>(p, pi)};
        }
      } else {
#pragma unroll
        for (int f = 0; f < NIN; ++f)
          if (f == p - NPROD)
#pragma unroll
            for (int e = 0; e < E; ++e) z[e] = v[f][e];
      }
      const gptr<cx<T>> __restrict__ dst =
          p < NPROD ? gmem(rb.buf[a.out_sel]) + (int64_t)(a.out_f0 + p) * field_stride(G, a.out_sel) +
                          row_offset(G, a.out_sel, row)
                    : gmem(rb.buf[a.out2_sel]) + (int64_t)(a.out2_f0 + p - NPROD) * field_stride(G, a.out2_sel) +
                          row_offset(G, a.out2_sel, row);
      const bool dst_work = (p < NPROD ? a.out_sel : a.out2_sel) == B_WORK;
      row_r2c_store<T, M, E, TX, Rfwd, WS, PAD>(z, tr, ldsf, xrow, sM, sh, dst, row_ok, dst_work ? kend_w : kend_c);
    }
  }
}

template <class T, int NX, int KIND, int NF, int VAR>
__global__ void __launch_bounds__(128)
xrow_kernel(RowArgs a, const RunBuf<T>* __restrict__ runs, const cx<T>* __restrict__ twM,
            const cx<T>* __restrict__ twh) {
  using SH = RowShape<T, NX, VAR>;
  __shared__ cx<T> xb[SH::LDS_CX];
  __shared__ T red[2 * 128];
  __shared__ cx<T> stw[2 * SH::M + 1];  // twiddles: M of the length-M transform, then M+1 of the real pack
  const int64_t item = xcd_item(blockIdx.x, (int64_t)gridDim.x);
  if (item >= a.blocks_per_run * a.nruns) return;
  xrow_item<T, NX, KIND, NF, VAR>(a, item, runs, twM, twh, xb, red, stw);
}

// ---------------------------------------------------------------------------
// spectral kernel: right-hand side at each mode + the Runge-Kutta stage update
// ---------------------------------------------------------------------------
enum SpecMode { SP_RK4 = 0, SP_LSRK54 = 1, SP_RHS = 2 };

struct SpecArgs {
  Geom g;
  int nruns;
  int mode;           // SpecMode
  int stage;
  int in_sel;         // stage input buffer (B_SOL / B_STIN / B_EXT_SOL)
  int mask_input;     // 1: the stage input is dealiased (masked modes read as 0)
  int zero_aliased;   // SP_RHS: write zeros into the aliased band of the input (dealias!)
  int add_force;      // add the forcing table (MHD only, as addforcing! follows the advection)
  int out_sel;        // SP_RHS: where N goes (B_EXT_N)
  int out_f0;         // SP_RHS: first field of N written (component calls)
  int f_lo, f_hi;     // SP_RHS: fields [f_lo, f_hi) written
  int prod_sel;       // products buffer (B_WORK or B_SCR)
  int self_sel;       // -1: diffusion on P_H(stage input); else R2C(vars) fields (B_SCR, at self_f0)
  int self_f0;
  double ca, cb;      // LSRK54 stage coefficients A_i, B_i (rounded to T in the kernel)
  int64_t nitems;     // nruns * nc
  int planes;         // 0: interior modes 1 <= kx < nx/2; 1: the kx = 0 and kx = nx/2 planes;
                      // 2: every mode, kx = 0 mirrors read from rb.plane (masked stage input)
  int gate;           // 1 inside a time step: skip runs whose clock reached its target
  int64_t draw_ctr;   // SP_RHS: draw counter of stochastic forcing (a step uses 8 step + stage)
  int modules;        // 1: a run has a calcF! module (dense field, negative damping, A99): the
                      // launch uses the kernel instance that evaluates them
};

template <class T>
struct ModeK {
  T kx, ky, kz, K, invK;
};

template <class T>
__device__ __forceinline__ ModeK<T> mode_k(const Geom& G, int kx, int l, int m) {
  ModeK<T> q;
  q.kx = T((double)kx * G.multx);
  q.ky = T((double)signed_index(l, G.ny) * G.multy);
  q.kz = T((double)signed_index(m, G.nz) * G.multz);
  // Krsq = kr^2 + l^2 + m^2 evaluated in T (utils.jl:73)
  q.K = (q.kx * q.kx + q.ky * q.ky) + q.kz * q.kz;
  q.invK = (kx == 0 && l == 0 && m == 0) ? T(0) : T(1) / q.K;
  return q;
}

// N += -(K^p) * coef * s, promoted to double and rounded back (MHDSolver.jl:94,168)
template <class T>
__device__ __forceinline__ cx<T> add_diffusion(cx<T> n, T Kp, double coef, cx<T> s) {
  const double c = -(double)Kp * coef;
  return cx<T>{T((double)n.re + c * (double)s.re), T((double)n.im + c * (double)s.im)};
}

template <class T>
__device__ __forceinline__ T ipow(T x, int n) {
  T r = x;
  for (int i = 1; i < n; ++i) r = r * x;
  return r;
}

// qp: the mode's offset in the products buffer (its planes may be padded)
template <class T, int PHYS>
__device__ __forceinline__ void rhs_at(const Geom& G, const RunBuf<T>& rb, const SpecArgs& a,
                                       int64_t qp, const ModeK<T>& k, const cx<T>* sH, cx<T>* R) {
  constexpr int NP = PHYS == PHYS_MHD ? 9 : 6;
  const gptr<const cx<T>> __restrict__ P = gmem<const cx<T>>(rb.buf[a.prod_sel]);
  const int64_t pfs = field_stride(G, a.prod_sel);
  cx<T> p[NP];
#pragma unroll
  for (int i = 0; i < NP; ++i) p[i] = ldst<true, T>(P + (int64_t)i * pfs + qp);
  const cx<T> zero{T(0), T(0)};
  if constexpr (PHYS == PHYS_MHD) {
    // D_j = sum_i i k_i T_ij ; N_u = D - k (k.D)/k^2  (UᵢUpdate!, MHDSolver.jl:63-83)
    const cx<T> d0 = mul_si<1>(p[0] * k.kx + p[1] * k.ky + p[2] * k.kz);
    const cx<T> d1 = mul_si<1>(p[1] * k.kx + p[3] * k.ky + p[4] * k.kz);
    const cx<T> d2 = mul_si<1>(p[2] * k.kx + p[4] * k.ky + p[5] * k.kz);
    const cx<T> kd = (d0 * k.kx + d1 * k.ky + d2 * k.kz) * k.invK;
    R[0] = d0 - kd * k.kx;
    R[1] = d1 - kd * k.ky;
    R[2] = d2 - kd * k.kz;
    // N_b_a = sum_{j != a} i k_j M_aj with M antisymmetric (BᵢUpdate!, MHDSolver.jl:147-158)
    R[3] = mul_si<1>(p[6] * k.ky + p[7] * k.kz);
    R[4] = mul_si<1>(p[8] * k.kz - p[6] * k.kx);
    R[5] = mul_si<1>(zero - p[7] * k.kx - p[8] * k.ky);
  } else {
    // HD: N_u = -P(i k . (u u))  (HDSolver.jl:58-74)
    const cx<T> d0 = mul_si<-1>(p[0] * k.kx + p[1] * k.ky + p[2] * k.kz);
    const cx<T> d1 = mul_si<-1>(p[1] * k.kx + p[3] * k.ky + p[4] * k.kz);
    const cx<T> d2 = mul_si<-1>(p[2] * k.kx + p[4] * k.ky + p[5] * k.kz);
    const cx<T> kd = (d0 * k.kx + d1 * k.ky + d2 * k.kz) * k.invK;
    R[0] = d0 - kd * k.kx;
    R[1] = d1 - kd * k.ky;
    R[2] = d2 - kd * k.kz;
  }
  constexpr int NF = PHYS == PHYS_MHD ? 6 : 3;
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    const bool isb = f >= 3;
    const double coef = isb ? rb.eta : rb.nu;
    const int pw = isb ? rb.neta : rb.nnu;
    R[f] = add_diffusion(R[f], k.K, coef, sH[f]);
    if (pw > 1) R[f] = add_diffusion(R[f], ipow(k.K, pw), coef, sH[f]);
  }
}

// ---------------------------------------------------------------------------
// calcF! modules evaluated per mode
// ---------------------------------------------------------------------------
// Counter-based uniform in [0, 1): splitmix64 of seed * 0x9E3779B97F4A7C15 + (counter << 32)
// + index + 1, top 53 bits (Float64) or top 24 bits (Float32, as rand(Float32)).
__device__ __forceinline__ uint64_t smix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
template <class T>
__device__ __forceinline__ T counter_u01(uint64_t seed, uint64_t counter, uint64_t index) {
  const uint64_t z = smix64(seed * 0x9E3779B97F4A7C15ull + (counter << 32) + 1 + index);
  if constexpr (sizeof(T) == 8) return (T)((double)(z >> 11) * (1.0 / 9007199254740992.0));
  else return (T)(float)(uint32_t)(z >> 40) * (1.0f / 16777216.0f);
}

// tanh of a complex Float64 by Kahan's method (tan / sinh form, as Julia's Base and FreeBSD's
// ctanh evaluate it): with t = tan(y), beta = 1 + t^2, s = sinh(x), rho = sqrt(1 + s^2),
// tanh(x + iy) = (beta rho s + i t) / (1 + beta s^2); |x| beyond the overflow bound gives +-1.
__device__ __forceinline__ void ctanh_d(double x, double y, double* re, double* im) {
  if (4.0 * fabs(x) > 710.4758600739439) {  // asinh(prevfloat(floatmax(Float64)))
    *re = copysign(1.0, x);
    *im = copysign(0.0, y * (isfinite(y) ? sin(2.0 * fabs(y)) : 1.0));
    return;
  }
  const double t = tan(y), beta = 1.0 + t * t, sh = sinh(x), rho = sqrt(1.0 + sh * sh);
  if (isinf(t)) {
    *re = rho / sh;
    *im = 1.0 / t;
    return;
  }
  const double den = 1.0 + beta * sh * sh;
  *re = beta * rho * sh / den;
  *im = t / den;
}

// principal square root of a complex T: rho = sqrt((|z| + |x|) / 2), then
// (rho, y / (2 rho)) for x >= 0 and (|y| / (2 rho), copysign(rho, y)) for x < 0
template <class T>
__device__ __forceinline__ cx<T> csqrt_t(cx<T> z) {
  if (z.re == T(0) && z.im == T(0)) return cx<T>{T(0), z.im};
  const T rho = sqrt((hypot(z.re, z.im) + fabs(z.re)) * T(0.5));
  const T eta = (z.im / rho) * T(0.5);
  if (z.re >= T(0)) return cx<T>{rho, eta};
  return cx<T>{fabs(eta), copysign(rho, z.im)};
}

// A99ForceDriving! (src/pgen/A99ForceDriving.jl:33-60) with SetUpFk (:93-127) at one mode:
// Fk = amp sqrt(exp(-(k - kf)^2 / s2) / 2 / pi) / k (zero on the kx = 0 plane), the unit
// vectors e1 = (ky, -kx, 0) / k_perp and e2 = (kx kz, ky kz, -k_perp^2) / (k_perp k), and four
// draws per mode and call: theta1, Phi (complex: rand! fills both parts of the Complex{T}
// buffer), theta2.  2pi and pi/2 are Float64 literals in the reference, so the phases and the
// tanh argument are evaluated in Float64 and rounded to T.
template <class T>
__device__ __forceinline__ void a99_add(const RunBuf<T>& rb, int kx, const ModeK<T>& k, uint64_t ctr, uint64_t q,
                                        cx<T>* R) {
  if (kx == 0) return;  // Fk[1,:,:] = 0 (A99ForceDriving.jl:110)
  const T kk = sqrt(k.K), kinv = sqrt(k.invK);
  const T d = kk - T(rb.a99_kf);
  const T g = exp(-(d * d) / T(rb.a99_s2));
  if (!(g > T(0))) return;  // Fk = 0: nothing is added
  const T Fk = T(rb.a99_amp) * sqrt(g / T(2) / T(3.14159265358979323846)) * kinv;
  const T kp = sqrt(k.kx * k.kx + k.ky * k.ky);
  const T e1x = k.ky / kp, e1y = -k.kx / kp;
  const T e2x = k.kx * k.kz / kp * kinv, e2y = k.ky * k.kz / kp * kinv, e2z = -kp * kinv;
  const uint64_t c4 = ctr << 2;
  const T u0 = counter_u01<T>(rb.a99_seed, c4, q), u1 = counter_u01<T>(rb.a99_seed, c4 + 1, q);
  const T u2 = counter_u01<T>(rb.a99_seed, c4 + 2, q), u3 = counter_u01<T>(rb.a99_seed, c4 + 3, q);
  const double twopi = 6.283185307179586, halfpi = 1.5707963267948966;
  double sn, cs;
  sincos((double)u0 * twopi, &sn, &cs);
  const cx<T> ph1{T(cs), T(sn)};
  const T pif = T(3.14159265358979323846);
  const cx<T> Phi{u1 * pif, u2 * pif};
  const double b = (double)T(rb.a99_b);
  double tr, ti;
  ctanh_d(b * ((double)Phi.re - halfpi), b * (double)Phi.im, &tr, &ti);
  const double den = rb.a99_den;
  const cx<T> gi{T(-tr / den), T(-ti / den)};
  const cx<T> gi2 = cmul(gi, gi);
  const cx<T> gj = csqrt_t(cx<T>{T(1) - gi2.re, -gi2.im});
  sincos((double)u3 * twopi, &sn, &cs);
  const cx<T> ph2{T(cs), T(sn)};
  const T AF = T(rb.a99_A) * Fk;
  const cx<T> w1 = cmul(ph1 * AF, gi);
  R[0] = R[0] + w1 * e1x;   // :48
  R[1] = R[1] + w1 * e1y;   // :49
  const cx<T> w2 = cmul(ph2 * AF, gj);
  R[0] = R[0] + w2 * e2x;   // :54
  R[1] = R[1] + w2 * e2y;   // :55
  R[2] = R[2] + w2 * e2z;   // :56
}

// MODS = 1 adds the calcF! modules (their registers would otherwise cost the plain launch
// occupancy: 61 -> 72 VGPRs for Float32 MHD)
template <class T, int PHYS, int PLANES, int MODS>
__device__ __forceinline__ void spectral_item(const SpecArgs& a, const RunBuf<T>* __restrict__ runs, uint32_t i,
                                              int m, int run) {
  constexpr int NF = PHYS == PHYS_MHD ? 6 : 3;
  const Geom& G = a.g;
  // grid: y = m, z = run.  Interior launch (PLANES == 0): x covers the modes
  // 1 <= kx < nx/2 of every l.  Plane launch (PLANES == 1): x covers the kx = 0
  // and kx = nx/2 planes, whose modes are handled in Hermitian pairs by the thread
  // of the lower index, so the pair's stage inputs are read before either is
  // overwritten.  Splitting keeps the interior waves free of the pair branch.
  // Whole launch (PLANES == 2, masked stage input only): x covers every mode of the
  // m slab; the mirror of a kx = 0 mode comes from the plane copy the inverse z-pass
  // made, which this launch does not write, and the masked kx = nx/2 plane is zero.
  int l, kx;
  if constexpr (PLANES == 2) {
    // dealiased step: a slab of aliased m holds no mode to update
    if (a.gate && !kept_z(G, m) && step_pruned(runs[run])) return;
    if (i >= (uint32_t)(G.ny * G.nkr)) return;
    l = (int)fastdiv(i, G.fd_nkr_m, G.fd_nkr_l);
    kx = (int)i - l * G.nkr;
  } else if constexpr (PLANES == 0) {
    const int ni = G.nkr - 2;
    if (i >= (uint32_t)(G.ny * ni)) return;
    l = (int)fastdiv(i, G.fd_nin_m, G.fd_nin_l);
    kx = (int)i - l * ni + 1;
  } else {
    if (i >= (uint32_t)(2 * G.ny)) return;
    const int p = i >= (uint32_t)G.ny ? 1 : 0;
    l = (int)i - p * G.ny;
    kx = p ? G.nx / 2 : 0;
  }
  const int64_t idx = ((int64_t)m * G.ny + l) * G.nkr + kx;
  const RunBuf<T>& rb = runs[run];
  if (gated_off(a.gate, rb)) return;

  // Every per-mode array below is indexed by compile-time constants only (the mode
  // loops have the fixed trip count NM and are unrolled), so it lives in registers;
  // a runtime index would place it in scratch memory and add its traffic to HBM.
  constexpr bool plane = PLANES == 1;
  constexpr int NM = plane ? 2 : 1;
  int nmodes = 1;
  int64_t qs[NM];
  int ls[NM], ms[NM];
#pragma unroll
  for (int j = 0; j < NM; ++j) { qs[j] = idx; ls[j] = l; ms[j] = m; }
  if constexpr (plane) {
    const int lm = (G.ny - l) % G.ny, mm = (G.nz - m) % G.nz;
    const int64_t mir = ((int64_t)mm * G.ny + lm) * G.nkr + kx;
    if (mir < idx) return;
    if (mir != idx) { nmodes = 2; qs[1] = mir; ls[1] = lm; ms[1] = mm; }
  }

  const gptr<const cx<T>> __restrict__ in = gmem<const cx<T>>(rb.buf[a.in_sel]);
  // Conditional reads are written as if-statements: a conditional operator over two
  // lvalues selects between their addresses and would move `zero` and `s` to scratch.
  const cx<T> zero{T(0), T(0)};
  cx<T> s[NM][NF];     // masked stage input at the (up to) two modes
  bool kept[NM];
#pragma unroll
  for (int j = 0; j < NM; ++j) {
    kept[j] = !a.mask_input || (kx < G.kx_keep && kept_y(G, ls[j]) && kept_z(G, ms[j]));
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      s[j][f] = zero;
      if (j < nmodes && kept[j] && a.self_sel < 0) s[j][f] = ldst<true, T>(in + (int64_t)f * G.nc + qs[j]);
    }
  }
  if constexpr (!plane) {
    if (a.gate && !kept[0] && step_pruned(rb)) return;  // pruned step: aliased modes are not updated
  }
  if (a.zero_aliased) {
    const gptr<cx<T>> __restrict__ io = gmem(rb.buf[a.in_sel]);
#pragma unroll
    for (int j = 0; j < NM; ++j)
      if (j < nmodes && !kept[j])
#pragma unroll
        for (int f = 0; f < NF; ++f) io[(int64_t)f * G.nc + qs[j]] = zero;
  }

#pragma unroll
  for (int j = 0; j < NM; ++j) {
    if (j >= nmodes) break;
    const int64_t q = qs[j];
    const ModeK<T> k = mode_k<T>(G, kx, ls[j], ms[j]);
    cx<T> sH[NF];
    if (a.self_sel >= 0) {
      const gptr<const cx<T>> __restrict__ sf = gmem<const cx<T>>(rb.buf[a.self_sel]);
#pragma unroll
      for (int f = 0; f < NF; ++f) sH[f] = sf[(int64_t)(a.self_f0 + f) * G.nc + q];
    } else if (plane) {
      // R2C(C2R(s)) = Hermitian projection on this plane; a self-conjugate mode is its own mirror
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        cx<T> mir = s[j][f];
        if (nmodes == 2) mir = s[NM - 1 - j][f];
        sH[f] = (s[j][f] + conj(mir)) * T(0.5);
      }
    } else if (PLANES == 2 && kx == 0) {
      const int lm = l == 0 ? 0 : G.ny - l, mm = m == 0 ? 0 : G.nz - m;
      const gptr<const cx<T>> pl = gmem<const cx<T>>(rb.plane) + (int64_t)mm * G.ny + lm;
      const int64_t pstride = (int64_t)G.ny * G.nz;
#pragma unroll
      for (int f = 0; f < NF; ++f) sH[f] = (s[j][f] + conj(ldg(pl + f * pstride))) * T(0.5);
    } else {
#pragma unroll
      for (int f = 0; f < NF; ++f) sH[f] = s[j][f];
    }
    cx<T> R[NF];
    const int64_t qp = ms[j] * plane_stride(G, a.prod_sel) + ls[j] * row_stride(G, a.prod_sel) + kx;
    rhs_at<T, PHYS>(G, rb, a, qp, k, sH, R);
    if (MODS && rb.vp) {
      // VP_UᵢUpdate! / VP_BᵢUpdate! (src/Solver/VPSolver.jl:21-56): N_a += -(delta_aj - k_j k_a k^-2) P_j
      // for the penalisation products P of u (and b), which follow the advective ones
      constexpr int NADV = PHYS == PHYS_MHD ? 9 : 6;
      const gptr<const cx<T>> Pv = gmem<const cx<T>>(rb.buf[a.prod_sel]);
      const int64_t pfs = field_stride(G, a.prod_sel);
      const T kv[3] = {k.kx, k.ky, k.kz};
#pragma unroll
      for (int g = 0; g < NF / 3; ++g)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const T kak2 = kv[c] * k.invK;
#pragma unroll
          for (int j = 0; j < 3; ++j) {
            const T coef = T(c == j ? 1 : 0) - kv[j] * kak2;
            const cx<T> P = ldst<true, T>(Pv + (int64_t)(NADV + 3 * g + j) * pfs + qp);
            R[3 * g + c] = R[3 * g + c] + cx<T>{-coef * P.re, -coef * P.im};
          }
        }
    }
    // addforcing! (src/pgen.jl:231-234) follows the MHD advection; HD discards it (:176-178)
    if (PHYS == PHYS_MHD && a.add_force) {
      if (rb.force) {  // seeded large-scale shell forcing
        const int K = rb.force_K;
        const int sl = signed_index(ls[j], G.ny), sm = signed_index(ms[j], G.nz);
        if (kx <= K && sl >= -K && sl <= K && sm >= -K && sm <= K) {
          const int64_t per = (int64_t)(K + 1) * (2 * K + 1) * (2 * K + 1);
          const int64_t fi = ((int64_t)(sm + K) * (2 * K + 1) + (sl + K)) * (K + 1) + kx;
#pragma unroll
          for (int c = 0; c < 3; ++c) R[c] = R[c] + rb.force[c * per + fi];
        }
      }
      if (MODS && rb.fdense) {  // N97ForceDriving! (src/pgen/TaylorGreenDynamo.jl:12-16): N_u += fh
        const gptr<const cx<T>> fdn = gmem<const cx<T>>(rb.fdense);
#pragma unroll
        for (int c = 0; c < 3; ++c) R[c] = R[c] + ldst<true, T>(fdn + (int64_t)c * G.nc + q);
      }
      if (MODS && rb.ndf) {  // NDForceDriving! (src/pgen/NegativeDamping.jl:40): N_u += A R2C(f u), A Float64
        const double A = rb.clk->nd_amp;
        const gptr<const cx<T>> Pn = gmem<const cx<T>>(rb.buf[a.prod_sel]);
        const int64_t pfs = field_stride(G, a.prod_sel);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const cx<T> F = ldst<true, T>(Pn + (int64_t)(9 + c) * pfs + qp);
          R[c] = cx<T>{T((double)R[c].re + A * (double)F.re), T((double)R[c].im + A * (double)F.im)};
        }
      }
      if (MODS && rb.a99) {
        const uint64_t ctr = a.mode == SP_RHS ? (uint64_t)a.draw_ctr : (uint64_t)(8 * rb.clk->step + a.stage);
        a99_add<T>(rb, kx, k, ctr, (uint64_t)q, R);
      }
    }

    if (a.mode == SP_RHS) {
      const gptr<cx<T>> __restrict__ out = gmem(rb.buf[a.out_sel]);
#pragma unroll
      for (int f = 0; f < NF; ++f)
        if (f >= a.f_lo && f < a.f_hi) out[(int64_t)(a.out_f0 + f - a.f_lo) * G.nc + q] = R[f];
      continue;
    }
    const T dt = T(rb.clk->dt);
    const gptr<cx<T>> __restrict__ sol = gmem(rb.buf[B_SOL]);
    const gptr<cx<T>> __restrict__ acc = gmem(rb.buf[B_ACC]);
    const gptr<cx<T>> __restrict__ stin = gmem(rb.buf[B_STIN]);
    if (a.mode == SP_RK4) {
      // RK4substeps! / RK4update! (DyeModule.jl:62-91): acc holds the running sum
      // RHS1/6 + RHS2/3 + RHS3/3 in the reference's left-to-right order.
      const T half = dt / T(2);
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        const int64_t o = (int64_t)f * G.nc + q;
        cx<T> base = zero;
        if (kept[j]) {
          if (a.stage == 0) base = s[j][f];
          else base = ldst<true, T>(sol + o);
        }
        if (a.stage == 0) {
          stst<true, T>(acc + o, cx<T>{R[f].re / T(6), R[f].im / T(6)});
          if (kept[j]) stst<true, T>(stin + o, base + R[f] * half);
        } else if (a.stage == 1 || a.stage == 2) {
          const cx<T> ac = ldst<true, T>(acc + o);
          stst<true, T>(acc + o, cx<T>{ac.re + R[f].re / T(3), ac.im + R[f].im / T(3)});
          if (kept[j]) stst<true, T>(stin + o, base + R[f] * (a.stage == 1 ? half : dt));
        } else {
          const cx<T> ac = ldst<true, T>(acc + o);
          const cx<T> tot{ac.re + R[f].re / T(6), ac.im + R[f].im / T(6)};
          stst<true, T>(sol + o, base + tot * dt);
        }
      }
    } else {
      // LSRK54: S = A_i S + dt RHS ; sol += B_i S (on the dealiased sol)
      const T A = T(a.ca), B = T(a.cb);
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        const int64_t o = (int64_t)f * G.nc + q;
        cx<T> Sold = zero;
        if (a.stage != 0) Sold = ldst<true, T>(acc + o);
        const cx<T> Snew = Sold * A + R[f] * dt;
        stst<true, T>(acc + o, Snew);
        stst<true, T>(sol + o, s[j][f] + Snew * B);
      }
    }
  }
}

template <class T, int PHYS, int PLANES, int MODS>
__global__ void __launch_bounds__(256)
spectral_kernel(SpecArgs a, const RunBuf<T>* __restrict__ runs) {
  spectral_item<T, PHYS, PLANES, MODS>(a, runs, blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y, blockIdx.z);
}

// ---------------------------------------------------------------------------
// small kernels: CFL time step, maxima reduction + clock advance, mask
// ---------------------------------------------------------------------------
// getCFL! (integrator.jl:158-198): vmax = sqrt(max u^2) in T, dt = min(Coef*dl/vmax, t_diff)
template <class T>
__global__ void cfl_kernel(const RunBuf<T>* __restrict__ runs, int nruns, int phys) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nruns) return;
  RunClock* c = runs[r].clk;
  if (!c->use_cfl || c->step >= c->target) return;
  T v = sqrt(T(c->maxu2));
  if (phys == PHYS_MHD) {
    const T va = sqrt(T(c->maxb2));
    v = (va > v || va != va) ? va : v;
  }
  const double cfl = c->cfl_coef * c->dl / (double)v;
  const double dt = (cfl != cfl) ? cfl : (cfl < c->t_diff ? cfl : c->t_diff);
  c->dt = (double)T(dt);
}

// reduce per-block maxima into the clock; optionally advance t and step
template <class T>
__global__ void __launch_bounds__(256)
reduce_max_kernel(const RunBuf<T>* __restrict__ runs, int64_t nblocks, int advance) {
  __shared__ T red[2][256];
  const RunBuf<T>& rb = runs[blockIdx.x];
  if (gated_off(advance, rb)) return;
  const int t = threadIdx.x;
  // eight independent accumulators keep eight loads in flight per thread (the maximum,
  // NaN included, does not depend on the order)
  constexpr int K = 8;
  T mus[K], mbs[K];
#pragma unroll
  for (int k = 0; k < K; ++k) mus[k] = mbs[k] = T(0);
  int64_t i = t;
  for (; i + (K - 1) * 256 < nblocks; i += K * 256) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      mus[k] = nanmax(mus[k], rb.partial[2 * (i + k * 256)]);
      mbs[k] = nanmax(mbs[k], rb.partial[2 * (i + k * 256) + 1]);
    }
  }
  for (; i < nblocks; i += 256) {
    mus[0] = nanmax(mus[0], rb.partial[2 * i]);
    mbs[0] = nanmax(mbs[0], rb.partial[2 * i + 1]);
  }
  T mu = mus[0], mb = mbs[0];
#pragma unroll
  for (int k = 1; k < K; ++k) {
    mu = nanmax(mu, mus[k]);
    mb = nanmax(mb, mbs[k]);
  }
  red[0][t] = mu;
  red[1][t] = mb;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) {
      red[0][t] = nanmax(red[0][t], red[0][t + s]);
      red[1][t] = nanmax(red[1][t], red[1][t + s]);
    }
    __syncthreads();
  }
  if (t == 0) {
    RunClock* c = rb.clk;
    c->maxu2 = (double)red[0][0];
    c->maxb2 = (double)red[1][0];
    c->nan_flag = (red[0][0] != red[0][0] || red[1][0] != red[1][0]) ? 1 : 0;
    if (advance) {
      c->t = (double)(T(c->t) + T(c->dt));
      c->step += 1;
    }
  }
}

// negative damping: A = P / (sum |u_i^2 f_i| dV) of the current right-hand side from the row
// kernel's per-block Float64 sums (fixed summation order), one workgroup per run
template <class T>
__global__ void __launch_bounds__(256)
nd_amp_kernel(const RunBuf<T>* __restrict__ runs, int64_t nblocks, double dV, int gate) {
  __shared__ double red[256];
  const RunBuf<T>& rb = runs[blockIdx.x];
  if (!rb.ndf || gated_off(gate, rb)) return;
  const int t = threadIdx.x;
  double s = 0.0;
  for (int64_t i = t; i < nblocks; i += 256) s += rb.ndpart[i];
  red[t] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (t < st) red[t] += red[t + st];
    __syncthreads();
  }
  if (t == 0) rb.clk->nd_amp = rb.nd_P / (red[0] * dV);
}

// per-block max of u^2 and b^2 straight over vars (after the initial condition is set)
template <class T>
__global__ void __launch_bounds__(256)
vars_max_kernel(const RunBuf<T>* __restrict__ runs, int64_t nr, int nf) {
  __shared__ T red[2][256];
  const RunBuf<T>& rb = runs[blockIdx.y];
  const int t = threadIdx.x;
  T mu = T(0), mb = T(0);
  for (int64_t i = (int64_t)blockIdx.x * 256 + t; i < nr; i += (int64_t)gridDim.x * 256) {
    for (int f = 0; f < nf; ++f) {
      const T v = rb.vars[(int64_t)f * nr + i];
      if (f < 3) mu = nanmax(mu, v * v);
      else mb = nanmax(mb, v * v);
    }
  }
  red[0][t] = mu;
  red[1][t] = mb;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) {
      red[0][t] = nanmax(red[0][t], red[0][t + s]);
      red[1][t] = nanmax(red[1][t], red[1][t + s]);
    }
    __syncthreads();
  }
  if (t == 0) {
    rb.partial[2 * blockIdx.x] = red[0][0];
    rb.partial[2 * blockIdx.x + 1] = red[1][0];
  }
}

template <class T>
__global__ void advance_clock_kernel(const RunBuf<T>* __restrict__ runs, int nruns) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nruns) return;
  RunClock* c = runs[r].clk;
  if (c->step >= c->target) return;
  c->t = (double)(T(c->t) + T(c->dt));
  c->step += 1;
}

// Before a call that advances the runs by nsteps steps: the step that ends the call (or
// reaches a run's target) is computed in full, the earlier ones are pruned (step_pruned).
template <class T>
__global__ void prune_mark_kernel(const RunBuf<T>* __restrict__ runs, int nruns, int64_t nsteps, int enable) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nruns) return;
  RunClock* c = runs[r].clk;
  const int64_t end = c->step + nsteps;
  c->prune_until = enable ? (end < c->target ? end : c->target) : 0;
}

// FP64 sums over the real fields: [0] sum |u|^2, [1] sum |b|^2, [2] sum u.b,
// [3] max |f0|, per block; final pass sums in a fixed order (deterministic).
template <class T>
__global__ void __launch_bounds__(256)
field_sums_kernel(const T* __restrict__ vars, int64_t nr, int nf, double* __restrict__ part) {
  __shared__ double red[3][256];
  const int t = threadIdx.x;
  double su = 0, sb = 0, sx = 0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + t; i < nr; i += (int64_t)gridDim.x * 256) {
    const double u0 = vars[i], u1 = vars[nr + i], u2 = vars[2 * nr + i];
    su += u0 * u0 + u1 * u1 + u2 * u2;
    if (nf == 6) {
      const double b0 = vars[3 * nr + i], b1 = vars[4 * nr + i], b2 = vars[5 * nr + i];
      sb += b0 * b0 + b1 * b1 + b2 * b2;
      sx += u0 * b0 + u1 * b1 + u2 * b2;
    }
  }
  red[0][t] = su; red[1][t] = sb; red[2][t] = sx;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) {
      red[0][t] += red[0][t + s];
      red[1][t] += red[1][t + s];
      red[2][t] += red[2][t + s];
    }
    __syncthreads();
  }
  if (t == 0) {
    part[3 * blockIdx.x] = red[0][0];
    part[3 * blockIdx.x + 1] = red[1][0];
    part[3 * blockIdx.x + 2] = red[2][0];
  }
}

// FP64 sum of products of pairs of real fields a_c * b_c (c < 3) and max |a0|:
// used for helicities (u.curl u, A.b) and max |div b|.
template <class T>
__global__ void __launch_bounds__(256)
dot_max_kernel(const T* __restrict__ a, const T* __restrict__ b, int64_t nr, int ncomp, double* __restrict__ part) {
  __shared__ double red[2][256];
  const int t = threadIdx.x;
  double s = 0, mx = 0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + t; i < nr; i += (int64_t)gridDim.x * 256) {
    if (b) {
      for (int c = 0; c < ncomp; ++c) s += (double)a[c * nr + i] * (double)b[c * nr + i];
    } else {
      const double v = fabs((double)a[i]);
      mx = (v > mx || v != v) ? v : mx;
    }
  }
  red[0][t] = s; red[1][t] = mx;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (t < st) {
      red[0][t] += red[0][t + st];
      const double o = red[1][t + st];
      red[1][t] = (o > red[1][t] || o != o) ? o : red[1][t];
    }
    __syncthreads();
  }
  if (t == 0) {
    part[2 * blockIdx.x] = red[0][0];
    part[2 * blockIdx.x + 1] = red[1][0];
  }
}

// Spectral operators for the diagnostics on a 3-component half spectrum X
// (fields 0..2 of src) written to dst fields 0..2 (curl / vector potential) or
// field 0 (divergence).  Wavenumbers in FP64.
// OP_CURL / OP_VECPOT / OP_DIV: diagnostics in Float64 wavenumbers.  OP_DX..OP_DZ, OP_INVLAP and
// OP_CURL_T / OP_DIV_T / OP_VECPOT_T: the reference's VectorCalculus.jl / MHDAnalysis.jl
// functions, evaluated in T with the grid's wavenumber arrays in T as those functions do.
enum DiagOp { OP_CURL = 0, OP_VECPOT = 1, OP_DIV = 2, OP_DX = 3, OP_DY = 4, OP_DZ = 5, OP_INVLAP = 6,
              OP_CURL_T = 7, OP_DIV_T = 8, OP_VECPOT_T = 9 };
// src and dst are fields of the work buffer (padded planes: field stride wnc, plane stride wps).
template <class T>
__global__ void __launch_bounds__(256)
diag_spectral_kernel(Geom G, const cx<T>* __restrict__ src, cx<T>* __restrict__ dst, int op) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= G.nc) return;
  const int kx = (int)(i % G.nkr);
  const int64_t rest = i / G.nkr;
  const int l = (int)(rest % G.ny), m = (int)(rest / G.ny);
  const int64_t idx = (int64_t)m * G.wps + (int64_t)l * G.wrs + kx;
  const double k0 = kx * G.multx, k1 = signed_index(l, G.ny) * G.multy, k2 = signed_index(m, G.nz) * G.multz;
  const cx<T> a0 = src[idx], a1 = src[G.wnc + idx], a2 = src[2 * G.wnc + idx];
  auto ik = [](double kk, cx<T> z) { return cx<T>{T(-kk * (double)z.im), T(kk * (double)z.re)}; };
  if (op >= OP_DX) {
    const ModeK<T> q = mode_k<T>(G, kx, l, m);
    // im * z: (-z.im, z.re) exactly
    auto im = [](cx<T> z) { return cx<T>{-z.im, z.re}; };
    auto sc = [](T c, cx<T> z) { return cx<T>{c * z.re, c * z.im}; };
    if (op == OP_DX || op == OP_DY || op == OP_DZ) {   // ∂i: im*k_i*B1h (VectorCalculus.jl:101-123)
      const T kk = op == OP_DX ? q.kx : (op == OP_DY ? q.ky : q.kz);
      dst[idx] = im(sc(kk, a0));
    } else if (op == OP_INVLAP) {                      // LaplaceSolver: Bh *= k^-2 (:155-168)
      dst[idx] = sc(q.invK, a0);
    } else if (op == OP_DIV_T) {                       // Div: im*(kx*B1h + ky*B2h + kz*B3h) (:81-92)
      dst[idx] = im(sc(q.kx, a0) + sc(q.ky, a1) + sc(q.kz, a2));
    } else {                                            // Curl (:27-50) / VectorPotential (MHDAnalysis.jl:156-180)
      cx<T> c0 = im(sc(q.ky, a2) - sc(q.kz, a1));
      cx<T> c1 = im(sc(q.kz, a0) - sc(q.kx, a2));
      cx<T> c2 = im(sc(q.kx, a1) - sc(q.ky, a0));
      if (op == OP_VECPOT_T) {
        c0 = sc(q.invK, c0);
        c1 = sc(q.invK, c1);
        c2 = sc(q.invK, c2);
      }
      dst[idx] = c0; dst[G.wnc + idx] = c1; dst[2 * G.wnc + idx] = c2;
    }
    return;
  }
  if (op == OP_DIV) {
    const cx<T> d = ik(k0, a0) + ik(k1, a1) + ik(k2, a2);
    dst[idx] = d;
    return;
  }
  double s = 1.0;
  if (op == OP_VECPOT) {
    const double K = k0 * k0 + k1 * k1 + k2 * k2;
    s = K == 0 ? 0.0 : 1.0 / K;
  }
  const cx<T> c0 = (ik(k1, a2) - ik(k2, a1)) * T(s);
  const cx<T> c1 = (ik(k2, a0) - ik(k0, a2)) * T(s);
  const cx<T> c2 = (ik(k0, a1) - ik(k1, a0)) * T(s);
  dst[idx] = c0; dst[G.wnc + idx] = c1; dst[2 * G.wnc + idx] = c2;
}

// ---------------------------------------------------------------------------
// analysis (src/utils/MHDAnalysis.jl): shell spectra and band filters on the work buffer
// ---------------------------------------------------------------------------
// |k| of a mode in T, as the reference's grid arrays give it: sqrt((kx^2 + ky^2) + kz^2)
template <class T>
__device__ __forceinline__ T mode_kabs(const Geom& G, int kx, int l, int m) {
  const T qx = T((double)kx * G.multx), qy = T((double)signed_index(l, G.ny) * G.multy);
  const T qz = T((double)signed_index(m, G.nz) * G.multz);
  return sqrt((qx * qx + qy * qy) + qz * qz);
}

// spectralline (MHDAnalysis.jl:236-255): shell r = round(|k|) (half to even) of every half-
// spectrum mode of ncomp work fields; out[c * nbins + r] += |A_k|^2 and out[ncomp * nbins + r]
// counts the modes.  Accumulated in Float64 through LDS and global atomics (the summation
// order, hence the last bits, is not fixed).  LDS: (ncomp + 1) * nbins doubles.
template <class T>
__global__ void __launch_bounds__(256)
shell_kernel(Geom G, const cx<T>* __restrict__ work, int ncomp, double* __restrict__ out, int nbins) {
  extern __shared__ double hist[];
  const int t = threadIdx.x;
  const int nh = (ncomp + 1) * nbins;
  for (int i = t; i < nh; i += 256) hist[i] = 0.0;
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * 256 + t; i < G.nc; i += (int64_t)gridDim.x * 256) {
    const int kx = (int)(i % G.nkr);
    const int64_t rest = i / G.nkr;
    const int l = (int)(rest % G.ny), m = (int)(rest / G.ny);
    const T k = mode_kabs<T>(G, kx, l, m);
    const int r = (int)rint((double)k);
    if (r < 0 || r >= nbins) continue;
    const int64_t idx = (int64_t)m * G.wps + (int64_t)l * G.wrs + kx;
    for (int c = 0; c < ncomp; ++c) {
      const cx<T> v = work[(int64_t)c * G.wnc + idx];
      atomicAdd(&hist[c * nbins + r], (double)v.re * (double)v.re + (double)v.im * (double)v.im);
    }
    atomicAdd(&hist[ncomp * nbins + r], 1.0);
  }
  __syncthreads();
  for (int i = t; i < nh; i += 256)
    if (hist[i] != 0.0) atomicAdd(&out[i], hist[i]);
}

// ScaleDecomposition (MHDAnalysis.jl:55-82): keep the modes with k1 <= |k| <= k2 of ncomp work
// fields (|k| in T), zero the rest, in place
template <class T>
__global__ void __launch_bounds__(256)
band_kernel(Geom G, cx<T>* __restrict__ work, int ncomp, double k1, double k2) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= G.nc) return;
  const int kx = (int)(i % G.nkr);
  const int64_t rest = i / G.nkr;
  const int l = (int)(rest % G.ny), m = (int)(rest / G.ny);
  const T k = mode_kabs<T>(G, kx, l, m);
  if ((double)k >= k1 && (double)k <= k2) return;
  const int64_t idx = (int64_t)m * G.wps + (int64_t)l * G.wrs + kx;
  for (int c = 0; c < ncomp; ++c) work[(int64_t)c * G.wnc + idx] = cx<T>{T(0), T(0)};
}

// DivVCorrection! / DivBCorrection! (src/Solver/VPSolver.jl:60-139) on three compact spectral
// fields: phi = -i (k . F) k^-2, F -= i k phi, in T in the reference's operation order.
template <class T>
__global__ void __launch_bounds__(256)
div_correction_kernel(Geom G, cx<T>* __restrict__ f) {
  const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= G.nc) return;
  const int kx = (int)(q % G.nkr);
  const int64_t rest = q / G.nkr;
  const int l = (int)(rest % G.ny), m = (int)(rest / G.ny);
  const ModeK<T> k = mode_k<T>(G, kx, l, m);
  cx<T> a = f[q], b = f[G.nc + q], c = f[2 * G.nc + q];
  // sum = kx a + ky b + kz c; phi = -im * sum; phi = phi * invKrsq
  const cx<T> sum{(k.kx * a.re + k.ky * b.re) + k.kz * c.re, (k.kx * a.im + k.ky * b.im) + k.kz * c.im};
  const cx<T> phi{sum.im * k.invK, -sum.re * k.invK};
  // F -= (im k) phi = (-k phi.im, k phi.re)
  a = cx<T>{a.re + k.kx * phi.im, a.im - k.kx * phi.re};
  b = cx<T>{b.re + k.ky * phi.im, b.im - k.ky * phi.re};
  c = cx<T>{c.re + k.kz * phi.im, c.im - k.kz * phi.re};
  f[q] = a;
  f[G.nc + q] = b;
  f[2 * G.nc + q] = c;
}

// Structure functions (src/utils/TurbStatTool.jl:67-160): SFC(V) = 2 (mean(V) - CF(V)) with
// CF(V) = real(ifft(|fft(V)|^2)), summed over the three components.  The power of the three
// transformed components goes into work field 0 (|X_c|^2 in T, summed in T: the transform is
// linear, so one inverse transform of the sum gives the sum of the CFs), and mean(V_c) =
// X_c(0) / N into means[c] (T values held in Float64).
template <class T>
__global__ void __launch_bounds__(256)
sf_power_kernel(Geom G, cx<T>* __restrict__ work, double* __restrict__ means) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= G.nc) return;
  const int kx = (int)(i % G.nkr);
  const int64_t rest = i / G.nkr;
  const int l = (int)(rest % G.ny), m = (int)(rest / G.ny);
  const int64_t idx = (int64_t)m * G.wps + (int64_t)l * G.wrs + kx;
  T p = T(0);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const cx<T> v = work[(int64_t)c * G.wnc + idx];
    p += v.re * v.re + v.im * v.im;
    if (i == 0) means[c] = (double)(v.re * T(G.inv_n));
  }
  work[idx] = cx<T>{p, T(0)};
}

// Binning of SFV(R) = sum_c 2 (mean(V_c) - CF_c(R)) over the lags R of the fftshifted cube, as
// the reference indexes it: the element at fftshifted 1-based position i has idx = i - nx/2,
// i.e. idx = s + 1 for the signed lag s in [-nx/2, nx/2) (likewise jdx, kdx).
//   radial (SF2 1D):    kk = round(|(idx, jdx, kdx)|), kk > 0 -> bin kk - 1 of nb
//   global (SFr_Global): kk = |(idx, jdx, kdx)| > 0, theta = angle(mean b, (idx, jdx, kdx)) in T,
//                         bin (round(|kk cos theta|), round(|kk sin theta|)), row length nbx
// hist[0, nb) sums SFV, hist[nb, 2 nb) counts.  Rounding is half to even (Julia round).
template <class T, bool GLOBAL>
__global__ void __launch_bounds__(256)
sf_bin_kernel(Geom G, const T* __restrict__ cf, double summean2, double mbx, double mby, double mbz,
              double* __restrict__ hist, int nb, int nbx) {
  extern __shared__ double lh[];  // radial: 2 nb doubles
  const int t = threadIdx.x;
  if (!GLOBAL) {
    for (int i = t; i < 2 * nb; i += 256) lh[i] = 0.0;
    __syncthreads();
  }
  const T bx = T(mbx), by = T(mby), bz = T(mbz);
  const T sb = bx * bx + by * by + bz * bz;
  for (int64_t p = (int64_t)blockIdx.x * 256 + t; p < G.nr; p += (int64_t)gridDim.x * 256) {
    const int px = (int)(p % G.nx);
    const int64_t rest = p / G.nx;
    const int py = (int)(rest % G.ny), pz = (int)(rest / G.ny);
    const int idx = (px < G.nx / 2 ? px : px - G.nx) + 1;
    const int jdx = (py < G.ny / 2 ? py : py - G.ny) + 1;
    const int kdx = (pz < G.nz / 2 ? pz : pz - G.nz) + 1;
    const int64_t ss = (int64_t)idx * idx + (int64_t)jdx * jdx + (int64_t)kdx * kdx;
    if (ss == 0) continue;
    const double kk = sqrt((double)ss);
    const T v = T(summean2) - T(2) * cf[p];
    int bin;
    if (!GLOBAL) {
      bin = (int)rint(kk) - 1;
    } else {
      T x = (bx * T(idx) + by * T(jdx) + bz * T(kdx)) / sqrt(sb * T(ss));
      x = x > T(1) ? T(1) : (x < T(-1) ? T(-1) : x);
      const T th = acos(x);
      const int rpar = (int)rint(fabs(kk * (double)cos(th)));
      const int rperp = (int)rint(fabs(kk * (double)sin(th)));
      bin = rperp * nbx + rpar;
    }
    if (bin < 0 || bin >= nb) continue;
    if (GLOBAL) {
      atomicAdd(&hist[bin], (double)v);
      atomicAdd(&hist[nb + bin], 1.0);
    } else {
      atomicAdd(&lh[bin], (double)v);
      atomicAdd(&lh[nb + bin], 1.0);
    }
  }
  if (!GLOBAL) {
    __syncthreads();
    for (int i = t; i < 2 * nb; i += 256)
      if (lh[i] != 0.0) atomicAdd(&hist[i], lh[i]);
  }
}

// scatter sparse half-spectrum entries (kx, l, m) of nf fields into x (already zeroed)
template <class T>
__global__ void __launch_bounds__(256)
scatter_modes_kernel(Geom G, cx<T>* __restrict__ x, int nf, int64_t nmodes, const int32_t* __restrict__ idx,
                     const cx<T>* __restrict__ vals) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= nmodes) return;
  const int kx = idx[3 * i], l = idx[3 * i + 1], m = idx[3 * i + 2];
  if (kx < 0 || kx >= G.nkr || l < 0 || l >= G.ny || m < 0 || m >= G.nz) return;
  const int64_t q = ((int64_t)m * G.ny + l) * G.nkr + kx;
  for (int f = 0; f < nf; ++f) x[(int64_t)f * G.nc + q] = vals[i * nf + f];
}

// gather sparse half-spectrum entries (kx, l, m) of nf fields of x into vals (field-fastest);
// entries outside the half spectrum read as zero
template <class T>
__global__ void __launch_bounds__(256)
gather_modes_kernel(Geom G, const cx<T>* __restrict__ x, int nf, int64_t nmodes, const int32_t* __restrict__ idx,
                    cx<T>* __restrict__ vals) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= nmodes) return;
  const int kx = idx[3 * i], l = idx[3 * i + 1], m = idx[3 * i + 2];
  const bool ok = kx >= 0 && kx < G.nkr && l >= 0 && l < G.ny && m >= 0 && m < G.nz;
  const int64_t q = ok ? ((int64_t)m * G.ny + l) * G.nkr + kx : 0;
  for (int f = 0; f < nf; ++f) vals[i * nf + f] = ok ? x[(int64_t)f * G.nc + q] : cx<T>{T(0), T(0)};
}

// copy one field of the work buffer (padded rows and planes) into the unpadded state layout
template <class T>
__global__ void __launch_bounds__(256)
unpad_kernel(Geom G, const cx<T>* __restrict__ src, cx<T>* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= G.nc) return;
  const int kx = (int)(i % G.nkr);
  const int64_t rest = i / G.nkr;
  const int l = (int)(rest % G.ny), m = (int)(rest / G.ny);
  dst[i] = src[(int64_t)m * G.wps + (int64_t)l * G.wrs + kx];
}

// zero n complex values (a fill kernel: usable where a memset node is not wanted)
template <class T>
__global__ void __launch_bounds__(256)
zero_kernel(cx<T>* __restrict__ x, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    x[i] = cx<T>{T(0), T(0)};
}

// zero the aliased band of nf fields in place (FourierFlows dealias!)
template <class T>
__global__ void __launch_bounds__(256)
dealias_kernel(Geom G, cx<T>* __restrict__ x, int nf) {
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= G.nc) return;
  const int kx = (int)(idx % G.nkr);
  const int64_t rest = idx / G.nkr;
  const int l = (int)(rest % G.ny), m = (int)(rest / G.ny);
  if (kx < G.kx_keep && kept_y(G, l) && kept_z(G, m)) return;
  for (int f = 0; f < nf; ++f) x[(int64_t)f * G.nc + idx] = cx<T>{T(0), T(0)};
}

}  // namespace mhd
