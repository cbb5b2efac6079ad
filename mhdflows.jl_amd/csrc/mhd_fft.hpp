// Register-resident Stockham FFT building blocks for gfx950 (CDNA4).
//
// A length-N transform is owned by TH = N/E threads, each holding E complex values
// in registers.  Pass p (radix R_p, Ns = R_0*...*R_{p-1}) has every thread run
// E/R_p butterflies b = t + i*TH; butterfly b reads positions b + r*(N/R_p),
// applies the twiddles w^(r*k), k = b mod Ns, runs a radix-R_p DFT and writes
// positions (b/Ns)*Ns*R_p + k + r*Ns.  Between passes the values go through LDS
// once; the first pass loads straight from the caller's source and the last pass
// leaves thread t holding positions t + i*TH + r*(N/R_last), so a following
// transform whose first radix equals R_last needs no exchange.
//
// Twiddles come from a table tw[m] = exp(-2*pi*i*m/N), m < N, generated on the
// host in extended precision and rounded to T; the inverse direction conjugates.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace mhd {

template <class T>
struct alignas(2 * sizeof(T)) cx {
  T re, im;
};

// HBM buffers reached through pointers stored in memory (per-run descriptors) are
// re-typed into the global address space: accesses then compile to global_* rather
// than flat_* instructions, which would also count against the LDS wait counter and
// serialise with LDS traffic.  (The host pass parses kernel bodies too; there the
// qualifier has no meaning and the pointers stay plain.)
#if defined(__HIP_DEVICE_COMPILE__)
template <class U> using gptr = __attribute__((address_space(1))) U*;
#else
template <class U> using gptr = U*;
#endif
template <class U> __device__ __forceinline__ gptr<U> gmem(U* p) { return (gptr<U>)p; }
template <class U> __device__ __forceinline__ U ldg(gptr<const U> p) { return *p; }

// Loads and stores of the hot kernels' streams (each element read or written once per
// launch).  NT = true adds the non-temporal hint.  Measured on MI355X: it speeds up the
// pointwise spectral launch (loads and stores) and the row loads, and slows the column
// passes; non-temporal row stores (1032 B rows, not line aligned) add partial-line writes.
template <bool NT, class T> __device__ __forceinline__ cx<T> ldst(gptr<const cx<T>> p) {
  if constexpr (NT) {
    typedef T v2 __attribute__((ext_vector_type(2)));
    const v2 v = __builtin_nontemporal_load((gptr<const v2>)p);
    return cx<T>{v.x, v.y};
  } else {
    return *p;
  }
}
template <bool NT, class T> __device__ __forceinline__ void stst(gptr<cx<T>> p, cx<T> x) {
  if constexpr (NT) {
    typedef T v2 __attribute__((ext_vector_type(2)));
    const v2 v = {x.re, x.im};
    __builtin_nontemporal_store(v, (gptr<v2>)p);
  } else {
    *p = x;
  }
}

template <class T> __device__ __forceinline__ cx<T> mk(T a, T b) { return cx<T>{a, b}; }
template <class T> __device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <class T> __device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <class T> __device__ __forceinline__ cx<T> operator*(cx<T> a, T s) { return {a.re * s, a.im * s}; }
template <class T> __device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <class T> __device__ __forceinline__ cx<T> cmulc(cx<T> a, cx<T> b) {  // a * conj(b)
  return {a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im};
}
template <class T> __device__ __forceinline__ cx<T> conj(cx<T> a) { return {a.re, -a.im}; }
// multiply by S*i (S = +1 or -1)
template <int S, class T> __device__ __forceinline__ cx<T> mul_si(cx<T> z) {
  return S > 0 ? cx<T>{-z.im, z.re} : cx<T>{z.im, -z.re};
}
// twiddle for direction S: table holds exp(-2 pi i m/N); S=+1 uses the conjugate
template <int S, class T> __device__ __forceinline__ cx<T> twmul(cx<T> a, cx<T> w) {
  return S < 0 ? cmul(a, w) : cmulc(a, w);
}

// ---------------------------------------------------------------------------
// radix-R DFTs, y_k = sum_j x_j exp(S*2*pi*i*j*k/R), in place on x[0..R)
// ---------------------------------------------------------------------------
template <class T, int R, int S> struct Bfly;

template <class T, int S> struct Bfly<T, 2, S> {
  static __device__ __forceinline__ void run(cx<T>* x) {
    cx<T> a = x[0], b = x[1];
    x[0] = a + b;
    x[1] = a - b;
  }
};

template <class T, int S> struct Bfly<T, 3, S> {
  static __device__ __forceinline__ void run(cx<T>* x) {
    const T c = T(0.86602540378443864676372317075293618);  // sqrt(3)/2
    cx<T> t1 = x[1] + x[2];
    cx<T> t2 = x[0] - t1 * T(0.5);
    cx<T> t3 = mul_si<S>((x[1] - x[2]) * c);
    x[0] = x[0] + t1;
    x[1] = t2 + t3;
    x[2] = t2 - t3;
  }
};

template <class T, int S> struct Bfly<T, 4, S> {
  static __device__ __forceinline__ void run(cx<T>* x) {
    cx<T> a = x[0] + x[2], b = x[0] - x[2];
    cx<T> c = x[1] + x[3], d = mul_si<S>(x[1] - x[3]);
    x[0] = a + c;
    x[2] = a - c;
    x[1] = b + d;
    x[3] = b - d;
  }
};

template <class T, int S> struct Bfly<T, 6, S> {
  static __device__ __forceinline__ void run(cx<T>* x) {
    // 6 = 3 x 2: A_j2[k1] = DFT3_j1(x[2 j1 + j2]); A_1[k1] *= w6^k1; y[k1 + 3 k2] = A_0 +- A_1
    const T c = T(0.86602540378443864676372317075293618);  // sqrt(3)/2
    cx<T> a[3] = {x[0], x[2], x[4]};
    cx<T> b[3] = {x[1], x[3], x[5]};
    Bfly<T, 3, S>::run(a);
    Bfly<T, 3, S>::run(b);
    const cx<T> w1 = {T(0.5), T(S) * c}, w2 = {T(-0.5), T(S) * c};
    b[1] = cmul(b[1], w1);
    b[2] = cmul(b[2], w2);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      x[k] = a[k] + b[k];
      x[k + 3] = a[k] - b[k];
    }
  }
};

template <class T, int S> struct Bfly<T, 8, S> {
  static __device__ __forceinline__ void run(cx<T>* x) {
    const T h = T(0.70710678118654752440084436210484904);  // sqrt(2)/2
    cx<T> e[4] = {x[0], x[2], x[4], x[6]};
    cx<T> o[4] = {x[1], x[3], x[5], x[7]};
    Bfly<T, 4, S>::run(e);
    Bfly<T, 4, S>::run(o);
    // o[k] *= w8^k, w8 = exp(S*2*pi*i/8) = h*(1 + S*i)
    cx<T> o1 = (o[1] + mul_si<S>(o[1])) * h;
    cx<T> o2 = mul_si<S>(o[2]);
    cx<T> o3 = (mul_si<S>(o[3]) - o[3]) * h;
    x[0] = e[0] + o[0];
    x[4] = e[0] - o[0];
    x[1] = e[1] + o1;
    x[5] = e[1] - o1;
    x[2] = e[2] + o2;
    x[6] = e[2] - o2;
    x[3] = e[3] + o3;
    x[7] = e[3] - o3;
  }
};

template <class T, int S> struct Bfly<T, 16, S> {
  static __device__ __forceinline__ void run(cx<T>* x) {
    // 16 = 4 x 4: A_j2[k1] = DFT4_j1(x[4 j1 + j2]); twiddle w16^(j2 k1); y[k1 + 4 k2] = DFT4_j2
    const T c1 = T(0.92387953251128675612818318939678829);  // cos(pi/8)
    const T s1 = T(0.38268343236508977172845998403039887);  // sin(pi/8)
    const T h = T(0.70710678118654752440084436210484904);
    cx<T> a[4][4];
#pragma unroll
    for (int j2 = 0; j2 < 4; ++j2) {
      cx<T> v[4] = {x[j2], x[4 + j2], x[8 + j2], x[12 + j2]};
      Bfly<T, 4, S>::run(v);
#pragma unroll
      for (int k1 = 0; k1 < 4; ++k1) a[j2][k1] = v[k1];
    }
    // twiddles w16^(j2*k1) with w16 = exp(S*2*pi*i/16) = c1 + S*i*s1
    const T sg = T(S);
    const cx<T> w1 = {c1, sg * s1}, w2 = {h, sg * h}, w3 = {s1, sg * c1};
    const cx<T> w4 = {T(0), sg}, w6 = {-h, sg * h}, w9 = {-c1, -sg * s1};
    a[1][1] = cmul(a[1][1], w1);
    a[1][2] = cmul(a[1][2], w2);
    a[1][3] = cmul(a[1][3], w3);
    a[2][1] = cmul(a[2][1], w2);
    a[2][2] = cmul(a[2][2], w4);
    a[2][3] = cmul(a[2][3], w6);
    a[3][1] = cmul(a[3][1], w3);
    a[3][2] = cmul(a[3][2], w6);
    a[3][3] = cmul(a[3][3], w9);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      cx<T> v[4] = {a[0][k1], a[1][k1], a[2][k1], a[3][k1]};
      Bfly<T, 4, S>::run(v);
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) x[k1 + 4 * k2] = v[k2];
    }
  }
};

// ---------------------------------------------------------------------------
// plans: elements per thread E and radix sequence, per length N
// ---------------------------------------------------------------------------
template <int... Rs> struct Radices {};

template <int N> struct Plan;  // column transforms (y/z passes)
#define MHD_PLAN(N_, E_, ...)                \
  template <> struct Plan<N_> {              \
    static constexpr int E = E_;             \
    using R = Radices<__VA_ARGS__>;          \
  };
MHD_PLAN(8, 8, 8)
MHD_PLAN(12, 12, 4, 3)
MHD_PLAN(16, 8, 8, 2)
MHD_PLAN(24, 12, 4, 3, 2)
MHD_PLAN(32, 8, 8, 4)
MHD_PLAN(48, 12, 4, 4, 3)
MHD_PLAN(64, 8, 8, 8)
MHD_PLAN(96, 12, 4, 4, 2, 3)
MHD_PLAN(128, 16, 16, 8)
MHD_PLAN(192, 12, 4, 4, 4, 3)
MHD_PLAN(256, 16, 16, 16)
MHD_PLAN(384, 12, 4, 4, 4, 2, 3)
MHD_PLAN(512, 8, 8, 8, 8)
MHD_PLAN(768, 12, 4, 4, 4, 4, 3)
MHD_PLAN(1024, 16, 16, 16, 4)
#undef MHD_PLAN

// Row transforms inside the fused x kernel hold six fields per thread, so they
// use fewer elements per thread.  Inverse radices run in the listed order; the
// forward transform of the products runs them reversed, so its first radix is
// the inverse's last one and the thread already owns its input positions.
template <int M> struct RowPlan;
#define MHD_ROWPLAN(M_, E_, ...)             \
  template <> struct RowPlan<M_> {           \
    static constexpr int E = E_;             \
    using R = Radices<__VA_ARGS__>;          \
  };
MHD_ROWPLAN(4, 4, 4)
MHD_ROWPLAN(8, 8, 8)
MHD_ROWPLAN(12, 6, 6, 2)
MHD_ROWPLAN(16, 8, 8, 2)
MHD_ROWPLAN(24, 6, 6, 2, 2)
MHD_ROWPLAN(32, 8, 8, 4)
MHD_ROWPLAN(48, 6, 6, 2, 2, 2)
MHD_ROWPLAN(64, 8, 8, 8)
MHD_ROWPLAN(96, 6, 6, 2, 2, 2, 2)
MHD_ROWPLAN(128, 8, 8, 8, 2)
MHD_ROWPLAN(192, 6, 6, 2, 2, 2, 2, 2)
MHD_ROWPLAN(256, 8, 8, 8, 4)
MHD_ROWPLAN(384, 6, 6, 2, 2, 2, 2, 2, 2)
MHD_ROWPLAN(512, 8, 8, 8, 8)
#undef MHD_ROWPLAN

// Float64 rows keep four (powers of two) or six complex values per field per
// thread, so six resident fields stay inside the architectural VGPRs.
template <class T, int M> struct RowPlanT : RowPlan<M> {};
#define MHD_ROWPLAN_D(M_, E_, ...)                   \
  template <> struct RowPlanT<double, M_> {          \
    static constexpr int E = E_;                     \
    using R = Radices<__VA_ARGS__>;                  \
  };
MHD_ROWPLAN_D(8, 4, 4, 2)
MHD_ROWPLAN_D(16, 4, 4, 4)
MHD_ROWPLAN_D(32, 4, 4, 4, 2)
MHD_ROWPLAN_D(64, 4, 4, 4, 4)
MHD_ROWPLAN_D(128, 4, 4, 4, 4, 2)
MHD_ROWPLAN_D(256, 4, 4, 4, 4, 4)
MHD_ROWPLAN_D(512, 4, 4, 4, 4, 4, 2)
#undef MHD_ROWPLAN_D

// Tunable row plans: variant 1 of Float32 power-of-two rows uses four elements
// per thread (twice the threads, half the registers); other cases have one plan.
template <class T, int M, int VAR> struct RowPlanV : RowPlanT<T, M> {};
template <int M> struct RowPlanV<float, M, 1> : RowPlanT<double, M> {};

// elements per thread of RowPlanV, for host-side launch shapes
inline int row_plan_E(int m, bool dbl, int var) {
  if (m == 4) return 4;
  if (m % 3 == 0) return 6;
  if (dbl) return 4;
  return var == 1 ? 4 : 8;
}

template <class L> struct RFirst;
template <int R0, int... Rs> struct RFirst<Radices<R0, Rs...>> { static constexpr int value = R0; };
template <class L> struct RLast;
template <int R0> struct RLast<Radices<R0>> { static constexpr int value = R0; };
template <int R0, int R1, int... Rs> struct RLast<Radices<R0, R1, Rs...>> {
  static constexpr int value = RLast<Radices<R1, Rs...>>::value;
};
template <class L, class Acc = Radices<>> struct RReverse;
template <int... As> struct RReverse<Radices<>, Radices<As...>> { using type = Radices<As...>; };
template <int R0, int... Rs, int... As> struct RReverse<Radices<R0, Rs...>, Radices<As...>> {
  using type = typename RReverse<Radices<Rs...>, Radices<R0, As...>>::type;
};

// ---------------------------------------------------------------------------
// Synchronisation of an LDS exchange.  WS = true when all threads of a transform
// belong to one wavefront: a wavefront executes its LDS operations in order, so a
// wavefront-scope release/acquire fence around a wave barrier (no cross-wave
// waiting) orders the exchange; otherwise a workgroup barrier is used.
// ---------------------------------------------------------------------------
template <bool WS>
__device__ __forceinline__ void xsync() {
  if constexpr (WS) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// The pass pipeline.  x[E] holds the thread's values in the input layout of the
// first pass on entry and in the output layout of the last pass on exit.
// Lds(pos) returns a reference to the exchange slot of position pos of this
// thread's transform.  A sync precedes every exchange write and follows every
// exchange read, so the same LDS area may be reused by the caller afterwards.
// ---------------------------------------------------------------------------
template <class T, int N, int TH, int E, int S, int Ns, bool WS, class Rad> struct Passes;

template <class T, int N, int TH, int E, int S, int Ns, bool WS> struct Passes<T, N, TH, E, S, Ns, WS, Radices<>> {
  template <class Lds> static __device__ __forceinline__ void run(cx<T>*, int, Lds&, const cx<T>*) {}
};

template <class T, int N, int TH, int E, int S, int Ns, bool WS, int R, int... Rest>
struct Passes<T, N, TH, E, S, Ns, WS, Radices<R, Rest...>> {
  template <class Lds>
  static __device__ __forceinline__ void run(cx<T>* x, int t, Lds& lds, const cx<T>* __restrict__ tw) {
    static_assert(E % R == 0, "radix must divide elements per thread");
    constexpr int NB = E / R;
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const int b = t + i * TH;
      cx<T>* v = x + i * R;
      if (Ns > 1) {
        const int k = b % Ns;
        const int m1 = k * (N / (Ns * R));
#pragma unroll
        for (int r = 1; r < R; ++r) v[r] = twmul<S>(v[r], tw[r * m1]);
      }
      Bfly<T, R, S>::run(v);
    }
    if constexpr (sizeof...(Rest) > 0) {
      constexpr int R2 = RFirst<Radices<Rest...>>::value;
      xsync<WS>();
#pragma unroll
      for (int i = 0; i < NB; ++i) {
        const int b = t + i * TH;
        const int j = (b / Ns) * Ns * R + b % Ns;
#pragma unroll
        for (int r = 0; r < R; ++r) lds(j + r * Ns) = x[i * R + r];
      }
      xsync<WS>();
#pragma unroll
      for (int i = 0; i < E / R2; ++i) {
#pragma unroll
        for (int r = 0; r < R2; ++r) x[i * R2 + r] = lds(t + i * TH + r * (N / R2));
      }
      Passes<T, N, TH, E, S, Ns * R, WS, Radices<Rest...>>::run(x, t, lds, tw);
    }
  }
};

// Whole transform: x in the pass-0 input layout, returns the last-pass output layout.
template <class T, int N, int E, int S, class Rad, bool WS = false, class Lds>
__device__ __forceinline__ void fft_regs(cx<T>* x, int t, Lds& lds, const cx<T>* __restrict__ tw) {
  Passes<T, N, N / E, E, S, 1, WS, Rad>::run(x, t, lds, tw);
}

}  // namespace mhd
