// Device-side FFT building blocks: radix-2/3/4/5/8 butterflies and a
// compile-time mixed-radix Stockham engine that keeps each thread's values in
// registers between passes.
//
// Register ownership: for a line of length N handled by TPL threads, thread tl
// holds elements tl + TPL*q (q = 0 .. N/TPL-1).  This ownership is both the
// input of the first Stockham pass and the output of the last one, so lines
// are loaded from and stored to global memory with lane-contiguous (coalesced)
// accesses; LDS is used only for the N-1 exchanges between passes.
#pragma once
#include "evx_common.hpp"

namespace evx {

template <typename T, bool INV>
__device__ __forceinline__ cpx<T> mul_mi(cpx<T> z) {  // times -i (forward) or +i (inverse)
  return INV ? cpx<T>{-z.y, z.x} : cpx<T>{z.y, -z.x};
}

template <typename T, bool INV>
__device__ __forceinline__ void dft2(cpx<T>* x) {
  const cpx<T> a = x[0], b = x[1];
  x[0] = cadd(a, b);
  x[1] = csub(a, b);
}

template <typename T, bool INV>
__device__ __forceinline__ void dft4(cpx<T>& a0, cpx<T>& a1, cpx<T>& a2, cpx<T>& a3) {
  const cpx<T> s0 = cadd(a0, a2), d0 = csub(a0, a2), s1 = cadd(a1, a3), d1 = mul_mi<T, INV>(csub(a1, a3));
  a0 = cadd(s0, s1);
  a2 = csub(s0, s1);
  a1 = cadd(d0, d1);
  a3 = csub(d0, d1);
}

template <typename T, bool INV>
__device__ __forceinline__ void dft3(cpx<T>* x) {
  const T h = T(0.86602540378443864676);  // sin(2 pi / 3)
  const cpx<T> a = cadd(x[1], x[2]), b = csub(x[1], x[2]);
  const cpx<T> y0 = cadd(x[0], a);
  const cpx<T> t = {x[0].x - T(0.5) * a.x, x[0].y - T(0.5) * a.y};
  const cpx<T> v = mul_mi<T, INV>(cscale(b, h));
  x[0] = y0;
  x[1] = cadd(t, v);
  x[2] = csub(t, v);
}

template <typename T, bool INV>
__device__ __forceinline__ void dft5(cpx<T>* x) {
  const T c1 = T(0.30901699437494742410), c2 = T(-0.80901699437494742410);  // cos(2pi/5), cos(4pi/5)
  const T s1 = T(0.95105651629515357212), s2 = T(0.58778525229247312917);   // sin(2pi/5), sin(4pi/5)
  const cpx<T> a1 = cadd(x[1], x[4]), b1 = csub(x[1], x[4]);
  const cpx<T> a2 = cadd(x[2], x[3]), b2 = csub(x[2], x[3]);
  const cpx<T> x0 = x[0];
  const cpx<T> t1 = {x0.x + c1 * a1.x + c2 * a2.x, x0.y + c1 * a1.y + c2 * a2.y};
  const cpx<T> t2 = {x0.x + c2 * a1.x + c1 * a2.x, x0.y + c2 * a1.y + c1 * a2.y};
  const cpx<T> u1 = mul_mi<T, INV>(cpx<T>{s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y});
  const cpx<T> u2 = mul_mi<T, INV>(cpx<T>{s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y});
  x[0] = cadd(x0, cadd(a1, a2));
  x[1] = cadd(t1, u1);
  x[4] = csub(t1, u1);
  x[2] = cadd(t2, u2);
  x[3] = csub(t2, u2);
}

template <typename T, bool INV>
__device__ __forceinline__ void dft8(cpx<T>* x) {
  const T h = T(0.70710678118654752440);  // sqrt(2)/2
  cpx<T> a0 = cadd(x[0], x[4]), b0 = csub(x[0], x[4]);
  cpx<T> a1 = cadd(x[1], x[5]), b1 = csub(x[1], x[5]);
  cpx<T> a2 = cadd(x[2], x[6]), b2 = csub(x[2], x[6]);
  cpx<T> a3 = cadd(x[3], x[7]), b3 = csub(x[3], x[7]);
  // b_r *= w8^r (forward w8 = e^{-i pi/4}; inverse uses the conjugate)
  b1 = INV ? cpx<T>{(b1.x - b1.y) * h, (b1.y + b1.x) * h} : cpx<T>{(b1.x + b1.y) * h, (b1.y - b1.x) * h};
  b2 = mul_mi<T, INV>(b2);
  b3 = INV ? cpx<T>{(-b3.x - b3.y) * h, (b3.x - b3.y) * h} : cpx<T>{(b3.y - b3.x) * h, (-b3.x - b3.y) * h};
  dft4<T, INV>(a0, a1, a2, a3);
  dft4<T, INV>(b0, b1, b2, b3);
  x[0] = a0; x[2] = a1; x[4] = a2; x[6] = a3;
  x[1] = b0; x[3] = b1; x[5] = b2; x[7] = b3;
}

template <typename T, bool INV, int R>
__device__ __forceinline__ void dft_fixed(cpx<T>* x) {
  if constexpr (R == 2) dft2<T, INV>(x);
  else if constexpr (R == 3) dft3<T, INV>(x);
  else if constexpr (R == 4) dft4<T, INV>(x[0], x[1], x[2], x[3]);
  else if constexpr (R == 5) dft5<T, INV>(x);
  else dft8<T, INV>(x);
}

// --------------------------------------------------------------------------
// Compile-time engine.  Lay::at(e) returns the LDS slot of element e of the
// calling thread's line; the engine writes pass outputs there, synchronises,
// and reads them back in register ownership order.  A layout with SPLIT set
// holds one real number per element: the exchange then runs twice (real
// parts, then imaginary parts) through half the LDS.
// --------------------------------------------------------------------------
template <int... R>
struct Rad {};

// Synchronisation of an exchange.  When every line's threads lie inside one
// wavefront (TPL divides 64 and a line's threads are contiguous), the
// exchange is private to that wavefront: a wavefront-scope fence and a
// code-motion barrier order its LDS writes before its reads, with no
// workgroup barrier.  Otherwise the whole workgroup synchronises.
template <bool WAVE>
__device__ __forceinline__ void xsync() {
  if constexpr (WAVE) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// Twiddles of a pass come from the plain table (PT < 0: tw[j] = w_N^j, read
// at r*k*N/(P*R)) or from per-pass tables (PT >= 0: pass tables laid out
// pass after pass from offset 0, entry (r-1)*P + k of a pass = w_N^(r*k*N/(P*R)),
// filled by PassTw), where lanes with consecutive k read consecutive entries
// instead of entries r*k*N/(P*R) apart (bank conflicts).
template <typename T, int N, int TPL, bool INV, int P, int PT, typename Lay, int R, int... Rest>
__device__ __forceinline__ void stockham(cpx<T> (&v)[N / TPL], const Lay& lay, int tl, const cpx<T>* __restrict__ tw) {
  constexpr int E = N / TPL;
  constexpr int S = E / R;
  constexpr int STR = N / (P * R);
  static_assert(E % R == 0, "each thread must own whole butterflies");
  cpx<T> y[S][R];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    cpx<T> x[R];
#pragma unroll
    for (int r = 0; r < R; ++r) x[r] = v[s + r * S];
    if constexpr (P > 1) {
      const int k = (tl + s * TPL) % P;
#pragma unroll
      for (int r = 1; r < R; ++r) {
        const cpx<T> w = PT < 0 ? tw[r * k * STR] : tw[(PT < 0 ? 0 : PT) + (r - 1) * P + k];
        x[r] = INV ? cmulc(x[r], w) : cmul(x[r], w);
      }
    }
    dft_fixed<T, INV, R>(x);
#pragma unroll
    for (int r = 0; r < R; ++r) y[s][r] = x[r];
  }
  if constexpr (sizeof...(Rest) == 0) {
    // Last pass: output slot i + r*P with P = N/R is register order s + r*S.
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
      for (int r = 0; r < R; ++r) v[s + r * S] = y[s][r];
  } else {
    if constexpr (Lay::SPLIT) {
      // Half-size exchange buffer: real parts, then imaginary parts.
      xsync<Lay::WAVE>();  // previous readers of the exchange buffer are done
#pragma unroll
      for (int s = 0; s < S; ++s) {
        const int i = tl + s * TPL;
        const int k = i % P;
        const int j = (i - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) *lay.at(j + r * P) = y[s][r].x;
      }
      xsync<Lay::WAVE>();
#pragma unroll
      for (int q = 0; q < E; ++q) v[q].x = *lay.template at_q<TPL>(tl, q);
      xsync<Lay::WAVE>();
#pragma unroll
      for (int s = 0; s < S; ++s) {
        const int i = tl + s * TPL;
        const int k = i % P;
        const int j = (i - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) *lay.at(j + r * P) = y[s][r].y;
      }
      xsync<Lay::WAVE>();
#pragma unroll
      for (int q = 0; q < E; ++q) v[q].y = *lay.template at_q<TPL>(tl, q);
    } else {
      xsync<Lay::WAVE>();  // previous readers of the exchange buffer are done
#pragma unroll
      for (int s = 0; s < S; ++s) {
        const int i = tl + s * TPL;
        const int k = i % P;
        const int j = (i - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) *lay.at(j + r * P) = y[s][r];
      }
      xsync<Lay::WAVE>();
#pragma unroll
      for (int q = 0; q < E; ++q) v[q] = *lay.template at_q<TPL>(tl, q);
    }
    stockham<T, N, TPL, INV, P * R, (PT < 0 ? -1 : PT + (P > 1 ? (R - 1) * P : 0)), Lay, Rest...>(v, lay, tl, tw);
  }
}

template <typename T, int N, int TPL, bool INV, typename Lay, int... R>
__device__ __forceinline__ void fft_regs(cpx<T> (&v)[N / TPL], const Lay& lay, int tl, const cpx<T>* tw, Rad<R...>) {
  stockham<T, N, TPL, INV, 1, -1, Lay, R...>(v, lay, tl, tw);
}
// The same with the per-pass twiddle tables of PassTw<N, 1, R...> at tw.
template <typename T, int N, int TPL, bool INV, typename Lay, int... R>
__device__ __forceinline__ void fft_regs_pt(cpx<T> (&v)[N / TPL], const Lay& lay, int tl, const cpx<T>* tw,
                                            Rad<R...>) {
  stockham<T, N, TPL, INV, 1, 0, Lay, R...>(v, lay, tl, tw);
}

// Per-pass twiddle tables of an N-point transform with radices R...: the
// pass of stride P and radix R holds (R-1)*P entries (none for P = 1).
template <int N, int P, int... R>
struct PassTw;
template <int N, int P>
struct PassTw<N, P> {
  static constexpr int SIZE = 0;
  template <typename T>
  __device__ static void fill(cpx<T>*, const cpx<T>*, int, int) {}
};
template <int N, int P, int R, int... Rest>
struct PassTw<N, P, R, Rest...> {
  static constexpr int HERE = P > 1 ? (R - 1) * P : 0;
  static constexpr int SIZE = HERE + PassTw<N, P * R, Rest...>::SIZE;
  // dst[i] from the plain table twg (twg[j] = w_N^j), threads tid of nt
  template <typename T>
  __device__ static void fill(cpx<T>* dst, const cpx<T>* twg, int tid, int nt) {
    constexpr int STR = N / (P * R);
    for (int i = tid; i < HERE; i += nt) {
      const int r = 1 + i / P, k = i % P;
      dst[i] = twg[r * k * STR];
    }
    PassTw<N, P * R, Rest...>::template fill<T>(dst + HERE, twg, tid, nt);
  }
};
template <int N, int... R>
constexpr int pass_tw_size(Rad<R...>) { return PassTw<N, 1, R...>::SIZE; }
template <typename T, int N, int... R>
__device__ __forceinline__ void pass_tw_fill(cpx<T>* dst, const cpx<T>* twg, int tid, int nt, Rad<R...>) {
  PassTw<N, 1, R...>::template fill<T>(dst, twg, tid, nt);
}

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X; remap
// the block index so that each XCD (and its L2) owns a contiguous range of
// tiles, keeping halos and shared cache lines inside one L2.
__host__ __device__ __forceinline__ unsigned xcd_tile(unsigned b, unsigned nb) {
  const unsigned xcd = b & 7u, local = b >> 3;
  const unsigned per = nb >> 3, rem = nb & 7u;
  return xcd * per + (xcd < rem ? xcd : rem) + local;
}

}  // namespace evx
