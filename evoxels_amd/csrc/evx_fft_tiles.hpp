// Size-specialised FFT tile bodies (z R2C / C2R line pairs, strided y / x
// bricks) and the float32 IMEX multiplier, shared by the one-grid kernels and
// the batched (many-grid) kernels.  Each body serves one tile with a
// caller-provided LDS buffer and a flat thread index.
#pragma once
#include "evx_common.hpp"
#include "evx_fft_engine.hpp"
#include "evx_sum.hpp"

#include <type_traits>

namespace evx {

// ---------------------------------------------------------- strided sweeps --
// Lines along y or x of the half spectrum.  Flat line index L:
//   outer = L / nh, k = L % nh, start = (outer / G) * GM + (outer % G) * OM + k
// y: G = 1, GM = ny*nh            (outer = c*nx + x)
// x: G = ny, GM = nx*ny*nh, OM = nh (outer = c*ny + y)
struct LineMap {
  int64_t nh, G, GM, OM, es;
  // x sweeps of a mirror-extended axis (N = 2n): 0 none; 1 element i >= n is
  // element N-1-i (even extension), 2 its negative (odd extension).  The
  // sweep then reads only the n stored planes and writes only them back.
  int32_t mirror = 0;
};

__device__ __forceinline__ int64_t line_start(const LineMap& m, int64_t L) {
  const int64_t outer = L / m.nh, k = L - (L / m.nh) * m.nh;
  return (outer / m.G) * m.GM + (outer % m.G) * m.OM + k;
}

// Line L0 + t of a brick whose first line L0 is block-uniform: the divisions
// resolve L0 once per wave; a lane's line carries k into o and o into g
// (k = L % nh, o = outer % G, g = outer / G, outer = L / nh), which takes a
// division only when nh is smaller than the lane offset t.  rel is the
// element offset of the lane's line from the brick's first line (>= 0:
// line starts grow with L since GM >= G*OM and OM >= nh or G == 1).
struct BrickLine {
  int64_t start0;
  uint32_t k, o, g, rel;
};

__host__ __device__ __forceinline__ BrickLine brick_line(const LineMap& m, uint32_t L0, uint32_t t) {
  const uint32_t nh = (uint32_t)m.nh, G = (uint32_t)m.G;
  const uint32_t outer0 = L0 / nh, k0 = L0 - outer0 * nh;
  const uint32_t g0 = outer0 / G, o0 = outer0 - g0 * G;
  BrickLine b;
  b.start0 = (int64_t)g0 * m.GM + (int64_t)o0 * m.OM + k0;
  uint32_t k = k0 + t, o = o0, g = g0;
  if (k >= nh) {
    uint32_t q = 1;
    k -= nh;
    if (k >= nh) {
      const uint32_t r = k / nh;
      q += r;
      k -= r * nh;
    }
    o += q;
    if (G == 1) {
      g += o;
      o = 0;
    } else if (o >= G) {
      const uint32_t r = o / G;
      g += r;
      o -= r * G;
    }
  }
  b.k = k;
  b.o = o;
  b.g = g;
  b.rel = (uint32_t)((int64_t)(g - g0) * m.GM + ((int64_t)o - (int64_t)o0) * m.OM + ((int64_t)k - (int64_t)k0));
  return b;
}

// A block-uniform pointer moved to scalar registers (the buffer descriptors
// below are scalar operands).
template <typename P>
__device__ __forceinline__ P* uniform_ptr(P* p) {
  const uint64_t a = reinterpret_cast<uint64_t>(p);
  const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(a));
  const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(a >> 32));
  return reinterpret_cast<P*>((static_cast<uint64_t>(hi) << 32) | lo);
}

// Raw (stride 0) buffer descriptor over the 4 GiB from p; the fourth word is
// the gfx9-family default (32-bit data format, no swizzle).  Loads and stores
// through it take a 32-bit per-lane offset plus a 32-bit scalar offset, so a
// brick needs one vector register of addressing for all its elements.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t raw_rsrc(void* p) {
  return __builtin_amdgcn_make_buffer_rsrc(p, 0, -1, 0x00020000);
}

// (The words are moved with memcpy: a bit_cast of a vector element reads the
// wrong element with this compiler.)
template <typename T>
__device__ __forceinline__ cpx<T> buf_load(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  cpx<T> v;
  if constexpr (sizeof(cpx<T>) == 8) {
    const auto w = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
    __builtin_memcpy(&v, &w, sizeof(v));
  } else {
    const auto w = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
    __builtin_memcpy(&v, &w, sizeof(v));
  }
  return v;
}

template <typename T>
__device__ __forceinline__ void buf_store(cpx<T> v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  if constexpr (sizeof(cpx<T>) == 8) {
    typedef uint32_t u2 __attribute__((ext_vector_type(2)));
    u2 w;
    __builtin_memcpy(&w, &v, sizeof(v));
    __builtin_amdgcn_raw_buffer_store_b64(w, r, voff, soff, 0);
  } else {
    typedef uint32_t u4 __attribute__((ext_vector_type(4)));
    u4 w;
    __builtin_memcpy(&w, &v, sizeof(v));
    __builtin_amdgcn_raw_buffer_store_b128(w, r, voff, soff, 0);
  }
}

// Host-side condition of the brick addressing: line indices in 32 bits, and
// every element reached through one of the brick's two descriptors (elements
// q < E/2 from the brick's first element, the rest from element E/2) at a
// lane offset rel + tl*es plus a scalar offset below E/2 * TPL*es, the sum
// below 4 GiB.  A brick of TL lines wraps k at most TL/nh + 1 times and g at
// most TL/(G*nh) + 1 times (never when all lines share one g).
template <typename T>
inline bool brick_offsets_fit(int64_t nlines, const LineMap& m, int n, int tpl, int tl) {
  if (nlines >= (int64_t(1) << 31) || m.nh >= (int64_t(1) << 31) || m.G >= (int64_t(1) << 31)) return false;
  const int64_t wraps_o = tl / m.nh + 1, wraps_g = nlines > m.G * m.nh ? tl / (m.G * m.nh) + 1 : 0;
  const int64_t lane = wraps_g * m.GM + wraps_o * m.OM + tl + (int64_t)(tpl - 1) * m.es;
  const int64_t e = n / tpl, half = (e + 1) / 2;
  const int64_t scalar = (half - 1) * (int64_t)tpl * m.es;
  return (lane + scalar) * (int64_t)sizeof(cpx<T>) < (int64_t(1) << 32);
}

// float32 IMEX multiplier of the reference: k^2 from fftfreq (float32), the
// symbol sc*k^2 (RD) or sc*(k^2)^2 (CH), prefac = reciprocal(1 - dt*sym)*dt.
struct Prefac {
  int ch;          // 1: Cahn-Hilliard k^4 symbol, 0: k^2 symbol
  float sc;        // float32 scalar of the symbol (e.g. -2*eps*D*A)
  float dt;        // float32(dt)
  float twopi;     // float32(2*pi)
  float inv_nd[3]; // float32(1/(n*d)) per axis
  int nx, ny, nz;
  double scale;    // 1/(nx*ny*nz) of the inverse transform
  const double* chan;  // optional per-channel {eps, D}; sc is then made per channel
  double A;
  // Exact-symbol mode (plan option): the multiplier in float64 from the
  // exact wavenumbers instead of the reference's float32 table.
  int exact;
  // 1: exponential Euler multiplier dt*phi_1(dt*sigma) instead of the
  // semi-implicit dt/(1 - dt*sigma).
  int etd;
  double scd, dtd, inv_ndd[3];
};

// float32 symbol scalar of channel c: the same double expression the host
// evaluates for scalar parameters (problem_definition.py:206,303).
__device__ __forceinline__ float symbol_scalar(const Prefac& p, int64_t c) {
  if (!p.chan) return p.sc;
  const double e = p.chan[2 * c], D = p.chan[2 * c + 1];
  return p.ch ? static_cast<float>(((-2.0 * e) * D) * p.A) : static_cast<float>((-D) * p.A);
}

// Exact-symbol multiplier dt/(1 - dt*sigma(k)) in float64, channel c.
__device__ __forceinline__ double prefac_f64(int jx, int jy, int jz, const Prefac& p, int64_t c) {
  double sc = p.scd;
  if (p.chan) {
    const double e = p.chan[2 * c], D = p.chan[2 * c + 1];
    sc = p.ch ? ((-2.0 * e) * D) * p.A : (-D) * p.A;
  }
  const int sx = jx < (p.nx + 1) / 2 ? jx : jx - p.nx;
  const int sy = jy < (p.ny + 1) / 2 ? jy : jy - p.ny;
  const double tau = 6.283185307179586476925286766559;
  const double kx = tau * (sx * p.inv_ndd[0]), ky = tau * (sy * p.inv_ndd[1]), kz = tau * (jz * p.inv_ndd[2]);
  const double k2 = kx * kx + ky * ky + kz * kz;
  const double sym = p.ch ? sc * (k2 * k2) : sc * k2;
  return p.dtd / (1.0 - p.dtd * sym);
}

// Exponential Euler (ETD1) multiplier of the reference's ExponentialEuler
// (timesteppers.py:136-199): z = dt*sigma(k) on the float32 k-table,
// phi_1(z) = (exp(z)-1)/z, replaced for |z| < 0.5 by the (6,6) Pade form of
// Hochbruck, Lubich & Selhofer (1998) evaluated by Horner's rule, all in
// float32 as the reference's torch expressions round; returns dt*phi_1(z).
__device__ __forceinline__ float etd1_f32(int jx, int jy, int jz, const Prefac& p, float sc) {
#pragma clang fp contract(off)
  const int sx = jx < (p.nx + 1) / 2 ? jx : jx - p.nx;
  const int sy = jy < (p.ny + 1) / 2 ? jy : jy - p.ny;
  const float kx = p.twopi * (static_cast<float>(sx) * p.inv_nd[0]);
  const float ky = p.twopi * (static_cast<float>(sy) * p.inv_nd[1]);
  const float kz = p.twopi * (static_cast<float>(jz) * p.inv_nd[2]);
  const float k2 = (kx * kx + ky * ky) + kz * kz;
  const float sym = p.ch ? sc * (k2 * k2) : sc * k2;
  const float z = p.dt * sym;
  float phi;
  if (fabsf(z) < 0.5f) {
    const float nc[7] = {1.0f, (float)(1.0 / 26), (float)(5.0 / 156), (float)(1.0 / 858), (float)(1.0 / 5720),
                         (float)(1.0 / 205920), (float)(1.0 / 8648640)};
    const float dc[7] = {1.0f, (float)(-6.0 / 13), (float)(5.0 / 52), (float)(-5.0 / 429), (float)(1.0 / 1144),
                         (float)(-1.0 / 25740), (float)(1.0 / 1235520)};
    float num = nc[6], den = dc[6];
#pragma unroll
    for (int k = 5; k >= 0; --k) {
      num = num * z + nc[k];
      den = den * z + dc[k];
    }
    phi = num / den;
  } else {
    phi = (expf(z) - 1.0f) / z;
  }
  return p.dt * phi;
}

__device__ __forceinline__ float prefac_f32(int jx, int jy, int jz, const Prefac& p, float sc) {
#pragma clang fp contract(off)
  const int sx = jx < (p.nx + 1) / 2 ? jx : jx - p.nx;
  const int sy = jy < (p.ny + 1) / 2 ? jy : jy - p.ny;
  const float kx = p.twopi * (static_cast<float>(sx) * p.inv_nd[0]);
  const float ky = p.twopi * (static_cast<float>(sy) * p.inv_nd[1]);
  const float kz = p.twopi * (static_cast<float>(jz) * p.inv_nd[2]);
  const float k2 = (kx * kx + ky * ky) + kz * kz;
  const float sym = p.ch ? sc * (k2 * k2) : sc * k2;
  const float den = 1.0f - p.dt * sym;
  return (1.0f / den) * p.dt;
}

// The float32 multiplier of one x line (jy, jz fixed): the per-line part of
// the k^2 sum is formed once, in the same float roundings as prefac_f32
// (k^2 = (kx*kx + ky*ky) + kz*kz).  FAST takes the hardware reciprocal (1 ulp)
// for 1/(1 - dt*sigma): used by float32 plans, whose FFT rounding is far
// larger; float64 plans keep the IEEE quotient, so their multiplier is the
// reference's float32 table value exactly.
struct PrefacLine {
  float ky2, kz2, ssc;
};

__device__ __forceinline__ PrefacLine prefac_line(int jy, int jz, const Prefac& p, float ssc) {
#pragma clang fp contract(off)
  const int sy = jy < (p.ny + 1) / 2 ? jy : jy - p.ny;
  const float ky = p.twopi * (static_cast<float>(sy) * p.inv_nd[1]);
  const float kz = p.twopi * (static_cast<float>(jz) * p.inv_nd[2]);
  return PrefacLine{ky * ky, kz * kz, ssc};
}

template <bool FAST>
__device__ __forceinline__ float prefac_f32_line(int jx, const PrefacLine& L, const Prefac& p) {
#pragma clang fp contract(off)
  const int sx = jx < (p.nx + 1) / 2 ? jx : jx - p.nx;
  const float kx = p.twopi * (static_cast<float>(sx) * p.inv_nd[0]);
  const float k2 = (kx * kx + L.ky2) + L.kz2;
  const float sym = p.ch ? L.ssc * (k2 * k2) : L.ssc * k2;
  const float den = 1.0f - p.dt * sym;
  if constexpr (FAST) return __builtin_amdgcn_rcpf(den) * p.dt;
  return (1.0f / den) * p.dt;
}

// --------------------------------------------------------- specialised sizes --
// Compile-time radix plans: L = lcm of the radices (minimum values per thread).
template <int N> struct FftSpec;
template <> struct FftSpec<8>    { using R = Rad<8>;          static constexpr int L = 8; };
template <> struct FftSpec<16>   { using R = Rad<4, 4>;       static constexpr int L = 4; };
template <> struct FftSpec<32>   { using R = Rad<4, 8>;       static constexpr int L = 8; };
template <> struct FftSpec<48>   { using R = Rad<3, 4, 4>;    static constexpr int L = 12; };
template <> struct FftSpec<64>   { using R = Rad<8, 8>;       static constexpr int L = 8; };
template <> struct FftSpec<96>   { using R = Rad<3, 4, 4, 2>; static constexpr int L = 12; };
// 100 = 2*5*2*5: ten values per thread (Rad<4, 5, 5> needs twenty), so a
// 100^3 grid's sweeps run twice the threads and blocks; the 100^3 float64
// CH step took 0.0581 instead of 0.0709 ms (MI355X).
template <> struct FftSpec<100>  { using R = Rad<2, 5, 2, 5>; static constexpr int L = 10; };
// 128 = 2*4*4*4: four values per thread instead of eight (Rad<2, 8, 8>):
// the 16 x 128^3 inversion gradient 0.3256 -> 0.3187 s, the 48-grid batched
// step 0.2655 -> 0.2649 ms (MI355X).  The same change for 32 and 64
// (Rad<4, 4, 2>, Rad<4, 4, 4>) made the batched step slower (0.2749 ms).
template <> struct FftSpec<128>  { using R = Rad<2, 4, 4, 4>; static constexpr int L = 4; };
template <> struct FftSpec<256>  { using R = Rad<4, 8, 8>;    static constexpr int L = 8; };
template <> struct FftSpec<512>  { using R = Rad<8, 8, 8>;    static constexpr int L = 8; };
template <> struct FftSpec<1024> { using R = Rad<2, 8, 8, 8>; static constexpr int L = 8; };
#define EVX_FAST_SIZES(X) X(8) X(16) X(32) X(48) X(64) X(96) X(100) X(128) X(256) X(512) X(1024)

// Every plan multiplies out to its size, and L is a multiple of each radix.
template <int... R>
constexpr int rad_product(Rad<R...>) { return (R * ... * 1); }
template <int L, int... R>
constexpr bool rad_divides(Rad<R...>) { return ((L % R == 0) && ...); }
#define X(N)                                                                                     \
  static_assert(rad_product(typename FftSpec<N>::R{}) == N && rad_divides<FftSpec<N>::L>(typename FftSpec<N>::R{}), \
                "radix plan of size " #N);
EVX_FAST_SIZES(X)
#undef X

// z sweeps: one line pair per TPL threads, TPL = N/L (each thread owns L values).
template <typename T, int N>
struct ZGeo {
  static constexpr int TPL = N / FftSpec<N>::L;
  static constexpr int LP = TPL >= 256 ? 1 : 256 / TPL;  // line pairs per workgroup
  static constexpr int THREADS = LP * TPL;
  static constexpr int LS = N + N / 16;                   // padded LDS line (line_slot)
  // a line pair's TPL threads (tid / TPL) share one wavefront
  static constexpr bool WAVE = 64 % TPL == 0;
  // first line (a) of pair pr of tile `tile`; b = a + 1
  __host__ __device__ static int64_t first_line(int64_t tile, int pr) { return (tile * LP + pr) * 2; }
  __host__ __device__ static int64_t blocks(int64_t nlines) { return ((nlines + 1) / 2 + LP - 1) / LP; }
};

// Slot of element e in a padded exchange line: one pad element after every
// 16 (a pad after every 8 measured no faster).  A line of N elements takes
// N + N/16 slots.
__host__ __device__ __forceinline__ int line_slot(int e) { return e + (e >> 4); }

// Layouts give element e's slot (at) and element tl + TPL*q's slot
// (at_q): for TPL a multiple of 16, line_slot(tl + TPL*q) = line_slot(tl) +
// q*(TPL + TPL/16), one lane base plus constant offsets.
template <typename T, int LS, bool W = false>
struct LineLay {
  static constexpr bool SPLIT = false;
  static constexpr bool WAVE = W;
  cpx<T>* base;
  __device__ __forceinline__ cpx<T>* at(int e) const { return base + line_slot(e); }
  template <int TPL>
  __device__ __forceinline__ cpx<T>* at_q(int tl, int q) const {
    if constexpr (TPL % 16 == 0) return base + line_slot(tl) + q * (TPL + TPL / 16);
    else return at(tl + TPL * q);
  }
};

template <typename T, int TL>
struct BrickLay {
  static constexpr bool SPLIT = false;
  static constexpr bool WAVE = false;
  cpx<T>* base;
  __device__ __forceinline__ cpx<T>* at(int e) const { return base + e * TL; }
  template <int TPL>
  __device__ __forceinline__ cpx<T>* at_q(int tl, int q) const { return at(tl + TPL * q); }
};

// Brick of real numbers (one component at a time; see the engine).
template <typename T, int TL>
struct BrickLaySplit {
  static constexpr bool SPLIT = true;
  static constexpr bool WAVE = false;
  T* base;
  __device__ __forceinline__ T* at(int e) const { return base + e * TL; }
  template <int TPL>
  __device__ __forceinline__ T* at_q(int tl, int q) const { return at(tl + TPL * q); }
};

// LDS of one z tile: twiddles plus LP padded lines, and one pair per
// wavefront for the block sum of the C2R epilogue.
template <typename T, int N>
struct ZTile {
  static constexpr size_t LINES_BYTES = sizeof(cpx<T>) * (N + ZGeo<T, N>::LP * ZGeo<T, N>::LS);
  static constexpr size_t LDS_BYTES = LINES_BYTES + sizeof(SumPair) * 16 + 16;
};

// z R2C of tile `tile` (LP line pairs): two real lines per complex FFT.
template <typename T, int N>
__device__ __forceinline__ void z_r2c_tile(const T* __restrict__ x, cpx<T>* __restrict__ spec, int64_t nlines,
                                           int64_t pitch, const cpx<T>* __restrict__ twg, int64_t tile, int tid,
                                           unsigned char* smem) {
  using G = ZGeo<T, N>;
  constexpr int TPL = G::TPL, E = N / TPL, NH = N / 2 + 1;
  cpx<T>* tw = reinterpret_cast<cpx<T>*>(smem);
  cpx<T>* lds = tw + N;
  const int pr = tid / TPL, tl = tid - (tid / TPL) * TPL;
  for (int i = tid; i < N; i += G::THREADS) tw[i] = twg[i];
  const int64_t la = G::first_line(tile, pr);
  const bool has_a = la < nlines, has_b = la + 1 < nlines;
  const T* xa = x + la * N;
  cpx<T> v[E];
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const int e = tl + TPL * q;
    v[q].x = has_a ? xa[e] : T(0);
    v[q].y = has_b ? xa[N + e] : T(0);
  }
  __syncthreads();  // twiddles
  const LineLay<T, G::LS, G::WAVE> lay{lds + pr * G::LS};
  fft_regs<T, N, TPL, false>(v, lay, tl, tw, typename FftSpec<N>::R{});
  xsync<G::WAVE>();
#pragma unroll
  for (int q = 0; q < E; ++q) *lay.at(tl + TPL * q) = v[q];
  xsync<G::WAVE>();
  cpx<T>* sa = spec + la * pitch;
#pragma unroll
  for (int q = 0; q < (NH + TPL - 1) / TPL; ++q) {
    const int k = tl + TPL * q;
    if (k < NH && has_a) {
      const cpx<T> Z = v[q];
      const cpx<T> Zc = *lay.at(k == 0 ? 0 : N - k);
      sa[k] = cpx<T>{(Z.x + Zc.x) * T(0.5), (Z.y - Zc.y) * T(0.5)};
      if (has_b) sa[pitch + k] = cpx<T>{(Z.y + Zc.y) * T(0.5), (Zc.x - Z.x) * T(0.5)};
    }
  }
}

// z C2R of tile `tile`: out = base + scale * irfft(spec) (base may be NULL,
// out may alias base).  With rec, the block also writes rec[tile]: the
// compensated float64 sum of its outputs and their NaN / Inf flags (every
// thread sums its own outputs in float64, the wavefronts merge with TwoSum,
// then the block in wavefront order).
template <typename T, int N>
__device__ __forceinline__ void z_c2r_tile(const cpx<T>* __restrict__ spec, T* out, const T* base, T scale,
                                           int64_t nlines, int64_t pitch, const cpx<T>* __restrict__ twg,
                                           int64_t tile, int tid, unsigned char* smem, SumRec* rec = nullptr) {
  using G = ZGeo<T, N>;
  constexpr int TPL = G::TPL, E = N / TPL, NH = N / 2 + 1;
  cpx<T>* tw = reinterpret_cast<cpx<T>*>(smem);
  cpx<T>* lds = tw + N;
  const int pr = tid / TPL, tl = tid - (tid / TPL) * TPL;
  for (int i = tid; i < N; i += G::THREADS) tw[i] = twg[i];
  const int64_t la = G::first_line(tile, pr);
  const bool has_a = la < nlines, has_b = la + 1 < nlines;
  const cpx<T>* sa = spec + la * pitch;
  const LineLay<T, G::LS, G::WAVE> lay{lds + pr * G::LS};
  cpx<T> v[E];
  // Each stored element of the two half spectra is read from HBM once and
  // parked in the pair's exchange line: slot k holds A[k] and slot N/2 + k
  // holds B[k] (0 < k < N/2); slot 0 holds (A[0], A[N/2]) and slot N/2 holds
  // (B[0], B[N/2]), real parts only (a real line's DC and Nyquist terms are
  // real: C2R drops their imaginary parts, as numpy's irfft does).
  {
    T* const d0 = &lay.at(0)->x;
    T* const dh = &lay.at(N / 2)->x;
#pragma unroll
    for (int q = 0; q < (NH + TPL - 1) / TPL; ++q) {
      const int k = tl + TPL * q;
      if (k < NH) {
        const cpx<T> A = has_a ? sa[k] : cpx<T>{T(0), T(0)};
        const cpx<T> B = has_b ? sa[pitch + k] : cpx<T>{T(0), T(0)};
        if (k == 0) {
          d0[0] = A.x;
          dh[0] = B.x;
        } else if (k == N / 2) {
          d0[1] = A.x;
          dh[1] = B.x;
        } else {
          *lay.at(k) = A;
          *lay.at(N / 2 + k) = B;
        }
      }
    }
  }
  xsync<G::WAVE>();
  // Z[k] = A[k] + i B[k] of the Hermitian extensions, A[N-k] = conj(A[k]).
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const int k = tl + TPL * q;
    const int j = k <= N / 2 ? k : N - k;
    if (j == 0 || j == N / 2) {
      const cpx<T> a0 = *lay.at(0), ah = *lay.at(N / 2);
      const T ax = j == 0 ? a0.x : a0.y, bx = j == 0 ? ah.x : ah.y;
      v[q] = cpx<T>{ax - T(0), T(0) + bx};
    } else {
      const cpx<T> a = *lay.at(j), b = *lay.at(N / 2 + j);
      v[q] = k < N / 2 ? cpx<T>{a.x - b.y, a.y + b.x} : cpx<T>{a.x + b.y, b.x - a.y};
    }
  }
  // The state added to the update is loaded now so its latency overlaps the FFT.
  const T* ba = base ? base + la * N : nullptr;
  cpx<T> bv[E];
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const int e = tl + TPL * q;
    bv[q].x = (ba && has_a) ? ba[e] : T(0);
    bv[q].y = (ba && has_b) ? ba[N + e] : T(0);
  }
  __syncthreads();  // twiddles
  fft_regs<T, N, TPL, true>(v, lay, tl, tw, typename FftSpec<N>::R{});
  T* oa = out + la * N;
  double acc = 0.0;
  bool bad = false;  // some output NaN or +-Inf
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const int e = tl + TPL * q;
    T oa_v = bv[q].x + scale * v[q].x, ob_v = bv[q].y + scale * v[q].y;
    if (has_a) oa[e] = oa_v;
    if (has_b) oa[N + e] = ob_v;
    if (rec) {
      oa_v = has_a ? oa_v : T(0);
      ob_v = has_b ? ob_v : T(0);
      acc += static_cast<double>(oa_v);
      acc += static_cast<double>(ob_v);
      bad = bad || !__builtin_isfinite(oa_v) || !__builtin_isfinite(ob_v);
    }
  }
  bool nan = false, inf = false;
  if (rec && __any(bad)) {  // rare: tell NaN from Inf over this thread's outputs again
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const T a = has_a ? bv[q].x + scale * v[q].x : T(0), b = has_b ? bv[q].y + scale * v[q].y : T(0);
      nan = nan || a != a || b != b;
      inf = inf || __builtin_isinf(a) || __builtin_isinf(b);
    }
  }
  if (rec) {  // uniform branch: the whole block writes its record
    SumPair* ws = reinterpret_cast<SumPair*>(smem + ZTile<T, N>::LINES_BYTES);
    int* wf = reinterpret_cast<int*>(ws + 16);
    const SumPair wp = wave_sum(SumPair{acc, 0.0});
    const int wflags = (__any(nan) ? 1 : 0) | (__any(inf) ? 2 : 0);
    const int w = tid >> 6, nw = (G::THREADS + 63) / 64;
    if ((tid & 63) == 0) {
      ws[w] = wp;
      wf[w] = wflags;
    }
    __syncthreads();
    if (tid == 0) {
      SumPair b = ws[0];
      int f = wf[0];
      for (int i = 1; i < nw; ++i) {
        b = sum_merge(b, ws[i]);
        f |= wf[i];
      }
      rec[tile] = SumRec{b.s, b.c, f & 1, (f >> 1) & 1, {0, 0}};
    }
  }
}

// Strided sweeps: TL adjacent lines (128-byte runs per element row), TPL
// threads per line with E values each.
constexpr int pick_strided_e(int n, int l, int tl0) {
  int e = l;
  while (n / e > 256 / tl0 && n % (e * 2) == 0) e *= 2;
  return e;
}

template <typename T, int N, int MODE = 0>
struct SGeo {
  static constexpr int TL0 = 128 / (int)sizeof(cpx<T>);
  // From 512 points up every sweep uses half the values per thread (twice
  // the threads per line): the forward*prefac*inverse sweeps stay off the
  // register limit, and the one-way sweeps reach eight waves per SIMD.
  static constexpr int E0 = pick_strided_e(N, FftSpec<N>::L, TL0);
  static constexpr int E1 = (N >= 512 && E0 >= 2 * FftSpec<N>::L) ? E0 / 2 : E0;
  // The float32 512-point forward*multiplier*inverse sweep halves it once
  // more (eight values, 1024 threads per brick): 56 VGPRs, eight waves per
  // SIMD instead of four, 2.7% faster on MI355X.
  static constexpr int E = (MODE == 2 && N == 512 && sizeof(T) == 4 && E1 >= 2 * FftSpec<N>::L) ? E1 / 2 : E1;
  static constexpr int TPL = N / E;
  static constexpr int TL = TL0 > 256 / TPL ? TL0 : 256 / TPL;
  static constexpr int THREADS = TL * TPL;
  // Bricks above 32 KB exchange real and imaginary parts in turn through a
  // half-size buffer, so that more workgroups fit a CU.
  static constexpr bool SPLIT = (size_t)TL * N * sizeof(cpx<T>) > 32768;
};

// Float32 x sweeps (MODE >= 2) whose lines take one wavefront each (TPL =
// 64), strided_x_pipe: after the coalesced brick load the block transposes
// the brick once through LDS, so that wavefront w owns line L0 + w and runs
// both transforms with wavefront-private exchanges, and transposes back for
// the coalesced store: three workgroup barriers per brick instead of one per
// exchange.
// Lines sit LSX complex values apart, LSX = 32/TL mod 32, so that the TL
// lines a half-wavefront writes in the transposes (64/TL consecutive
// elements each) start 64/TL banks apart.
template <typename T, int N, int MODE>
struct SWave {
  using G = SGeo<T, N, MODE>;
  static constexpr bool ON = MODE >= 2 && sizeof(T) == 4 && G::TPL == 64 && G::TL * 64 == G::THREADS;
  static constexpr int LS0 = N + N / 16;  // line_slot(N - 1) + 1
  static constexpr int LSX = LS0 + ((32 / G::TL - LS0 % 32) + 32) % 32;
  static_assert(!ON || G::THREADS / G::TL == 64, "64 threads per line: tl < 64");
};

// The forward*prefac*inverse sweeps (MODE 2, 3) are register-bound, not
// LDS-bound, so they keep the full-size exchange.
template <typename T, int N, int MODE>
struct STile {
  static constexpr bool SPLIT = SGeo<T, N, MODE>::SPLIT && MODE < 2;
  static constexpr size_t LDS_BYTES =
      SWave<T, N, MODE>::ON
          ? sizeof(cpx<T>) * (N + (size_t)SGeo<T, N, MODE>::TL * SWave<T, N, MODE>::LSX)
          : sizeof(cpx<T>) * N + (SPLIT ? sizeof(T) : sizeof(cpx<T>)) * SGeo<T, N, MODE>::TL * N;
};

// Global side of one brick (TL adjacent lines): lane t serves line L0 + t,
// lanes past the last line read line L0 (always present) and store nothing.
// Element q of a lane sits at descriptor q < H ? lo : hi, scalar offset
// (q mod H)*TPL*es and lane offset rel + tl*es (bytes; the host checks the
// 4 GiB bound with brick_offsets_fit).
// MIR: -1 the mirror mode of the LineMap decides at run time; 0 never
// mirrored (no mirror code); 1 mirrored (the sign still from the LineMap).
template <typename T, int N, int MODE, int MIR = -1>
struct BrickIo {
  using G = SGeo<T, N, MODE>;
  static constexpr int TPL = G::TPL, TL = G::TL, E = N / TPL, H = (E + 1) / 2;
  __amdgpu_buffer_rsrc_t rlo, rhi;
  uint32_t voff, voff_hi, qstep;
  int32_t mirror;
  bool active;
  BrickLine bl;

  __device__ __forceinline__ int32_t mir() const {
    if constexpr (MIR == 0) return 0;
    else return mirror;
  }

  __device__ __forceinline__ void setup(cpx<T>* spec, int64_t nlines, const LineMap& lm, uint32_t tile, int t,
                                        int tl) {
    const uint32_t L0 = tile * TL;
    active = (int64_t)L0 + t < nlines;
    bl = brick_line(lm, L0, active ? t : 0);
    cpx<T>* const base = uniform_ptr(spec + bl.start0);
    qstep = (uint32_t)TPL * (uint32_t)lm.es * (uint32_t)sizeof(cpx<T>);
    mirror = MIR == 0 ? 0 : lm.mirror;
    rlo = raw_rsrc(base);
    voff = (bl.rel + (uint32_t)tl * (uint32_t)lm.es) * (uint32_t)sizeof(cpx<T>);
    if (mir()) {
      // element tl + TPL*q (q >= H) mirrors element N-1-tl-TPL*q =
      // (TPL-1-tl) + TPL*(E-1-q), which lies in the lower half
      rhi = rlo;
      voff_hi = (bl.rel + (uint32_t)(TPL - 1 - tl) * (uint32_t)lm.es) * (uint32_t)sizeof(cpx<T>);
    } else {
      rhi = raw_rsrc(base + (int64_t)H * TPL * lm.es);
      voff_hi = MIR == 0 ? 0u : voff;  // the unmirrored loads and stores use voff
    }
  }
  __device__ __forceinline__ void load(cpx<T> (&v)[E]) const {
#pragma unroll
    for (int q = 0; q < H; ++q) v[q] = buf_load<T>(rlo, voff, (uint32_t)q * qstep);
    if (mir()) {
      static_assert(E % 2 == 0, "mirrored sweeps need an even number of values per thread");
#pragma unroll
      for (int q = H; q < E; ++q) v[q] = buf_load<T>(rhi, voff_hi, (uint32_t)(E - 1 - q) * qstep);
      if (mir() == 2) {
#pragma unroll
        for (int q = H; q < E; ++q) v[q] = {-v[q].x, -v[q].y};
      }
    } else {
#pragma unroll
      for (int q = H; q < E; ++q) v[q] = buf_load<T>(rhi, MIR == 0 ? voff : voff_hi, (uint32_t)(q - H) * qstep);
    }
  }
  __device__ __forceinline__ void store(const cpx<T> (&v)[E]) const {
    if (active) {
#pragma unroll
      for (int q = 0; q < H; ++q) buf_store<T>(v[q], rlo, voff, (uint32_t)q * qstep);
      if (!mir()) {  // a mirrored sweep keeps only the stored (lower) half
#pragma unroll
        for (int q = H; q < E; ++q) buf_store<T>(v[q], rhi, MIR == 0 ? voff : voff_hi, (uint32_t)(q - H) * qstep);
      }
    }
  }
};

// LDS exchange layout of a brick: line t's elements TL apart.
template <typename T, int N, int MODE>
using BrickLayout = typename std::conditional<STile<T, N, MODE>::SPLIT, BrickLaySplit<T, SGeo<T, N, MODE>::TL>,
                                              BrickLay<T, SGeo<T, N, MODE>::TL>>::type;

template <typename T, int N, int MODE>
__device__ __forceinline__ BrickLayout<T, N, MODE> brick_layout(cpx<T>* lds, int t) {
  BrickLayout<T, N, MODE> lay;
  if constexpr (STile<T, N, MODE>::SPLIT) lay.base = reinterpret_cast<T*>(lds) + t;
  else lay.base = lds + t;
  return lay;
}

// In-register transform of a brick: forward (MODE 0, 2-4) or inverse (1);
// MODE >= 2 multiplies by the step's multiplier and transforms back.  The
// caller has synchronised the block since the exchange buffer's last use.
template <typename T, int N, int MODE, bool PTW = false, int E, typename Lay>
__device__ __forceinline__ void brick_transform(cpx<T> (&v)[E], const Lay& lay, int tl, const cpx<T>* tw,
                                                const BrickLine& bl, const Prefac& pf) {
  constexpr int TPL = SGeo<T, N, MODE>::TPL;
  static_assert(E == N / TPL, "one brick's values per thread");
  // PTW: tw holds the per-pass twiddle tables (pass_tw_fill), else the plain one
  if constexpr (PTW) fft_regs_pt<T, N, TPL, MODE == 1>(v, lay, tl, tw, typename FftSpec<N>::R{});
  else fft_regs<T, N, TPL, MODE == 1>(v, lay, tl, tw, typename FftSpec<N>::R{});
  if constexpr (MODE >= 2) {
    // x lines: G = ny, so (k, o, g) = (jz, jy, channel)
    const int jz = (int)bl.k;
    const int jy = (int)bl.o;
    const T sc = static_cast<T>(pf.scale);
    const int64_t ch = bl.g;
    if constexpr (MODE == 3) {
#pragma unroll
      for (int q = 0; q < E; ++q)
        v[q] = cscale(v[q], static_cast<T>(prefac_f64(tl + TPL * q, jy, jz, pf, ch) * pf.scale));
    } else if constexpr (MODE == 4) {
      const float ssc = symbol_scalar(pf, ch);
#pragma unroll
      for (int q = 0; q < E; ++q) v[q] = cscale(v[q], static_cast<T>(etd1_f32(tl + TPL * q, jy, jz, pf, ssc)) * sc);
    } else {
      const PrefacLine L = prefac_line(jy, jz, pf, symbol_scalar(pf, ch));
#pragma unroll
      for (int q = 0; q < E; ++q)
        v[q] = cscale(v[q], static_cast<T>(prefac_f32_line<sizeof(T) == 4>(tl + TPL * q, L, pf)) * sc);
    }
    if constexpr (PTW) fft_regs_pt<T, N, TPL, true>(v, lay, tl, tw, typename FftSpec<N>::R{});
    else fft_regs<T, N, TPL, true>(v, lay, tl, tw, typename FftSpec<N>::R{});
  }
}

// Strided sweep of brick `tile` (TL adjacent lines).  MODE 0: forward,
// 1: inverse, 2: forward * prefac * inverse (x axis only) with the float32
// multiplier, 3: the same with the exact-symbol float64 multiplier, 4: the
// same with the exponential Euler multiplier.
template <typename T, int N, int MODE>
__device__ __forceinline__ void strided_tile(cpx<T>* __restrict__ spec, int64_t nlines, const LineMap& lm,
                                             const cpx<T>* __restrict__ twg, const Prefac& pf, int64_t tile,
                                             int tid, unsigned char* smem) {
  using G = SGeo<T, N, MODE>;
  constexpr int TL = G::TL, E = N / G::TPL;
  cpx<T>* tw = reinterpret_cast<cpx<T>*>(smem);
  cpx<T>* lds = tw + N;
  const int t = tid % TL, tl = tid / TL;
  for (int i = tid; i < N; i += G::THREADS) tw[i] = twg[i];
  static_assert(!SWave<T, N, MODE>::ON, "the wavefront-per-line x sweeps run strided_x_pipe");
  BrickIo<T, N, MODE> io;
  io.setup(spec, nlines, lm, (uint32_t)tile, t, tl);
  cpx<T> v[E];
  io.load(v);
  __syncthreads();
  brick_transform<T, N, MODE>(v, brick_layout<T, N, MODE>(lds, t), tl, tw, io.bl, pf);
  io.store(v);
}

// The wavefront-per-line x sweep, persistent: a block walks bricks first,
// first + step, ... and loads the next brick into registers while it
// transforms the current one from LDS, so that its loads and its transforms
// overlap instead of every CU alternating between the two (a thread holds
// elements tl + 64q of line t of a brick, at exchange slot line_slot(tl +
// 64q) = line_slot(tl) + 68q).  Wavefront w transforms line L0 + w; line
// L0 stands in past the last line, and its results are not stored.
// MIR (0 or 1): the brick IO's compile-time mirror mode.
template <typename T, int N, int MODE, int MIR = 0>
__device__ __forceinline__ void strided_x_pipe(cpx<T>* __restrict__ spec, int64_t nlines, const LineMap& lm,
                                               const cpx<T>* __restrict__ twg, const Prefac& pf, uint32_t nbricks,
                                               uint32_t first, uint32_t step, int tid, unsigned char* smem) {
  using G = SGeo<T, N, MODE>;
  static_assert(SWave<T, N, MODE>::ON, "wavefront-per-line x sweeps only");
  constexpr int TL = G::TL, E = N / G::TPL, LSX = SWave<T, N, MODE>::LSX;
  // per-pass twiddle tables (conflict-free reads) in the twiddle region
  static_assert(pass_tw_size<N>(typename FftSpec<N>::R{}) <= N, "pass tables fit the twiddle region");
  cpx<T>* tw = reinterpret_cast<cpx<T>*>(smem);
  cpx<T>* lds = tw + N;
  const int t = tid % TL, tl = tid / TL;
  pass_tw_fill<T, N>(tw, twg, tid, G::THREADS, typename FftSpec<N>::R{});
  if (first >= nbricks) return;  // uniform in the block
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6), l = tid & 63;
  cpx<T>* const bt = lds + t * LSX + line_slot(tl);
  cpx<T>* const wl = lds + w * LSX + line_slot(l);
  const LineLay<T, LSX, true> lay{lds + w * LSX};
  uint32_t tile = first;
  BrickIo<T, N, MODE, MIR> io;
  io.setup(spec, nlines, lm, tile, t, tl);
  cpx<T> v[E];
  io.load(v);
  while (true) {
#pragma unroll
    for (int q = 0; q < E; ++q) bt[68 * q] = v[q];
    __syncthreads();  // the brick (and the twiddles) are in LDS
    const uint32_t next = tile + step;
    const bool more = next < nbricks;  // uniform in the block
    BrickIo<T, N, MODE, MIR> io2;
    if (more) {
      io2.setup(spec, nlines, lm, next, t, tl);
      io2.load(v);  // in flight during the transform below
    }
    cpx<T> x[E];
#pragma unroll
    for (int q = 0; q < E; ++q) x[q] = wl[68 * q];
    const uint32_t L0 = tile * TL;
    const BrickLine bl = brick_line(lm, L0, (int64_t)L0 + w < nlines ? (uint32_t)w : 0u);
    brick_transform<T, N, MODE, true>(x, lay, l, tw, bl, pf);
    xsync<true>();  // the wavefront's last exchange reads are done
#pragma unroll
    for (int q = 0; q < E; ++q) wl[68 * q] = x[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < E; ++q) x[q] = bt[68 * q];
    io.store(x);
    if (!more) break;
    __syncthreads();  // every read of this brick's LDS is done before the next one is written
    io = io2;
    tile = next;
  }
}

}  // namespace evx
