// Real 3-D FFTs and the semi-implicit spectral steps.
//
// Spectrum layout: (channels, nx, ny, nh) complex, nh = nz/2 + 1, kz fastest,
// the same index order as torch.fft.rfftn's output (voxelgrid.py:203-204).
// Kernels take the row pitch of the spectrum separately from the row length
// (nh); the steps' workspace spectrum uses pitch = nh.
//
// Transforms are mixed-radix Stockham (self-sorting, decimation in time)
// passes on lines staged in LDS.  A workgroup owns `lpw` lines laid out
// element-major (LDS index e*lpw + line) and `tpl` threads per line; thread
// `tid` works on line tid % lpw.  Radices 2, 3, 4, 5 and 8 have closed-form
// butterflies; any other prime factor up to 31 uses a generic O(R^2) DFT.
//
// Sweeps of one IMEX step (each reads and writes HBM once):
//   z : RHS lines (pairs packed as one complex line) -> half spectra
//   y : forward C2C along y
//   x : forward C2C along x, times prefac(k)/N, inverse C2C along x
//   y : inverse C2C along y
//   z : C2R of line pairs, u_out = u + update
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "evx_common.hpp"
#include "evx_fft_engine.hpp"
#include "evx_fft_tiles.hpp"
#include "evx_rz_tile.hpp"

namespace evx {

int stencil_rhs_ch_periodic(int32_t dtype, const void* u, void* out, const Dims& d, const double div_dx[3],
                            const double div_dx2[3], double eps, double D, const double* chan, hipStream_t s,
                            int64_t xa = 0, int64_t xb = -1);
int stencil_ch_xchunk(int32_t dtype, const Dims& d);
size_t ch_rhs_vjp_ws_bytes(int32_t dtype, const Dims& d);
int ch_rhs_vjp_launch(int32_t dtype, const void* u, const void* w, const void* base, void* grad,
                      double* dparams, const Dims& d, const double dd[3], const double dd2[3], double eps,
                      double D, const double* chan, const void* u1, double dt, void* ws, hipStream_t s);
int stencil_rhs_rd_periodic(int32_t dtype, const void* u, void* out, const Dims& d, const double div_dx2[3],
                            double D, hipStream_t s);

namespace {

constexpr int MAXPASS = 24;
constexpr int EMAX = 32;           // max complex values a thread holds in one pass
constexpr size_t LDS_BUDGET = 64 * 1024;

template <typename T>
struct Axis {
  int n;
  int npass;
  int tpl;
  int radix[MAXPASS];
  const cpx<T>* tw;  // n roots of unity exp(-2 pi i m / n)
};

// One Stockham pass of radix R over every line of the workgroup.
// p = product of the radices already applied.  Butterfly i (0 <= i < n/R)
// reads elements i + r*n/R, twiddles by w_{pR}^{r*(i mod p)}, and writes
// (i - i mod p)*R + i mod p + r*p.
template <typename T, bool INV, int R>
__device__ __forceinline__ void pass_fixed(cpx<T>* buf, int lpw, int line, int tl, int tpl, int n, int p,
                                           const cpx<T>* tw, bool active) {
  constexpr int SMAX = EMAX / R;
  const int m = n / R;
  const int stride = n / (p * R);
  cpx<T> y[SMAX][R];
#pragma unroll
  for (int s = 0; s < SMAX; ++s) {
    const int i = tl + s * tpl;
    if (active && i < m) {
      const int k = i % p;
      cpx<T> x[R];
#pragma unroll
      for (int r = 0; r < R; ++r) x[r] = buf[(i + r * m) * lpw + line];
      if (p > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          const cpx<T> w = tw[r * k * stride];
          x[r] = INV ? cmulc(x[r], w) : cmul(x[r], w);
        }
      }
      dft_fixed<T, INV, R>(x);
#pragma unroll
      for (int r = 0; r < R; ++r) y[s][r] = x[r];
    }
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < SMAX; ++s) {
    const int i = tl + s * tpl;
    if (active && i < m) {
      const int k = i % p;
      const int j = (i - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[(j + r * p) * lpw + line] = y[s][r];
    }
  }
  __syncthreads();
}

// Generic odd-prime radix: y_q = sum_r (x_r w^{rk}) w_R^{rq}, O(R^2).
template <typename T, bool INV>
__device__ __noinline__ void pass_generic(cpx<T>* buf, int lpw, int line, int tl, int tpl, int n, int p, int R,
                                          const cpx<T>* tw, bool active) {
  const int m = n / R;
  const int stride = n / (p * R);
  const int nr = n / R;
  cpx<T> y[EMAX];
  int cnt = 0;
  for (int i = tl; active && i < m; i += tpl) {
    const int k = i % p;
    for (int q = 0; q < R; ++q) {
      cpx<T> acc = {T(0), T(0)};
      for (int r = 0; r < R; ++r) {
        cpx<T> xr = buf[(i + r * m) * lpw + line];
        const cpx<T> w = tw[r * k * stride];
        xr = INV ? cmulc(xr, w) : cmul(xr, w);
        const cpx<T> wq = tw[((r * q) % R) * nr];
        acc = cadd(acc, INV ? cmulc(xr, wq) : cmul(xr, wq));
      }
      y[cnt++] = acc;
    }
  }
  __syncthreads();
  cnt = 0;
  for (int i = tl; active && i < m; i += tpl) {
    const int k = i % p;
    const int j = (i - k) * R + k;
    for (int q = 0; q < R; ++q) buf[(j + q * p) * lpw + line] = y[cnt++];
  }
  __syncthreads();
}

template <typename T, bool INV>
__device__ void fft_lines(cpx<T>* buf, int lpw, int line, int tl, const Axis<T>& a, const cpx<T>* tw,
                          bool active) {
  int p = 1;
  for (int s = 0; s < a.npass; ++s) {
    const int R = a.radix[s];
    switch (R) {
      case 8: pass_fixed<T, INV, 8>(buf, lpw, line, tl, a.tpl, a.n, p, tw, active); break;
      case 4: pass_fixed<T, INV, 4>(buf, lpw, line, tl, a.tpl, a.n, p, tw, active); break;
      case 2: pass_fixed<T, INV, 2>(buf, lpw, line, tl, a.tpl, a.n, p, tw, active); break;
      case 3: pass_fixed<T, INV, 3>(buf, lpw, line, tl, a.tpl, a.n, p, tw, active); break;
      case 5: pass_fixed<T, INV, 5>(buf, lpw, line, tl, a.tpl, a.n, p, tw, active); break;
      default: pass_generic<T, INV>(buf, lpw, line, tl, a.tpl, a.n, p, R, tw, active); break;
    }
    p *= R;
  }
}

template <typename T>
__device__ __forceinline__ void load_twiddles(cpx<T>* tw, const Axis<T>& a) {
  for (int i = threadIdx.x; i < a.n; i += blockDim.x) tw[i] = a.tw[i];
}

// --------------------------------------------------------------- z sweeps --
// Real lines of length n along z; line pairs (2q, 2q+1) form one complex line.
template <typename T>
__global__ void z_r2c_kernel(const T* __restrict__ x, cpx<T>* __restrict__ spec, int64_t nlines, int nh,
                             int64_t pitch, Axis<T> a, int lpw) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<T>* tw = reinterpret_cast<cpx<T>*>(smem_raw);
  cpx<T>* buf = tw + a.n;
  const int n = a.n, tid = threadIdx.x;
  load_twiddles(tw, a);
  const int64_t l0 = (int64_t)blockIdx.x * lpw * 2;
  const int nl = (int)min((int64_t)2 * lpw, nlines - l0);
  const int np = (nl + 1) / 2;
  const T* src = x + l0 * n;
  for (int q = tid; q < 2 * np * n; q += blockDim.x) {
    const int l = q / n, e = q - (q / n) * n;
    const T v = l < nl ? src[q] : T(0);
    T* z = reinterpret_cast<T*>(&buf[e * lpw + (l >> 1)]);
    z[l & 1] = v;
  }
  __syncthreads();
  const int line = tid % lpw, tl = tid / lpw;
  fft_lines<T, false>(buf, lpw, line, tl, a, tw, line < np && tl < a.tpl);
  cpx<T>* dst = spec + l0 * pitch;
  for (int q = tid; q < nl * nh; q += blockDim.x) {
    const int l = q / nh, k = q - (q / nh) * nh;
    const cpx<T> Z = buf[k * lpw + (l >> 1)];
    const cpx<T> Zc = buf[((n - k) % n) * lpw + (l >> 1)];
    cpx<T> o;
    if ((l & 1) == 0) o = {(Z.x + Zc.x) * T(0.5), (Z.y - Zc.y) * T(0.5)};
    else o = {(Z.y + Zc.y) * T(0.5), (Zc.x - Z.x) * T(0.5)};
    dst[l * pitch + k] = o;
  }
}

template <typename T>
__device__ __forceinline__ cpx<T> herm_load(const cpx<T>* line, int k, int n, int nh) {
  cpx<T> v = k < nh ? line[k] : cpx<T>{line[n - k].x, -line[n - k].y};
  if (k == 0 || 2 * k == n) v.y = T(0);  // C2R ignores Im of the DC and Nyquist bins
  return v;
}

// out = base + scale * irfft(spec) per line (base may be NULL; out may alias base).
template <typename T>
__global__ void z_c2r_kernel(const cpx<T>* __restrict__ spec, T* out, const T* base, T scale, int64_t nlines,
                             int nh, int64_t pitch, Axis<T> a, int lpw) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<T>* tw = reinterpret_cast<cpx<T>*>(smem_raw);
  cpx<T>* buf = tw + a.n;
  const int n = a.n, tid = threadIdx.x;
  load_twiddles(tw, a);
  const int64_t l0 = (int64_t)blockIdx.x * lpw * 2;
  const int nl = (int)min((int64_t)2 * lpw, nlines - l0);
  const int np = (nl + 1) / 2;
  for (int q = tid; q < np * n; q += blockDim.x) {
    const int pr = q / n, k = q - (q / n) * n;
    const cpx<T>* la = spec + (l0 + 2 * pr) * pitch;
    const cpx<T> A = herm_load(la, k, n, nh);
    const cpx<T> B = (2 * pr + 1 < nl) ? herm_load(la + pitch, k, n, nh) : cpx<T>{T(0), T(0)};
    buf[k * lpw + pr] = {A.x - B.y, A.y + B.x};
  }
  __syncthreads();
  const int line = tid % lpw, tl = tid / lpw;
  fft_lines<T, true>(buf, lpw, line, tl, a, tw, line < np && tl < a.tpl);
  T* o = out + l0 * n;
  const T* b = base ? base + l0 * n : nullptr;
  for (int q = tid; q < nl * n; q += blockDim.x) {
    const int l = q / n, e = q - (q / n) * n;
    const T* z = reinterpret_cast<const T*>(&buf[e * lpw + (l >> 1)]);
    const T v = scale * z[l & 1];
    o[q] = b ? b[q] + v : v;
  }
}


// MODE 0: forward, 1: inverse, 2/3: forward * prefac * inverse (x axis only;
// 3 with the exact-symbol multiplier)
template <typename T, int MODE>
__global__ void strided_kernel(cpx<T>* spec, int64_t nlines, LineMap lm, Axis<T> a, int lpw, Prefac pf) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<T>* tw = reinterpret_cast<cpx<T>*>(smem_raw);
  cpx<T>* buf = tw + a.n;
  const int n = a.n, tid = threadIdx.x;
  load_twiddles(tw, a);
  const int64_t L0 = (int64_t)blockIdx.x * lpw;
  const int nl = (int)min((int64_t)lpw, nlines - L0);
  // blockDim.x is a multiple of lpw, so each thread always serves line tid % lpw.
  const int line = tid % lpw, tl = tid / lpw, jstep = blockDim.x / lpw;
  const int64_t start = line < nl ? line_start(lm, L0 + line) : 0;
  if (line < nl)
    for (int j = tl; j < n; j += jstep) buf[j * lpw + line] = spec[start + j * lm.es];
  __syncthreads();
  const bool active = line < nl && tl < a.tpl;
  if (MODE == 1) {
    fft_lines<T, true>(buf, lpw, line, tl, a, tw, active);
  } else {
    fft_lines<T, false>(buf, lpw, line, tl, a, tw, active);
  }
  if (MODE >= 2) {
    const T sc = static_cast<T>(pf.scale);
    if (line < nl) {
      const int64_t L = L0 + line;
      const int64_t outer = L / lm.nh;
      const int jz = (int)(L - outer * lm.nh);
      const int jy = (int)(outer % pf.ny);
      const float ssc = symbol_scalar(pf, outer / pf.ny);
      for (int j = tl; j < n; j += jstep) {
        const T f = pf.etd     ? static_cast<T>(etd1_f32(j, jy, jz, pf, ssc)) * sc
                    : pf.exact ? static_cast<T>(prefac_f64(j, jy, jz, pf, outer / pf.ny) * pf.scale)
                               : static_cast<T>(prefac_f32(j, jy, jz, pf, ssc)) * sc;
        buf[j * lpw + line] = cscale(buf[j * lpw + line], f);
      }
    }
    __syncthreads();
    fft_lines<T, true>(buf, lpw, line, tl, a, tw, active);
  }
  if (line < nl)
    for (int j = tl; j < n; j += jstep) spec[start + j * lm.es] = buf[j * lpw + line];
}


template <typename T, int N>
__global__ __launch_bounds__((ZGeo<T, N>::THREADS)) void z_r2c_fast(const T* __restrict__ x, cpx<T>* __restrict__ spec,
                                                                  int64_t nlines, int64_t pitch,
                                                                  const cpx<T>* __restrict__ twg) {
  __shared__ __align__(16) unsigned char smem[ZTile<T, N>::LDS_BYTES];
  z_r2c_tile<T, N>(x, spec, nlines, pitch, twg, blockIdx.x, threadIdx.x, smem);
}

template <typename T, int N>
__global__ __launch_bounds__((ZGeo<T, N>::THREADS)) void z_c2r_fast(const cpx<T>* __restrict__ spec, T* out,
                                                                  const T* base, T scale, int64_t nlines,
                                                                  int64_t pitch, const cpx<T>* __restrict__ twg,
                                                                  SumRec* rec) {
  __shared__ __align__(16) unsigned char smem[ZTile<T, N>::LDS_BYTES];
  z_c2r_tile<T, N>(spec, out, base, scale, nlines, pitch, twg, blockIdx.x, threadIdx.x, smem, rec);
}

// Fused CH right-hand side + z R2C, float32, nz = 512 (evx_rz_tile.hpp).
__global__ __launch_bounds__(RzGeo::THREADS) void ch_rhs_z_r2c_512(const float* __restrict__ u,
                                                                   cpx<float>* __restrict__ spec, RzMap m,
                                                                   Geo<float> g, ChW<float> W,
                                                                   const double* __restrict__ chan,
                                                                   const cpx<float>* __restrict__ twg) {
  __shared__ __align__(16) float lds[RzGeo::LDS_FLOATS];
  rz_tile_body(u, spec, m, g, W, chan, twg, xcd_tile(blockIdx.x, gridDim.x), threadIdx.x, lds);
}

template <typename T, int N, int MODE>
__global__ __launch_bounds__((SGeo<T, N, MODE>::THREADS)) void strided_fast(cpx<T>* __restrict__ spec, int64_t nlines,
                                                                    LineMap lm, const cpx<T>* __restrict__ twg,
                                                                    Prefac pf) {
  __shared__ __align__(16) unsigned char smem[STile<T, N, MODE>::LDS_BYTES];
  strided_tile<T, N, MODE>(spec, nlines, lm, twg, pf, xcd_tile(blockIdx.x, gridDim.x), threadIdx.x, smem);
}

// The persistent x sweep: one resident round of blocks (86 VGPRs, one
// 1024-thread block per CU) walking the bricks.
// MIR = 1: mirrored x sweep (a separate instantiation, so that the
// unmirrored sweep carries no mirror code).
template <typename T, int N, int MODE, int MIR = 0>
__global__ __launch_bounds__((SGeo<T, N, MODE>::THREADS), 4) void strided_x_pipe_kernel(
    cpx<T>* __restrict__ spec, int64_t nlines, LineMap lm, const cpx<T>* __restrict__ twg, Prefac pf,
    uint32_t nbricks) {
  __shared__ __align__(16) unsigned char smem[STile<T, N, MODE>::LDS_BYTES];
  strided_x_pipe<T, N, MODE, MIR>(spec, nlines, lm, twg, pf, nbricks, blockIdx.x, gridDim.x, threadIdx.x, smem);
}

template <typename T, int N, int MODE, int MIR>
int64_t x_pipe_resident() {
  static const int64_t resident = [] {
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, strided_x_pipe_kernel<T, N, MODE, MIR>,
                                                     SGeo<T, N, MODE>::THREADS, 0) != hipSuccess ||
        per < 1 || cus < 1)
      return (int64_t)256;
    return (int64_t)per * cus;
  }();
  return resident;
}

// Sizes with size-specialised kernels.
bool fast_size(int n) {
  switch (n) {
#define X(NN) case NN:
    EVX_FAST_SIZES(X)
#undef X
    return true;
    default:
      return false;
  }
}

template <typename T>
bool fast_z_r2c(int n, const void* x, void* spec, int64_t nlines, int64_t pitch, const void* tw, hipStream_t s) {
  switch (n) {
#define X(NN)                                                                                               \
  case NN: {                                                                                                \
    using G = ZGeo<T, NN>;                                                                                  \
    const int64_t blocks = G::blocks(nlines);                                                               \
    hipLaunchKernelGGL((z_r2c_fast<T, NN>), dim3((unsigned)blocks), dim3(G::THREADS), 0, s,                 \
                       static_cast<const T*>(x), static_cast<cpx<T>*>(spec), nlines, pitch,                 \
                       static_cast<const cpx<T>*>(tw));                                                     \
    return true;                                                                                            \
  }
    EVX_FAST_SIZES(X)
#undef X
    default:
      return false;
  }
}

template <typename T>
bool fast_z_c2r(int n, const void* spec, void* out, const void* base, double scale, int64_t nlines,
                int64_t pitch, const void* tw, hipStream_t s, SumRec* rec = nullptr) {
  switch (n) {
#define X(NN)                                                                                               \
  case NN: {                                                                                                \
    using G = ZGeo<T, NN>;                                                                                  \
    const int64_t blocks = G::blocks(nlines);                                                               \
    hipLaunchKernelGGL((z_c2r_fast<T, NN>), dim3((unsigned)blocks), dim3(G::THREADS), 0, s,                 \
                       static_cast<const cpx<T>*>(spec), static_cast<T*>(out), static_cast<const T*>(base), \
                       static_cast<T>(scale), nlines, pitch, static_cast<const cpx<T>*>(tw), rec);          \
    return true;                                                                                            \
  }
    EVX_FAST_SIZES(X)
#undef X
    default:
      return false;
  }
}

template <typename T, int MODE>
bool fast_strided(int n, void* spec, int64_t nlines, const LineMap& lm, const void* tw, const Prefac& pf,
                  hipStream_t s) {
  switch (n) {
#define X(NN)                                                                                               \
  case NN: {                                                                                                \
    using G = SGeo<T, NN, MODE>;                                                                            \
    if (!brick_offsets_fit<T>(nlines, lm, NN, G::TPL, G::TL)) return false;                                 \
    const int64_t blocks = (nlines + G::TL - 1) / G::TL;                                                    \
    if constexpr (SWave<T, NN, MODE>::ON) {                                                                 \
      const int64_t res = lm.mirror ? x_pipe_resident<T, NN, MODE, 1>() : x_pipe_resident<T, NN, MODE, 0>();  \
      const int64_t grid = blocks < res ? blocks : res;                                                     \
      if (lm.mirror)                                                                                        \
        hipLaunchKernelGGL((strided_x_pipe_kernel<T, NN, MODE, 1>), dim3((unsigned)grid), dim3(G::THREADS), 0, \
                           s, static_cast<cpx<T>*>(spec), nlines, lm, static_cast<const cpx<T>*>(tw), pf,    \
                           (uint32_t)blocks);                                                               \
      else                                                                                                  \
        hipLaunchKernelGGL((strided_x_pipe_kernel<T, NN, MODE, 0>), dim3((unsigned)grid), dim3(G::THREADS), 0, \
                           s, static_cast<cpx<T>*>(spec), nlines, lm, static_cast<const cpx<T>*>(tw), pf,    \
                           (uint32_t)blocks);                                                               \
    } else {                                                                                                \
      hipLaunchKernelGGL((strided_fast<T, NN, MODE>), dim3((unsigned)blocks), dim3(G::THREADS), 0, s,       \
                         static_cast<cpx<T>*>(spec), nlines, lm, static_cast<const cpx<T>*>(tw), pf);        \
    }                                                                                                       \
    return true;                                                                                            \
  }
    EVX_FAST_SIZES(X)
#undef X
    default:
      return false;
  }
}


// ------------------------------------------------------------------- plan --
std::vector<int> factorize(int n) {
  std::vector<int> r;
  while (n % 8 == 0) { r.push_back(8); n /= 8; }
  while (n % 4 == 0) { r.push_back(4); n /= 4; }
  while (n % 2 == 0) { r.push_back(2); n /= 2; }
  for (int p = 3; p <= n; p += 2)
    while (n % p == 0) { r.push_back(p); n /= p; }
  return r;
}

// Threads per line such that no pass needs more than EMAX values per thread.
int min_tpl(int n, const std::vector<int>& rad) {
  int tpl = 1;
  for (;;) {
    bool ok = true;
    for (int R : rad) {
      const int m = n / R;
      if (((m + tpl - 1) / tpl) * R > EMAX) { ok = false; break; }
    }
    if (ok) return tpl;
    ++tpl;
  }
}

// Largest threads-per-line <= tmax for which every pass is balanced (each
// thread gets the same butterfly count); tpl_min if none is.
int pick_tpl(int n, const std::vector<int>& rad, int tpl_min, int tmax) {
  if (rad.empty()) return 1;
  for (int t = tmax; t >= tpl_min; --t) {
    bool ok = true;
    for (int R : rad)
      if ((n / R) % t) { ok = false; break; }
    if (ok) return t;
  }
  return tpl_min;
}

struct AxisHost {
  int n = 1;
  std::vector<int> rad;
  int tpl_min = 1;
  void* tw = nullptr;  // device
};

}  // namespace
}  // namespace evx

struct evx_fft_plan {
  int dtype;
  evx::Dims d;
  int64_t nh;
  int64_t pitch;  // row pitch (complex values) of the workspace spectrum, >= nh
  evx::AxisHost ax[3];  // x, y, z (z: complex length nz)
  // One-shot stage trace: events recorded around the kernels of the next step.
  // One-shot launch trace of the next step: event 0 before the first kernel,
  // event i after the i-th kernel, whose kind is kinds[i-1].
  hipEvent_t trace[64] = {};
  int ntrace = 0, tpos = 0, nkinds = 0;
  int kinds[64] = {};
  bool fast = true;  // use the size-specialised kernels where one exists
  bool exact_symbol = false;  // float64 multiplier from exact wavenumbers
  int slabs = 0;               // > 1: semi-implicit CH steps run x-slab by x-slab (one-channel plans)
  bool fused_rhs = true;       // CH steps: fused right-hand side + z R2C where it applies
  bool step_sums = true;       // one-channel steps: mass and NaN/Inf records from the z C2R epilogue
  const void* sums_ws = nullptr;  // workspace holding the records of the last recording step
};

namespace evx {
namespace {

template <typename T>
Axis<T> axis_dev(const AxisHost& h, int tpl) {
  Axis<T> a;
  a.n = h.n;
  a.npass = (int)h.rad.size();
  for (int i = 0; i < a.npass; ++i) a.radix[i] = h.rad[i];
  a.tpl = tpl;
  a.tw = static_cast<const cpx<T>*>(h.tw);
  return a;
}

template <typename T>
int make_twiddles(AxisHost& h) {
  std::vector<cpx<T>> t(h.n);
  const long double tau = 6.283185307179586476925286766559005768L;
  for (int m = 0; m < h.n; ++m) {
    const long double ang = tau * (long double)m / (long double)h.n;
    t[m].x = static_cast<T>(cosl(ang));
    t[m].y = static_cast<T>(-sinl(ang));
  }
  if (hipMalloc(&h.tw, sizeof(cpx<T>) * h.n) != hipSuccess) {
    h.tw = nullptr;
    return fail(EVX_ERR_NOMEM, "hipMalloc of %d twiddles failed", h.n);
  }
  if (hipMemcpy(h.tw, t.data(), sizeof(cpx<T>) * h.n, hipMemcpyHostToDevice) != hipSuccess)
    return fail(EVX_ERR_HIP, "twiddle upload failed");
  return EVX_OK;
}

// Launch geometry: lines per workgroup and threads per line for an axis.
struct Geom {
  int lpw, tpl, threads;
  size_t lds;
};

template <typename T>
Geom z_geom(const AxisHost& h, int64_t npairs) {
  Geom g;
  g.tpl = pick_tpl(h.n, h.rad, h.tpl_min, 64);
  // Prefer ~256 threads per workgroup within the LDS budget.
  const size_t line_bytes = sizeof(cpx<T>) * h.n;
  int lpw = 256 / g.tpl;
  if (lpw < 1) lpw = 1;
  while (lpw > 1 && line_bytes * (lpw + 1) > LDS_BUDGET) --lpw;
  if (lpw > npairs) lpw = (int)(npairs < 1 ? 1 : npairs);
  g.lpw = lpw;
  g.threads = lpw * g.tpl;
  if (g.threads > 1024) g.threads = 1024;
  g.lds = line_bytes * (lpw + 1);
  return g;
}

template <typename T>
Geom strided_geom(const AxisHost& h, int64_t nlines) {
  Geom g;
  g.tpl = pick_tpl(h.n, h.rad, h.tpl_min, 16);
  const size_t line_bytes = sizeof(cpx<T>) * h.n;
  int lpw = 16;  // 16 adjacent lines: 128 B (f32) / 256 B (f64) contiguous runs
  if (lpw * g.tpl > 1024) lpw = 1024 / g.tpl;
  if (lpw < 1) lpw = 1;
  while (lpw * g.tpl < 256 && line_bytes * (2 * lpw + 1) <= LDS_BUDGET && 2 * lpw * g.tpl <= 1024) lpw *= 2;
  while (lpw > 1 && line_bytes * (lpw + 1) > LDS_BUDGET) --lpw;
  if (lpw > nlines) lpw = (int)(nlines < 1 ? 1 : nlines);
  g.lpw = lpw;
  g.threads = lpw * g.tpl;
  g.lds = line_bytes * (lpw + 1);
  return g;
}

template <typename T>
// xa/xb (one-channel plans only) restrict a z or y sweep to the x-slab
// [xa, xb); the pointers are offset to the slab's first line.
int run_z_r2c(evx_fft_plan* p, const void* x, void* spec, int64_t pitch, hipStream_t s, int64_t xa = 0,
              int64_t xb = -1) {
  const AxisHost& h = p->ax[2];
  int64_t nlines = p->d.c * p->d.nx * p->d.ny;
  if (xb >= 0) {
    x = static_cast<const T*>(x) + xa * p->d.ny * p->d.nz;
    spec = static_cast<cpx<T>*>(spec) + xa * p->d.ny * pitch;
    nlines = (xb - xa) * p->d.ny;
  }
  if (p->fast && fast_z_r2c<T>(h.n, x, spec, nlines, pitch, h.tw, s)) return check_launch("z_r2c_fast");
  const Geom g = z_geom<T>(h, (nlines + 1) / 2);
  const int64_t blocks = (nlines + 2 * g.lpw - 1) / (2 * g.lpw);
  hipLaunchKernelGGL(z_r2c_kernel<T>, dim3((unsigned)blocks), dim3(g.threads), g.lds, s, static_cast<const T*>(x),
                     static_cast<cpx<T>*>(spec), nlines, (int)p->nh, pitch, axis_dev<T>(h, g.tpl), g.lpw);
  return check_launch("z_r2c_kernel");
}

template <typename T>
int run_z_c2r(evx_fft_plan* p, const void* spec, void* out, const void* base, double scale, int64_t pitch,
              hipStream_t s, int64_t xa = 0, int64_t xb = -1, SumRec* rec = nullptr) {
  const AxisHost& h = p->ax[2];
  int64_t nlines = p->d.c * p->d.nx * p->d.ny;
  if (xb >= 0) {
    const int64_t off = xa * p->d.ny * p->d.nz;
    out = static_cast<T*>(out) + off;
    if (base) base = static_cast<const T*>(base) + off;
    spec = static_cast<const cpx<T>*>(spec) + xa * p->d.ny * pitch;
    nlines = (xb - xa) * p->d.ny;
  }
  if (p->fast && fast_z_c2r<T>(h.n, spec, out, base, scale, nlines, pitch, h.tw, s, rec))
    return check_launch("z_c2r_fast");
  if (rec) return fail(EVX_ERR_INVALID, "step sums need the size-specialised z C2R");
  const Geom g = z_geom<T>(h, (nlines + 1) / 2);
  const int64_t blocks = (nlines + 2 * g.lpw - 1) / (2 * g.lpw);
  hipLaunchKernelGGL(z_c2r_kernel<T>, dim3((unsigned)blocks), dim3(g.threads), g.lds, s,
                     static_cast<const cpx<T>*>(spec), static_cast<T*>(out), static_cast<const T*>(base),
                     static_cast<T>(scale), nlines, (int)p->nh, pitch, axis_dev<T>(h, g.tpl), g.lpw);
  return check_launch("z_c2r_kernel");
}

// mirror (x sweeps of the size-specialised kernels only): see LineMap.
template <typename T, int MODE>
int run_strided(evx_fft_plan* p, int axis, void* spec, const Prefac& pf, int64_t pitch, hipStream_t s,
                int64_t xa = 0, int64_t xb = -1, int mirror = 0) {
  const AxisHost& h = p->ax[axis];
  const Dims& d = p->d;
  LineMap lm;
  lm.nh = p->nh;
  lm.mirror = mirror;
  int64_t nlines;
  if (axis == 1) {
    nlines = d.c * d.nx * p->nh;
    if (xb >= 0) {  // y lines of the x-slab [xa, xb)
      spec = static_cast<cpx<T>*>(spec) + xa * d.ny * pitch;
      nlines = (xb - xa) * p->nh;
    }
    lm.G = 1;
    lm.GM = d.ny * pitch;
    lm.OM = 0;
    lm.es = pitch;
  } else {
    nlines = d.c * d.ny * p->nh;
    lm.G = d.ny;
    lm.GM = d.nx * d.ny * pitch;
    lm.OM = pitch;
    lm.es = d.ny * pitch;
  }
  if (h.n == 1 && MODE < 2) return EVX_OK;  // length-1 transform is the identity
  if (p->fast && fast_strided<T, MODE>(h.n, spec, nlines, lm, h.tw, pf, s)) return check_launch("strided_fast");
  if (mirror) return fail(EVX_ERR_UNSUPPORTED, "mirrored sweeps need a size-specialised transform");
  const Geom g = strided_geom<T>(h, nlines);
  const int64_t blocks = (nlines + g.lpw - 1) / g.lpw;
  hipLaunchKernelGGL((strided_kernel<T, MODE>), dim3((unsigned)blocks), dim3(g.threads), g.lds, s,
                     static_cast<cpx<T>*>(spec), nlines, lm, axis_dev<T>(h, g.tpl), g.lpw, pf);
  return check_launch("strided_kernel");
}

Prefac make_prefac(const evx_fft_plan* p, int ch, double sym_scalar, double dt, const double spacing[3]) {
  Prefac pf;
  pf.ch = ch;
  pf.sc = static_cast<float>(sym_scalar);
  pf.dt = static_cast<float>(dt);
  pf.twopi = static_cast<float>(2.0 * M_PI);
  const int64_t n[3] = {p->d.nx, p->d.ny, p->d.nz};
  for (int a = 0; a < 3; ++a) pf.inv_nd[a] = static_cast<float>(1.0 / ((double)n[a] * spacing[a]));
  pf.nx = (int)p->d.nx;
  pf.ny = (int)p->d.ny;
  pf.nz = (int)p->d.nz;
  pf.scale = 1.0 / ((double)p->d.nx * (double)p->d.ny * (double)p->d.nz);
  pf.chan = nullptr;
  pf.A = 0.0;
  pf.exact = p->exact_symbol ? 1 : 0;
  pf.etd = 0;
  pf.scd = sym_scalar;
  pf.dtd = dt;
  for (int a = 0; a < 3; ++a) pf.inv_ndd[a] = 1.0 / ((double)n[a] * spacing[a]);
  return pf;
}

size_t align_up(size_t v) { return (v + 255) & ~size_t(255); }

size_t real_bytes(const evx_fft_plan* p) {
  return align_up((size_t)p->d.c * p->d.vol() * (p->dtype == EVX_F32 ? 4 : 8));
}
// Bytes of the workspace spectrum (rows of p->pitch complex values).
size_t spec_bytes(const evx_fft_plan* p) {
  return align_up((size_t)p->d.c * p->d.nx * p->d.ny * p->pitch * (p->dtype == EVX_F32 ? 8 : 16));
}

// Records of the z C2R epilogue: one per block of the size-specialised z C2R
// of a one-channel plan whose blocks are whole wavefronts.
template <typename T>
int64_t rec_blocks_t(const evx_fft_plan* p) {
  const int64_t nlines = p->d.c * p->d.nx * p->d.ny;
  switch (p->ax[2].n) {
#define X(NN)                                                              \
  case NN:                                                                 \
    return ZGeo<T, NN>::THREADS % 64 == 0 ? ZGeo<T, NN>::blocks(nlines) : 0;
    EVX_FAST_SIZES(X)
#undef X
    default:
      return 0;
  }
}
int64_t rec_blocks(const evx_fft_plan* p) {
  if (p->d.c != 1) return 0;
  return p->dtype == EVX_F32 ? rec_blocks_t<float>(p) : rec_blocks_t<double>(p);
}
size_t rec_bytes(const evx_fft_plan* p) { return align_up((size_t)rec_blocks(p) * sizeof(SumRec)); }
// Bytes of the forward workspace: right-hand-side staging, spectrum, records.
size_t fwd_ws_bytes(const evx_fft_plan* p) { return real_bytes(p) + spec_bytes(p) + rec_bytes(p); }
// Records of a step on workspace ws, or nullptr when this step keeps none.
SumRec* step_recs(evx_fft_plan* p, void* ws) {
  if (!p->step_sums || !p->fast || p->slabs > 1 || rec_blocks(p) == 0) return nullptr;
  return reinterpret_cast<SumRec*>(static_cast<char*>(ws) + real_bytes(p) + spec_bytes(p));
}

__global__ __launch_bounds__(256) void fold_recs(const SumRec* __restrict__ rec, int64_t n, double* out,
                                                 int32_t* nonfinite) {
  __shared__ SumPair wsum[4];
  __shared__ int wflag[4];
  SumPair acc{0.0, 0.0};
  int f = 0;
  for (int64_t i = threadIdx.x; i < n; i += 256) {
    acc = sum_merge(acc, SumPair{rec[i].s, rec[i].c});
    f |= rec[i].nan | (rec[i].inf << 1);
  }
  acc = wave_sum(acc);
  const int wf = (__any(f & 1) ? 1 : 0) | (__any(f & 2) ? 2 : 0);
  if ((threadIdx.x & 63) == 0) {
    wsum[threadIdx.x >> 6] = acc;
    wflag[threadIdx.x >> 6] = wf;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    SumPair b = wsum[0];
    int g = wflag[0];
    for (int i = 1; i < 4; ++i) {
      b = sum_merge(b, wsum[i]);
      g |= wflag[i];
    }
    out[0] = b.s + b.c;
    // set only, as evx_sum_f64 does: the caller clears the flags
    if (nonfinite && (g & 1)) nonfinite[0] = 1;
    if (nonfinite && (g & 2)) nonfinite[1] = 1;
  }
}

// The x sweep: forward x transform, spectral multiplier, inverse x transform.
template <typename T>
int run_x_multiplier(evx_fft_plan* p, void* spec, const Prefac& pf, int64_t pt, hipStream_t s, int mirror = 0) {
  if (pf.etd) return run_strided<T, 4>(p, 0, spec, pf, pt, s, 0, -1, mirror);
  if (pf.exact) return run_strided<T, 3>(p, 0, spec, pf, pt, s, 0, -1, mirror);
  return run_strided<T, 2>(p, 0, spec, pf, pt, s, 0, -1, mirror);
}

// Launch trace (armed by evx_fft_plan_trace_next_step).  Kernel kinds:
enum { TK_RHS = 0, TK_Z_R2C = 1, TK_Y_FWD = 2, TK_X = 3, TK_Y_INV = 4, TK_Z_C2R = 5, TK_RHS_Z_R2C = 6,
       TK_OTHER = 7 };

int trace_start(evx_fft_plan* p, hipStream_t s) {
  p->tpos = 0;
  p->nkinds = 0;
  if (p->ntrace > 0 && p->trace[0]) {
    if (hipEventRecord(p->trace[0], s) != hipSuccess) return fail(EVX_ERR_HIP, "hipEventRecord failed");
    p->tpos = 1;
  }
  return EVX_OK;
}

int trace_tick(evx_fft_plan* p, int kind, hipStream_t s) {
  if (p->tpos > 0 && p->tpos < p->ntrace && p->trace[p->tpos]) {
    if (hipEventRecord(p->trace[p->tpos], s) != hipSuccess) return fail(EVX_ERR_HIP, "hipEventRecord failed");
    p->kinds[p->tpos - 1] = kind;
    ++p->tpos;
  }
  return EVX_OK;
}

void trace_end(evx_fft_plan* p) {
  p->nkinds = p->tpos > 0 ? p->tpos - 1 : 0;
  p->ntrace = 0;
  p->tpos = 0;
}

// rhs == nullptr: the z half spectrum of the right-hand side is already in
// spec (the fused first sweep).
template <typename T>
int spectral_update(evx_fft_plan* p, const void* rhs, void* spec, const void* u, void* u_out, const Prefac& pf,
                    hipStream_t s, SumRec* rec = nullptr) {
  const int64_t pt = p->pitch;
  if (rhs) {
    EVX_TRY(run_z_r2c<T>(p, rhs, spec, pt, s));
    EVX_TRY(trace_tick(p, TK_Z_R2C, s));
  }
  EVX_TRY((run_strided<T, 0>(p, 1, spec, pf, pt, s)));
  EVX_TRY(trace_tick(p, TK_Y_FWD, s));
  EVX_TRY(run_x_multiplier<T>(p, spec, pf, pt, s));
  EVX_TRY(trace_tick(p, TK_X, s));
  EVX_TRY((run_strided<T, 1>(p, 1, spec, pf, pt, s)));
  EVX_TRY(trace_tick(p, TK_Y_INV, s));
  EVX_TRY(run_z_c2r<T>(p, spec, u_out, u, 1.0, pt, s, 0, -1, rec));
  EVX_TRY(trace_tick(p, TK_Z_C2R, s));
  trace_end(p);
  return EVX_OK;
}

// Semi-implicit CH step x-slab by x-slab: for each slab, the right-hand side,
// the z R2C and the forward y sweep; then the x sweep over the whole
// spectrum; then, for each slab, the inverse y sweep and the z C2R + add.
// The z and y sweeps couple only lines inside one x-slab, so a slab's
// staging and spectrum (a few tens of MB) are re-read from the last-level
// cache instead of HBM; only the x sweep and the field itself stream from
// HBM.
template <typename T>
int ch_slab_step(evx_fft_plan* p, const void* u, void* u_out, void* rhs, void* spec, const Prefac& pf,
                 const double dd[3], const double dd2[3], const evx_imex_params* prm, hipStream_t s) {
  const Dims& d = p->d;
  const int64_t pt = p->pitch;
  const int64_t xc = stencil_ch_xchunk(p->dtype, d);
  int64_t w = (d.nx + p->slabs - 1) / p->slabs;
  w = (w + xc - 1) / xc * xc;  // slab width: a whole number of CH tile chunks
  EVX_TRY(trace_start(p, s));
  for (int64_t xa = 0; xa < d.nx; xa += w) {
    const int64_t xb = xa + w < d.nx ? xa + w : d.nx;
    EVX_TRY(stencil_rhs_ch_periodic(p->dtype, u, rhs, d, dd, dd2, prm->eps, prm->D, prm->chan_params, s, xa, xb));
    EVX_TRY(trace_tick(p, TK_RHS, s));
    EVX_TRY(run_z_r2c<T>(p, rhs, spec, pt, s, xa, xb));
    EVX_TRY(trace_tick(p, TK_Z_R2C, s));
    EVX_TRY((run_strided<T, 0>(p, 1, spec, pf, pt, s, xa, xb)));
    EVX_TRY(trace_tick(p, TK_Y_FWD, s));
  }
  EVX_TRY(run_x_multiplier<T>(p, spec, pf, pt, s));
  EVX_TRY(trace_tick(p, TK_X, s));
  for (int64_t xa = 0; xa < d.nx; xa += w) {
    const int64_t xb = xa + w < d.nx ? xa + w : d.nx;
    EVX_TRY((run_strided<T, 1>(p, 1, spec, pf, pt, s, xa, xb)));
    EVX_TRY(trace_tick(p, TK_Y_INV, s));
    EVX_TRY(run_z_c2r<T>(p, spec, u_out, u, 1.0, pt, s, xa, xb));
    EVX_TRY(trace_tick(p, TK_Z_C2R, s));
  }
  trace_end(p);
  return EVX_OK;
}

int check_plan(const evx_fft_plan* p) {
  if (!p) return fail(EVX_ERR_INVALID, "null plan");
  return EVX_OK;
}

// Launch geometry of the fused right-hand side + z R2C sweep: about one
// resident round of blocks (one per CU), tiles of 16 rows by x-chunks.
RzMap rz_map(const evx_fft_plan* p, unsigned* blocks) {
  static const int resident = [] {
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, ch_rhs_z_r2c_512, RzGeo::THREADS, 0) != hipSuccess ||
        per < 1 || cus < 1)
      return 256;
    return per * cus;
  }();
  return rz_make_map(p->d, p->pitch, resident, blocks);
}

// Whether a CH step of this plan takes the fused first sweep: float32,
// nz = 512 with the specialised kernels, ny a multiple of the tile's 16
// rows, a 16-byte aligned state and index ranges that fit the kernel's
// 32-bit tile numbering.
bool rz_applies(const evx_fft_plan* p, const void* u) {
  if (!p->fused_rhs || !p->fast || p->dtype != EVX_F32 || p->d.nz != RzGeo::N) return false;
  if (p->d.ny % RzGeo::BY != 0 || (reinterpret_cast<uintptr_t>(u) & 15u) != 0) return false;
  const int64_t tiles = p->d.c * (p->d.ny / RzGeo::BY) * p->d.nx;
  return tiles < (int64_t(1) << 31) && p->ax[2].tw != nullptr;
}

int run_rz(evx_fft_plan* p, const void* u, void* spec, const double dd[3], const double dd2[3],
           const evx_imex_params* prm, hipStream_t s) {
  unsigned blocks = 0;
  const RzMap m = rz_map(p, &blocks);
  const Geo<float> g = make_geo_from<float>(dd, dd2);
  const ChCoef<float> k{static_cast<float>(18.0 / prm->eps), static_cast<float>(2.0 * prm->eps),
                        static_cast<float>(prm->D)};
  hipLaunchKernelGGL(ch_rhs_z_r2c_512, dim3(blocks), dim3(RzGeo::THREADS), 0, s, static_cast<const float*>(u),
                     static_cast<cpx<float>*>(spec), m, g, ch_weights(g, k), prm->chan_params,
                     static_cast<const cpx<float>*>(p->ax[2].tw));
  return check_launch("ch_rhs_z_r2c_512");
}

int imex_common(evx_fft_plan* p, const void* u, void* u_out, void* ws, const evx_imex_params* prm) {
  EVX_TRY(check_plan(p));
  if (!u || !u_out || !ws || !prm) return fail(EVX_ERR_INVALID, "null argument to imex step");
  for (int a = 0; a < 3; ++a)
    if (!(prm->spacing[a] > 0.0)) return fail(EVX_ERR_INVALID, "spacing must be positive");
  return EVX_OK;
}

// div_dx = 1/float(spacing), div_dx2 = 1/float(spacing)^2 in the working
// precision, as VoxelGrid.__init__ computes them (voxelgrid.py:29-30).
void geometry_from_spacing(int dtype, const double sp[3], double dd[3], double dd2[3]) {
  for (int a = 0; a < 3; ++a) {
    if (dtype == EVX_F32) {
      const float s = static_cast<float>(sp[a]);
      dd[a] = 1.0f / s;
      dd2[a] = 1.0f / (s * s);
    } else {
      dd[a] = 1.0 / sp[a];
      dd2[a] = 1.0 / (sp[a] * sp[a]);
    }
  }
}

}  // namespace
}  // namespace evx

using namespace evx;

extern "C" int evx_fft_plan_create(evx_fft_plan** out, int32_t dtype, int64_t channels, int64_t nx, int64_t ny,
                                   int64_t nz) {
  if (!out) return fail(EVX_ERR_INVALID, "null plan out-pointer");
  *out = nullptr;
  if (dtype != EVX_F32 && dtype != EVX_F64) return fail(EVX_ERR_INVALID, "bad dtype");
  if (channels < 1 || nx < 1 || ny < 1 || nz < 1) return fail(EVX_ERR_INVALID, "shape must be positive");
  if (nx > (1 << 20) || ny > (1 << 20) || nz > (1 << 20)) return fail(EVX_ERR_UNSUPPORTED, "axis too long");
  evx_fft_plan* p = new evx_fft_plan();
  p->dtype = dtype;
  p->d = Dims{channels, nx, ny, nz};
  p->nh = nz / 2 + 1;
  // Workspace spectrum rows are unpadded.  Rows padded to 128-byte
  // multiples (272 complex values at nz = 512) with bricks numbered over the
  // padded rows, so that every y/x brick starts 128-byte aligned, cut the y
  // sweeps' counted HBM traffic from 1.17x to 1.03x of the algorithmic bytes
  // but measured slower on MI355X at 512^3 float32: y sweeps 0.204 -> 0.218
  // and 0.205 -> 0.234 ms, z C2R 0.331 -> 0.354 ms, x sweep 0.268 -> 0.258
  // ms, step 1.356 -> 1.412 ms.  The odd row stride spreads a brick's rows
  // over the memory channels better than the aligned one.
  p->pitch = p->nh;
  const int64_t n[3] = {nx, ny, nz};
  for (int a = 0; a < 3; ++a) {
    AxisHost& h = p->ax[a];
    h.n = (int)n[a];
    h.rad = factorize(h.n);
    if (h.n == 1) h.rad.clear();
    for (int R : h.rad)
      if (R > 31) {
        evx_fft_plan_destroy(p);
        return fail(EVX_ERR_UNSUPPORTED, "axis length %d has a prime factor %d > 31", h.n, R);
      }
    if ((int)h.rad.size() > MAXPASS) {
      evx_fft_plan_destroy(p);
      return fail(EVX_ERR_UNSUPPORTED, "axis length %d needs too many passes", h.n);
    }
    h.tpl_min = min_tpl(h.n, h.rad);
    if (h.tpl_min > 1024) {
      evx_fft_plan_destroy(p);
      return fail(EVX_ERR_UNSUPPORTED, "axis length %d needs more than 1024 threads per line", h.n);
    }
    const size_t elem = dtype == EVX_F32 ? sizeof(cpx<float>) : sizeof(cpx<double>);
    if (elem * h.n * 2 > LDS_BUDGET) {
      evx_fft_plan_destroy(p);
      return fail(EVX_ERR_UNSUPPORTED, "axis length %d exceeds the LDS line budget", h.n);
    }
    const int st = dtype == EVX_F32 ? make_twiddles<float>(h) : make_twiddles<double>(h);
    if (st != EVX_OK) {
      evx_fft_plan_destroy(p);
      return st;
    }
  }
  *out = p;
  return EVX_OK;
}

extern "C" int evx_fft_plan_destroy(evx_fft_plan* p) {
  if (!p) return EVX_OK;
  for (auto& h : p->ax)
    if (h.tw) (void)hipFree(h.tw);
  delete p;
  return EVX_OK;
}

extern "C" int evx_fft_plan_workspace_bytes(const evx_fft_plan* p, size_t* bytes) {
  EVX_TRY(check_plan(p));
  if (!bytes) return fail(EVX_ERR_INVALID, "null out-pointer");
  *bytes = fwd_ws_bytes(p);
  return EVX_OK;
}

extern "C" int evx_fft_plan_spectrum_elems(const evx_fft_plan* p, int64_t* elems) {
  EVX_TRY(check_plan(p));
  if (!elems) return fail(EVX_ERR_INVALID, "null out-pointer");
  *elems = p->d.c * p->d.nx * p->d.ny * p->nh;
  return EVX_OK;
}

extern "C" int evx_rfftn(evx_fft_plan* p, const void* x, void* spec, void* stream) {
  EVX_TRY(check_plan(p));
  if (!x || !spec) return fail(EVX_ERR_INVALID, "null buffer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  Prefac pf{};
  if (p->dtype == EVX_F32) {
    EVX_TRY(run_z_r2c<float>(p, x, spec, p->nh, s));
    EVX_TRY((run_strided<float, 0>(p, 1, spec, pf, p->nh, s)));
    return run_strided<float, 0>(p, 0, spec, pf, p->nh, s);
  }
  EVX_TRY(run_z_r2c<double>(p, x, spec, p->nh, s));
  EVX_TRY((run_strided<double, 0>(p, 1, spec, pf, p->nh, s)));
  return run_strided<double, 0>(p, 0, spec, pf, p->nh, s);
}

extern "C" int evx_irfftn(evx_fft_plan* p, void* spec, void* y, void* stream) {
  EVX_TRY(check_plan(p));
  if (!y || !spec) return fail(EVX_ERR_INVALID, "null buffer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  Prefac pf{};
  const double scale = 1.0 / ((double)p->d.nx * (double)p->d.ny * (double)p->d.nz);
  if (p->dtype == EVX_F32) {
    EVX_TRY((run_strided<float, 1>(p, 0, spec, pf, p->nh, s)));
    EVX_TRY((run_strided<float, 1>(p, 1, spec, pf, p->nh, s)));
    return run_z_c2r<float>(p, spec, y, nullptr, scale, p->nh, s);
  }
  EVX_TRY((run_strided<double, 1>(p, 0, spec, pf, p->nh, s)));
  EVX_TRY((run_strided<double, 1>(p, 1, spec, pf, p->nh, s)));
  return run_z_c2r<double>(p, spec, y, nullptr, scale, p->nh, s);
}

namespace evx {
namespace {

// One semi-implicit (etd = 0) or exponential Euler (etd = 1) CH step.
int ch_spectral_step(evx_fft_plan* p, const void* u, void* u_out, void* ws, const evx_imex_params* prm,
                     hipStream_t s, int etd) {
  EVX_TRY(imex_common(p, u, u_out, ws, prm));
  if (!prm->chan_params && !(prm->eps != 0.0)) return fail(EVX_ERR_INVALID, "eps must be nonzero");
  double dd[3], dd2[3];
  geometry_from_spacing(p->dtype, prm->spacing, dd, dd2);
  void* rhs = ws;
  void* spec = static_cast<char*>(ws) + real_bytes(p);
  // problem_definition.py:303 -> ((( -2*eps)*D)*A) as a Python double
  Prefac pf = make_prefac(p, 1, ((-2.0 * prm->eps) * prm->D) * prm->A, prm->dt, prm->spacing);
  pf.chan = prm->chan_params;
  pf.A = prm->A;
  pf.etd = etd;
  if (etd) pf.exact = 0;
  if (p->slabs > 1 && p->d.c == 1)
    return p->dtype == EVX_F32 ? ch_slab_step<float>(p, u, u_out, rhs, spec, pf, dd, dd2, prm, s)
                               : ch_slab_step<double>(p, u, u_out, rhs, spec, pf, dd, dd2, prm, s);
  SumRec* rec = step_recs(p, ws);
  p->sums_ws = nullptr;
  EVX_TRY(trace_start(p, s));
  if (rz_applies(p, u)) {
    EVX_TRY(run_rz(p, u, spec, dd, dd2, prm, s));
    EVX_TRY(trace_tick(p, TK_RHS_Z_R2C, s));
    EVX_TRY(spectral_update<float>(p, nullptr, spec, u, u_out, pf, s, rec));
  } else {
    EVX_TRY(stencil_rhs_ch_periodic(p->dtype, u, rhs, p->d, dd, dd2, prm->eps, prm->D, prm->chan_params, s));
    EVX_TRY(trace_tick(p, TK_RHS, s));
    EVX_TRY(p->dtype == EVX_F32 ? spectral_update<float>(p, rhs, spec, u, u_out, pf, s, rec)
                                : spectral_update<double>(p, rhs, spec, u, u_out, pf, s, rec));
  }
  if (rec) p->sums_ws = ws;
  return EVX_OK;
}

int rd_spectral_step(evx_fft_plan* p, const void* u, void* u_out, void* ws, const evx_imex_params* prm,
                     hipStream_t s, int etd) {
  EVX_TRY(imex_common(p, u, u_out, ws, prm));
  if (prm->chan_params) return fail(EVX_ERR_UNSUPPORTED, "per-channel parameters are implemented for CH only");
  double dd[3], dd2[3];
  geometry_from_spacing(p->dtype, prm->spacing, dd, dd2);
  void* rhs = ws;
  void* spec = static_cast<char*>(ws) + real_bytes(p);
  EVX_TRY(trace_start(p, s));
  EVX_TRY(stencil_rhs_rd_periodic(p->dtype, u, rhs, p->d, dd2, prm->D, s));
  EVX_TRY(trace_tick(p, TK_RHS, s));
  // problem_definition.py:206 -> (-D)*A as a Python double
  Prefac pf = make_prefac(p, 0, (-prm->D) * prm->A, prm->dt, prm->spacing);
  pf.etd = etd;
  if (etd) pf.exact = 0;
  SumRec* rec = step_recs(p, ws);
  p->sums_ws = nullptr;
  EVX_TRY(p->dtype == EVX_F32 ? spectral_update<float>(p, rhs, spec, u, u_out, pf, s, rec)
                              : spectral_update<double>(p, rhs, spec, u, u_out, pf, s, rec));
  if (rec) p->sums_ws = ws;
  return EVX_OK;
}

// Mirror extension along x (boundary_conditions.py:61-71,133-144), applied
// to the z/y half spectrum of the right-hand side: planes [0, n) hold the
// transformed field; the kernel writes the extension planes after them.
// mode 0: [f, flip f] (zero flux), 1: [f, -flip f] (Dirichlet), 2:
// [f, 0, -flip f, 0] (staggered_x Dirichlet).  The z and y transforms act
// plane by plane and commute exactly with a sign flip, so this equals
// transforming the mirror-extended field.  Block (i, j) copies a chunk of
// extension plane j; `plane` counts complex elements.
template <typename T>
__global__ void mirror_spec_x_kernel(cpx<T>* __restrict__ spec, int64_t n, int64_t plane, int mode) {
  const int64_t j = blockIdx.y;
  cpx<T>* const dst = spec + (n + j) * plane;
  const bool zero = mode == 2 && (j == 0 || j == n + 1);
  const cpx<T>* const src = spec + (zero ? 0 : mode == 2 ? n - j : n - 1 - j) * plane;
  const T sg = mode == 0 ? T(1) : T(-1);
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < plane; r += (int64_t)gridDim.x * blockDim.x) {
    cpx<T> v = {T(0), T(0)};
    if (!zero) {
      const cpx<T> f = src[r];
      v = {sg * f.x, sg * f.y};
    }
    dst[r] = v;
  }
}

// Semi-implicit / exponential step with one non-periodic axis, x
// (timesteppers.py:85-89,195-199 with pad_fft_bc): the plan covers the
// mirror-extended x length (2n, or 2n+2 on staggered_x grids); the
// right-hand side of the n physical planes is evaluated with the ghost rules
// of `bc` and transformed along z and y on those planes only, the half
// spectrum is mirror-extended along x, the x sweep runs over the extended
// length, and only the n physical planes are transformed back and added
// to u.
int mirror_spectral_step(evx_fft_plan* p, int ch, const void* u, void* u_out, void* ws, const evx_imex_params* prm,
                         const evx_bc* bc, int staggered, hipStream_t s, int etd) {
  EVX_TRY(imex_common(p, u, u_out, ws, prm));
  p->sums_ws = nullptr;
  if (!bc) return fail(EVX_ERR_INVALID, "null bc");
  if (prm->chan_params) return fail(EVX_ERR_UNSUPPORTED, "per-channel parameters need a periodic grid");
  if (p->d.c != 1) return fail(EVX_ERR_UNSUPPORTED, "the mirrored spectral step handles one channel");
  if (ch && !(prm->eps != 0.0)) return fail(EVX_ERR_INVALID, "eps must be nonzero");
  const int kx = bc->kind[0];
  const bool dir = kx == EVX_BC_DIRICHLET || kx == EVX_BC_NODE_VALUE;
  const bool neu = kx == EVX_BC_NEUMANN || kx == EVX_BC_NODE_EXTRAP;
  if (bc->kind[1] != EVX_BC_PERIODIC || bc->kind[2] != EVX_BC_PERIODIC || !(dir || neu))
    return fail(EVX_ERR_UNSUPPORTED, "the mirrored spectral step needs x non-periodic and y, z periodic");
  if (staggered && !dir)
    return fail(EVX_ERR_UNSUPPORTED, "staggered_x spectral steps implement the Dirichlet extension only");
  const int64_t NX = p->d.nx;
  const int64_t n = staggered ? (NX - 2) / 2 : NX / 2;
  if (n < 1 || (staggered ? NX != 2 * n + 2 : NX != 2 * n))
    return fail(EVX_ERR_INVALID, "the plan's x length must be the mirror extension of the field");
  double dd[3], dd2[3];
  geometry_from_spacing(p->dtype, prm->spacing, dd, dd2);
  void* rhs = ws;
  void* spec = static_cast<char*>(ws) + real_bytes(p);
  const evx_grid g{1, n, p->d.ny, p->d.nz, {dd[0], dd[1], dd[2]}, {dd2[0], dd2[1], dd2[2]}};
  const evx_epilogue e{EVX_OUT_RHS, 0.0, nullptr, nullptr, nullptr};
  EVX_TRY(trace_start(p, s));
  EVX_TRY(ch ? evx_ch_rhs(p->dtype, u, rhs, &g, bc, prm->eps, prm->D, &e, s)
             : evx_rd_rhs(p->dtype, u, rhs, &g, bc, prm->D, &e, s));
  EVX_TRY(trace_tick(p, TK_RHS, s));
  const int mode = staggered ? 2 : (dir ? 1 : 0);
  // Cell-centred extensions (N = 2n) on a size-specialised x transform: the
  // x sweep reads the upper half as the mirror image of the lower half and
  // stores the lower half only; otherwise the half spectrum is extended first.
  const bool xmir = !staggered && p->fast && fast_size((int)NX);
  const int xmode = xmir ? (dir ? 2 : 1) : 0;
  const int64_t ext = staggered ? n + 2 : n;  // extension planes
  const int64_t cplane = p->d.ny * p->pitch;   // complex elements per spectrum plane
  if (ext > 65535) return fail(EVX_ERR_UNSUPPORTED, "the mirrored spectral step takes at most 65533 x planes");
  const dim3 mgrid((unsigned)std::min<int64_t>((cplane + 255) / 256, 1024), (unsigned)ext);
  // problem_definition.py:206,303 -> (-D)*A or ((-2*eps)*D)*A as Python doubles
  Prefac pf = make_prefac(p, ch, ch ? ((-2.0 * prm->eps) * prm->D) * prm->A : (-prm->D) * prm->A, prm->dt,
                          prm->spacing);
  pf.A = prm->A;
  pf.etd = etd;
  if (etd) pf.exact = 0;
  const int64_t pt = p->pitch;
  if (p->dtype == EVX_F32) {
    EVX_TRY(run_z_r2c<float>(p, rhs, spec, pt, s, 0, n));
    EVX_TRY(trace_tick(p, TK_Z_R2C, s));
    EVX_TRY((run_strided<float, 0>(p, 1, spec, pf, pt, s, 0, n)));
    EVX_TRY(trace_tick(p, TK_Y_FWD, s));
    if (!xmir) {
      hipLaunchKernelGGL(mirror_spec_x_kernel<float>, mgrid, dim3(256), 0, s, static_cast<cpx<float>*>(spec), n,
                         cplane, mode);
      EVX_TRY(check_launch("mirror_spec_x_kernel"));
      EVX_TRY(trace_tick(p, TK_OTHER, s));
    }
    EVX_TRY(run_x_multiplier<float>(p, spec, pf, pt, s, xmode));
    EVX_TRY(trace_tick(p, TK_X, s));
    EVX_TRY((run_strided<float, 1>(p, 1, spec, pf, pt, s, 0, n)));
    EVX_TRY(trace_tick(p, TK_Y_INV, s));
    EVX_TRY(run_z_c2r<float>(p, spec, u_out, u, 1.0, pt, s, 0, n));
  } else {
    EVX_TRY(run_z_r2c<double>(p, rhs, spec, pt, s, 0, n));
    EVX_TRY(trace_tick(p, TK_Z_R2C, s));
    EVX_TRY((run_strided<double, 0>(p, 1, spec, pf, pt, s, 0, n)));
    EVX_TRY(trace_tick(p, TK_Y_FWD, s));
    if (!xmir) {
      hipLaunchKernelGGL(mirror_spec_x_kernel<double>, mgrid, dim3(256), 0, s, static_cast<cpx<double>*>(spec), n,
                         cplane, mode);
      EVX_TRY(check_launch("mirror_spec_x_kernel"));
      EVX_TRY(trace_tick(p, TK_OTHER, s));
    }
    EVX_TRY(run_x_multiplier<double>(p, spec, pf, pt, s, xmode));
    EVX_TRY(trace_tick(p, TK_X, s));
    EVX_TRY((run_strided<double, 1>(p, 1, spec, pf, pt, s, 0, n)));
    EVX_TRY(trace_tick(p, TK_Y_INV, s));
    EVX_TRY(run_z_c2r<double>(p, spec, u_out, u, 1.0, pt, s, 0, n));
  }
  EVX_TRY(trace_tick(p, TK_Z_C2R, s));
  trace_end(p);
  return EVX_OK;
}

}  // namespace
}  // namespace evx

using namespace evx;

extern "C" int evx_ch_imex_step_mirror(evx_fft_plan* p, const void* u, void* u_out, void* ws,
                                       const evx_imex_params* prm, const evx_bc* bc, int32_t staggered, int32_t etd,
                                       void* stream) {
  return mirror_spectral_step(p, 1, u, u_out, ws, prm, bc, staggered, static_cast<hipStream_t>(stream), etd);
}

extern "C" int evx_rd_imex_step_mirror(evx_fft_plan* p, const void* u, void* u_out, void* ws,
                                       const evx_imex_params* prm, const evx_bc* bc, int32_t staggered, int32_t etd,
                                       void* stream) {
  return mirror_spectral_step(p, 0, u, u_out, ws, prm, bc, staggered, static_cast<hipStream_t>(stream), etd);
}

extern "C" int evx_ch_imex_step(evx_fft_plan* p, const void* u, void* u_out, void* ws, const evx_imex_params* prm,
                                void* stream) {
  return ch_spectral_step(p, u, u_out, ws, prm, static_cast<hipStream_t>(stream), 0);
}

extern "C" int evx_rd_imex_step(evx_fft_plan* p, const void* u, void* u_out, void* ws, const evx_imex_params* prm,
                                void* stream) {
  return rd_spectral_step(p, u, u_out, ws, prm, static_cast<hipStream_t>(stream), 0);
}

extern "C" int evx_ch_etd1_step(evx_fft_plan* p, const void* u, void* u_out, void* ws, const evx_imex_params* prm,
                                void* stream) {
  return ch_spectral_step(p, u, u_out, ws, prm, static_cast<hipStream_t>(stream), 1);
}

extern "C" int evx_rd_etd1_step(evx_fft_plan* p, const void* u, void* u_out, void* ws, const evx_imex_params* prm,
                                void* stream) {
  return rd_spectral_step(p, u, u_out, ws, prm, static_cast<hipStream_t>(stream), 1);
}

extern "C" int evx_ch_imex_vjp_workspace_bytes(const evx_fft_plan* p, size_t* bytes) {
  EVX_TRY(check_plan(p));
  if (!bytes) return fail(EVX_ERR_INVALID, "null argument");
  *bytes = fwd_ws_bytes(p) + ch_rhs_vjp_ws_bytes(p->dtype, p->d);
  return EVX_OK;
}

// lam0 = lam1 + J^T L(lam1) and the parameter gradients; w = L(lam1) goes
// through the same five spectral sweeps as the forward step with no base.
extern "C" int evx_ch_imex_step_vjp(evx_fft_plan* p, const void* u0, const void* u1, const void* lam1, void* lam0,
                                    double* dparams, void* ws, const evx_imex_params* prm, void* stream) {
  EVX_TRY(imex_common(p, u0, lam0, ws, prm));
  p->sums_ws = nullptr;
  if (!u1 || !lam1 || !dparams) return fail(EVX_ERR_INVALID, "null argument to imex step vjp");
  if (!prm->chan_params && !(prm->eps != 0.0 && prm->D != 0.0))
    return fail(EVX_ERR_INVALID, "eps and D must be nonzero");
  if (!(prm->dt != 0.0)) return fail(EVX_ERR_INVALID, "dt must be nonzero");
  if (lam0 == u0 || lam0 == u1) return fail(EVX_ERR_INVALID, "lam0 must not alias u0 or u1");
  hipStream_t s = static_cast<hipStream_t>(stream);
  double dd[3], dd2[3];
  geometry_from_spacing(p->dtype, prm->spacing, dd, dd2);
  void* w = ws;
  void* spec = static_cast<char*>(ws) + real_bytes(p);
  void* adj = static_cast<char*>(ws) + fwd_ws_bytes(p);
  Prefac pf = make_prefac(p, 1, ((-2.0 * prm->eps) * prm->D) * prm->A, prm->dt, prm->spacing);
  pf.chan = prm->chan_params;
  pf.A = prm->A;
  p->ntrace = 0;
  p->tpos = 0;
  EVX_TRY(p->dtype == EVX_F32 ? spectral_update<float>(p, lam1, spec, nullptr, w, pf, s)
                              : spectral_update<double>(p, lam1, spec, nullptr, w, pf, s));
  return ch_rhs_vjp_launch(p->dtype, u0, w, lam1, lam0, dparams, p->d, dd, dd2, prm->eps, prm->D,
                           prm->chan_params, u1, prm->dt, adj, s);
}

extern "C" int evx_fft_plan_trace_next_step(evx_fft_plan* p, void* const* events, int n_events) {
  EVX_TRY(check_plan(p));
  if (n_events < 0 || n_events > 64 || (n_events > 0 && !events)) return fail(EVX_ERR_INVALID, "bad trace events");
  for (int i = 0; i < n_events; ++i) p->trace[i] = static_cast<hipEvent_t>(events[i]);
  p->ntrace = n_events;
  return EVX_OK;
}

extern "C" int evx_fft_plan_trace_kinds(const evx_fft_plan* p, int32_t* kinds, int32_t* n) {
  EVX_TRY(check_plan(p));
  if (!n) return fail(EVX_ERR_INVALID, "null argument");
  if (kinds)
    for (int i = 0; i < p->nkinds; ++i) kinds[i] = p->kinds[i];
  *n = p->nkinds;
  return EVX_OK;
}

extern "C" int evx_fft_plan_step_sums(evx_fft_plan* p, const void* ws, double* out, int32_t* nonfinite,
                                      void* stream) {
  EVX_TRY(check_plan(p));
  if (!ws || !out) return fail(EVX_ERR_INVALID, "null argument");
  if (p->sums_ws == nullptr || p->sums_ws != ws)
    return fail(EVX_ERR_INVALID, "no step records on this workspace: the last semi-implicit step of this plan "
                                 "on it kept none (multi-channel, slab or mirrored steps, step sums off)");
  const SumRec* rec = reinterpret_cast<const SumRec*>(static_cast<const char*>(ws) + real_bytes(p) + spec_bytes(p));
  hipLaunchKernelGGL(fold_recs, dim3(1), dim3(256), 0, static_cast<hipStream_t>(stream), rec, rec_blocks(p), out,
                     nonfinite);
  return check_launch("fold_recs");
}

extern "C" int evx_fft_plan_set_option(evx_fft_plan* p, int option, int value) {
  EVX_TRY(check_plan(p));
  switch (option) {
    case EVX_OPT_SPECIALIZED: p->fast = value != 0; return EVX_OK;
    case EVX_OPT_EXACT_SYMBOL: p->exact_symbol = value != 0; return EVX_OK;
    case EVX_OPT_SLABS:
      if (value < 0 || value > 4096) return fail(EVX_ERR_INVALID, "slab count must be in [0, 4096]");
      p->slabs = value;
      return EVX_OK;
    case EVX_OPT_FUSED_RHS: p->fused_rhs = value != 0; return EVX_OK;
    case EVX_OPT_STEP_SUMS:
      p->step_sums = value != 0;
      p->sums_ws = nullptr;
      return EVX_OK;
    default: return fail(EVX_ERR_INVALID, "unknown plan option %d", option);
  }
}
