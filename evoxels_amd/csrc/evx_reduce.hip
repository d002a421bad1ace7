// Compensated float64 reductions: mass (sum of a field) and the discrete
// Cahn-Hilliard free energy, both per channel, plus a NaN/Inf flag.
//
// Each thread keeps a Neumaier pair (sum, compensation); pairs are merged
// with the error-free TwoSum inside a wavefront (64 lanes, xor shuffles),
// across the wavefronts of a block through LDS, and finally across a fixed
// number of per-block partials by one block per channel.  The order of all
// merges is fixed, so results are bitwise reproducible run to run.
#include <mutex>

#include "evx_common.hpp"

#pragma clang fp contract(off)

namespace evx {
namespace {

constexpr int RB = 256;     // threads per block
constexpr int NBLK = 1024;  // partial blocks per channel

struct Pair {
  double s, c;
};

__device__ __forceinline__ void neumaier(Pair& p, double x) {
  const double t = p.s + x;
  if (fabs(p.s) >= fabs(x)) p.c += (p.s - t) + x;
  else p.c += (x - t) + p.s;
  p.s = t;
}

__device__ __forceinline__ Pair merge(Pair a, Pair b) {
  const double s = a.s + b.s;
  const double bb = s - a.s;
  const double e = (a.s - (s - bb)) + (b.s - bb);
  return Pair{s, (a.c + b.c) + e};
}

__device__ Pair block_reduce(Pair p) {
  __shared__ double ss[RB / 64], sc[RB / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    Pair q{__shfl_xor(p.s, o, 64), __shfl_xor(p.c, o, 64)};
    p = merge(p, q);
  }
  const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
  if (lane == 0) {
    ss[w] = p.s;
    sc[w] = p.c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    Pair r{ss[0], sc[0]};
    for (int i = 1; i < RB / 64; ++i) r = merge(r, Pair{ss[i], sc[i]});
    p = r;
  }
  return p;
}

template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
  using type = float4;
  static constexpr int W = 4;
  __device__ static double at(const float4& v, int j) { return j == 0 ? v.x : j == 1 ? v.y : j == 2 ? v.z : v.w; }
};
template <>
struct Vec16<double> {
  using type = double2;
  static constexpr int W = 2;
  __device__ static double at(const double2& v, int j) { return j == 0 ? v.x : v.y; }
};

struct Flags {
  bool nan = false, inf = false;
  __device__ void see(double v) {
    nan |= v != v;
    inf |= fabs(v) == __builtin_inf();
  }
};

// Per channel, NBLK blocks stream the field with 16-byte loads (U vectors in
// flight per thread) into W independent Neumaier pairs, one per vector lane,
// merged in a fixed order at the end; the few elements before the first
// 16-byte boundary and after the last whole vector are added by the first
// threads of the grid.  The split of elements over threads depends only on n
// and the channel's address modulo 16, so results are reproducible.  Two
// vectors in flight per thread with 1024 blocks measured fastest on MI355X
// (512^3 float32, partial and fold kernels: 5.3 TB/s; four in flight 4.9,
// eight 4.3, one 4.6; 2048 blocks with four 4.5, 256 blocks 3.7).
template <typename T>
__global__ __launch_bounds__(RB) void sum_partial_kernel(const T* __restrict__ x, int64_t n, Pair* partial,
                                                         int32_t* nonfinite) {
  using VT = Vec16<T>;
  using V = typename VT::type;
  constexpr int W = VT::W, U = 2;
  const int64_t ch = blockIdx.y;
  const T* xc = x + ch * n;
  const int64_t mis = (int64_t)((reinterpret_cast<uintptr_t>(xc) & 15u) / sizeof(T));
  const int64_t head = mis ? (W - mis < n ? W - mis : n) : 0;
  const int64_t nv = (n - head) / W;
  const V* xv = reinterpret_cast<const V*>(xc + head);
  Pair p[W];
#pragma unroll
  for (int j = 0; j < W; ++j) p[j] = Pair{0.0, 0.0};
  Flags f;
  const int64_t stride = (int64_t)NBLK * RB;
  const int64_t g = (int64_t)blockIdx.x * RB + threadIdx.x;
  int64_t i = g;
  for (; i + (U - 1) * stride < nv; i += U * stride) {
    V v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = xv[i + u * stride];
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int j = 0; j < W; ++j) {
        const double d = VT::at(v[u], j);
        f.see(d);
        neumaier(p[j], d);
      }
  }
  for (; i < nv; i += stride) {
    const V v = xv[i];
#pragma unroll
    for (int j = 0; j < W; ++j) {
      const double d = VT::at(v, j);
      f.see(d);
      neumaier(p[j], d);
    }
  }
  const int64_t tail0 = head + nv * W;
  if (g < head) {
    const double d = static_cast<double>(xc[g]);
    f.see(d);
    neumaier(p[0], d);
  }
  if (g < n - tail0) {
    const double d = static_cast<double>(xc[tail0 + g]);
    f.see(d);
    neumaier(p[W - 1], d);
  }
  Pair q = p[0];
#pragma unroll
  for (int j = 1; j < W; ++j) q = merge(q, p[j]);
  // Every writer stores the same value, so plain stores suffice.
  if (nonfinite && f.nan) nonfinite[0] = 1;
  if (nonfinite && f.inf) nonfinite[1] = 1;
  q = block_reduce(q);
  if (threadIdx.x == 0) partial[ch * NBLK + blockIdx.x] = q;
}

__global__ __launch_bounds__(RB) void sum_final_kernel(const Pair* partial, double scale, double* out) {
  const int64_t ch = blockIdx.x;
  Pair p{0.0, 0.0};
  for (int i = threadIdx.x; i < NBLK; i += RB) p = merge(p, partial[ch * NBLK + i]);
  p = block_reduce(p);
  if (threadIdx.x == 0) out[ch] = (p.s + p.c) * scale;
}

// Energy density eps*|grad_center c|^2 + 9/eps*c^2(1-c)^2 on pad_bc(c),
// gradients as fd_stencils.py:44-60 evaluated in float64.
template <typename T>
__global__ __launch_bounds__(RB) void energy_partial_kernel(const T* __restrict__ c, Dims d, Bc bc, double idx,
                                                            double idy, double idz, double eps, double s9,
                                                            Pair* partial) {
  const int64_t ch = blockIdx.y;
  const T* cc = c + ch * d.vol();
  const int64_t n = d.vol();
  Pair p{0.0, 0.0};
  auto at = [&](int64_t x, int64_t y, int64_t z) -> double {
    Resolved<T> r = resolve<T>(x, y, z, d, bc);
    return static_cast<double>(apply_rules(cc[r.off], r.o, bc));
  };
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)NBLK * RB) {
    const int64_t z = i % d.nz, t = i / d.nz;
    const int64_t y = t % d.ny, x = t / d.ny;
    const double gx = (0.5 * (at(x + 1, y, z) - at(x - 1, y, z))) * idx;
    const double gy = (0.5 * (at(x, y + 1, z) - at(x, y - 1, z))) * idy;
    const double gz = (0.5 * (at(x, y, z + 1) - at(x, y, z - 1))) * idz;
    const double v = static_cast<double>(cc[i]);
    const double g2 = (gx * gx + gy * gy) + gz * gz;
    const double dens = eps * g2 + ((s9 * (v * v)) * ((1.0 - v) * (1.0 - v)));
    neumaier(p, dens);
  }
  p = block_reduce(p);
  if (threadIdx.x == 0) partial[ch * NBLK + blockIdx.x] = p;
}

}  // namespace

// Scratch of evx_sum_f64_noalloc: the per-block partials of up to
// EVX_NOALLOC_MAX_CHANNELS channels (1 MiB), held by each GPU the library
// runs on until the process ends.  Host code never passes `scratch` itself
// to a kernel: its device address on the current GPU comes from
// hipGetSymbolAddress (noalloc_scratch below).
static __device__ __align__(256) Pair scratch[EVX_NOALLOC_MAX_CHANNELS * NBLK];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current GPU, looked up once per device
// (with that device current) and kept.
int noalloc_scratch(Pair** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0) return fail(EVX_ERR_HIP, "hipGetDevice failed");
  if (dev >= kMaxDevices) return fail(EVX_ERR_UNSUPPORTED, "device index %d beyond %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(EVX_ERR_HIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev, hipGetErrorString(e));
    g_scratch_ptr[dev] = ptr;
  }
  *out = static_cast<Pair*>(g_scratch_ptr[dev]);
  return EVX_OK;
}
}  // namespace
}  // namespace evx

using namespace evx;

extern "C" int evx_reduce_workspace_bytes(int64_t channels, size_t* bytes) {
  if (!bytes || channels < 1) return fail(EVX_ERR_INVALID, "bad arguments");
  *bytes = (size_t)channels * NBLK * sizeof(Pair);
  return EVX_OK;
}

extern "C" int evx_sum_f64(int32_t dtype, const void* x, int64_t channels, int64_t n, double* out,
                           int32_t* nonfinite, void* workspace, void* stream) {
  if (dtype != EVX_F32 && dtype != EVX_F64) return fail(EVX_ERR_INVALID, "bad dtype");
  if (!x || !out || !workspace || channels < 1 || n < 1) return fail(EVX_ERR_INVALID, "bad arguments");
  if (channels > 65535) return fail(EVX_ERR_UNSUPPORTED, "too many channels");
  hipStream_t s = static_cast<hipStream_t>(stream);
  Pair* part = static_cast<Pair*>(workspace);
  if (dtype == EVX_F32)
    hipLaunchKernelGGL(sum_partial_kernel<float>, dim3(NBLK, (unsigned)channels), dim3(RB), 0, s,
                       static_cast<const float*>(x), n, part, nonfinite);
  else
    hipLaunchKernelGGL(sum_partial_kernel<double>, dim3(NBLK, (unsigned)channels), dim3(RB), 0, s,
                       static_cast<const double*>(x), n, part, nonfinite);
  EVX_TRY(check_launch("sum_partial_kernel"));
  hipLaunchKernelGGL(sum_final_kernel, dim3((unsigned)channels), dim3(RB), 0, s, part, 1.0, out);
  return check_launch("sum_final_kernel");
}

extern "C" int evx_sum_f64_noalloc(int32_t dtype, const void* x, int64_t channels, int64_t n, double* out,
                                   int32_t* nonfinite, void* stream) {
  if (dtype != EVX_F32 && dtype != EVX_F64) return fail(EVX_ERR_INVALID, "bad dtype");
  if (!x || !out || channels < 1 || n < 1) return fail(EVX_ERR_INVALID, "bad arguments");
  if (channels > EVX_NOALLOC_MAX_CHANNELS)
    return fail(EVX_ERR_INVALID, "evx_sum_f64_noalloc: %lld channels exceed its static scratch of %d channels",
                (long long)channels, EVX_NOALLOC_MAX_CHANNELS);
  Pair* part = nullptr;
  EVX_TRY(noalloc_scratch(&part));
  // Every block of every channel overwrites its partial, so the scratch left
  // by an earlier call needs no clearing.
  return evx_sum_f64(dtype, x, channels, n, out, nonfinite, part, stream);
}

extern "C" int evx_ch_energy(int32_t dtype, const void* c, const evx_grid* g, const evx_bc* bc, double eps,
                             double dv, double* out, void* workspace, void* stream) {
  Dims d;
  Bc b;
  if (dtype != EVX_F32 && dtype != EVX_F64) return fail(EVX_ERR_INVALID, "bad dtype");
  if (!c || !out || !workspace || !bc) return fail(EVX_ERR_INVALID, "bad arguments");
  EVX_TRY(make_dims(g, &d));
  EVX_TRY(make_bc(bc, &b));
  if (staggered_x(b)) return fail(EVX_ERR_UNSUPPORTED, "the energy implements the cell_center rules only");
  if (!(eps != 0.0)) return fail(EVX_ERR_INVALID, "eps must be nonzero");
  if (d.c > 65535) return fail(EVX_ERR_UNSUPPORTED, "too many channels");
  hipStream_t s = static_cast<hipStream_t>(stream);
  Pair* part = static_cast<Pair*>(workspace);
  const double s9 = 9.0 / eps;
  if (dtype == EVX_F32)
    hipLaunchKernelGGL(energy_partial_kernel<float>, dim3(NBLK, (unsigned)d.c), dim3(RB), 0, s,
                       static_cast<const float*>(c), d, b, g->div_dx[0], g->div_dx[1], g->div_dx[2], eps, s9, part);
  else
    hipLaunchKernelGGL(energy_partial_kernel<double>, dim3(NBLK, (unsigned)d.c), dim3(RB), 0, s,
                       static_cast<const double*>(c), d, b, g->div_dx[0], g->div_dx[1], g->div_dx[2], eps, s9,
                       part);
  EVX_TRY(check_launch("energy_partial_kernel"));
  hipLaunchKernelGGL(sum_final_kernel, dim3((unsigned)d.c), dim3(RB), 0, s, part, dv, out);
  return check_launch("sum_final_kernel");
}
