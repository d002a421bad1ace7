// Reverse mode of the periodic Cahn-Hilliard right-hand side
// (CahnHilliard.rhs, problem_definition.py:328-371):
//
//   c = clip(u, 0, 1)
//   mu_j = f(c_j) - kappa * lap(c)_j,          f = 18/eps c(1-c)(1-2c), kappa = 2 eps
//   F at the face between lo and hi along axis a:
//        F = m(cf) * (mu_hi - mu_lo) * id_a,    cf = (c_lo + c_hi)/2, m = cf (1 - cf)
//   r_j = D * sum_a id_a (F_{j+a/2} - F_{j-a/2})
//
// Given w = dJ/dr, the transpose is evaluated face by face:
//   GF  = dJ/dF      = D id_a (w_lo - w_hi)
//   H   = dJ/dg      = GF * m(cf)                 (g = (mu_hi - mu_lo) id_a)
//   Kf  = dJ/dcf     = GF * g * (1 - 2 cf)
//   Gmu_j = dJ/dmu_j = sum_a id_a (H_{j-a/2} - H_{j+a/2})
//   dJ/dc_j = f'(c_j) Gmu_j - kappa lap(Gmu)_j + 1/2 sum_faces(j) Kf
// (the periodic 7-point Laplacian is symmetric), and dJ/du_j = dJ/dc_j where
// 0 <= u_j <= 1 (torch.clip's backward), else 0.  Parameter gradients:
//   dJ/dD   = sum over faces F * GF / D
//   dJ/deps = sum_j Gmu_j * (-f(c_j)/eps - 2 lap(c)_j)
//
// Three stencil kernels (mu; Gmu and the pointwise part; the output with
// lap(Gmu)) and one per-channel reduction.  Every block writes its partial
// parameter sums to its own slot and a last kernel folds them in a fixed
// order, so results are deterministic.
#include "evx_ch_math.hpp"
#include "evx_common.hpp"

namespace evx {
namespace {

constexpr int AZ = 64, AY = 4, AT = AZ * AY;

struct AdjDims {
  int nx, ny, nz, c;
  int gx, gy;  // blocks along z and y per plane
};

template <typename T>
struct AdjCoef {
  T s18, kappa, D, eps;
};

template <typename T>
__device__ __forceinline__ AdjCoef<T> adj_coef(const double* chan, double eps, double D, int c) {
  if (chan) {
    eps = chan[2 * c];
    D = chan[2 * c + 1];
  }
  return {static_cast<T>(18.0 / eps), static_cast<T>(2.0 * eps), static_cast<T>(D), static_cast<T>(eps)};
}

// Offsets of a voxel and its six periodic neighbours inside one channel.
struct Nb7 {
  int64_t base;  // channel offset
  int j, xm, xp, ym, yp, zm, zp;
  bool valid;
  int c, blk;  // channel, block index inside the channel
};

__device__ __forceinline__ Nb7 neighbours(const AdjDims& d) {
  Nb7 n;
  const int z = blockIdx.x * AZ + threadIdx.x;
  const int y = blockIdx.y * AY + threadIdx.y;
  const int zc = blockIdx.z;  // = c * nx + x
  const int c = zc / d.nx, x = zc - (zc / d.nx) * d.nx;
  n.c = c;
  n.blk = (x * d.gy + (int)blockIdx.y) * d.gx + (int)blockIdx.x;
  n.valid = z < d.nz && y < d.ny;
  const int zz = n.valid ? z : 0, yy = n.valid ? y : 0;
  const int plane = d.ny * d.nz;
  n.base = (int64_t)c * d.nx * plane;
  n.j = (x * d.ny + yy) * d.nz + zz;
  n.xm = n.j + (x == 0 ? (d.nx - 1) * plane : -plane);
  n.xp = n.j + (x == d.nx - 1 ? -(d.nx - 1) * plane : plane);
  n.ym = n.j + (yy == 0 ? (d.ny - 1) * d.nz : -d.nz);
  n.yp = n.j + (yy == d.ny - 1 ? -(d.ny - 1) * d.nz : d.nz);
  n.zm = n.j + (zz == 0 ? d.nz - 1 : -1);
  n.zp = n.j + (zz == d.nz - 1 ? -(d.nz - 1) : 1);
  return n;
}

template <typename T>
__device__ __forceinline__ T f_hom(T c, T s18) {
  return ((s18 * c) * (T(1) - c)) * (T(1) - T(2) * c);
}

// Deterministic block sum of NV doubles per thread; thread 0 gets the totals.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV]) {
  __shared__ double red[AT / 64][NV];
#pragma unroll
  for (int k = 0; k < NV; ++k)
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off, 64);
  const int tid = threadIdx.y * AZ + threadIdx.x, wid = tid >> 6;
  if ((tid & 63) == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) red[wid][k] = v[k];
  __syncthreads();
  if (tid == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = 0.0;
      for (int w = 0; w < AT / 64; ++w) s += red[w][k];
      v[k] = s;
    }
  }
}

// mu = f(c) - kappa*lap(c), c = clip(u).
template <typename T>
__global__ __launch_bounds__(AT) void adj_mu_kernel(const T* __restrict__ u, T* __restrict__ mu, AdjDims d,
                                                    Geo<T> g, const double* __restrict__ chan, double eps,
                                                    double D) {
  const Nb7 n = neighbours(d);
  if (!n.valid) return;
  const AdjCoef<T> k = adj_coef<T>(chan, eps, D, n.c);
  const T* uc = u + n.base;
  const T c = clip01(uc[n.j]);
  const T lap = Ops<false>::lap7(c, clip01(uc[n.xp]), clip01(uc[n.xm]), clip01(uc[n.yp]), clip01(uc[n.ym]),
                                 clip01(uc[n.zp]), clip01(uc[n.zm]), g);
  mu[n.base + n.j] = f_hom(c, k.s18) - k.kappa * lap;
}

// Gmu and P = f'(c) Gmu + 1/2 sum Kf; partial parameter sums (dD, deps).
template <typename T>
__global__ __launch_bounds__(AT) void adj_faces_kernel(const T* __restrict__ u, const T* __restrict__ w,
                                                       const T* __restrict__ mu, T* __restrict__ gmu,
                                                       T* __restrict__ pw, double* __restrict__ part, AdjDims d,
                                                       Geo<T> g, const double* __restrict__ chan, double eps,
                                                       double D, int nblk) {
  const Nb7 n = neighbours(d);
  const AdjCoef<T> k = adj_coef<T>(chan, eps, D, n.c);
  double acc[2] = {0.0, 0.0};  // dJ/dD, dJ/deps
  if (n.valid) {
    const T* uc = u + n.base;
    const T* wc = w + n.base;
    const T* mc = mu + n.base;
    const T cj = clip01(uc[n.j]), wj = wc[n.j], mj = mc[n.j];
    const int lo[3] = {n.xm, n.ym, n.zm}, hi[3] = {n.xp, n.yp, n.zp};
    const T id[3] = {g.idx, g.idy, g.idz};
    T cn[3][2];
    T gm = T(0), ks = T(0), dd = T(0);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const T cl = clip01(uc[lo[a]]), ch = clip01(uc[hi[a]]);
      cn[a][0] = cl;
      cn[a][1] = ch;
      // face (j-a, j)
      T cf = T(0.5) * (cl + cj);
      T m = cf * (T(1) - cf);
      T gg = (mj - mc[lo[a]]) * id[a];
      T gf = k.D * id[a] * (wc[lo[a]] - wj);
      const T hm = gf * m;
      ks += gf * gg * (T(1) - T(2) * cf);
      // face (j, j+a)
      cf = T(0.5) * (cj + ch);
      m = cf * (T(1) - cf);
      gg = (mc[hi[a]] - mj) * id[a];
      const T wd = wj - wc[hi[a]];
      gf = k.D * id[a] * wd;
      const T hp = gf * m;
      ks += gf * gg * (T(1) - T(2) * cf);
      dd += (m * gg) * (id[a] * wd);
      gm += id[a] * (hm - hp);
    }
    const T fp = k.s18 * ((T(1) - T(6) * cj) + T(6) * cj * cj);
    gmu[n.base + n.j] = gm;
    pw[n.base + n.j] = fp * gm + T(0.5) * ks;
    const T lapc = Ops<false>::lap7(cj, cn[0][1], cn[0][0], cn[1][1], cn[1][0], cn[2][1], cn[2][0], g);
    const T dmu_deps = -f_hom(cj, k.s18) / k.eps - T(2) * lapc;
    acc[0] = static_cast<double>(dd);
    acc[1] = static_cast<double>(gm) * static_cast<double>(dmu_deps);
  }
  block_sum<2>(acc);
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    double* p = part + 2 * ((int64_t)n.c * nblk + n.blk);
    p[0] = acc[0];
    p[1] = acc[1];
  }
}

// out = base + mask(u) * (P - kappa lap(Gmu)); with u1 != NULL also the
// partial of sum (w - dt*base) * (u1 - u) for the prefactor gradient.
template <typename T>
__global__ __launch_bounds__(AT) void adj_out_kernel(const T* __restrict__ u, const T* __restrict__ gmu,
                                                     const T* __restrict__ pw, const T* base, T* out,
                                                     const T* __restrict__ w, const T* __restrict__ u1,
                                                     double dt, double* __restrict__ part, AdjDims d, Geo<T> g,
                                                     const double* __restrict__ chan, double eps, double D,
                                                     int nblk) {
  const Nb7 n = neighbours(d);
  const AdjCoef<T> k = adj_coef<T>(chan, eps, D, n.c);
  double acc[1] = {0.0};
  if (n.valid) {
    const T* gc = gmu + n.base;
    const int64_t i = n.base + n.j;
    const T uj = u[i];
    const T lapg = Ops<false>::lap7(gc[n.j], gc[n.xp], gc[n.xm], gc[n.yp], gc[n.ym], gc[n.zp], gc[n.zm], g);
    const T gu = (uj >= T(0) && uj <= T(1)) ? pw[i] - k.kappa * lapg : T(0);
    const T b = base ? base[i] : T(0);
    out[i] = b + gu;
    if (u1) acc[0] = static_cast<double>(w[i] - static_cast<T>(dt) * b) * static_cast<double>(u1[i] - uj);
  }
  if (u1) {
    block_sum<1>(acc);
    if (threadIdx.x == 0 && threadIdx.y == 0) part[(int64_t)n.c * nblk + n.blk] = acc[0];
  }
}

// Per channel: fold the block partials (Neumaier, fixed order) and add the
// gradients to dparams = {dJ/deps, dJ/dD}.  With pre != NULL the prefactor
// term S/(dt*theta) is added as well.
__global__ __launch_bounds__(256) void adj_final_kernel(const double* __restrict__ part,
                                                        const double* __restrict__ pre, int nblk,
                                                        double* dparams, const double* __restrict__ chan,
                                                        double eps, double D, double dt) {
  const int c = blockIdx.x, tid = threadIdx.x;
  double s[3] = {0, 0, 0}, e[3] = {0, 0, 0};
  auto add = [&](int k, double v) {
    const double t = s[k] + v;
    e[k] += fabs(s[k]) >= fabs(v) ? (s[k] - t) + v : (v - t) + s[k];
    s[k] = t;
  };
  for (int b = tid; b < nblk; b += 256) {
    const double* p = part + 2 * ((int64_t)c * nblk + b);
    add(0, p[0]);
    add(1, p[1]);
    if (pre) add(2, pre[(int64_t)c * nblk + b]);
  }
  __shared__ double sh[3][256];
  for (int k = 0; k < 3; ++k) sh[k][tid] = s[k] + e[k];
  __syncthreads();
  if (tid == 0) {
    double t[3] = {0, 0, 0}, te[3] = {0, 0, 0};
    for (int k = 0; k < 3; ++k)
      for (int i = 0; i < 256; ++i) {
        const double v = sh[k][i], q = t[k] + v;
        te[k] += fabs(t[k]) >= fabs(v) ? (t[k] - q) + v : (v - q) + t[k];
        t[k] = q;
      }
    if (chan) {
      eps = chan[2 * c];
      D = chan[2 * c + 1];
    }
    double g_d = t[0] + te[0], g_e = t[1] + te[1];
    if (pre) {
      const double S = t[2] + te[2];
      g_e += S / (dt * eps);
      g_d += S / (dt * D);
    }
    dparams[2 * c] += g_e;
    dparams[2 * c + 1] += g_d;
  }
}

size_t align_up(size_t v) { return (v + 255) & ~size_t(255); }

int adj_dims(const Dims& d, AdjDims* a) {
  if (d.nx * d.c > 65535) return fail(EVX_ERR_UNSUPPORTED, "adjoint kernels need channels*nx <= 65535");
  if (d.vol() >= (int64_t(1) << 31)) return fail(EVX_ERR_UNSUPPORTED, "adjoint kernels need nx*ny*nz < 2^31");
  a->nx = (int)d.nx;
  a->ny = (int)d.ny;
  a->nz = (int)d.nz;
  a->c = (int)d.c;
  a->gx = (int)((d.nz + AZ - 1) / AZ);
  a->gy = (int)((d.ny + AY - 1) / AY);
  return EVX_OK;
}

}  // namespace

// Workspace: mu, Gmu, P (three fields) and the block partials.
size_t ch_rhs_vjp_ws_bytes(int32_t dtype, const Dims& d) {
  const size_t field = align_up((size_t)d.c * d.vol() * (dtype == EVX_F32 ? 4 : 8));
  const int64_t nblk = ((d.nz + AZ - 1) / AZ) * ((d.ny + AY - 1) / AY) * d.nx;
  return 3 * field + align_up((size_t)d.c * nblk * 3 * sizeof(double));
}

// dparams += ...; grad = base + J^T w.  With u1 != NULL (semi-implicit step)
// the prefactor term is added with S = sum (w - dt*base)(u1 - u).
int ch_rhs_vjp_launch(int32_t dtype, const void* u, const void* w, const void* base, void* grad,
                      double* dparams, const Dims& d, const double dd[3], const double dd2[3], double eps,
                      double D, const double* chan, const void* u1, double dt, void* ws, hipStream_t s) {
  AdjDims a;
  EVX_TRY(adj_dims(d, &a));
  const size_t es = dtype == EVX_F32 ? 4 : 8;
  const size_t field = align_up((size_t)d.c * d.vol() * es);
  char* p = static_cast<char*>(ws);
  void* mu = p;
  void* gmu = p + field;
  void* pw = p + 2 * field;
  const int nblk = a.gx * a.gy * a.nx;
  double* part = reinterpret_cast<double*>(p + 3 * field);
  double* pre = part + 2 * (int64_t)d.c * nblk;
  const dim3 grid(a.gx, a.gy, a.nx * a.c), block(AZ, AY);
  evx_grid eg{d.c, d.nx, d.ny, d.nz, {dd[0], dd[1], dd[2]}, {dd2[0], dd2[1], dd2[2]}};
  if (dtype == EVX_F32) {
    using T = float;
    const Geo<T> g = make_geo_from<T>(eg.div_dx, eg.div_dx2);
    hipLaunchKernelGGL(adj_mu_kernel<T>, grid, block, 0, s, static_cast<const T*>(u), static_cast<T*>(mu), a, g,
                       chan, eps, D);
    EVX_TRY(check_launch("adj_mu_kernel"));
    hipLaunchKernelGGL(adj_faces_kernel<T>, grid, block, 0, s, static_cast<const T*>(u), static_cast<const T*>(w),
                       static_cast<const T*>(mu), static_cast<T*>(gmu), static_cast<T*>(pw), part, a, g, chan, eps,
                       D, nblk);
    EVX_TRY(check_launch("adj_faces_kernel"));
    hipLaunchKernelGGL(adj_out_kernel<T>, grid, block, 0, s, static_cast<const T*>(u), static_cast<const T*>(gmu),
                       static_cast<const T*>(pw), static_cast<const T*>(base), static_cast<T*>(grad),
                       static_cast<const T*>(w), static_cast<const T*>(u1), dt, pre, a, g, chan, eps, D, nblk);
    EVX_TRY(check_launch("adj_out_kernel"));
  } else {
    using T = double;
    const Geo<T> g = make_geo_from<T>(eg.div_dx, eg.div_dx2);
    hipLaunchKernelGGL(adj_mu_kernel<T>, grid, block, 0, s, static_cast<const T*>(u), static_cast<T*>(mu), a, g,
                       chan, eps, D);
    EVX_TRY(check_launch("adj_mu_kernel"));
    hipLaunchKernelGGL(adj_faces_kernel<T>, grid, block, 0, s, static_cast<const T*>(u), static_cast<const T*>(w),
                       static_cast<const T*>(mu), static_cast<T*>(gmu), static_cast<T*>(pw), part, a, g, chan, eps,
                       D, nblk);
    EVX_TRY(check_launch("adj_faces_kernel"));
    hipLaunchKernelGGL(adj_out_kernel<T>, grid, block, 0, s, static_cast<const T*>(u), static_cast<const T*>(gmu),
                       static_cast<const T*>(pw), static_cast<const T*>(base), static_cast<T*>(grad),
                       static_cast<const T*>(w), static_cast<const T*>(u1), dt, pre, a, g, chan, eps, D, nblk);
    EVX_TRY(check_launch("adj_out_kernel"));
  }
  hipLaunchKernelGGL(adj_final_kernel, dim3(a.c), dim3(256), 0, s, part, u1 ? pre : nullptr, nblk, dparams, chan,
                     eps, D, dt);
  return check_launch("adj_final_kernel");
}

}  // namespace evx

using namespace evx;

extern "C" int evx_ch_rhs_vjp_workspace_bytes(int32_t dtype, const evx_grid* g, size_t* bytes) {
  if (!g || !bytes) return fail(EVX_ERR_INVALID, "null argument");
  if (dtype != EVX_F32 && dtype != EVX_F64) return fail(EVX_ERR_INVALID, "dtype must be EVX_F32 or EVX_F64");
  Dims d;
  EVX_TRY(make_dims(g, &d));
  *bytes = ch_rhs_vjp_ws_bytes(dtype, d);
  return EVX_OK;
}

extern "C" int evx_ch_rhs_vjp(int32_t dtype, const void* u, const void* w, const void* base, void* grad_u,
                              double* dparams, const evx_grid* g, double eps, double D, const double* chan_params,
                              void* workspace, void* stream) {
  if (!u || !w || !grad_u || !dparams || !g || !workspace) return fail(EVX_ERR_INVALID, "null argument");
  if (dtype != EVX_F32 && dtype != EVX_F64) return fail(EVX_ERR_INVALID, "dtype must be EVX_F32 or EVX_F64");
  if (!chan_params && !(eps != 0.0)) return fail(EVX_ERR_INVALID, "eps must be nonzero");
  if (grad_u == u || grad_u == w) return fail(EVX_ERR_INVALID, "grad_u must not alias u or w");
  Dims d;
  EVX_TRY(make_dims(g, &d));
  return ch_rhs_vjp_launch(dtype, u, w, base, grad_u, dparams, d, g->div_dx, g->div_dx2, eps, D, chan_params,
                           nullptr, 0.0, workspace, static_cast<hipStream_t>(stream));
}
