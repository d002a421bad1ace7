// Finite-difference right-hand sides with in-kernel boundary conditions.
//
// Every kernel streams along x through a ring of (y,z) planes staged in LDS
// (2.5-D tiling): a block owns a BY x 64 tile of (y,z) and an x-chunk, loads
// each x-plane of the tile (plus halo) once, and keeps the planes the stencil
// needs resident.  Ghost values never exist in HBM: out-of-grid points are
// resolved on load (periodic wrap, Neumann copy, Dirichlet 2v-u, applied x
// then y then z as in boundary_conditions.py:33-59).
//
// Floating-point contraction is off in this file so that every product and
// sum rounds where the reference's eager torch expression rounds.
#include "evx_stencil_tiles.hpp"

#pragma clang fp contract(off)

namespace evx {
namespace {


// ---------------------------------------------------------------------------
// Reaction-diffusion (f == 0): r = D * laplace(pad_bc(u))
// (problem_definition.py:226-229, fd_stencils.py:62-75)
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(NT) void rd_rhs_kernel(const T* __restrict__ u, T* __restrict__ out, Dims d, Bc bc,
                                                    Geo<T> g, T D, Epi ep, TileGrid tg) {
  constexpr int PY = TileCfg<T>::PY, BY = NTY * PY;
  constexpr int W = BZ + 2, ROWS = BY + 2, NR = 4;
  __shared__ T cs[NR][ROWS][W];
  const int tz = threadIdx.x, ty = threadIdx.y, tid = ty * BZ + tz;
  const TileOrigin o = tile_origin<BY>(tg, d, xcd_tile(blockIdx.x, gridDim.x));
  const int64_t z0 = o.z0, y0 = o.y0, x0 = o.x0, x1 = o.x1;
  const T* uc = u + o.ch * d.vol();
  T* oc = out + o.ch * d.vol();
  PlaneLoader<T, 1, ROWS, W, false, true> pl;
  pl.init(y0, z0, d, bc, tid);
  for (int64_t xg = x0 - 1; xg <= x0 + 1; ++xg) {
    pl.fetch(uc, xg, d, bc);
    pl.commit(&cs[ring<NR>(xg - x0)][0][0], bc, tid);
  }
  if (x0 + 2 <= x1) pl.fetch(uc, x0 + 2, d, bc);
  __syncthreads();
  const int64_t plane = d.ny * d.nz;
  const int64_t gz = z0 + tz;
  for (int64_t x = x0; x < x1; ++x) {
    // plane x+2 (fetched one iteration ago) goes to the slot of plane x-2
    if (x + 2 <= x1) pl.commit(&cs[ring<NR>(x + 2 - x0)][0][0], bc, tid);
    if (x + 3 <= x1) pl.fetch(uc, x + 3, d, bc);
    const T(*Cm)[W] = cs[ring<NR>(x - 1 - x0)];
    const T(*C0)[W] = cs[ring<NR>(x - x0)];
    const T(*Cp)[W] = cs[ring<NR>(x + 1 - x0)];
#pragma unroll
    for (int q = 0; q < PY; ++q) {
      const int yy = ty + NTY * q;
      const int64_t gy = y0 + yy;
      if (gy >= d.ny || gz >= d.nz) continue;
      const int a = yy + 1, b = tz + 1;
      const T c0 = C0[a][b];
      const T lap = lap7(c0, Cp[a][b], Cm[a][b], C0[a + 1][b], C0[a - 1][b], C0[a][b + 1], C0[a][b - 1], g);
      epilogue_v(ep, c0, oc, x * plane + gy * d.nz + gz, D * lap);
    }
    __syncthreads();
  }
}


// ---------------------------------------------------------------------------
// Two species with the Gray-Scott interaction a*b^2
// (problem_definition.py:612-633):
//   r_a = D_a*lap(a) - a*b^2 + F*(1 - a),   r_b = D_b*lap(b) + a*b^2 - k*b.
// One block streams both species through their own plane rings, so the
// interaction reads each value once; the tile grid has no channel dimension.
// ---------------------------------------------------------------------------
template <typename T>
struct GsCoef {
  T Da, Db, feed, kill;
};

template <typename T>
__global__ __launch_bounds__(NT) void gs_rhs_kernel(const T* __restrict__ u, T* __restrict__ out, Dims d, Bc bc,
                                                    Geo<T> g, GsCoef<T> k, Epi ep, TileGrid tg) {
  constexpr int PY = TileCfg<T>::PY, BY = NTY * PY;
  constexpr int W = BZ + 2, ROWS = BY + 2, NR = 4;
  __shared__ T sa[NR][ROWS][W];
  __shared__ T sb[NR][ROWS][W];
  const int tz = threadIdx.x, ty = threadIdx.y, tid = ty * BZ + tz;
  const TileOrigin o = tile_origin<BY>(tg, d, xcd_tile(blockIdx.x, gridDim.x));
  const int64_t z0 = o.z0, y0 = o.y0, x0 = o.x0, x1 = o.x1;
  const int64_t vol = d.vol();
  const T* ua = u;
  const T* ub = u + vol;
  PlaneLoader<T, 1, ROWS, W, false, true> pa, pb;
  pa.init(y0, z0, d, bc, tid);
  pb.init(y0, z0, d, bc, tid);
  for (int64_t xg = x0 - 1; xg <= x0 + 1; ++xg) {
    pa.fetch(ua, xg, d, bc);
    pb.fetch(ub, xg, d, bc);
    pa.commit(&sa[ring<NR>(xg - x0)][0][0], bc, tid);
    pb.commit(&sb[ring<NR>(xg - x0)][0][0], bc, tid);
  }
  if (x0 + 2 <= x1) {
    pa.fetch(ua, x0 + 2, d, bc);
    pb.fetch(ub, x0 + 2, d, bc);
  }
  __syncthreads();
  const int64_t plane = d.ny * d.nz;
  const int64_t gz = z0 + tz;
  for (int64_t x = x0; x < x1; ++x) {
    if (x + 2 <= x1) {
      pa.commit(&sa[ring<NR>(x + 2 - x0)][0][0], bc, tid);
      pb.commit(&sb[ring<NR>(x + 2 - x0)][0][0], bc, tid);
    }
    if (x + 3 <= x1) {
      pa.fetch(ua, x + 3, d, bc);
      pb.fetch(ub, x + 3, d, bc);
    }
    const T(*Am)[W] = sa[ring<NR>(x - 1 - x0)];
    const T(*A0)[W] = sa[ring<NR>(x - x0)];
    const T(*Ap)[W] = sa[ring<NR>(x + 1 - x0)];
    const T(*Bm)[W] = sb[ring<NR>(x - 1 - x0)];
    const T(*B0)[W] = sb[ring<NR>(x - x0)];
    const T(*Bp)[W] = sb[ring<NR>(x + 1 - x0)];
#pragma unroll
    for (int q = 0; q < PY; ++q) {
      const int yy = ty + NTY * q;
      const int64_t gy = y0 + yy;
      if (gy >= d.ny || gz >= d.nz) continue;
      const int a = yy + 1, b = tz + 1;
      const T av = A0[a][b], bv = B0[a][b];
      const T la = lap7(av, Ap[a][b], Am[a][b], A0[a + 1][b], A0[a - 1][b], A0[a][b + 1], A0[a][b - 1], g);
      const T lb = lap7(bv, Bp[a][b], Bm[a][b], B0[a + 1][b], B0[a - 1][b], B0[a][b + 1], B0[a][b - 1], g);
      const T inter = av * (bv * bv);
      const T ra = (k.Da * la - inter) + k.feed * (T(1) - av);
      const T rb = (k.Db * lb + inter) - k.kill * bv;
      const int64_t i = x * plane + gy * d.nz + gz;
      epilogue_v(ep, av, out, i, ra);
      epilogue_v(ep, bv, out, vol + i, rb);
    }
    __syncthreads();
  }
}

template <typename T, bool STRICT>
__global__ __launch_bounds__(NT) void ch_rhs_kernel(const T* __restrict__ u, T* __restrict__ out, Dims d, Bc bc,
                                                    Geo<T> g, ChCoef<T> k, Epi ep, TileGrid tg,
                                                    const double* __restrict__ chan, unsigned tile_base) {
  __shared__ __align__(16) unsigned char smem[ChTile<T>::LDS_BYTES];
  ch_rhs_tile<T, STRICT>(u, out, d, bc, g, k, ep, tg, chan, tile_base + xcd_tile(blockIdx.x, gridDim.x),
                         threadIdx.y * BZ + threadIdx.x, smem);
}

// ---------------------------------------------------------------------------
// Two-phase Allen-Cahn (problem_definition.py:440-447) with the 19-point
// normal Laplacian of fd_stencils.py:77-103.  Edge ghosts are used here, so
// the x-then-y-then-z composition of the ghost rules matters.
// ---------------------------------------------------------------------------
template <typename T>
struct AcCoef {
  T s18;       // 18/eps
  T eps;       // eps
  T curv;      // curvature
  T one_m_curv;// 1-curvature
  T gab, M, force;
  T s3;        // 3/eps
};

template <typename T, int W>
__device__ __forceinline__ T normal_lap_at(const T (*Pm)[W], const T (*P0)[W], const T (*Pp)[W], int a,
                                           int b, const Geo<T>& g) {
  const T C = P0[a][b];
  const T R = Pp[a][b], L = Pm[a][b];
  const T Tp = P0[a + 1][b], Bm = P0[a - 1][b];
  const T F = P0[a][b + 1], K = P0[a][b - 1];
  const T gx = (T(0.5) * (R - L)) * g.idx;
  const T gy = (T(0.5) * (Tp - Bm)) * g.idy;
  const T gz = (T(0.5) * (F - K)) * g.idz;
  const T t1 = ((gx * gx) * ((R - T(2) * C) + L)) * g.idx2;
  const T t2 = ((gy * gy) * ((Tp - T(2) * C) + Bm)) * g.idy2;
  const T t3 = ((gz * gz) * ((F - T(2) * C) + K)) * g.idz2;
  const T cxy = ((Pp[a + 1][b] + Pm[a - 1][b]) - Pm[a + 1][b]) - Pp[a - 1][b];
  const T cxz = ((Pp[a][b + 1] + Pm[a][b - 1]) - Pm[a][b + 1]) - Pp[a][b - 1];
  const T cyz = ((P0[a + 1][b + 1] + P0[a - 1][b - 1]) - P0[a - 1][b + 1]) - P0[a + 1][b - 1];
  const T t4 = ((((T(0.5) * gx) * gy) * cxy) * g.idx) * g.idy;
  const T t5 = ((((T(0.5) * gx) * gz) * cxz) * g.idx) * g.idz;
  const T t6 = ((((T(0.5) * gy) * gz) * cyz) * g.idy) * g.idz;
  const T num = ((((t1 + t2) + t3) + t4) + t5) + t6;
  T n2 = ((gx * gx) + (gy * gy)) + (gz * gz);
  if (n2 <= static_cast<T>(1e-7)) n2 = T(1);
  return num / n2;
}

template <typename T>
__global__ __launch_bounds__(NT) void ac_rhs_kernel(const T* __restrict__ u, T* __restrict__ out, Dims d, Bc bc,
                                                    Geo<T> g, AcCoef<T> k, Epi ep, TileGrid tg) {
  constexpr int PY = TileCfg<T>::PY, BY = NTY * PY;
  constexpr int W = BZ + 2, ROWS = BY + 2, NR = 4;
  __shared__ T cs[NR][ROWS][W];
  const int tz = threadIdx.x, ty = threadIdx.y, tid = ty * BZ + tz;
  const TileOrigin o = tile_origin<BY>(tg, d, xcd_tile(blockIdx.x, gridDim.x));
  const int64_t z0 = o.z0, y0 = o.y0, x0 = o.x0, x1 = o.x1;
  const T* uc = u + o.ch * d.vol();
  T* oc = out + o.ch * d.vol();
  PlaneLoader<T, 1, ROWS, W, true, true> pl;
  pl.init(y0, z0, d, bc, tid);
  for (int64_t xg = x0 - 1; xg <= x0 + 1; ++xg) {
    pl.fetch(uc, xg, d, bc);
    pl.commit(&cs[ring<NR>(xg - x0)][0][0], bc, tid);
  }
  if (x0 + 2 <= x1) pl.fetch(uc, x0 + 2, d, bc);
  __syncthreads();
  const int64_t plane = d.ny * d.nz;
  const int64_t gz = z0 + tz;
  for (int64_t x = x0; x < x1; ++x) {
    if (x + 2 <= x1) pl.commit(&cs[ring<NR>(x + 2 - x0)][0][0], bc, tid);
    if (x + 3 <= x1) pl.fetch(uc, x + 3, d, bc);
    const T(*Pm)[W] = cs[ring<NR>(x - 1 - x0)];
    const T(*P0)[W] = cs[ring<NR>(x - x0)];
    const T(*Pp)[W] = cs[ring<NR>(x + 1 - x0)];
#pragma unroll
    for (int q = 0; q < PY; ++q) {
      const int yy = ty + NTY * q;
      const int64_t gy = y0 + yy;
      if (gy >= d.ny || gz >= d.nz) continue;
      const int a = yy + 1, b = tz + 1;
      const T phi = P0[a][b];
      const T pot = ((k.s18 * phi) * (T(1) - phi)) * (T(1) - T(2) * phi);
      const T lap = k.curv * lap7(phi, Pp[a][b], Pm[a][b], P0[a + 1][b], P0[a - 1][b], P0[a][b + 1],
                                  P0[a][b - 1], g);
      const T nlap = k.one_m_curv * normal_lap_at<T, W>(Pm, P0, Pp, a, b, g);
      const T df = k.gab * ((lap + nlap) - (pot / T(2)) / k.eps) +
                   ((k.s3 * phi) * (T(1) - phi)) * k.force;
      epilogue(ep, uc, oc, x * plane + gy * d.nz + gz, k.M * df);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Operator-level kernels for the generic path
// ---------------------------------------------------------------------------
template <typename T>
__global__ void pad_bc_kernel(const T* __restrict__ u, T* __restrict__ out, Dims d, Bc bc) {
  const int64_t px = d.nx + 2, py = d.ny + 2, pz = d.nz + 2;
  const int64_t n = d.c * px * py * pz;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    int64_t k = i % pz, t = i / pz;
    int64_t j = t % py;
    t /= py;
    int64_t xi = t % px, ch = t / px;
    Resolved<T> r = resolve<T>(xi - 1, j - 1, k - 1, d, bc);
    const T* uc = u + ch * d.vol();
    if (r.o[0] != 0 && staggered_x(bc)) {
      // staggered_x ghost plane (boundary_conditions.py:96-121); y, z are periodic
      if (bc.kind[0] == EVX_BC_NODE_VALUE) {
        out[i] = static_cast<T>(r.o[0] < 0 ? bc.node_lo : bc.node_hi);
      } else {
        const int64_t inward = r.o[0] < 0 ? d.ny * d.nz : -d.ny * d.nz;
        out[i] = static_cast<T>(4.0 / 3.0) * uc[r.off] - static_cast<T>(1.0 / 3.0) * uc[r.off + inward];
      }
      continue;
    }
    out[i] = apply_rules(uc[r.off], r.o, bc);
  }
}

template <typename T, bool NORMAL>
__global__ void padded_stencil_kernel(const T* __restrict__ p, T* __restrict__ out, Dims d, Geo<T> g) {
  const int64_t py = d.ny + 2, pz = d.nz + 2, pplane = py * pz, pvol = (d.nx + 2) * pplane;
  const int64_t n = d.c * d.vol();
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    int64_t k = i % d.nz, t = i / d.nz;
    int64_t j = t % d.ny;
    t /= d.ny;
    int64_t xi = t % d.nx, ch = t / d.nx;
    const T* q = p + ch * pvol + (xi + 1) * pplane + (j + 1) * pz + (k + 1);
    auto at = [&](int dx, int dy, int dz) { return q[dx * pplane + dy * pz + dz]; };
    if (!NORMAL) {
      out[i] = lap7(at(0, 0, 0), at(1, 0, 0), at(-1, 0, 0), at(0, 1, 0), at(0, -1, 0), at(0, 0, 1),
                    at(0, 0, -1), g);
    } else {
      T Pm[3][3], P0[3][3], Pp[3][3];
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
          Pm[a][b] = at(-1, a - 1, b - 1);
          P0[a][b] = at(0, a - 1, b - 1);
          Pp[a][b] = at(1, a - 1, b - 1);
        }
      out[i] = normal_lap_at<T, 3>(Pm, P0, Pp, 1, 1, g);
    }
  }
}

// ---------------------------------------------------------------------------
// Multiphase Allen-Cahn (problem_definition.py:508-573) in three launches:
// simplex projection per voxel, the ghost pad of the projected phases, and the
// right-hand side from the padded phases.  Expressions follow the reference's
// torch order with contraction off.
// ---------------------------------------------------------------------------
constexpr int MAX_PHASES = 8;

template <typename T>
__device__ __forceinline__ T clip_min0(T v) { return v != v ? v : (v < T(0) ? T(0) : v); }

// Sloppy projection clip(phis, 0) / sum (problem_definition.py:508-512) or the
// Euclidean one through sort / cumsum (problem_definition.py:514-530).
template <typename T, bool EUCLID>
__global__ void simplex_project_kernel(const T* __restrict__ in, T* __restrict__ out, int nph, int64_t vol) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < vol; i += (int64_t)gridDim.x * blockDim.x) {
    T v[MAX_PHASES];
    for (int p = 0; p < nph; ++p) v[p] = in[p * vol + i];
    if (!EUCLID) {
      T sum = T(0);
      for (int p = 0; p < nph; ++p) {
        v[p] = clip_min0(v[p]);
        sum = p == 0 ? v[p] : sum + v[p];
      }
      for (int p = 0; p < nph; ++p) out[p * vol + i] = v[p] / sum;
    } else {
      T u[MAX_PHASES];
      for (int p = 0; p < nph; ++p) u[p] = v[p];
      for (int a = 1; a < nph; ++a)  // insertion sort, descending
        for (int b = a; b > 0 && u[b - 1] < u[b]; --b) {
          const T tmp = u[b];
          u[b] = u[b - 1];
          u[b - 1] = tmp;
        }
      T csum = T(0), theta_num = T(0);
      int active = 0;
      T cssv[MAX_PHASES];
      for (int p = 0; p < nph; ++p) {
        csum = p == 0 ? u[0] : csum + u[p];
        cssv[p] = csum - T(1);
        if (u[p] - cssv[p] / static_cast<T>(p + 1) > T(0)) ++active;
      }
      if (active < 1) active = 1;
      theta_num = cssv[active - 1];
      const T theta = theta_num / static_cast<T>(active);
      for (int p = 0; p < nph; ++p) out[p * vol + i] = clip_min0(v[p] - theta);
    }
  }
}

template <typename T>
struct MpCoef {
  T neg_curv;     // -curvature
  T one_m_curv;   // 1 - curvature
  T pot_factor;   // 9/(2 eps^2)
  T neg_mgab;     // -M*gab
  T n_phases;     // the mean over phases divides by it
  bool normal;    // curvature != 1: the normal-Laplacian term is evaluated
};

template <typename T>
__global__ void multiphase_rhs_kernel(const T* __restrict__ phis, const T* __restrict__ pad, T* __restrict__ out,
                                      Dims d, Geo<T> g, MpCoef<T> k) {
  const int nph = (int)d.c;
  const int64_t py = d.ny + 2, pz = d.nz + 2, pplane = py * pz, pvol = (d.nx + 2) * pplane, vol = d.vol();
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < vol; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t z = i % d.nz, t = i / d.nz, y = t % d.ny, x = t / d.ny;
    const int64_t c0 = (x + 1) * pplane + (y + 1) * pz + (z + 1);
    T ph[MAX_PHASES], df[MAX_PHASES];
    T sum_sq = T(0);
    for (int p = 0; p < nph; ++p) {
      ph[p] = phis[p * vol + i];
      const T sq = ph[p] * ph[p];
      sum_sq = p == 0 ? sq : sum_sq + sq;
    }
    T dsum = T(0);
    for (int p = 0; p < nph; ++p) {
      const T* q = pad + p * pvol + c0;
      auto at = [&](int dx, int dy, int dz) { return q[dx * pplane + dy * pz + dz]; };
      T grad = k.neg_curv * lap7(at(0, 0, 0), at(1, 0, 0), at(-1, 0, 0), at(0, 1, 0), at(0, -1, 0), at(0, 0, 1),
                                 at(0, 0, -1), g);
      if (k.normal) {
        T Pm[3][3], P0[3][3], Pp[3][3];
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) {
            Pm[a][b] = at(-1, a - 1, b - 1);
            P0[a][b] = at(0, a - 1, b - 1);
            Pp[a][b] = at(1, a - 1, b - 1);
          }
        grad = grad - k.one_m_curv * normal_lap_at<T, 3>(Pm, P0, Pp, 1, 1, g);
      }
      const T phi = ph[p];
      const T pot = k.pot_factor * ((((T(3) * phi) * (sum_sq - phi * phi)) + (phi * phi) * phi) - phi);
      df[p] = grad + pot;
      dsum = p == 0 ? df[0] : dsum + df[p];
    }
    const T mean = dsum / k.n_phases;
    for (int p = 0; p < nph; ++p) out[p * vol + i] = k.neg_mgab * (df[p] - mean);
  }
}

// ---------------------------------------------------------------------------
// Smoothed-boundary reaction-diffusion (problem_definition.py:273-287) from
// the padded field z and padded mask m:
//   r = D * sum_a grad_a( grad_a(z) - to_face_a(z/m) * grad_a(m) ) + norm*bc_flux
// with face operators of fd_stencils.py:20-42, in the reference's order.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void sbm_rhs_kernel(const T* __restrict__ zp, const T* __restrict__ mp, const T* __restrict__ norm,
                               T norm_scalar, T* __restrict__ out, Dims d, Geo<T> g, T D, T bc_flux) {
  const int64_t py = d.ny + 2, pz = d.nz + 2, pplane = py * pz, pvol = (d.nx + 2) * pplane, vol = d.vol();
  const int64_t n = d.c * vol;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t ch = i / vol, v = i - ch * vol;
    const int64_t z = v % d.nz, t = v / d.nz, y = t % d.ny, x = t / d.ny;
    const int64_t c0 = (x + 1) * pplane + (y + 1) * pz + (z + 1);
    const T* zc = zp + ch * pvol + c0;
    const T* mc = mp + c0;
    // face flux between padded points lo and lo + s along an axis with weight id
    auto face = [&](int64_t lo, int64_t st, T id) {
      const T zl = zc[lo], zh = zc[lo + st], ml = mc[lo], mh = mc[lo + st];
      const T gz = (zh - zl) * id;
      const T tf = T(0.5) * (zh / mh + zl / ml);
      const T gm = (mh - ml) * id;
      return gz - tf * gm;
    };
    const T dx = (face(0, pplane, g.idx) - face(-pplane, pplane, g.idx)) * g.idx;
    const T dy = (face(0, pz, g.idy) - face(-pz, pz, g.idy)) * g.idy;
    const T dz = (face(0, 1, g.idz) - face(-1, 1, g.idz)) * g.idz;
    const T div = (dx + dy) + dz;
    const T nv = norm ? norm[v] : norm_scalar;
    out[i] = D * div + nv * bc_flux;
  }
}

template <typename T>
__global__ void axpy_kernel(const T* __restrict__ a, const T* __restrict__ b, T c, T* __restrict__ out,
                            int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = a[i] + c * b[i];
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
int check_common(int32_t dtype, const void* u, const void* out, const evx_grid* g, const evx_bc* bc,
                 Dims* d, Bc* b) {
  if (dtype != EVX_F32 && dtype != EVX_F64) return fail(EVX_ERR_INVALID, "dtype must be EVX_F32 or EVX_F64");
  if (!u || !out) return fail(EVX_ERR_INVALID, "null field pointer");
  EVX_TRY(make_dims(g, d));
  if (bc) {
    EVX_TRY(make_bc(bc, b));
    if (b->kind[0] == EVX_BC_NODE_EXTRAP && d->nx < 2)
      return fail(EVX_ERR_INVALID, "the staggered_x zero-flux rule needs at least two interior x planes");
  }
  return EVX_OK;
}

int check_epi(const evx_epilogue* e, const void* u, const void* out, Epi* ep) {
  Epi x{EVX_OUT_RHS, 0.0, nullptr, nullptr, nullptr};
  if (e) {
    x.mode = e->mode;
    x.coef = e->coef;
    x.base = e->base;
    x.acc_in = e->acc_in;
    x.acc_out = e->acc_out;
  }
  if (x.mode < EVX_OUT_RHS || x.mode > EVX_OUT_RK4_LAST) return fail(EVX_ERR_INVALID, "bad epilogue mode %d", x.mode);
  if ((x.mode == EVX_OUT_RK4_FIRST || x.mode == EVX_OUT_RK4_MID) && !x.acc_out)
    return fail(EVX_ERR_INVALID, "RK4 stage needs acc_out");
  if ((x.mode == EVX_OUT_RK4_MID || x.mode == EVX_OUT_RK4_LAST) && !x.acc_in)
    return fail(EVX_ERR_INVALID, "RK4 stage needs acc_in");
  if (u == out) return fail(EVX_ERR_INVALID, "stencil output must not alias its input");
  *ep = x;
  return EVX_OK;
}

// Blocks per launch of the streaming kernels: one resident round, as many
// blocks as the device holds at once (blocks per CU from the occupancy query
// for the kernel, times the CU count).  That gives the longest x-chunks, the
// fewest warm-up planes and no partial tail round; a fractional number of
// rounds leaves CUs idle at the end (measured on MI355X).
constexpr int64_t XCHUNK_BLOCKS = 1024;  // fallback when the occupancy query fails

template <typename K>
int64_t resident_round(K kernel) {
  int dev = 0, cus = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, NT, 0) != hipSuccess || per < 1 || cus < 1)
    return XCHUNK_BLOCKS;
  return (int64_t)per * cus;
}

// Cached per kernel instantiation (the first launch queries the device; all
// devices of a process are taken to be alike).
#define EVX_RESIDENT_ROUND(kernel)                                   \
  ([]() {                                                            \
    static const int64_t r = resident_round(kernel);                 \
    return r;                                                        \
  }())

template <typename T>
dim3 stencil_grid(const Dims& d, TileGrid* tg, int64_t target = XCHUNK_BLOCKS) {
  return dim3(stencil_tiles<T>(d, tg, target));
}

template <typename T>
int launch_rd(const void* u, void* out, const Dims& d, const Bc& bc, const evx_grid* g, double D,
              const Epi& ep, hipStream_t s) {
  TileGrid tg;
  dim3 grid = stencil_grid<T>(d, &tg, EVX_RESIDENT_ROUND(rd_rhs_kernel<T>));
  hipLaunchKernelGGL(rd_rhs_kernel<T>, grid, dim3(BZ, NTY), 0, s, static_cast<const T*>(u),
                     static_cast<T*>(out), d, bc, make_geo<T>(g), static_cast<T>(D), ep, tg);
  return check_launch("rd_rhs_kernel");
}

template <typename T>
int launch_gs(const void* u, void* out, const Dims& d, const Bc& bc, const evx_grid* g, double Da, double Db,
              double feed, double kill, const Epi& ep, hipStream_t s) {
  TileGrid tg;
  Dims d1 = d;
  d1.c = 1;  // one block covers both species of its tile
  dim3 grid = stencil_grid<T>(d1, &tg, EVX_RESIDENT_ROUND(gs_rhs_kernel<T>));
  GsCoef<T> k{static_cast<T>(Da), static_cast<T>(Db), static_cast<T>(feed), static_cast<T>(kill)};
  hipLaunchKernelGGL(gs_rhs_kernel<T>, grid, dim3(BZ, NTY), 0, s, static_cast<const T*>(u),
                     static_cast<T*>(out), d1, bc, make_geo<T>(g), k, ep, tg);
  return check_launch("gs_rhs_kernel");
}

constexpr int SLAB_XCHUNK = 16;

template <typename T, bool STRICT = true>
int launch_ch(const void* u, void* out, const Dims& d, const Bc& bc, const evx_grid* g, double eps,
              double D, const Epi& ep, hipStream_t s, const double* chan = nullptr, int64_t xa = 0,
              int64_t xb = -1) {
  TileGrid tg;
  dim3 grid = stencil_grid<T>(d, &tg, EVX_RESIDENT_ROUND((ch_rhs_kernel<T, STRICT>)));
  unsigned base = 0;
  if (xb >= 0) {  // x-slab [xa, xb) of a one-channel field, at tile-chunk granularity
    tg.xc = SLAB_XCHUNK;
    tg.nxc = (int)((d.nx + tg.xc - 1) / tg.xc);
    const int64_t per_chunk = (int64_t)tg.ntz * tg.nty;
    const int64_t ca = xa / tg.xc, cb = (xb + tg.xc - 1) / tg.xc;
    base = (unsigned)(ca * per_chunk);
    grid = dim3((unsigned)((cb - ca) * per_chunk));
  }
  ChCoef<T> k{static_cast<T>(18.0 / eps), static_cast<T>(2.0 * eps), static_cast<T>(D)};
  hipLaunchKernelGGL((ch_rhs_kernel<T, STRICT>), grid, dim3(BZ, NTY), 0, s, static_cast<const T*>(u),
                     static_cast<T*>(out), d, bc, make_geo<T>(g), k, ep, tg, chan, base);
  return check_launch("ch_rhs_kernel");
}

template <typename T>
int launch_ac(const void* u, void* out, const Dims& d, const Bc& bc, const evx_grid* g, double eps,
              double gab, double M, double force, double curv, const Epi& ep, hipStream_t s) {
  TileGrid tg;
  dim3 grid = stencil_grid<T>(d, &tg, EVX_RESIDENT_ROUND(ac_rhs_kernel<T>));
  AcCoef<T> k{static_cast<T>(18.0 / eps), static_cast<T>(eps), static_cast<T>(curv),
              static_cast<T>(1.0 - curv), static_cast<T>(gab), static_cast<T>(M),
              static_cast<T>(force), static_cast<T>(3.0 / eps)};
  hipLaunchKernelGGL(ac_rhs_kernel<T>, grid, dim3(BZ, NTY), 0, s, static_cast<const T*>(u),
                     static_cast<T*>(out), d, bc, make_geo<T>(g), k, ep, tg);
  return check_launch("ac_rhs_kernel");
}

inline unsigned flat_blocks(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b > 65535LL * 16) b = 65535LL * 16;
  return (unsigned)(b < 1 ? 1 : b);
}

}  // namespace
}  // namespace evx

namespace evx {
// Fully periodic right-hand sides written to a plain buffer; used by the
// semi-implicit spectral steps (timesteppers.py:85-86).
int stencil_rhs_ch_periodic(int32_t dtype, const void* u, void* out, const Dims& d, const double div_dx[3],
                            const double div_dx2[3], double eps, double D, const double* chan, hipStream_t s,
                            int64_t xa, int64_t xb) {
  evx_grid g{d.c, d.nx, d.ny, d.nz, {div_dx[0], div_dx[1], div_dx[2]}, {div_dx2[0], div_dx2[1], div_dx2[2]}};
  Bc b{};
  const Epi ep{EVX_OUT_RHS, 0.0, nullptr, nullptr, nullptr};
  if (d.ny * d.nz * (dtype == EVX_F32 ? 4 : 8) >= (int64_t(1) << 32))
    return fail(EVX_ERR_UNSUPPORTED, "a (y,z) plane of more than 4 GiB");
  return dtype == EVX_F32 ? launch_ch<float, false>(u, out, d, b, &g, eps, D, ep, s, chan, xa, xb)
                          : launch_ch<double, false>(u, out, d, b, &g, eps, D, ep, s, chan, xa, xb);
}

// x-planes per CH tile chunk of slab launches (slabs start on a multiple of
// it); small enough that one slab still spreads over every CU.
int stencil_ch_xchunk(int32_t, const Dims&) { return SLAB_XCHUNK; }

int stencil_rhs_rd_periodic(int32_t dtype, const void* u, void* out, const Dims& d, const double div_dx2[3],
                            double D, hipStream_t s) {
  evx_grid g{d.c, d.nx, d.ny, d.nz, {0.0, 0.0, 0.0}, {div_dx2[0], div_dx2[1], div_dx2[2]}};
  Bc b{};
  const Epi ep{EVX_OUT_RHS, 0.0, nullptr, nullptr, nullptr};
  return dtype == EVX_F32 ? launch_rd<float>(u, out, d, b, &g, D, ep, s)
                          : launch_rd<double>(u, out, d, b, &g, D, ep, s);
}
}  // namespace evx

using namespace evx;

extern "C" int evx_pad_bc(int32_t dtype, const void* u, void* out, const evx_grid* g, const evx_bc* bc,
                          void* stream) {
  Dims d;
  Bc b;
  if (!bc) return fail(EVX_ERR_INVALID, "null bc");
  EVX_TRY(check_common(dtype, u, out, g, bc, &d, &b));
  const int64_t n = d.c * (d.nx + 2) * (d.ny + 2) * (d.nz + 2);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == EVX_F32)
    hipLaunchKernelGGL(pad_bc_kernel<float>, dim3(flat_blocks(n)), dim3(256), 0, s,
                       static_cast<const float*>(u), static_cast<float*>(out), d, b);
  else
    hipLaunchKernelGGL(pad_bc_kernel<double>, dim3(flat_blocks(n)), dim3(256), 0, s,
                       static_cast<const double*>(u), static_cast<double*>(out), d, b);
  return check_launch("pad_bc_kernel");
}

template <bool NORMAL>
static int padded_stencil(int32_t dtype, const void* p, void* out, const evx_grid* g, void* stream) {
  Dims d;
  Bc b;
  EVX_TRY(check_common(dtype, p, out, g, nullptr, &d, &b));
  const int64_t n = d.c * d.vol();
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == EVX_F32)
    hipLaunchKernelGGL((padded_stencil_kernel<float, NORMAL>), dim3(flat_blocks(n)), dim3(256), 0, s,
                       static_cast<const float*>(p), static_cast<float*>(out), d, make_geo<float>(g));
  else
    hipLaunchKernelGGL((padded_stencil_kernel<double, NORMAL>), dim3(flat_blocks(n)), dim3(256), 0, s,
                       static_cast<const double*>(p), static_cast<double*>(out), d, make_geo<double>(g));
  return check_launch("padded_stencil_kernel");
}

extern "C" int evx_laplace_padded(int32_t dtype, const void* padded, void* out, const evx_grid* g,
                                  void* stream) {
  return padded_stencil<false>(dtype, padded, out, g, stream);
}

extern "C" int evx_normal_laplace_padded(int32_t dtype, const void* padded, void* out, const evx_grid* g,
                                         void* stream) {
  return padded_stencil<true>(dtype, padded, out, g, stream);
}

extern "C" int evx_rd_rhs(int32_t dtype, const void* u, void* out, const evx_grid* g, const evx_bc* bc,
                          double D, const evx_epilogue* e, void* stream) {
  Dims d;
  Bc b;
  Epi ep;
  if (!bc) return fail(EVX_ERR_INVALID, "null bc");
  EVX_TRY(check_common(dtype, u, out, g, bc, &d, &b));
  EVX_TRY(check_epi(e, u, out, &ep));
  hipStream_t s = static_cast<hipStream_t>(stream);
  return dtype == EVX_F32 ? launch_rd<float>(u, out, d, b, g, D, ep, s)
                          : launch_rd<double>(u, out, d, b, g, D, ep, s);
}

extern "C" int evx_coupled_rd_rhs(int32_t dtype, const void* u, void* out, const evx_grid* g, const evx_bc* bc,
                                  double D_A, double D_B, double feed, double kill, const evx_epilogue* e,
                                  void* stream) {
  Dims d;
  Bc b;
  Epi ep;
  if (!bc) return fail(EVX_ERR_INVALID, "null bc");
  EVX_TRY(check_common(dtype, u, out, g, bc, &d, &b));
  EVX_TRY(check_epi(e, u, out, &ep));
  if (d.c != 2) return fail(EVX_ERR_INVALID, "the two-species right-hand side needs exactly two channels");
  hipStream_t s = static_cast<hipStream_t>(stream);
  return dtype == EVX_F32 ? launch_gs<float>(u, out, d, b, g, D_A, D_B, feed, kill, ep, s)
                          : launch_gs<double>(u, out, d, b, g, D_A, D_B, feed, kill, ep, s);
}

extern "C" int evx_ch_rhs(int32_t dtype, const void* u, void* out, const evx_grid* g, const evx_bc* bc,
                          double eps, double D, const evx_epilogue* e, void* stream) {
  Dims d;
  Bc b;
  Epi ep;
  if (!bc) return fail(EVX_ERR_INVALID, "null bc");
  EVX_TRY(check_common(dtype, u, out, g, bc, &d, &b));
  EVX_TRY(check_epi(e, u, out, &ep));
  if (!(eps != 0.0)) return fail(EVX_ERR_INVALID, "eps must be nonzero");
  if (staggered_x(b))
    return fail(EVX_ERR_UNSUPPORTED, "the fused Cahn-Hilliard kernel implements the cell_center rules only");
  hipStream_t s = static_cast<hipStream_t>(stream);
  return dtype == EVX_F32 ? launch_ch<float>(u, out, d, b, g, eps, D, ep, s)
                          : launch_ch<double>(u, out, d, b, g, eps, D, ep, s);
}

extern "C" int evx_ac_rhs(int32_t dtype, const void* u, void* out, const evx_grid* g, const evx_bc* bc,
                          double eps, double gab, double M, double force, double curvature,
                          const evx_epilogue* e, void* stream) {
  Dims d;
  Bc b;
  Epi ep;
  if (!bc) return fail(EVX_ERR_INVALID, "null bc");
  EVX_TRY(check_common(dtype, u, out, g, bc, &d, &b));
  EVX_TRY(check_epi(e, u, out, &ep));
  if (!(eps != 0.0)) return fail(EVX_ERR_INVALID, "eps must be nonzero");
  hipStream_t s = static_cast<hipStream_t>(stream);
  return dtype == EVX_F32 ? launch_ac<float>(u, out, d, b, g, eps, gab, M, force, curvature, ep, s)
                          : launch_ac<double>(u, out, d, b, g, eps, gab, M, force, curvature, ep, s);
}

extern "C" int evx_simplex_project(int32_t dtype, const void* phis, void* out, int32_t n_phases, int64_t voxels,
                                   int32_t euclidean, void* stream) {
  if (dtype != EVX_F32 && dtype != EVX_F64) return fail(EVX_ERR_INVALID, "bad dtype");
  if (n_phases < 1 || n_phases > MAX_PHASES)
    return fail(EVX_ERR_UNSUPPORTED, "the fused projection handles 1 to %d phases", MAX_PHASES);
  if (voxels < 0 || (voxels > 0 && (!phis || !out))) return fail(EVX_ERR_INVALID, "bad projection arguments");
  if (voxels == 0) return EVX_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const dim3 grid(flat_blocks(voxels)), block(256);
  if (dtype == EVX_F32) {
    if (euclidean)
      hipLaunchKernelGGL((simplex_project_kernel<float, true>), grid, block, 0, s, static_cast<const float*>(phis),
                         static_cast<float*>(out), (int)n_phases, voxels);
    else
      hipLaunchKernelGGL((simplex_project_kernel<float, false>), grid, block, 0, s, static_cast<const float*>(phis),
                         static_cast<float*>(out), (int)n_phases, voxels);
  } else {
    if (euclidean)
      hipLaunchKernelGGL((simplex_project_kernel<double, true>), grid, block, 0, s, static_cast<const double*>(phis),
                         static_cast<double*>(out), (int)n_phases, voxels);
    else
      hipLaunchKernelGGL((simplex_project_kernel<double, false>), grid, block, 0, s, static_cast<const double*>(phis),
                         static_cast<double*>(out), (int)n_phases, voxels);
  }
  return check_launch("simplex_project_kernel");
}

template <typename T>
static int launch_multiphase(const void* phis, const void* padded, void* out, const Dims& d, const evx_grid* g,
                             double eps, double gab, double M, double curvature, hipStream_t s) {
  MpCoef<T> k{static_cast<T>(-curvature), static_cast<T>(1.0 - curvature), static_cast<T>(9.0 / (2.0 * eps * eps)),
              static_cast<T>((-M) * gab), static_cast<T>(d.c), curvature != 1.0};
  hipLaunchKernelGGL(multiphase_rhs_kernel<T>, dim3(flat_blocks(d.vol())), dim3(256), 0, s,
                     static_cast<const T*>(phis), static_cast<const T*>(padded), static_cast<T*>(out), d,
                     make_geo<T>(g), k);
  return check_launch("multiphase_rhs_kernel");
}

extern "C" int evx_multiphase_ac_rhs_padded(int32_t dtype, const void* phis, const void* padded, void* out,
                                            const evx_grid* g, double eps, double gab, double M, double curvature,
                                            void* stream) {
  Dims d;
  Bc b;
  EVX_TRY(check_common(dtype, phis, out, g, nullptr, &d, &b));
  if (!padded) return fail(EVX_ERR_INVALID, "null padded field");
  if (d.c < 1 || d.c > MAX_PHASES)
    return fail(EVX_ERR_UNSUPPORTED, "the fused multiphase right-hand side handles 1 to %d phases", MAX_PHASES);
  if (!(eps != 0.0)) return fail(EVX_ERR_INVALID, "eps must be nonzero");
  hipStream_t s = static_cast<hipStream_t>(stream);
  return dtype == EVX_F32 ? launch_multiphase<float>(phis, padded, out, d, g, eps, gab, M, curvature, s)
                          : launch_multiphase<double>(phis, padded, out, d, g, eps, gab, M, curvature, s);
}

template <typename T>
static int launch_sbm(const void* z, const void* m, const void* norm, double norm_scalar, void* out, const Dims& d,
                      const evx_grid* g, double D, double bc_flux, hipStream_t s) {
  hipLaunchKernelGGL(sbm_rhs_kernel<T>, dim3(flat_blocks(d.c * d.vol())), dim3(256), 0, s,
                     static_cast<const T*>(z), static_cast<const T*>(m), static_cast<const T*>(norm),
                     static_cast<T>(norm_scalar), static_cast<T*>(out), d, make_geo<T>(g), static_cast<T>(D),
                     static_cast<T>(bc_flux));
  return check_launch("sbm_rhs_kernel");
}

extern "C" int evx_sbm_rhs_padded(int32_t dtype, const void* z_padded, const void* mask_padded, const void* norm,
                                  double norm_scalar, void* out, const evx_grid* g, double D, double bc_flux,
                                  void* stream) {
  Dims d;
  Bc b;
  EVX_TRY(check_common(dtype, z_padded, out, g, nullptr, &d, &b));
  if (!mask_padded) return fail(EVX_ERR_INVALID, "null padded mask");
  hipStream_t s = static_cast<hipStream_t>(stream);
  return dtype == EVX_F32 ? launch_sbm<float>(z_padded, mask_padded, norm, norm_scalar, out, d, g, D, bc_flux, s)
                          : launch_sbm<double>(z_padded, mask_padded, norm, norm_scalar, out, d, g, D, bc_flux, s);
}

extern "C" int evx_axpy(int32_t dtype, const void* a, const void* b, double coef, void* out, int64_t n,
                        void* stream) {
  if (dtype != EVX_F32 && dtype != EVX_F64) return fail(EVX_ERR_INVALID, "bad dtype");
  if (n < 0 || (n > 0 && (!a || !b || !out))) return fail(EVX_ERR_INVALID, "bad axpy arguments");
  if (n == 0) return EVX_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == EVX_F32)
    hipLaunchKernelGGL(axpy_kernel<float>, dim3(flat_blocks(n)), dim3(256), 0, s,
                       static_cast<const float*>(a), static_cast<const float*>(b), static_cast<float>(coef),
                       static_cast<float*>(out), n);
  else
    hipLaunchKernelGGL(axpy_kernel<double>, dim3(flat_blocks(n)), dim3(256), 0, s,
                       static_cast<const double*>(a), static_cast<const double*>(b), coef,
                       static_cast<double*>(out), n);
  return check_launch("axpy_kernel");
}
