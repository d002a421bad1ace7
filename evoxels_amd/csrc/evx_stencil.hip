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
#include "evx_common.hpp"
#include "evx_fft_engine.hpp"
#include "evx_ch_math.hpp"

#pragma clang fp contract(off)

namespace evx {
namespace {

constexpr int BZ = 64;   // threads (and tile width) along z
constexpr int NTY = 4;   // thread rows along y
constexpr int NT = BZ * NTY;

template <typename T>
struct TileCfg;
template <>
struct TileCfg<float> { static constexpr int PY = 4; };   // BY = 16
template <>
struct TileCfg<double> { static constexpr int PY = 2; };  // BY = 8

struct Epi {
  int mode;
  double coef;
  const void* base;
  const void* acc_in;
  void* acc_out;
};

// Writes the fused output for right-hand side value r at flat index i.
template <typename T>
__device__ __forceinline__ void epilogue(const Epi& ep, const T* __restrict__ u,
                                         T* __restrict__ out, int64_t i, T r) {
  const T c = static_cast<T>(ep.coef);
  switch (ep.mode) {
    case EVX_OUT_RHS:
      out[i] = r;
      break;
    case EVX_OUT_EULER:
      out[i] = u[i] + c * r;
      break;
    case EVX_OUT_RK4_FIRST: {
      const T* b = ep.base ? static_cast<const T*>(ep.base) : u;
      out[i] = b[i] + c * r;
      static_cast<T*>(ep.acc_out)[i] = r;
      break;
    }
    case EVX_OUT_RK4_MID: {
      const T* b = ep.base ? static_cast<const T*>(ep.base) : u;
      const T a = static_cast<const T*>(ep.acc_in)[i];
      out[i] = b[i] + c * r;
      static_cast<T*>(ep.acc_out)[i] = a + T(2) * r;
      break;
    }
    default: {  // EVX_OUT_RK4_LAST
      const T* b = ep.base ? static_cast<const T*>(ep.base) : u;
      const T a = static_cast<const T*>(ep.acc_in)[i];
      out[i] = b[i] + c * (a + r);
      break;
    }
  }
}

// As epilogue(), with the state value at i supplied by the caller (no reload).
template <typename T>
__device__ __forceinline__ void epilogue_v(const Epi& ep, T u_i, T* __restrict__ out, int64_t i, T r) {
  const T c = static_cast<T>(ep.coef);
  switch (ep.mode) {
    case EVX_OUT_RHS:
      out[i] = r;
      break;
    case EVX_OUT_EULER:
      out[i] = u_i + c * r;
      break;
    case EVX_OUT_RK4_FIRST: {
      const T b = ep.base ? static_cast<const T*>(ep.base)[i] : u_i;
      out[i] = b + c * r;
      static_cast<T*>(ep.acc_out)[i] = r;
      break;
    }
    case EVX_OUT_RK4_MID: {
      const T b = ep.base ? static_cast<const T*>(ep.base)[i] : u_i;
      const T a = static_cast<const T*>(ep.acc_in)[i];
      out[i] = b + c * r;
      static_cast<T*>(ep.acc_out)[i] = a + T(2) * r;
      break;
    }
    default: {
      const T b = ep.base ? static_cast<const T*>(ep.base)[i] : u_i;
      const T a = static_cast<const T*>(ep.acc_in)[i];
      out[i] = b + c * (a + r);
      break;
    }
  }
}

template <typename T>
__host__ Geo<T> make_geo(const evx_grid* g) {
  return make_geo_from<T>(g->div_dx, g->div_dx2);
}

// Register-staged loader of one x-plane tile [y0-H, y0+BY+H) x [z0-H, z0+BZ+H).
// init() resolves, once per block, the y/z part of every element this thread
// loads: periodic axes wrap, non-periodic ones clamp and remember the ghost
// rule.  fetch() then only resolves the (uniform) x index of the plane and
// issues one load per element; commit() applies clip and the ghost rules and
// writes the element to LDS slot tid + k*NT of the plane.
template <typename T, int H, int ROWS, int W, bool CLIP>
struct PlaneLoader {
  static constexpr int NPT = (ROWS * W + NT - 1) / NT;
  int off[NPT];      // in-plane offset y*nz + z of the source point, -1 if unused
  uint8_t yzf[NPT];  // y/z ghost flags: 4/8 = y low/high, 16/32 = z low/high
  T v[NPT];
  uint8_t xf;        // x ghost flag of the fetched plane: 1 = low, 2 = high

  __device__ __forceinline__ void init(int64_t y0, int64_t z0, const Dims& d, const Bc& bc, int tid) {
#pragma unroll
    for (int k = 0; k < NPT; ++k) {
      const int i = tid + k * NT;
      off[k] = -1;
      yzf[k] = 0;
      if (i < ROWS * W) {
        int64_t gy = y0 - H + i / W, gz = z0 - H + i % W;
        uint8_t f = 0;
        if (gy < 0 || gy >= d.ny) {
          if (bc.kind[1] == EVX_BC_PERIODIC) gy = wrap_near(gy, d.ny);
          else { f |= gy < 0 ? 4 : 8; gy = gy < 0 ? 0 : d.ny - 1; }
        }
        if (gz < 0 || gz >= d.nz) {
          if (bc.kind[2] == EVX_BC_PERIODIC) gz = wrap_near(gz, d.nz);
          else { f |= gz < 0 ? 16 : 32; gz = gz < 0 ? 0 : d.nz - 1; }
        }
        off[k] = (int)(gy * d.nz + gz);
        yzf[k] = f;
      }
    }
  }

  __device__ __forceinline__ void fetch(const T* __restrict__ uc, int64_t xg, const Dims& d, const Bc& bc) {
    xf = 0;
    if (xg < 0 || xg >= d.nx) {
      if (bc.kind[0] == EVX_BC_PERIODIC) xg = wrap_near(xg, d.nx);
      else { xf = xg < 0 ? 1 : 2; xg = xg < 0 ? 0 : d.nx - 1; }
    }
    const T* pl = uc + xg * (d.ny * d.nz);
#pragma unroll
    for (int k = 0; k < NPT; ++k)
      if (off[k] >= 0) v[k] = pl[off[k]];
  }

  __device__ __forceinline__ void commit(T* flat, const Bc& bc, int tid) const {
#pragma unroll
    for (int k = 0; k < NPT; ++k) {
      if (off[k] >= 0) {
        T x = CLIP ? clip01(v[k]) : v[k];
        const int f = xf | yzf[k];
        if (f) {
          const int8_t o[3] = {(int8_t)((f & 1) ? -1 : (f & 2) ? 1 : 0), (int8_t)((f & 4) ? -1 : (f & 8) ? 1 : 0),
                               (int8_t)((f & 16) ? -1 : (f & 32) ? 1 : 0)};
          x = apply_rules(x, o, bc);
        }
        flat[tid + k * NT] = x;
      }
    }
  }
};

// Ring slot of a plane at small relative index rel >= -3 (32-bit, constant modulus).
template <int NS>
__device__ __forceinline__ int ring(int64_t rel) { return (int)((unsigned)((int)rel + 3 * NS) % NS); }

// Flat launch: tiles ordered z-tile fastest, then y-tile, then x-chunk and
// channel; xcd_tile() gives each XCD a contiguous range so y/z halos of
// neighbouring tiles are shared through one L2.
struct TileGrid {
  int ntz, nty, nxc, xc;
};

struct TileOrigin {
  int64_t z0, y0, x0, x1, ch;
};

template <int BY>
__device__ __forceinline__ TileOrigin tile_origin(const TileGrid& tg, const Dims& d) {
  unsigned t = xcd_tile(blockIdx.x, gridDim.x);
  TileOrigin o;
  o.z0 = (int64_t)(t % tg.ntz) * BZ;
  t /= tg.ntz;
  o.y0 = (int64_t)(t % tg.nty) * BY;
  t /= tg.nty;
  o.x0 = (int64_t)(t % tg.nxc) * tg.xc;
  o.ch = t / tg.nxc;
  o.x1 = min(o.x0 + (int64_t)tg.xc, d.nx);
  return o;
}

__host__ inline int pick_xchunk(const Dims& d, int64_t tiles_yz) {
  // Enough blocks to fill 256 CUs a few times over, but at least 8 planes
  // per block so the warm-up planes stay a small fraction of the traffic.
  const int64_t target = 2048;
  int64_t chunks = (target + tiles_yz * d.c - 1) / (tiles_yz * d.c);
  if (chunks < 1) chunks = 1;
  int64_t xc = (d.nx + chunks - 1) / chunks;
  if (xc < 8) xc = 8;
  if (xc > d.nx) xc = d.nx;
  return (int)xc;
}

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
  const TileOrigin o = tile_origin<BY>(tg, d);
  const int64_t z0 = o.z0, y0 = o.y0, x0 = o.x0, x1 = o.x1;
  const T* uc = u + o.ch * d.vol();
  T* oc = out + o.ch * d.vol();
  PlaneLoader<T, 1, ROWS, W, false> pl;
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
// Cahn-Hilliard (problem_definition.py:328-371):
//   c = clip(u,0,1); mu = ((18/eps*c)*(1-c))*(1-2c) - (2eps)*lap(cB)
//   r = D * sum_axes grad_face( (cf*(1-cf)) * grad_face(muB) )
// mu lives in a 4-plane LDS ring with a one-voxel (y,z) halo; c in a 6-plane
// ring with a two-voxel halo.  mu ghosts (non-periodic faces) are the BC rule
// applied to mu at the adjacent interior voxel, as pad_bc(mu) does.
// ---------------------------------------------------------------------------
// Thread (ty, tz) owns PY consecutive rows r0 = ty*PY .. r0+PY-1 of column tz.
// Per output plane it evaluates mu for its own rows (sharing the column loads
// between rows) while threads 0..H-1 evaluate the one-voxel mu halo with the
// full ghost logic; the flux through the face between two own rows is
// computed once, and the flux through the x face x+1/2 is carried in a
// register to the next plane.  The c ring holds 5 planes (x-2..x+2), the mu
// ring 3 (x-1..x+1); the plane x+3 is fetched into registers during the
// arithmetic of plane x.
template <typename T, bool STRICT>
__global__ __launch_bounds__(NT) void ch_rhs_kernel(const T* __restrict__ u, T* __restrict__ out, Dims d, Bc bc,
                                                    Geo<T> g, ChCoef<T> k, Epi ep, TileGrid tg) {
  using Op = Ops<STRICT>;
  constexpr int PY = TileCfg<T>::PY, BY = NTY * PY;
  constexpr int CW = BZ + 4, CR = BY + 4, NC = 5;
  constexpr int MW = BZ + 2, MR = BY + 2, NM = 3;
  constexpr int HALO = 2 * MW + 2 * BY;  // mu halo points: rows -1, BY and columns -1, BZ
  __shared__ T cs[NC][CR][CW];
  __shared__ T ms[NM][MR][MW];
  const int tz = threadIdx.x, ty = threadIdx.y, tid = ty * BZ + tz;
  const TileOrigin o = tile_origin<BY>(tg, d);
  const int64_t z0 = o.z0, y0 = o.y0, x0 = o.x0, x1 = o.x1;
  const T* uc = u + o.ch * d.vol();
  T* oc = out + o.ch * d.vol();
  const bool xper = bc.kind[0] == EVX_BC_PERIODIC;
  const bool yper = bc.kind[1] == EVX_BC_PERIODIC, zper = bc.kind[2] == EVX_BC_PERIODIC;
  const int r0 = ty * PY;
  const int64_t gz = z0 + tz;

  // mu at tile point (yy, zz) of x-plane xm with every ghost rule applied.
  auto mu_general = [&](int64_t xm, int yy, int zz) -> T {
    int8_t oo[3] = {0, 0, 0};
    int64_t xs = xm;
    if (!xper && xm < 0) { oo[0] = -1; xs = 0; }
    if (!xper && xm >= d.nx) { oo[0] = 1; xs = d.nx - 1; }
    const T(*Cm)[CW] = cs[ring<NC>(xs - 1 - x0 + 2)];
    const T(*C0)[CW] = cs[ring<NC>(xs - x0 + 2)];
    const T(*Cp)[CW] = cs[ring<NC>(xs + 1 - x0 + 2)];
    const int64_t gy = y0 + yy, gzz = z0 + zz;
    int a = yy + 2, b = zz + 2;
    if (!yper && (gy < 0 || gy >= d.ny)) {
      oo[1] = gy < 0 ? -1 : 1;
      a += gy < 0 ? 1 : -1;
    }
    if (!zper && (gzz < 0 || gzz >= d.nz)) {
      oo[2] = gzz < 0 ? -1 : 1;
      b += gzz < 0 ? 1 : -1;
    }
    T m = Op::mu(C0[a][b], Cp[a][b], Cm[a][b], C0[a + 1][b], C0[a - 1][b], C0[a][b + 1], C0[a][b - 1], g, k);
    if (oo[0] | oo[1] | oo[2]) m = apply_rules(m, oo, bc);
    return m;
  };

  // Tiles that cut a non-periodic y/z boundary evaluate every mu point with
  // the ghost logic; all other tiles only their halo.
  const bool tile_partial = (!yper && y0 + BY > d.ny) || (!zper && z0 + BZ > d.nz);
  // This thread's point of the one-voxel mu halo (rows -1, BY and columns -1,
  // BZ, corners excluded), resolved once: c coordinates (ha, hb) of the source
  // point and its y/z ghost flags.
  constexpr int NHALO = 2 * MW + 2 * BY - 4;
  int hy = 0, hz = 0, ha = 0, hb = 0;
  int8_t hoy = 0, hoz = 0;
  if (tid < NHALO) {
    int h = tid;
    if (h < 2 * (MW - 2)) {
      hy = h < MW - 2 ? -1 : BY;
      hz = h < MW - 2 ? h : h - (MW - 2);
    } else {
      h -= 2 * (MW - 2);
      hy = h % BY;
      hz = h < BY ? -1 : BZ;
    }
    ha = hy + 2;
    hb = hz + 2;
    const int64_t gy = y0 + hy, gzz = z0 + hz;
    if (!yper && (gy < 0 || gy >= d.ny)) { hoy = gy < 0 ? -1 : 1; ha += gy < 0 ? 1 : -1; }
    if (!zper && (gzz < 0 || gzz >= d.nz)) { hoz = gzz < 0 ? -1 : 1; hb += gzz < 0 ? 1 : -1; }
  }

  auto compute_mu = [&](int64_t xm) {
    T(*M)[MW] = ms[ring<NM>(xm - x0 + 1)];
    const bool general = tile_partial || (!xper && (xm < 0 || xm >= d.nx));
    if (!general) {
      const T(*Cm)[CW] = cs[ring<NC>(xm - 1 - x0 + 2)];
      const T(*C0)[CW] = cs[ring<NC>(xm - x0 + 2)];
      const T(*Cp)[CW] = cs[ring<NC>(xm + 1 - x0 + 2)];
      T col[PY + 2];
#pragma unroll
      for (int j = 0; j < PY + 2; ++j) col[j] = C0[r0 + j + 1][tz + 2];
#pragma unroll
      for (int q = 0; q < PY; ++q) {
        const int a = r0 + q + 2, b = tz + 2;
        M[r0 + q + 1][tz + 1] =
            Op::mu(col[q + 1], Cp[a][b], Cm[a][b], col[q + 2], col[q], C0[a][b + 1], C0[a][b - 1], g, k);
      }
      if (tid < NHALO) {
        T m = Op::mu(C0[ha][hb], Cp[ha][hb], Cm[ha][hb], C0[ha + 1][hb], C0[ha - 1][hb], C0[ha][hb + 1],
                     C0[ha][hb - 1], g, k);
        if (hoy | hoz) {
          const int8_t oo[3] = {0, hoy, hoz};
          m = apply_rules(m, oo, bc);
        }
        M[hy + 1][hz + 1] = m;
      }
    } else {
      for (int p = tid; p < BY * BZ + NHALO; p += NT) {
        int yy, zz;
        if (p < BY * BZ) {
          yy = p / BZ;
          zz = p % BZ;
        } else {
          int h = p - BY * BZ;
          if (h < 2 * (MW - 2)) {
            yy = h < MW - 2 ? -1 : BY;
            zz = h < MW - 2 ? h : h - (MW - 2);
          } else {
            h -= 2 * (MW - 2);
            yy = h % BY;
            zz = h < BY ? -1 : BZ;
          }
        }
        M[yy + 1][zz + 1] = mu_general(xm, yy, zz);
      }
    }
  };

  PlaneLoader<T, 2, CR, CW, true> pl;
  pl.init(y0, z0, d, bc, tid);
  for (int64_t xg = x0 - 2; xg <= x0 + 2; ++xg) {
    pl.fetch(uc, xg, d, bc);
    pl.commit(&cs[ring<NC>(xg - x0 + 2)][0][0], bc, tid);
  }
  if (x0 + 3 <= x1 + 1) pl.fetch(uc, x0 + 3, d, bc);
  __syncthreads();
  compute_mu(x0 - 1);
  compute_mu(x0);
  __syncthreads();  // the first commit below reuses the slot mu(x0-1) read
  const int64_t plane = d.ny * d.nz;
  T carry[PY];  // flux through the x face x-1/2 of each own row
  for (int64_t x = x0; x < x1; ++x) {
    // plane x+3 (fetched one iteration ago) goes to the slot of plane x-2;
    // it feeds mu(x+2), needed while x+1 < x1
    if (x + 3 <= x1 + 1) pl.commit(&cs[ring<NC>(x + 3 - x0 + 2)][0][0], bc, tid);
    if (x + 4 <= x1 + 1) pl.fetch(uc, x + 4, d, bc);
    compute_mu(x + 1);
    __syncthreads();
    const T(*C0)[CW] = cs[ring<NC>(x - x0 + 2)];
    const T(*Cp)[CW] = cs[ring<NC>(x + 1 - x0 + 2)];
    const T(*M0)[MW] = ms[ring<NM>(x - x0 + 1)];
    const T(*Mp)[MW] = ms[ring<NM>(x + 1 - x0 + 1)];
    const int b = tz + 2, bm = tz + 1;
    T cc[PY + 2], mc[PY + 2], fy[PY + 1];
#pragma unroll
    for (int j = 0; j < PY + 2; ++j) {
      cc[j] = C0[r0 + j + 1][b];
      mc[j] = M0[r0 + j][bm];
    }
#pragma unroll
    for (int j = 0; j <= PY; ++j) fy[j] = Op::flux(cc[j], cc[j + 1], mc[j], mc[j + 1], g.idy);
    if (x == x0) {
      const T(*Cm)[CW] = cs[ring<NC>(x - 1 - x0 + 2)];
      const T(*Mm)[MW] = ms[ring<NM>(x - 1 - x0 + 1)];
#pragma unroll
      for (int q = 0; q < PY; ++q) carry[q] = Op::flux(Cm[r0 + q + 2][b], cc[q + 1], Mm[r0 + q + 1][bm], mc[q + 1], g.idx);
    }
#pragma unroll
    for (int q = 0; q < PY; ++q) {
      const int a = r0 + q + 2, am = r0 + q + 1;
      const T c0 = cc[q + 1], m0 = mc[q + 1];
      const T fxp = Op::flux(c0, Cp[a][b], m0, Mp[am][bm], g.idx);
      const T fzm = Op::flux(C0[a][b - 1], c0, M0[am][bm - 1], m0, g.idz);
      const T fzp = Op::flux(c0, C0[a][b + 1], m0, M0[am][bm + 1], g.idz);
      const T r = Op::div(carry[q], fxp, fy[q], fy[q + 1], fzm, fzp, g, k.D);
      carry[q] = fxp;
      const int64_t gy = y0 + r0 + q;
      if (gy < d.ny && gz < d.nz) epilogue(ep, uc, oc, x * plane + gy * d.nz + gz, r);
    }
    __syncthreads();
  }
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
  const TileOrigin o = tile_origin<BY>(tg, d);
  const int64_t z0 = o.z0, y0 = o.y0, x0 = o.x0, x1 = o.x1;
  const T* uc = u + o.ch * d.vol();
  T* oc = out + o.ch * d.vol();
  PlaneLoader<T, 1, ROWS, W, true> pl;
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
    out[i] = apply_rules(u[ch * d.vol() + r.off], r.o, bc);
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
  if (bc) EVX_TRY(make_bc(bc, b));
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

template <typename T>
dim3 stencil_grid(const Dims& d, TileGrid* tg) {
  constexpr int BY = NTY * TileCfg<T>::PY;
  const int64_t ty = (d.ny + BY - 1) / BY, tzn = (d.nz + BZ - 1) / BZ;
  tg->xc = pick_xchunk(d, ty * tzn);
  const int64_t nxc = (d.nx + tg->xc - 1) / tg->xc;
  tg->ntz = (int)tzn;
  tg->nty = (int)ty;
  tg->nxc = (int)nxc;
  return dim3((unsigned)(tzn * ty * nxc * d.c));
}

template <typename T>
int launch_rd(const void* u, void* out, const Dims& d, const Bc& bc, const evx_grid* g, double D,
              const Epi& ep, hipStream_t s) {
  TileGrid tg;
  dim3 grid = stencil_grid<T>(d, &tg);
  hipLaunchKernelGGL(rd_rhs_kernel<T>, grid, dim3(BZ, NTY), 0, s, static_cast<const T*>(u),
                     static_cast<T*>(out), d, bc, make_geo<T>(g), static_cast<T>(D), ep, tg);
  return check_launch("rd_rhs_kernel");
}

template <typename T, bool STRICT = true>
int launch_ch(const void* u, void* out, const Dims& d, const Bc& bc, const evx_grid* g, double eps,
              double D, const Epi& ep, hipStream_t s) {
  TileGrid tg;
  dim3 grid = stencil_grid<T>(d, &tg);
  ChCoef<T> k{static_cast<T>(18.0 / eps), static_cast<T>(2.0 * eps), static_cast<T>(D)};
  hipLaunchKernelGGL((ch_rhs_kernel<T, STRICT>), grid, dim3(BZ, NTY), 0, s, static_cast<const T*>(u),
                     static_cast<T*>(out), d, bc, make_geo<T>(g), k, ep, tg);
  return check_launch("ch_rhs_kernel");
}

template <typename T>
int launch_ac(const void* u, void* out, const Dims& d, const Bc& bc, const evx_grid* g, double eps,
              double gab, double M, double force, double curv, const Epi& ep, hipStream_t s) {
  TileGrid tg;
  dim3 grid = stencil_grid<T>(d, &tg);
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
                            const double div_dx2[3], double eps, double D, hipStream_t s) {
  evx_grid g{d.c, d.nx, d.ny, d.nz, {div_dx[0], div_dx[1], div_dx[2]}, {div_dx2[0], div_dx2[1], div_dx2[2]}};
  Bc b{};
  const Epi ep{EVX_OUT_RHS, 0.0, nullptr, nullptr, nullptr};
  return dtype == EVX_F32 ? launch_ch<float, false>(u, out, d, b, &g, eps, D, ep, s)
                          : launch_ch<double, false>(u, out, d, b, &g, eps, D, ep, s);
}

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

extern "C" int evx_ch_rhs(int32_t dtype, const void* u, void* out, const evx_grid* g, const evx_bc* bc,
                          double eps, double D, const evx_epilogue* e, void* stream) {
  Dims d;
  Bc b;
  Epi ep;
  if (!bc) return fail(EVX_ERR_INVALID, "null bc");
  EVX_TRY(check_common(dtype, u, out, g, bc, &d, &b));
  EVX_TRY(check_epi(e, u, out, &ep));
  if (!(eps != 0.0)) return fail(EVX_ERR_INVALID, "eps must be nonzero");
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
