// Tile machinery of the 2.5-D streaming stencils and the Cahn-Hilliard
// right-hand-side tile body, shared by the one-grid kernels and the batched
// (many-grid) kernels.
#pragma once
#include "evx_ch_math.hpp"
#include "evx_common.hpp"
#include "evx_fft_engine.hpp"

#include <type_traits>

namespace evx {

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
#pragma clang fp contract(off)
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
#pragma clang fp contract(off)
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
// writes the element to LDS slot tid + k*NT of the plane.  With STAG the
// staggered_x x rules are available too: an x ghost plane is then the
// boundary node value or the zero-flux extrapolation of the first two
// interior planes (of the clipped field when CLIP), formed in fetch().
template <typename T, int H, int ROWS, int W, bool CLIP, bool STAG = false>
struct PlaneLoader {
  static constexpr int NPT = (ROWS * W + NT - 1) / NT;
  int off[NPT];      // in-plane offset y*nz + z of the source point, -1 if unused
  uint8_t yzf[NPT];  // y/z ghost flags: 4/8 = y low/high, 16/32 = z low/high
  T v[NPT];
  uint8_t xf;        // x ghost flag of the fetched plane: 1 = low, 2 = high, 64 = final staggered value

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

  // staggered_x ghost plane (boundary_conditions.py:96-121), in the order
  // of the reference's expression fac1*u1 - fac2*u2 with the factors rounded
  // to the working precision.
  __device__ __forceinline__ void fetch_node(const T* __restrict__ uc, bool low, const Dims& d, const Bc& bc) {
    xf = 64;
    if (bc.kind[0] == EVX_BC_NODE_VALUE) {
      const T v0 = static_cast<T>(low ? bc.node_lo : bc.node_hi);
#pragma unroll
      for (int k = 0; k < NPT; ++k) v[k] = v0;
      return;
    }
    const int64_t plane = d.ny * d.nz;
    const T* p1 = uc + (low ? 0 : d.nx - 1) * plane;
    const T* p2 = uc + (low ? 1 : d.nx - 2) * plane;
    const T f1 = static_cast<T>(4.0 / 3.0), f2 = static_cast<T>(1.0 / 3.0);
#pragma unroll
    for (int k = 0; k < NPT; ++k) {
      if (off[k] >= 0) {
        T a = p1[off[k]], b = p2[off[k]];
        if (CLIP) {
          a = clip01(a);
          b = clip01(b);
        }
        v[k] = f1 * a - f2 * b;
      }
    }
  }

  __device__ __forceinline__ void fetch(const T* __restrict__ uc, int64_t xg, const Dims& d, const Bc& bc) {
    xf = 0;
    if (xg < 0 || xg >= d.nx) {
      if (bc.kind[0] == EVX_BC_PERIODIC) xg = wrap_near(xg, d.nx);
      else if (STAG && staggered_x(bc)) { fetch_node(uc, xg < 0, d, bc); return; }
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
        if (STAG && (xf & 64)) {
          flat[tid + k * NT] = v[k];
          continue;
        }
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

// Loader of the fully periodic tiles (the semi-implicit path): every
// coordinate wraps, so there are no ghost rules.  Source points are kept as
// 32-bit byte offsets from the plane base, which lets the loads use a scalar
// base address plus a 32-bit vector offset (no 64-bit address arithmetic per
// element).  Requires ny*nz*sizeof(T) < 2^32.
template <typename T, int H, int ROWS, int W, bool CLIP>
struct PeriodicLoader {
  static constexpr int NPT = (ROWS * W + NT - 1) / NT;
  uint32_t boff[NPT];
  T v[NPT];
  bool last_ok;  // whether this thread's last element lies in the tile

  // Byte offset in its x-plane of the tile point this thread loads as
  // element k (clamped to the tile's last point past the end).
  __host__ __device__ static uint32_t src_byte(int tid, int k, int64_t y0, int64_t z0, const Dims& d) {
    const int i = tid + k * NT < ROWS * W - 1 ? tid + k * NT : ROWS * W - 1;
    const int64_t gy = wrap_near(y0 - H + i / W, d.ny), gz = wrap_near(z0 - H + i % W, d.nz);
    return static_cast<uint32_t>((gy * d.nz + gz) * (int64_t)sizeof(T));
  }

  __device__ __forceinline__ void init(int64_t y0, int64_t z0, const Dims& d, int tid) {
    last_ok = tid + (NPT - 1) * NT < ROWS * W;
#pragma unroll
    for (int k = 0; k < NPT; ++k) boff[k] = src_byte(tid, k, y0, z0, d);
  }

  __device__ __forceinline__ bool valid(int k) const { return k < NPT - 1 || last_ok; }

  __device__ __forceinline__ void fetch(const T* __restrict__ uc, int64_t xg, const Dims& d) {
    const char* pl = reinterpret_cast<const char*>(uc + wrap_near(xg, d.nx) * (d.ny * d.nz));
#pragma unroll
    for (int k = 0; k < NPT; ++k)
      if (valid(k)) v[k] = *reinterpret_cast<const T*>(pl + boff[k]);
  }

  __device__ __forceinline__ void commit(T* flat, int tid) const {
#pragma unroll
    for (int k = 0; k < NPT; ++k)
      if (valid(k)) flat[tid + k * NT] = CLIP ? clip01(v[k]) : v[k];
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
__host__ __device__ __forceinline__ TileOrigin tile_origin(const TileGrid& tg, const Dims& d, unsigned t) {
  TileOrigin o;
  o.z0 = (int64_t)(t % tg.ntz) * BZ;
  t /= tg.ntz;
  o.y0 = (int64_t)(t % tg.nty) * BY;
  t /= tg.nty;
  o.x0 = (int64_t)(t % tg.nxc) * tg.xc;
  o.ch = t / tg.nxc;
  o.x1 = o.x0 + (int64_t)tg.xc < d.nx ? o.x0 + (int64_t)tg.xc : d.nx;
  return o;
}

__host__ inline int pick_xchunk(const Dims& d, int64_t tiles_yz, int64_t target = 2048) {
  // About `target` blocks (a few per CU of the 256), but at least 4 planes
  // per block so the warm-up planes stay a bounded fraction of the work.
  // The minimum only binds for grids of a few tens of MB, which the
  // last-level cache holds; there a lower minimum means more blocks and
  // shorter per-block chains (100^3 float64 CH step: 0.0743 ms/step with at
  // least 8 planes, 0.0712 with 4, 0.0708 with 2; MI355X).
  int64_t chunks = (target + tiles_yz * d.c - 1) / (tiles_yz * d.c);
  if (chunks < 1) chunks = 1;
  int64_t xc = (d.nx + chunks - 1) / chunks;
  if (xc < 4) xc = 4;
  if (xc > d.nx) xc = d.nx;
  return (int)xc;
}

// Tile grid of a streaming stencil launch of about `target` blocks; returns
// the block count (tiles over z, y, x-chunks and channels).
template <typename T>
__host__ inline unsigned stencil_tiles(const Dims& d, TileGrid* tg, int64_t target) {
  constexpr int BY = NTY * TileCfg<T>::PY;
  const int64_t ty = (d.ny + BY - 1) / BY, tzn = (d.nz + BZ - 1) / BZ;
  tg->xc = pick_xchunk(d, ty * tzn, target);
  const int64_t nxc = (d.nx + tg->xc - 1) / tg->xc;
  tg->ntz = (int)tzn;
  tg->nty = (int)ty;
  tg->nxc = (int)nxc;
  return (unsigned)(tzn * ty * nxc * d.c);
}

// ---------------------------------------------------------------------------
// Cahn-Hilliard (problem_definition.py:328-371):
//   c = clip(u,0,1); mu = ((18/eps*c)*(1-c))*(1-2c) - (2eps)*lap(cB)
//   r = D * sum_axes grad_face( (cf*(1-cf)) * grad_face(muB) )
// mu lives in an LDS plane ring with a one-voxel (y,z) halo, c in one with a
// two-voxel halo.  mu ghosts (non-periodic faces) are the BC rule applied to
// mu at the adjacent interior voxel, as pad_bc(mu) does.
// ---------------------------------------------------------------------------
// Thread (ty, tz) owns PY consecutive rows r0 = ty*PY .. r0+PY-1 of column tz.
// Per output plane it evaluates mu for its own rows (sharing the column loads
// between rows) while threads 0..H-1 evaluate the one-voxel mu halo with the
// full ghost logic; the flux through the face between two own rows is
// computed once, and the flux through the x face x+1/2 is carried in a
// register to the next plane.  The c ring holds 4 planes (x-1..x+2, the
// slot of x-1 being refilled with x+3 during the output of plane x), the mu
// ring 2 (x, x+1); plane x+4 is fetched into registers during the output of
// plane x.
template <typename T>
struct ChTile {
  static constexpr int PY = TileCfg<T>::PY, BY = NTY * PY;
  static constexpr int CW = BZ + 4, CR = BY + 4, NC = 4;
  static constexpr int MW = BZ + 2, MR = BY + 2, NM = 2;
  static constexpr size_t LDS_BYTES = sizeof(T) * (NC * CR * CW + NM * MR * MW);
};

// One (y,z) tile x x-chunk of the CH right-hand side.  tid = flat thread index
// in a 256-thread block, tile = flat tile index (see tile_origin), smem =
// ChTile<T>::LDS_BYTES of LDS.  The contracted flavour (!STRICT) serves the
// fully periodic semi-implicit path only: it has no ghost logic, loads
// through PeriodicLoader and stores the plain right-hand side with 32-bit
// in-plane offsets whatever ep says (its callers pass EVX_OUT_RHS).
template <typename T, bool STRICT>
__device__ __forceinline__ void ch_rhs_tile(const T* __restrict__ u, T* __restrict__ out, const Dims& d,
                                            const Bc& bc, const Geo<T>& g, ChCoef<T> k, const Epi& ep,
                                            const TileGrid& tg, const double* __restrict__ chan, unsigned tile,
                                            int tid, unsigned char* smem) {
#pragma clang fp contract(off)
  using Op = Ops<STRICT>;
  using TC = ChTile<T>;
  constexpr int PY = TC::PY, BY = TC::BY;
  constexpr int CW = TC::CW, CR = TC::CR, NC = TC::NC;
  constexpr int MW = TC::MW, MR = TC::MR, NM = TC::NM;
  constexpr int HALO = 2 * MW + 2 * BY;  // mu halo points: rows -1, BY and columns -1, BZ
  T(*cs)[CR][CW] = reinterpret_cast<T(*)[CR][CW]>(smem);
  T(*ms)[MR][MW] = reinterpret_cast<T(*)[MR][MW]>(smem + sizeof(T) * NC * CR * CW);
  const int tz = tid % BZ, ty = tid / BZ;
  const TileOrigin o = tile_origin<BY>(tg, d, tile);
  const int64_t z0 = o.z0, y0 = o.y0, x0 = o.x0, x1 = o.x1;
  const T* uc = u + o.ch * d.vol();
  T* oc = out + o.ch * d.vol();
  if (chan) {  // per-channel (eps, D), rounded as the host rounds the scalars
    const double e = chan[2 * o.ch], dd = chan[2 * o.ch + 1];
    k = ChCoef<T>{static_cast<T>(18.0 / e), static_cast<T>(2.0 * e), static_cast<T>(dd)};
  }
  constexpr bool PER = !STRICT;
  const bool xper = PER || bc.kind[0] == EVX_BC_PERIODIC;
  const bool yper = PER || bc.kind[1] == EVX_BC_PERIODIC, zper = PER || bc.kind[2] == EVX_BC_PERIODIC;
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

  using Loader = typename std::conditional<PER, PeriodicLoader<T, 2, CR, CW, true>,
                                           PlaneLoader<T, 2, CR, CW, true>>::type;
  Loader pl;
  auto pl_init = [&]() {
    if constexpr (PER) pl.init(y0, z0, d, tid);
    else pl.init(y0, z0, d, bc, tid);
  };
  auto pl_fetch = [&](int64_t xg) {
    if constexpr (PER) pl.fetch(uc, xg, d);
    else pl.fetch(uc, xg, d, bc);
  };
  auto pl_commit = [&](int slot) {
    if constexpr (PER) pl.commit(&cs[slot][0][0], tid);
    else pl.commit(&cs[slot][0][0], bc, tid);
  };
  pl_init();
  const int b = tz + 2, bm = tz + 1;
  // Warm-up: planes x0-2 .. x0+1, mu(x0-1) and mu(x0), and the flux through
  // the x face x0-1/2 (afterwards carried in registers, so neither c(x-1)
  // nor mu(x-1) is read again and the rings hold 4 c planes and 2 mu planes).
  for (int64_t xg = x0 - 2; xg <= x0 + 1; ++xg) {
    pl_fetch(xg);
    pl_commit(ring<NC>(xg - x0 + 2));
  }
  if (x0 + 2 <= x1 + 1) pl_fetch(x0 + 2);
  __syncthreads();
  compute_mu(x0 - 1);
  compute_mu(x0);
  __syncthreads();
  T carry[PY];  // flux through the x face x-1/2 of each own row
  {
    const T(*Cm)[CW] = cs[ring<NC>(-1 + 2)];
    const T(*C0)[CW] = cs[ring<NC>(0 + 2)];
    const T(*Mm)[MW] = ms[ring<NM>(-1 + 1)];
    const T(*M0)[MW] = ms[ring<NM>(0 + 1)];
#pragma unroll
    for (int q = 0; q < PY; ++q)
      carry[q] = Op::flux(Cm[r0 + q + 2][b], C0[r0 + q + 2][b], Mm[r0 + q + 1][bm], M0[r0 + q + 1][bm], g.idx);
  }
  // plane x0+2 goes to the slot of x0-2 (last read by mu(x0-1))
  if (x0 + 2 <= x1 + 1) pl_commit(ring<NC>(4));
  if (x0 + 3 <= x1 + 1) pl_fetch(x0 + 3);
  __syncthreads();  // mu(x0+1) below reuses the slot of mu(x0-1) read above
  const int64_t plane = d.ny * d.nz;
  // plain right-hand-side stores of the periodic flavour: 32-bit byte offset
  // of this thread's first row in the plane, rows nz apart
  const bool full_tile = y0 + BY <= d.ny && z0 + BZ <= d.nz;
  const uint32_t obyte0 = static_cast<uint32_t>(((y0 + r0) * d.nz + gz) * (int64_t)sizeof(T));
  const uint32_t orow_bytes = static_cast<uint32_t>(d.nz * (int64_t)sizeof(T));
  for (int64_t x = x0; x < x1; ++x) {
    // mu(x+1) from c(x), c(x+1), c(x+2), into the slot of mu(x-1)
    compute_mu(x + 1);
    __syncthreads();
    const T(*C0)[CW] = cs[ring<NC>(x - x0 + 2)];
    const T(*Cp)[CW] = cs[ring<NC>(x + 1 - x0 + 2)];
    const T(*M0)[MW] = ms[ring<NM>(x - x0 + 1)];
    const T(*Mp)[MW] = ms[ring<NM>(x + 1 - x0 + 1)];
    T cc[PY + 2], mc[PY + 2], fy[PY + 1], qc[PY + 1];
#pragma unroll
    for (int j = 0; j < PY + 2; ++j) {
      cc[j] = C0[r0 + j + 1][b];
      mc[j] = M0[r0 + j][bm];
    }
    if constexpr (PER) {  // q = 1 - c of each own point, shared by its faces
#pragma unroll
      for (int j = 0; j <= PY; ++j) qc[j] = T(1) - cc[j];
    }
#pragma unroll
    for (int j = 0; j <= PY; ++j) {
      if constexpr (PER) fy[j] = Op::flux_q(qc[j], cc[j + 1], mc[j], mc[j + 1]);
      else fy[j] = Op::flux(cc[j], cc[j + 1], mc[j], mc[j + 1], g.idy);
    }
#pragma unroll
    for (int q = 0; q < PY; ++q) {
      const int a = r0 + q + 2, am = r0 + q + 1;
      const T c0 = cc[q + 1], m0 = mc[q + 1];
      T fxp, fzm, fzp;
      if constexpr (PER) {
        const T q0 = qc[q + 1];
        fxp = Op::flux_q(q0, Cp[a][b], m0, Mp[am][bm]);
        fzm = Op::flux_q(q0, C0[a][b - 1], M0[am][bm - 1], m0);
        fzp = Op::flux_q(q0, C0[a][b + 1], m0, M0[am][bm + 1]);
      } else {
        fxp = Op::flux(c0, Cp[a][b], m0, Mp[am][bm], g.idx);
        fzm = Op::flux(C0[a][b - 1], c0, M0[am][bm - 1], m0, g.idz);
        fzp = Op::flux(c0, C0[a][b + 1], m0, M0[am][bm + 1], g.idz);
      }
      const T r = Op::div(carry[q], fxp, fy[q], fy[q + 1], fzm, fzp, g, k.D);
      carry[q] = fxp;
      const int64_t gy = y0 + r0 + q;
      if constexpr (PER) {  // the contracted flavour's callers all take the plain right-hand side
        if (full_tile || (gy < d.ny && gz < d.nz))
          *reinterpret_cast<T*>(reinterpret_cast<char*>(oc + x * plane) + (obyte0 + q * orow_bytes)) = r;
      } else if (gy < d.ny && gz < d.nz) {
        epilogue(ep, uc, oc, x * plane + gy * d.nz + gz, r);
      }
    }
    // plane x+3 (fetched one iteration ago) goes to the slot of c(x-1), which
    // neither the output above nor mu(x+1) reads; it feeds mu(x+2)
    if (x + 3 <= x1 + 1) pl_commit(ring<NC>(x + 3 - x0 + 2));
    if (x + 4 <= x1 + 1) pl_fetch(x + 4);
    __syncthreads();
  }
}

}  // namespace evx
