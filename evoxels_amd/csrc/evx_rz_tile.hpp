// Fused Cahn-Hilliard right-hand side + z R2C sweep of the semi-implicit step
// (float32, fully periodic, nz = 512): the first sweep of the five-sweep
// pipeline.  It reads the state c and writes the z half spectrum of
// rhs(c) (problem_definition.py:328-371 followed by the first axis of
// rfftn, timesteppers.py:86); the right-hand side never goes to HBM.
//
// Layout.  A block of NW = 8 wavefronts owns BY = 16 y-rows of full z lines
// (the whole z extent) of one channel and walks an x-chunk.  Wavefront w
// owns the line pair (rows 2w, 2w+1) - the pair the z R2C packs into one
// complex line - and lane l owns z = 4l..4l+3 and 256+4l..256+4l+3 of both
// rows (two 16-byte chunks per row), so global loads and the LDS plane
// reads are 16-byte vectors.  Per output plane x each thread holds c at
// x, x+1, x+2 and mu at x, x+1 of its 16 points in registers; the LDS keeps
// three c planes (rows -1..BY: the y neighbours of a wavefront's rows) and
// the mu plane of x, so a point's x and z neighbours come from registers
// (z across chunk edges from the neighbouring lane) and its y neighbours
// from LDS.  mu at the two halo rows -1 and BY is evaluated by the block
// itself (two points per thread) from the outer rows -2 and BY+1.
//
// Per plane: mu(x+1), the face fluxes and the divergence at x, one
// workgroup barrier; then the new c and mu planes go to LDS and each
// wavefront moves its two rows of rhs(x) into its private exchange line in
// the free c slot and transforms them there (the 512-point complex FFT
// of the packed pair, exchanges synchronised within the wavefront), unpacks
// the two half spectra and stores them; a second barrier closes the plane.
//
// Every global and LDS index is computed by the RzMap helpers below, which
// also run on the host in the bounds emulation of the test suite.
#pragma once
#include "evx_ch_math.hpp"
#include "evx_fft_tiles.hpp"

namespace evx {

struct RzGeo {
  static constexpr int N = 512, NH = N / 2 + 1;
  static constexpr int BY = 16;             // y rows per tile
  static constexpr int NW = BY / 2;         // wavefronts = line pairs
  static constexpr int THREADS = 64 * NW;
  static constexpr int ROWS = BY + 2;       // LDS rows -1 .. BY
  static constexpr int SLOT = ROWS * N;     // floats per LDS plane
  static constexpr int LS = N + N / 16;     // padded complex exchange line (line_slot)
  static constexpr int NSLOT = 4;           // three c planes and the mu plane
  static constexpr size_t LDS_FLOATS = (size_t)NSLOT * SLOT + 2 * N;  // + twiddles
  static constexpr size_t LDS_BYTES = sizeof(float) * LDS_FLOATS;
};
static_assert(2 * RzGeo::NW * RzGeo::LS <= RzGeo::SLOT, "the exchange lines fit one c slot");

// Tile of a block: channel, first row y0, x-chunk [x0, x1).
struct RzTile {
  int64_t ch, y0, x0, x1;
};

struct RzMap {
  Dims d;
  int64_t pitch;  // complex values per spectrum row (>= N/2+1)
  int nty, nxc, xc;

  // Blocks take tiles y-fastest; with xcd_tile each XCD then owns whole
  // x-chunks, so the y halos of neighbouring tiles meet in its L2.
  __host__ __device__ RzTile tile(unsigned t) const {
    RzTile o;
    o.y0 = (int64_t)(t % (unsigned)nty) * RzGeo::BY;
    t /= (unsigned)nty;
    o.x0 = (int64_t)(t % (unsigned)nxc) * xc;
    o.ch = t / (unsigned)nxc;
    o.x1 = o.x0 + xc < d.nx ? o.x0 + xc : d.nx;
    return o;
  }
  // z of element e (0..7) of lane l: two 4-wide chunks, N/2 apart.
  __host__ __device__ static int z_of(int l, int e) { return (e < 4 ? 4 * l : RzGeo::N / 2 + 4 * l) + (e & 3); }
  // Element offset of plane x of channel ch and of row y inside a plane (x,
  // y taken modulo the grid: periodic wrap of halo and warm-up planes); a
  // line starts at plane + row.
  __host__ __device__ int64_t plane(int64_t ch, int64_t x) const {
    return (ch * d.nx + wrap_near(x, d.nx)) * d.ny * (int64_t)RzGeo::N;
  }
  __host__ __device__ int64_t row(int64_t y) const { return wrap_near(y, d.ny) * (int64_t)RzGeo::N; }
  __host__ __device__ int64_t line(int64_t ch, int64_t x, int64_t y) const { return plane(ch, x) + row(y); }
  // Spectrum element k of the line of row y (inside the grid) at plane x.
  __host__ __device__ int64_t spec(int64_t ch, int64_t x, int64_t y, int k) const {
    return ((ch * d.nx + x) * d.ny + y) * pitch + k;
  }
  // LDS float offset of row r (-1..BY) at z in c slot s (0..2) or the mu
  // slot (s = 3).
  __host__ __device__ static int lds(int s, int r, int z) { return s * RzGeo::SLOT + (r + 1) * RzGeo::N + z; }
  // Exchange line of wavefront w inside slot s: complex element e.
  __host__ __device__ static int xline(int s, int w, int e) {
    return s * RzGeo::SLOT + 2 * (w * RzGeo::LS + line_slot(e));
  }
  __host__ __device__ static int twiddle(int i) { return RzGeo::NSLOT * RzGeo::SLOT + 2 * i; }
  // Halo points of thread (w, l): z = 64w + l on rows -1 and BY.
  __host__ __device__ static int halo_z(int w, int l) { return 64 * w + l; }
};

// Row offsets of a wavefront's loads, fixed for the whole tile: its own rows
// 2w, 2w+1, the halo rows -1, BY and the outer rows -2, BY+1.
struct RzRows {
  int64_t a, b, h0, h1, o0, o1;
};
__host__ __device__ inline RzRows rz_rows(const RzMap& m, int64_t y0, int w) {
  return RzRows{m.row(y0 + 2 * w), m.row(y0 + 2 * w + 1), m.row(y0 - 1),
                m.row(y0 + RzGeo::BY),  m.row(y0 - 2),         m.row(y0 + RzGeo::BY + 1)};
}

// Planes the main loop loads at iteration x: `next` = plane x+3 (own and
// halo rows), `outer` = plane x+2 (outer rows).  Wrapped once at set-up,
// then advanced by one plane per iteration.
struct RzPlanes {
  int64_t chb, stride, nx, xn, next, outer;
  __host__ __device__ RzPlanes(const RzMap& m, int64_t ch, int64_t x)
      : chb(ch * m.d.nx * m.d.ny * (int64_t)RzGeo::N), stride(m.d.ny * (int64_t)RzGeo::N), nx(m.d.nx) {
    xn = wrap_near(x + 3, nx);
    next = chb + xn * stride;
    outer = m.plane(ch, x + 2);
  }
  __host__ __device__ void advance() {
    outer = next;
    xn = xn + 1 == nx ? 0 : xn + 1;
    next = chb + xn * stride;
  }
};

// Launch geometry: about `resident` blocks (one resident round), tiles of
// BY rows by x-chunks; returns the map and the block count.
inline RzMap rz_make_map(const Dims& d, int64_t pitch, int64_t resident, unsigned* blocks) {
  RzMap m;
  m.d = d;
  m.pitch = pitch;
  m.nty = (int)(d.ny / RzGeo::BY);
  const int64_t per_x = (int64_t)m.nty * d.c;
  int64_t nxc = (resident + per_x - 1) / per_x;
  if (nxc < 1) nxc = 1;
  const int64_t xc = (d.nx + nxc - 1) / nxc;
  m.xc = (int)xc;
  m.nxc = (int)((d.nx + xc - 1) / xc);
  *blocks = (unsigned)(per_x * m.nxc);
  return m;
}

// Value of v at lane (l + 1) mod 64 / (l - 1) mod 64 of the wavefront
// (ds_bpermute; measured faster here than the DPP wavefront rotates).
__device__ __forceinline__ float lane_next(float v) { return __shfl(v, (int)((threadIdx.x + 1) & 63), 64); }
__device__ __forceinline__ float lane_prev(float v) { return __shfl(v, (int)((threadIdx.x + 63) & 63), 64); }

// One row's 8 values: e = 0..3 at z = 4l.., e = 4..7 at z = 256 + 4l..
struct Row8 {
  float v[8];
};

// z+1 and z-1 neighbours of a row's elements: inside a chunk from the
// register to the right/left, across a chunk edge from the neighbouring
// lane; lane 63's z = 4*63+3 = 255 continues at z = 256 (lane 0, chunk 2),
// and z = 511 wraps to z = 0 (periodic).
__device__ __forceinline__ void z_neighbours(const Row8& c, float (&zp)[8], float (&zm)[8], int l) {
  const float n0 = lane_next(c.v[0]), n4 = lane_next(c.v[4]);
  const float p3 = lane_prev(c.v[3]), p7 = lane_prev(c.v[7]);
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    if ((e & 3) != 3) zp[e] = c.v[e + 1];
    if ((e & 3) != 0) zm[e] = c.v[e - 1];
  }
  zp[3] = l == 63 ? n4 : n0;
  zp[7] = l == 63 ? n0 : n4;
  zm[0] = l == 0 ? p7 : p3;
  zm[4] = l == 0 ? p3 : p7;
}

// z+1 neighbours only.
__device__ __forceinline__ void z_next(const Row8& c, float (&zp)[8], int l) {
  const float n0 = lane_next(c.v[0]), n4 = lane_next(c.v[4]);
#pragma unroll
  for (int e = 0; e < 8; ++e)
    if ((e & 3) != 3) zp[e] = c.v[e + 1];
  zp[3] = l == 63 ? n4 : n0;
  zp[7] = l == 63 ? n0 : n4;
}

__device__ __forceinline__ void load_row(const float* __restrict__ p, Row8& r, int l) {
  const float4 a = *reinterpret_cast<const float4*>(p + 4 * l);
  const float4 b = *reinterpret_cast<const float4*>(p + RzGeo::N / 2 + 4 * l);
  r.v[0] = a.x; r.v[1] = a.y; r.v[2] = a.z; r.v[3] = a.w;
  r.v[4] = b.x; r.v[5] = b.y; r.v[6] = b.z; r.v[7] = b.w;
}

__device__ __forceinline__ void clip_row(Row8& r) {
#pragma unroll
  for (int e = 0; e < 8; ++e) r.v[e] = clip01(r.v[e]);
}

__device__ __forceinline__ void store_row_lds(float* s, const Row8& r, int l) {
  *reinterpret_cast<float4*>(s + 4 * l) = make_float4(r.v[0], r.v[1], r.v[2], r.v[3]);
  *reinterpret_cast<float4*>(s + RzGeo::N / 2 + 4 * l) = make_float4(r.v[4], r.v[5], r.v[6], r.v[7]);
}

__device__ __forceinline__ void load_row_lds(const float* s, Row8& r, int l) {
  const float4 a = *reinterpret_cast<const float4*>(s + 4 * l);
  const float4 b = *reinterpret_cast<const float4*>(s + RzGeo::N / 2 + 4 * l);
  r.v[0] = a.x; r.v[1] = a.y; r.v[2] = a.z; r.v[3] = a.w;
  r.v[4] = b.x; r.v[5] = b.y; r.v[6] = b.z; r.v[7] = b.w;
}

// The two rows (a = 2w, b = 2w+1) of a thread at one x-plane.
struct Pair8 {
  Row8 a, b;
};

// mu(x+1) of the thread's 16 points (Ops<false>::mu): centre cB, x
// neighbours cA (x) and cC (x+2), y neighbours from the own other row and
// the c(x+1) plane in LDS (rows a-1 and b+1).
__device__ __forceinline__ void mu_pair(const Pair8& cA, const Pair8& cB, const Pair8& cC, const float* lds, int s,
                                        int w, int l, const ChW<float>& W, Pair8& m) {
  float zp[8], zm[8];
  {  // row a: y-1 is row a-1 of the plane, y+1 the own row b
    Row8 ym;
    load_row_lds(lds + RzMap::lds(s, 2 * w - 1, 0), ym, l);
    z_neighbours(cB.a, zp, zm, l);
#pragma unroll
    for (int e = 0; e < 8; ++e)
      m.a.v[e] = mu_w(cB.a.v[e], cC.a.v[e], cA.a.v[e], cB.b.v[e], ym.v[e], zp[e], zm[e], W);
  }
  {  // row b: y-1 is the own row a, y+1 row b+1 of the plane
    Row8 yp;
    load_row_lds(lds + RzMap::lds(s, 2 * w + 2, 0), yp, l);
    z_neighbours(cB.b, zp, zm, l);
#pragma unroll
    for (int e = 0; e < 8; ++e)
      m.b.v[e] = mu_w(cB.b.v[e], cC.b.v[e], cA.b.v[e], yp.v[e], cB.a.v[e], zp[e], zm[e], W);
  }
}

// rhs at plane x of the thread's 16 points from c(x), c(x+1), mu(x),
// mu(x+1) (registers), the c(x) and mu(x) planes in LDS (rows a-1, b+1) and
// the carried x-face fluxes x-1/2; updates the carry to x+1/2.
__device__ __forceinline__ void div_pair(const Pair8& cA, const Pair8& cB, const Pair8& mA, const Pair8& mB,
                                         const float* lds, int s, int w, int l, const ChW<float>& W,
                                         Pair8& carry, Pair8& r) {
  using Op = Ops<false>;
  float fab[8];  // face a|b: row a's y+1/2 face is row b's y-1/2 face
#pragma unroll
  for (int row = 0; row < 2; ++row) {
    const Row8& c = row ? cA.b : cA.a;
    const Row8& m = row ? mA.b : mA.a;
    // the row's outer y neighbour (row a-1 for a, b+1 for b) from the c(x)
    // and mu(x) planes
    Row8 cy, my;
    load_row_lds(lds + RzMap::lds(s, row ? 2 * w + 2 : 2 * w - 1, 0), cy, l);
    load_row_lds(lds + RzMap::lds(3, row ? 2 * w + 2 : 2 * w - 1, 0), my, l);
    float cz[8], mz[8];
    z_next(c, cz, l);
    z_next(m, mz, l);
    float fz[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) fz[e] = Op::flux_q(1.0f - c.v[e], cz[e], m.v[e], mz[e]);  // face z+1/2
    // face z-1/2 of each point is face z+1/2 of the point below it in z
    float fzm[8];
    const float f3 = lane_prev(fz[3]), f7 = lane_prev(fz[7]);
#pragma unroll
    for (int e = 0; e < 8; ++e)
      if ((e & 3) != 0) fzm[e] = fz[e - 1];
    fzm[0] = l == 0 ? f7 : f3;
    fzm[4] = l == 0 ? f3 : f7;
    Row8& out = row ? r.b : r.a;
    Row8& cr = row ? carry.b : carry.a;
    const Row8& cp = row ? cB.b : cB.a;
    const Row8& mp = row ? mB.b : mB.a;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const float q = 1.0f - c.v[e];
      const float fxp = Op::flux_q(q, cp.v[e], m.v[e], mp.v[e]);
      float fym, fyp;
      if (row == 0) {  // faces a-1|a and a|b
        fym = Op::flux_q(1.0f - cy.v[e], c.v[e], my.v[e], m.v[e]);
        fyp = Op::flux_q(q, cA.b.v[e], m.v[e], mA.b.v[e]);
        fab[e] = fyp;
      } else {  // faces a|b and b|b+1
        fym = fab[e];
        fyp = Op::flux_q(q, cy.v[e], m.v[e], my.v[e]);
      }
      out.v[e] = div_w(cr.v[e], fxp, fym, fyp, fzm[e], fz[e], W);
      cr.v[e] = fxp;
    }
  }
}

// mu(x+1) at the thread's two halo points (rows -1 and BY at z = 64w + l):
// centre hB, x neighbours hA, hC, outer-row y neighbour hO, the inner y
// neighbour and the z neighbours from the c(x+1) plane in LDS.
__device__ __forceinline__ void mu_halo(const float (&hA)[2], const float (&hB)[2], const float (&hC)[2],
                                        const float (&hO)[2], const float* lds, int s, int z, const ChW<float>& W,
                                        float (&mh)[2]) {
  constexpr int N = RzGeo::N, BY = RzGeo::BY;
  const int zp = (z + 1) & (N - 1), zm = (z + N - 1) & (N - 1);
  // row -1: y-1 is the outer row -2, y+1 is row 0
  mh[0] = mu_w(hB[0], hC[0], hA[0], lds[RzMap::lds(s, 0, z)], hO[0], lds[RzMap::lds(s, -1, zp)],
               lds[RzMap::lds(s, -1, zm)], W);
  // row BY: y-1 is row BY-1, y+1 the outer row BY+1
  mh[1] = mu_w(hB[1], hC[1], hA[1], hO[1], lds[RzMap::lds(s, BY - 1, z)], lds[RzMap::lds(s, BY, zp)],
               lds[RzMap::lds(s, BY, zm)], W);
}

// One tile (block) of the fused sweep.  u: state (C, nx, ny, 512), spec:
// half spectrum with row pitch m.pitch, twg: the 512 twiddles.
__device__ __forceinline__ void rz_tile_body(const float* __restrict__ u, cpx<float>* __restrict__ spec,
                                             const RzMap& m, const Geo<float>& g, ChW<float> W,
                                             const double* __restrict__ chan, const cpx<float>* __restrict__ twg,
                                             unsigned tile, int tid, float* lds) {
  constexpr int N = RzGeo::N, NH = RzGeo::NH, BY = RzGeo::BY;
  // the wavefront index is uniform in a wavefront: keep it (and the row
  // addresses derived from it) in scalar registers
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6), l = tid & 63;
  const RzTile t = m.tile(tile);
  if (chan) {  // per-channel (eps, D), rounded as the host rounds the scalars
    const double e = chan[2 * t.ch], dd = chan[2 * t.ch + 1];
    W = ch_weights(g, ChCoef<float>{static_cast<float>(18.0 / e), static_cast<float>(2.0 * e),
                                    static_cast<float>(dd)});
  }
  // per-pass twiddle tables of the 512-point transform (conflict-free reads)
  static_assert(pass_tw_size<N>(FftSpec<N>::R{}) <= N, "pass tables fit the twiddle region");
  cpx<float>* tw = reinterpret_cast<cpx<float>*>(lds + RzMap::twiddle(0));
  pass_tw_fill<float, N>(tw, twg, tid, RzGeo::THREADS, FftSpec<N>::R{});
  const int ra = 2 * w, rb = 2 * w + 1;
  const int hz = RzMap::halo_z(w, l);
  const int64_t y0 = t.y0;

  // loads of plane offset p (RzMap::plane)
  const RzRows rw = rz_rows(m, y0, w);
  auto load_pair = [&](int64_t p, Pair8& q) {
    load_row(u + p + rw.a, q.a, l);
    load_row(u + p + rw.b, q.b, l);
  };
  auto load_halo = [&](int64_t p, float (&h)[2]) {  // rows -1, BY
    h[0] = u[p + rw.h0 + hz];
    h[1] = u[p + rw.h1 + hz];
  };
  auto load_outer = [&](int64_t p, float (&h)[2]) {  // rows -2, BY+1
    h[0] = u[p + rw.o0 + hz];
    h[1] = u[p + rw.o1 + hz];
  };
  auto plane = [&](int64_t x) { return m.plane(t.ch, x); };
  auto clip_pair = [&](Pair8& p) {
    clip_row(p.a);
    clip_row(p.b);
  };
  auto clip2 = [&](float (&h)[2]) {
    h[0] = clip01(h[0]);
    h[1] = clip01(h[1]);
  };
  // c plane into slot s: own rows and the halo rows
  auto put_c = [&](int s, const Pair8& p, const float (&h)[2]) {
    store_row_lds(lds + RzMap::lds(s, ra, 0), p.a, l);
    store_row_lds(lds + RzMap::lds(s, rb, 0), p.b, l);
    lds[RzMap::lds(s, -1, hz)] = h[0];
    lds[RzMap::lds(s, BY, hz)] = h[1];
  };
  auto put_mu = [&](const Pair8& p, const float (&h)[2]) {
    store_row_lds(lds + RzMap::lds(3, ra, 0), p.a, l);
    store_row_lds(lds + RzMap::lds(3, rb, 0), p.b, l);
    lds[RzMap::lds(3, -1, hz)] = h[0];
    lds[RzMap::lds(3, BY, hz)] = h[1];
  };

  // ---- warm-up: planes x0-2 .. x0+2, mu(x0-1) (own), mu(x0) (own + halo),
  // the x-face flux x0-1/2, and the LDS state of the loop's first plane.
  const int64_t x0 = t.x0;
  Pair8 cA, cB, cC, mA, mB, carry, r;
  float hA[2], hB[2], hC[2], hO[2];
  {
    Pair8 cP;
    float hP[2], hOA[2];
    load_pair(plane(x0 - 1), cP);
    load_pair(plane(x0), cA);
    load_pair(plane(x0 + 1), cB);
    load_halo(plane(x0 - 1), hP);
    load_halo(plane(x0), hA);
    load_halo(plane(x0 + 1), hB);
    load_outer(plane(x0), hOA);
    clip_pair(cP); clip_pair(cA); clip_pair(cB);
    clip2(hP); clip2(hA); clip2(hB); clip2(hOA);
    // slot 2 <- c(x0-1), slot 0 <- c(x0), slot 1 <- c(x0+1)
    put_c(2, cP, hP);
    put_c(0, cA, hA);
    put_c(1, cB, hB);
    __syncthreads();
    Pair8 mP;
    {
      Pair8 c2;
      load_pair(plane(x0 - 2), c2);
      clip_pair(c2);
      mu_pair(c2, cP, cA, lds, 2, w, l, W, mP);  // mu(x0-1), own points
    }
    mu_pair(cP, cA, cB, lds, 0, w, l, W, mA);    // mu(x0)
    float mh[2];
    mu_halo(hP, hA, hB, hOA, lds, 0, hz, W, mh);
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      carry.a.v[e] = Ops<false>::flux_q(1.0f - cP.a.v[e], cA.a.v[e], mP.a.v[e], mA.a.v[e]);
      carry.b.v[e] = Ops<false>::flux_q(1.0f - cP.b.v[e], cA.b.v[e], mP.b.v[e], mA.b.v[e]);
    }
    load_pair(plane(x0 + 2), cC);
    load_halo(plane(x0 + 2), hC);
    load_outer(plane(x0 + 1), hO);
    clip_pair(cC);
    clip2(hC);
    clip2(hO);
    __syncthreads();  // slot 2 is free again
    put_mu(mA, mh);
  }
  __syncthreads();

  // ---- main loop: slot sA holds c(x), sB c(x+1), sF is free; slot 3 mu(x).
  int sA = 0, sB = 1, sF = 2;
  Pair8 pre;
  float preH[2], preO[2];
  RzPlanes pl(m, t.ch, x0);
  cpx<float>* sa = spec + m.spec(t.ch, x0, y0 + ra, 0);
  const int64_t sstep = m.d.ny * m.pitch;
  for (int64_t x = x0; x < t.x1; ++x) {
    // next plane's loads: c(x+3) (own and halo rows), outer rows of c(x+2)
    load_pair(pl.next, pre);
    load_halo(pl.next, preH);
    load_outer(pl.outer, preO);
    // mu(x+1), own points and halo points
    mu_pair(cA, cB, cC, lds, sB, w, l, W, mB);
    float mh[2];
    mu_halo(hA, hB, hC, hO, lds, sB, hz, W, mh);
    // rhs at x
    div_pair(cA, cB, mA, mB, lds, sA, w, l, W, carry, r);
    __syncthreads();  // every read of c(x) and mu(x) is done
    put_mu(mB, mh);
    put_c(sF, cC, hC);
    // the wavefront's two rows of rhs(x) into its exchange line (slot sA)
    cpx<float>* xl = reinterpret_cast<cpx<float>*>(lds + RzMap::xline(sA, w, 0));
    const LineLay<float, RzGeo::LS, true> xlay{xl};
#pragma unroll
    for (int e = 0; e < 8; ++e) *xlay.at(RzMap::z_of(l, e)) = cpx<float>{r.a.v[e], r.b.v[e]};
    xsync<true>();
    cpx<float> v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = *xlay.at(l + 64 * q);
    fft_regs_pt<float, N, 64, false>(v, xlay, l, tw, typename FftSpec<N>::R{});
    xsync<true>();
#pragma unroll
    for (int q = 0; q < 8; ++q) *xlay.at(l + 64 * q) = v[q];
    xsync<true>();
#pragma unroll
    for (int q = 0; q < (NH + 63) / 64; ++q) {
      const int kk = l + 64 * q;
      if (kk < NH) {
        const cpx<float> Z = v[q];
        const cpx<float> Zc = *xlay.at(kk == 0 ? 0 : N - kk);
        sa[kk] = cpx<float>{(Z.x + Zc.x) * 0.5f, (Z.y - Zc.y) * 0.5f};
        sa[m.pitch + kk] = cpx<float>{(Z.y + Zc.y) * 0.5f, (Zc.x - Z.x) * 0.5f};
      }
    }
    // rotate registers and slots
    cA = cB;
    cB = cC;
    cC = pre;
    clip_pair(cC);
    hA[0] = hB[0]; hA[1] = hB[1];
    hB[0] = hC[0]; hB[1] = hC[1];
    hC[0] = clip01(preH[0]); hC[1] = clip01(preH[1]);
    hO[0] = clip01(preO[0]); hO[1] = clip01(preO[1]);
    mA = mB;
    const int s0 = sA;
    sA = sB;
    sB = sF;
    sF = s0;
    pl.advance();
    sa += sstep;
    __syncthreads();  // the new c and mu planes are visible; the exchange is done
  }
}

}  // namespace evx
