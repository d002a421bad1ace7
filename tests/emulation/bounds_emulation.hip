// Host emulation of the index and address arithmetic of the semi-implicit
// Cahn-Hilliard step's first sweep, for every block, thread and x-plane:
//   * the fused right-hand side + z R2C kernel (evx_rz_tile.hpp), through
//     the same RzMap / rz_make_map / xcd_tile helpers the kernel calls;
//   * the two-kernel path it replaces: the periodic CH tile
//     (PeriodicLoader::src_byte, tile_origin, stencil_tiles) and the z R2C
//     and C2R line-pair tiles (ZGeo::first_line, ZGeo::blocks).
// Every global element and LDS word an access touches must lie inside its
// allocation, and every spectrum (or right-hand-side) element must be
// written exactly once.  Prints "OK <case>" per case; exits non-zero with
// the first violation.
//
// Usage: bounds_emulation  (runs the built-in cases)
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "evx_rz_tile.hpp"
#include "evx_stencil_tiles.hpp"

using namespace evx;

namespace {

int failures = 0;

#define CHECK(cond, ...)                         \
  do {                                           \
    if (!(cond)) {                               \
      std::fprintf(stderr, "FAIL: " __VA_ARGS__); \
      std::fprintf(stderr, "\n");                \
      ++failures;                                \
      return false;                              \
    }                                            \
  } while (0)

// ---------------------------------------------------------------- fused ----
bool emulate_rz(int64_t C, int64_t nx, int64_t ny, int64_t pitch, int64_t resident) {
  constexpr int N = RzGeo::N, NH = RzGeo::NH, BY = RzGeo::BY, NW = RzGeo::NW;
  Dims d{C, nx, ny, N};
  unsigned blocks = 0;
  const RzMap m = rz_make_map(d, pitch, resident, &blocks);
  const int64_t field = C * nx * ny * N;
  const int64_t spec_elems = C * nx * ny * pitch;
  const int64_t lds_floats = (int64_t)RzGeo::LDS_FLOATS;
  std::vector<uint8_t> written((size_t)(C * nx * ny * NH), 0);
  std::vector<uint8_t> tile_seen(blocks, 0);
  auto gload = [&](int64_t off, int width) -> bool {
    CHECK(off >= 0 && off + width <= field, "rz load %lld+%d outside field %lld", (long long)off, width,
          (long long)field);
    CHECK(width != 4 || off % 4 == 0, "rz 16-byte load at element %lld", (long long)off);
    return true;
  };
  auto lds = [&](int64_t off, int width) -> bool {
    CHECK(off >= 0 && off + width <= lds_floats, "rz LDS %lld+%d outside %lld", (long long)off, width,
          (long long)lds_floats);
    return true;
  };
  for (unsigned b = 0; b < blocks; ++b) {
    const unsigned tile = xcd_tile(b, blocks);
    CHECK(tile < blocks && !tile_seen[tile], "block %u maps to tile %u twice or out of range", b, tile);
    tile_seen[tile] = 1;
    const RzTile t = m.tile(tile);
    CHECK(t.ch >= 0 && t.ch < C && t.y0 >= 0 && t.y0 + BY <= ny && t.x0 >= 0 && t.x0 < t.x1 && t.x1 <= nx,
          "tile %u: ch %lld y0 %lld x [%lld, %lld)", tile, (long long)t.ch, (long long)t.y0, (long long)t.x0,
          (long long)t.x1);
    for (int tid = 0; tid < RzGeo::THREADS; ++tid) {
      const int w = tid >> 6, l = tid & 63;
      const int ra = 2 * w, rb = 2 * w + 1, hz = RzMap::halo_z(w, l);
      CHECK(hz >= 0 && hz < N, "halo z %d", hz);
      // the kernel's loads: plane offset p plus the tile's fixed row offsets
      const RzRows rw = rz_rows(m, t.y0, w);
      CHECK(rw.a == m.row(t.y0 + ra) && rw.b == m.row(t.y0 + rb) && rw.h0 == m.row(t.y0 - 1) &&
                rw.h1 == m.row(t.y0 + BY) && rw.o0 == m.row(t.y0 - 2) && rw.o1 == m.row(t.y0 + BY + 1),
            "row offsets of tile %u wavefront %d", tile, w);
      auto pair_at = [&](int64_t p) -> bool {
        for (int64_t r : {rw.a, rw.b})
          if (!gload(p + r + 4 * l, 4) || !gload(p + r + N / 2 + 4 * l, 4)) return false;
        return true;
      };
      auto pair = [&](int64_t x) -> bool { return pair_at(m.plane(t.ch, x)); };
      auto halo_at = [&](int64_t p, int64_t r0, int64_t r1) -> bool {
        return gload(p + r0 + hz, 1) && gload(p + r1 + hz, 1);
      };
      auto halo = [&](int64_t x, int r0, int r1) -> bool {
        return gload(m.line(t.ch, x, t.y0 + r0) + hz, 1) && gload(m.line(t.ch, x, t.y0 + r1) + hz, 1);
      };
      auto lds_row = [&](int s, int r) -> bool {
        return lds(RzMap::lds(s, r, 0) + 4 * l, 4) && lds(RzMap::lds(s, r, 0) + N / 2 + 4 * l, 4);
      };
      auto put_c = [&](int s) -> bool {
        return lds_row(s, ra) && lds_row(s, rb) && lds(RzMap::lds(s, -1, hz), 1) && lds(RzMap::lds(s, BY, hz), 1);
      };
      auto mu_pair = [&](int s) -> bool { return lds_row(s, 2 * w - 1) && lds_row(s, 2 * w + 2); };
      auto mu_halo = [&](int s) -> bool {
        const int zp = (hz + 1) & (N - 1), zm = (hz + N - 1) & (N - 1);
        return lds(RzMap::lds(s, 0, hz), 1) && lds(RzMap::lds(s, -1, zp), 1) && lds(RzMap::lds(s, -1, zm), 1) &&
               lds(RzMap::lds(s, BY - 1, hz), 1) && lds(RzMap::lds(s, BY, zp), 1) &&
               lds(RzMap::lds(s, BY, zm), 1);
      };
      for (int i = tid; i < N; i += RzGeo::THREADS)
        if (!lds(RzMap::twiddle(i), 2)) return false;
      // warm-up
      const int64_t x0 = t.x0;
      for (int64_t x : {x0 - 1, x0, x0 + 1, x0 - 2, x0 + 2})
        if (!pair(x)) return false;
      for (int64_t x : {x0 - 1, x0, x0 + 1, x0 + 2})
        if (!halo(x, -1, BY)) return false;
      for (int64_t x : {x0, x0 + 1})
        if (!halo(x, -2, BY + 1)) return false;
      if (!put_c(2) || !put_c(0) || !put_c(1) || !mu_pair(2) || !mu_pair(0) || !mu_halo(0) || !put_c(3))
        return false;
      int sA = 0, sB = 1, sF = 2;
      RzPlanes pl(m, t.ch, t.x0);
      int64_t sa = m.spec(t.ch, t.x0, t.y0 + ra, 0);
      const int64_t sstep = m.d.ny * m.pitch;
      for (int64_t x = t.x0; x < t.x1; ++x) {
        CHECK(pl.next == m.plane(t.ch, x + 3) && pl.outer == m.plane(t.ch, x + 2), "plane walk at x %lld",
              (long long)x);
        CHECK(sa == m.spec(t.ch, x, t.y0 + ra, 0), "spectrum pointer at x %lld", (long long)x);
        if (!pair_at(pl.next) || !halo_at(pl.next, rw.h0, rw.h1) || !halo_at(pl.outer, rw.o0, rw.o1)) return false;
        if (!mu_pair(sB) || !mu_halo(sB)) return false;
        if (!mu_pair(sA) || !mu_pair(3)) return false;  // div_pair: c(x) and mu(x) planes
        if (!put_c(3) || !put_c(sF)) return false;
        // exchange line of wavefront w in slot sA: every element index the
        // writes, the Stockham exchanges, the reads and the unpacking use
        for (int e = 0; e < N; ++e)
          if (!lds(RzMap::xline(sA, w, e), 2)) return false;
        for (int e8 = 0; e8 < 8; ++e8) {
          const int z = RzMap::z_of(l, e8);
          CHECK(z >= 0 && z < N, "z_of(%d,%d) = %d", l, e8, z);
        }
        CHECK(RzMap::xline(sA, w, N - 1) + 2 <= RzMap::lds(sA + 1, -1, 0), "exchange line leaves slot %d", sA);
        for (int q = 0; q < (NH + 63) / 64; ++q) {
          const int kk = l + 64 * q;
          if (kk >= NH) continue;
          for (int r : {ra, rb}) {
            const int64_t off = sa + (r - ra) * m.pitch + kk;
            CHECK(off >= 0 && off < spec_elems, "spectrum store %lld outside %lld", (long long)off,
                  (long long)spec_elems);
            const int64_t o = ((t.ch * nx + x) * ny + t.y0 + r) * NH + kk;
            CHECK(!written[o], "spectrum element written twice (ch %lld x %lld y %lld k %d)", (long long)t.ch,
                  (long long)x, (long long)(t.y0 + r), kk);
            written[o] = 1;
          }
          // the unpack of this element reads element N - kk of the line
          if (!lds(RzMap::xline(sA, w, kk == 0 ? 0 : N - kk), 2)) return false;
        }
        const int s0 = sA;
        sA = sB;
        sB = sF;
        sF = s0;
        pl.advance();
        sa += sstep;
      }
    }
  }
  for (size_t i = 0; i < written.size(); ++i) CHECK(written[i], "spectrum element %zu never written", i);
  return true;
}

// ------------------------------------------------- two-kernel reference ----
// Periodic CH tile (float32): loads of PeriodicLoader and plain stores.
bool emulate_ch_tile(int64_t C, int64_t nx, int64_t ny, int64_t nz, int64_t target) {
  using T = float;
  using TC = ChTile<T>;
  constexpr int BY = TC::BY, CW = TC::CW, CR = TC::CR;
  using Loader = PeriodicLoader<T, 2, CR, CW, true>;
  Dims d{C, nx, ny, nz};
  TileGrid tg;
  const unsigned blocks = stencil_tiles<T>(d, &tg, target);
  const int64_t plane_bytes = ny * nz * (int64_t)sizeof(T);
  std::vector<uint8_t> written((size_t)(C * nx * ny * nz), 0);
  for (unsigned b = 0; b < blocks; ++b) {
    const TileOrigin o = tile_origin<BY>(tg, d, xcd_tile(b, blocks));
    CHECK(o.ch < C && o.x0 < o.x1 && o.x1 <= nx, "ch tile %u out of range", b);
    for (int tid = 0; tid < NT; ++tid) {
      const bool last_ok = tid + (Loader::NPT - 1) * NT < CR * CW;
      for (int k = 0; k < Loader::NPT; ++k) {
        if (!(k < Loader::NPT - 1 || last_ok)) continue;
        const uint32_t off = Loader::src_byte(tid, k, o.y0, o.z0, d);
        CHECK(off + sizeof(T) <= (uint64_t)plane_bytes, "ch load byte %u outside the plane", off);
        CHECK(tid + k * NT < CR * CW, "ch LDS slot %d outside the plane ring", tid + k * NT);
      }
      // planes fetched: x0-2 .. x1+1, each wrapped into [0, nx)
      for (int64_t xg = o.x0 - 2; xg <= o.x1 + 1; ++xg) {
        const int64_t xw = wrap_near(xg, nx);
        CHECK(xw >= 0 && xw < nx, "ch plane %lld wraps to %lld", (long long)xg, (long long)xw);
      }
      const int tz = tid % BZ, ty = tid / BZ, r0 = ty * TileCfg<T>::PY;
      const int64_t gz = o.z0 + tz;
      const bool full = o.y0 + BY <= ny && o.z0 + BZ <= nz;
      for (int64_t x = o.x0; x < o.x1; ++x)
        for (int q = 0; q < TileCfg<T>::PY; ++q) {
          const int64_t gy = o.y0 + r0 + q;
          if (!(full || (gy < ny && gz < nz))) continue;
          const int64_t i = ((o.ch * nx + x) * ny + gy) * nz + gz;
          CHECK(gy < ny && gz < nz && i < C * nx * ny * nz, "ch store (%lld, %lld, %lld) outside", (long long)x,
                (long long)gy, (long long)gz);
          CHECK(!written[i], "ch store twice at %lld", (long long)i);
          written[i] = 1;
        }
    }
  }
  for (size_t i = 0; i < written.size(); ++i) CHECK(written[i], "ch element %zu never written", i);
  return true;
}

// z R2C / C2R line-pair tiles of size N over nlines real lines.
template <int N>
bool emulate_z_pairs(int64_t nlines, int64_t pitch) {
  using G = ZGeo<float, N>;
  constexpr int TPL = G::TPL, E = N / TPL, NH = N / 2 + 1;
  const int64_t blocks = G::blocks(nlines);
  const int64_t lds_elems = (int64_t)N + G::LP * G::LS;
  std::vector<uint8_t> seen((size_t)nlines, 0);
  for (int64_t b = 0; b < blocks; ++b)
    for (int tid = 0; tid < G::THREADS; ++tid) {
      const int pr = tid / TPL, tl = tid % TPL;
      const int64_t la = G::first_line(b, pr);
      const bool has_a = la < nlines, has_b = la + 1 < nlines;
      for (int q = 0; q < E; ++q) {
        const int e = tl + TPL * q;
        if (has_a) CHECK(la * N + e < nlines * N, "z load a");
        if (has_b) CHECK((la + 1) * N + e < nlines * N, "z load b");
        const int64_t slot = N + pr * G::LS + line_slot(e);
        CHECK(slot < lds_elems, "z LDS %lld outside %lld", (long long)slot, (long long)lds_elems);
      }
      for (int q = 0; q < (NH + TPL - 1) / TPL; ++q) {
        const int k = tl + TPL * q;
        if (k >= NH) continue;
        if (has_a) CHECK(la * pitch + k < nlines * pitch, "z spectrum a");
        if (has_b) CHECK((la + 1) * pitch + k < nlines * pitch, "z spectrum b");
      }
      if (tl == 0) {
        if (has_a) seen[la] += 1;
        if (has_b) seen[la + 1] += 1;
      }
    }
  for (int64_t i = 0; i < nlines; ++i) CHECK(seen[i] == 1, "z line %lld covered %d times", (long long)i, seen[i]);
  return true;
}

}  // namespace

int main() {
  struct RzCase {
    int64_t C, nx, ny, resident;
  };
  const RzCase rz[] = {{1, 512, 512, 256}, {1, 12, 16, 256}, {3, 5, 32, 256},  {1, 4, 16, 256},
                       {2, 1, 48, 7},      {1, 9, 16, 1024}, {1, 512, 512, 8}, {1, 2, 16, 256},
                       {2, 3, 32, 4}};
  for (const RzCase& c : rz) {
    const bool ok = emulate_rz(c.C, c.nx, c.ny, RzGeo::NH, c.resident);
    std::printf("%s fused rhs+z-r2c (%lld,%lld,%lld,512) resident %lld\n", ok ? "OK" : "FAILED", (long long)c.C,
                (long long)c.nx, (long long)c.ny, (long long)c.resident);
  }
  struct ChCase {
    int64_t C, nx, ny, nz, target;
  };
  const ChCase ch[] = {{1, 4, 4, 1024, 1280}, {1, 12, 8, 512, 1280}, {1, 512, 512, 512, 1280},
                       {1, 4, 4, 1024, 7}};
  for (const ChCase& c : ch) {
    const bool ok = emulate_ch_tile(c.C, c.nx, c.ny, c.nz, c.target);
    std::printf("%s periodic CH tile (%lld,%lld,%lld,%lld) target %lld\n", ok ? "OK" : "FAILED", (long long)c.C,
                (long long)c.nx, (long long)c.ny, (long long)c.nz, (long long)c.target);
  }
  {
    const bool ok = emulate_z_pairs<1024>(4 * 4, 1024 / 2 + 1) && emulate_z_pairs<1024>(15, 513);
    std::printf("%s z line pairs N=1024 (4,4,1024) and 15 lines\n", ok ? "OK" : "FAILED");
  }
  {
    const bool ok = emulate_z_pairs<512>(12 * 8, 257) && emulate_z_pairs<512>(512 * 512, 257);
    std::printf("%s z line pairs N=512 (12,8,512) and 512^2 lines\n", ok ? "OK" : "FAILED");
  }
  return failures ? 1 : 0;
}
