// Batched semi-implicit Cahn-Hilliard steps for many independent periodic
// grids of different shapes (float32).
//
// A step of one grid is six sweeps (CH right-hand side, z R2C, y forward,
// x forward * prefac * inverse, y inverse, z C2R + add).  Each sweep is cut
// into tiles; an item is (grid, sweep, tile).  A step is six launches, one
// per sweep, each covering the items of every grid (one workgroup per item,
// whatever the grid's shape), so small grids share the machine instead of
// each paying its own launches.  The host keeps the item list sweep-major.
// The same item list also runs as one cooperative launch per step: every
// block resident, the blocks walking each sweep's items and meeting at the
// runtime's grid barrier between sweeps (one_launch schedule).
// Two tile-shape parameters are chosen by the autotuner: the x-chunk of the
// CH tiles, and the FFT group - how many consecutive z tiles or y/x bricks
// of one grid an FFT item covers (a workgroup runs them back to back).
//
// The tile bodies are the ones of the single-grid kernels.
#include <cmath>
#include <cstring>
#include <vector>

#include <hip/hip_cooperative_groups.h>

#include "evx_fft_tiles.hpp"
#include "evx_stencil_tiles.hpp"

namespace evx {
namespace {

using T = float;
using C = cpx<float>;

constexpr int NSTAGE = 6;
constexpr int BT = 256;  // threads of every batched workgroup

// FFT sizes the batched kernels are built for (all give 256-thread tiles).
#define EVX_BATCH_SIZES(X) X(16) X(32) X(48) X(64) X(96) X(128)

constexpr size_t cmax(size_t a, size_t b) { return a > b ? a : b; }

template <int N>
constexpr size_t fft_lds() {
  return cmax(ZTile<T, N>::LDS_BYTES, cmax(STile<T, N, 0>::LDS_BYTES, STile<T, N, 2>::LDS_BYTES));
}
constexpr size_t BATCH_LDS =
    cmax(ChTile<T>::LDS_BYTES,
         cmax(fft_lds<16>(), cmax(fft_lds<32>(), cmax(fft_lds<48>(), cmax(fft_lds<64>(), cmax(fft_lds<96>(),
                                                                                           fft_lds<128>()))))));
static_assert(BATCH_LDS <= 160 * 1024 / 3, "three batched workgroups must fit one CU's LDS");

#define X(N)                                                                                 \
  static_assert(ZGeo<T, N>::THREADS == BT && SGeo<T, N, 0>::THREADS == BT && SGeo<T, N, 2>::THREADS == BT && \
                    SGeo<T, N, 0>::TL == SGeo<T, N, 2>::TL,                                       \
                "256-thread tiles");
EVX_BATCH_SIZES(X)
#undef X

struct BInst {
  int nx, ny, nz, nh;
  int64_t u_off;  // floats: field in the packed state
  int64_t r_off;  // floats: right-hand side staging in the workspace
  int64_t s_off;  // complex: half spectrum in the workspace
  int tw[3];      // complex offsets of the x, y, z twiddle tables
  int nt[NSTAGE]; // tiles per sweep
  Geo<T> g;
  ChCoef<T> k;
  Prefac pf;
  TileGrid tg;
};

struct BArgs {
  const BInst* inst;
  const int4* items;  // {grid, sweep, first tile, tile count}
  int n_items;
  int n_inst;
  T* u;
  T* rhs;
  C* spec;
  const C* tw;
  int first[NSTAGE + 1];  // items of sweep s: first[s] .. first[s+1]-1
};

// Tiles first .. first+count-1 of one sweep run back to back, a workgroup
// barrier between them (each tile body reuses the LDS).  The loop sits
// inside each size case, so only that case's set-up stays live across it.
template <typename F>
__device__ __forceinline__ void tile_run(int first, int count, F&& body) {
#pragma unroll 1
  for (int j = 0; j < count; ++j) {
    if (j > 0) __syncthreads();
    body(first + j);
  }
}

template <int MODE>
__device__ __forceinline__ void strided_dispatch(const BInst& b, C* spec, const C* tw, int axis, int tile, int count,
                                                 int tid, unsigned char* smem) {
  LineMap lm;
  lm.nh = b.nh;
  int64_t nlines;
  int n;
  if (axis == 1) {
    nlines = (int64_t)b.nx * b.nh;
    lm.G = 1;
    lm.GM = (int64_t)b.ny * b.nh;
    lm.OM = 0;
    lm.es = b.nh;
    n = b.ny;
  } else {
    nlines = (int64_t)b.ny * b.nh;
    lm.G = b.ny;
    lm.GM = (int64_t)b.nx * b.ny * b.nh;
    lm.OM = b.nh;
    lm.es = (int64_t)b.ny * b.nh;
    n = b.nx;
  }
  // Grids of at most 128 points per axis keep every brick offset far below
  // the 4 GiB of strided_tile's 32-bit lane offsets.
  C* s = spec + b.s_off;
  const C* t = tw + b.tw[axis];
  switch (n) {
#define X(N)                                                            \
  case N:                                                               \
    tile_run(tile, count, [&](int k) { strided_tile<T, N, MODE>(s, nlines, lm, t, b.pf, k, tid, smem); }); \
    break;
    EVX_BATCH_SIZES(X)
#undef X
    default:
      break;
  }
}

// One item: it = {grid, sweep, first tile, tile count}.  GROUP = false:
// items of one tile (the count is not read, and the tile loops vanish at
// compile time); the grouped kernel is a separate instantiation, so its
// register needs never touch the one-tile schedule.
// Sweep ST of one item: it = {grid, sweep ST, first tile, tile count}.
// GROUP = false: items of one tile (the count is not read, and the tile
// loops vanish at compile time); the grouped kernels are separate
// instantiations, so their register needs never touch the one-tile schedule.
template <int ST, bool GROUP>
__device__ __forceinline__ void run_sweep(const BArgs& a, int4 it, int tid, unsigned char* smem) {
  const BInst& b = a.inst[it.x];
  const int tile = it.z, count = GROUP ? it.w : 1;
  if constexpr (ST == 0) {
    const Dims d{1, b.nx, b.ny, b.nz};
    Bc bc{};
    const Epi ep{EVX_OUT_RHS, 0.0, nullptr, nullptr, nullptr};
    ch_rhs_tile<T, false>(a.u + b.u_off, a.rhs + b.r_off, d, bc, b.g, b.k, ep, b.tg, nullptr, tile, tid, smem);
  } else if constexpr (ST == 1) {
    const int64_t nlines = (int64_t)b.nx * b.ny;
    switch (b.nz) {
#define X(N)                                                                                                   \
  case N:                                                                                                      \
    tile_run(tile, count, [&](int k) { z_r2c_tile<T, N>(a.rhs + b.r_off, a.spec + b.s_off, nlines, b.nh, a.tw + b.tw[2], k, tid, smem); }); \
    break;
      EVX_BATCH_SIZES(X)
#undef X
      default:
        break;
    }
  } else if constexpr (ST == 2) {
    strided_dispatch<0>(b, a.spec, a.tw, 1, tile, count, tid, smem);
  } else if constexpr (ST == 3) {
    strided_dispatch<2>(b, a.spec, a.tw, 0, tile, count, tid, smem);
  } else if constexpr (ST == 4) {
    strided_dispatch<1>(b, a.spec, a.tw, 1, tile, count, tid, smem);
  } else {
    const int64_t nlines = (int64_t)b.nx * b.ny;
    T* u = a.u + b.u_off;
    switch (b.nz) {
#define X(N)                                                                                              \
  case N:                                                                                                 \
    tile_run(tile, count, [&](int k) { z_c2r_tile<T, N>(a.spec + b.s_off, u, u, T(1), nlines, b.nh, a.tw + b.tw[2], k, tid, smem); }); \
    break;
      EVX_BATCH_SIZES(X)
#undef X
      default:
        break;
    }
  }
}

// One item of any sweep.
template <bool GROUP>
__device__ __forceinline__ void run_item(const BArgs& a, int4 it, int tid, unsigned char* smem) {
  switch (it.y) {
    case 0: run_sweep<0, GROUP>(a, it, tid, smem); break;
    case 1: run_sweep<1, GROUP>(a, it, tid, smem); break;
    case 2: run_sweep<2, GROUP>(a, it, tid, smem); break;
    case 3: run_sweep<3, GROUP>(a, it, tid, smem); break;
    case 4: run_sweep<4, GROUP>(a, it, tid, smem); break;
    default: run_sweep<5, GROUP>(a, it, tid, smem); break;
  }
}

template <bool GROUP>
__global__ __launch_bounds__(BT) void batch_stage(BArgs a, int first) {
  __shared__ __align__(16) unsigned char smem[BATCH_LDS];
  run_item<GROUP>(a, a.items[first + (int)blockIdx.x], threadIdx.x, smem);
}

// The whole step in one cooperative launch (all blocks resident): block b
// takes items b, b + grid, ... of each sweep, and every block passes the
// grid barrier between two sweeps the same number of times (no early exit),
// so the barrier always completes.
template <int ST, bool GROUP>
__device__ __forceinline__ void coop_sweep(const BArgs& a, unsigned char* smem) {
  for (int i = a.first[ST] + (int)blockIdx.x; i < a.first[ST + 1]; i += (int)gridDim.x) {
    // the thread index through an opaque copy: the tile bodies' lane offsets
    // of every size are then recomputed per item instead of being hoisted
    // out of the loop and held in registers all at once
    int tid = (int)threadIdx.x;
    asm volatile("" : "+v"(tid));
    run_sweep<ST, GROUP>(a, a.items[i], tid, smem);
    __syncthreads();  // the next item's tile body rewrites the LDS
  }
}

template <bool GROUP>
__global__ __launch_bounds__(BT) void batch_step_coop(BArgs a) {
  __shared__ __align__(16) unsigned char smem[BATCH_LDS];
  cooperative_groups::grid_group grid = cooperative_groups::this_grid();
  // one loop per sweep (only that sweep's code live in it); every block
  // reaches each of the five grid barriers exactly once
  coop_sweep<0, GROUP>(a, smem);
  grid.sync();  // the sweep's outputs are visible to every block
  coop_sweep<1, GROUP>(a, smem);
  grid.sync();
  coop_sweep<2, GROUP>(a, smem);
  grid.sync();
  coop_sweep<3, GROUP>(a, smem);
  grid.sync();
  coop_sweep<4, GROUP>(a, smem);
  grid.sync();
  coop_sweep<5, GROUP>(a, smem);
}

size_t align_up(size_t v) { return (v + 255) & ~size_t(255); }

bool batch_size_ok(int64_t n) {
  switch (n) {
#define X(N) case N:
    EVX_BATCH_SIZES(X)
#undef X
    return true;
    default:
      return false;
  }
}

template <int N>
constexpr int zt_lp() { return ZGeo<T, N>::LP; }
template <int N>
constexpr int st_tl() { return SGeo<T, N>::TL; }

int lines_per_z_tile(int n) {
  switch (n) {
#define X(N) \
  case N:    \
    return 2 * zt_lp<N>();
    EVX_BATCH_SIZES(X)
#undef X
    default:
      return 0;
  }
}
int lines_per_brick(int n) {
  switch (n) {
#define X(N) \
  case N:    \
    return st_tl<N>();
    EVX_BATCH_SIZES(X)
#undef X
    default:
      return 0;
  }
}

int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace
}  // namespace evx

struct evx_batch_plan {
  int n;
  std::vector<evx::BInst> inst;
  std::vector<int> tw_off_of_size;  // per grid: twiddle offsets (host copy in inst)
  int64_t state_elems = 0;
  size_t ws_bytes = 0;
  // schedule
  int xchunk = 8;   // x planes per CH tile
  int fft_group = 1;  // consecutive tiles per FFT item
  bool one_launch = false;  // the step as one cooperative launch
  int n_items = 0;
  int items_cap = 0;  // capacity of d_items: the item count of the finest tiling (one x plane per CH tile)
  int stage_first[evx::NSTAGE + 1] = {0};
  bool params_set = false;
  // device
  evx::BInst* d_inst = nullptr;
  int4* d_items = nullptr;
  evx::C* d_tw = nullptr;
};

namespace evx {
namespace {

void free_plan(evx_batch_plan* p) {
  if (!p) return;
  (void)hipFree(p->d_inst);
  (void)hipFree(p->d_items);
  (void)hipFree(p->d_tw);
  delete p;
}

// Tile counts of every sweep for the current x-chunk, and the item list
// (FFT sweeps: groups of fft_group consecutive tiles per item).
int build_items(evx_batch_plan* p) {
  std::vector<int4> items;
  for (int s = 0; s < NSTAGE; ++s) {
    p->stage_first[s] = (int)items.size();
    for (int i = 0; i < p->n; ++i) {
      BInst& b = p->inst[i];
      int64_t nt = 0;
      switch (s) {
        case 0: {
          constexpr int BY = NTY * TileCfg<T>::PY;
          b.tg.ntz = (int)cdiv(b.nz, BZ);
          b.tg.nty = (int)cdiv(b.ny, BY);
          b.tg.xc = p->xchunk < b.nx ? p->xchunk : b.nx;
          b.tg.nxc = (int)cdiv(b.nx, b.tg.xc);
          nt = (int64_t)b.tg.ntz * b.tg.nty * b.tg.nxc;
          break;
        }
        case 1:
        case 5:
          nt = cdiv((int64_t)b.nx * b.ny, lines_per_z_tile(b.nz));
          break;
        case 2:
        case 4:
          nt = cdiv((int64_t)b.nx * b.nh, lines_per_brick(b.ny));
          break;
        default:
          nt = cdiv((int64_t)b.ny * b.nh, lines_per_brick(b.nx));
          break;
      }
      if (nt > (1 << 30)) return fail(EVX_ERR_UNSUPPORTED, "too many tiles in one sweep");
      b.nt[s] = (int)nt;
      const int g = s == 0 ? 1 : p->fft_group;
      for (int t = 0; t < (int)nt; t += g) items.push_back(int4{i, s, t, (int)nt - t < g ? (int)nt - t : g});
    }
  }
  p->stage_first[NSTAGE] = (int)items.size();
  p->n_items = (int)items.size();
  if (p->d_items == nullptr) {
    // Allocated once, for the finest tiling, and kept until the plan is
    // destroyed: a later evx_batch_set_tiles rewrites the list in place, so a
    // graph captured before it never reads freed memory (it does replay the
    // old schedule against the new list, which the caller must not do).
    int64_t cap = (int64_t)items.size();
    for (int i = 0; i < p->n; ++i) {
      const BInst& b = p->inst[i];
      cap += (int64_t)b.tg.ntz * b.tg.nty * (b.nx - b.tg.nxc);
    }
    if (cap > (1 << 30)) return fail(EVX_ERR_UNSUPPORTED, "too many tiles in one sweep");
    if (hipMalloc(&p->d_items, sizeof(int4) * cap) != hipSuccess)
      return fail(EVX_ERR_NOMEM, "hipMalloc of the batched item list failed");
    p->items_cap = (int)cap;
  }
  if (p->n_items > p->items_cap) return fail(EVX_ERR_INVALID, "batched item list exceeds its capacity");
  if (hipMemcpy(p->d_items, items.data(), sizeof(int4) * items.size(), hipMemcpyHostToDevice) != hipSuccess)
    return fail(EVX_ERR_HIP, "upload of the batched item list failed");
  if (hipMemcpy(p->d_inst, p->inst.data(), sizeof(BInst) * p->n, hipMemcpyHostToDevice) != hipSuccess)
    return fail(EVX_ERR_HIP, "upload of the batched grid table failed");
  return EVX_OK;
}

BArgs make_args(evx_batch_plan* p, void* state, void* ws) {
  BArgs a;
  a.inst = p->d_inst;
  a.items = p->d_items;
  a.n_items = p->n_items;
  a.n_inst = p->n;
  a.u = static_cast<T*>(state);
  a.rhs = static_cast<T*>(ws);
  a.spec = reinterpret_cast<C*>(static_cast<char*>(ws) + align_up(sizeof(T) * p->state_elems));
  a.tw = p->d_tw;
  for (int st = 0; st <= NSTAGE; ++st) a.first[st] = p->stage_first[st];
  return a;
}

// Blocks of the cooperative step: at most two per CU (fewer if fewer fit).
// The grid barrier's cost grows with the number of blocks: on the 48-grid
// draw the step took 0.95 ms with all four resident blocks per CU, 0.68 ms
// with two and 0.71 ms with one (MI355X).
template <bool GROUP>
int coop_resident() {
  static const int resident = [] {
    int dev = 0, cus = 0, per = 0, coop = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) != hipSuccess || !coop ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, batch_step_coop<GROUP>, BT, 0) != hipSuccess)
      return 0;
    return (per < 2 ? per : 2) * cus;
  }();
  return resident;
}

template <bool GROUP>
int launch_coop(const BArgs& a, hipStream_t s) {
  const int res = coop_resident<GROUP>();
  if (res < 1) return fail(EVX_ERR_UNSUPPORTED, "cooperative launches are not available on this device");
  BArgs arg = a;
  void* params[] = {&arg};
  if (hipLaunchCooperativeKernel(reinterpret_cast<const void*>(batch_step_coop<GROUP>), dim3((unsigned)res),
                                 dim3(BT), params, 0, s) != hipSuccess)
    return fail(EVX_ERR_HIP, "cooperative launch of the batched step failed: %s", hipGetErrorString(hipGetLastError()));
  return EVX_OK;
}

int launch_step(evx_batch_plan* p, void* state, void* ws, hipStream_t s) {
  const BArgs a = make_args(p, state, ws);
  if (p->one_launch) return p->fft_group > 1 ? launch_coop<true>(a, s) : launch_coop<false>(a, s);
  for (int st = 0; st < NSTAGE; ++st) {
    const int n = p->stage_first[st + 1] - p->stage_first[st];
    if (n == 0) continue;
    if (st > 0 && p->fft_group > 1)
      hipLaunchKernelGGL(batch_stage<true>, dim3(n), dim3(BT), 0, s, a, p->stage_first[st]);
    else
      hipLaunchKernelGGL(batch_stage<false>, dim3(n), dim3(BT), 0, s, a, p->stage_first[st]);
    EVX_TRY(check_launch("batch_stage"));
  }
  return EVX_OK;
}

}  // namespace
}  // namespace evx

using namespace evx;

extern "C" int evx_batch_plan_create(evx_batch_plan** out, int32_t dtype, int32_t n_inst, const int64_t* shapes,
                                     const double* spacing) {
  if (!out || !shapes || !spacing) return fail(EVX_ERR_INVALID, "null argument");
  *out = nullptr;
  if (dtype != EVX_F32) return fail(EVX_ERR_UNSUPPORTED, "batched steps are built for float32");
  if (n_inst < 1 || n_inst > 4096) return fail(EVX_ERR_INVALID, "n_inst must be in [1, 4096]");
  auto* p = new evx_batch_plan();
  p->n = n_inst;
  p->inst.resize(n_inst);
  // twiddle tables, one per distinct size
  std::vector<int> sizes, offs;
  std::vector<C> tw;
  auto tw_of = [&](int n) -> int {
    for (size_t i = 0; i < sizes.size(); ++i)
      if (sizes[i] == n) return offs[i];
    sizes.push_back(n);
    offs.push_back((int)tw.size());
    const long double tau = 6.283185307179586476925286766559005768L;
    for (int m = 0; m < n; ++m) {
      const long double ang = tau * (long double)m / (long double)n;
      tw.push_back(C{static_cast<T>(cosl(ang)), static_cast<T>(-sinl(ang))});
    }
    return offs.back();
  };
  int64_t u_off = 0, s_off = 0;
  for (int i = 0; i < n_inst; ++i) {
    BInst& b = p->inst[i];
    const int64_t nx = shapes[3 * i], ny = shapes[3 * i + 1], nz = shapes[3 * i + 2];
    if (!batch_size_ok(nx) || !batch_size_ok(ny) || !batch_size_ok(nz)) {
      free_plan(p);
      return fail(EVX_ERR_UNSUPPORTED, "grid %d: edges must be in {16, 32, 48, 64, 96, 128}, got (%lld, %lld, %lld)",
                  i, (long long)nx, (long long)ny, (long long)nz);
    }
    for (int a = 0; a < 3; ++a)
      if (!(spacing[3 * i + a] > 0.0)) {
        free_plan(p);
        return fail(EVX_ERR_INVALID, "grid %d: spacing must be positive", i);
      }
    b.nx = (int)nx;
    b.ny = (int)ny;
    b.nz = (int)nz;
    b.nh = (int)(nz / 2 + 1);
    b.u_off = u_off;
    b.r_off = u_off;
    b.s_off = s_off;
    u_off += (int64_t)(align_up(sizeof(T) * nx * ny * nz) / sizeof(T));
    s_off += (int64_t)(align_up(sizeof(C) * nx * ny * b.nh) / sizeof(C));
    b.tw[0] = tw_of(b.nx);
    b.tw[1] = tw_of(b.ny);
    b.tw[2] = tw_of(b.nz);
    double dd[3], dd2[3];
    for (int a = 0; a < 3; ++a) {
      const float sp = static_cast<float>(spacing[3 * i + a]);
      dd[a] = 1.0f / sp;
      dd2[a] = 1.0f / (sp * sp);
    }
    b.g = make_geo_from<T>(dd, dd2);
    Prefac& pf = b.pf;
    std::memset(&pf, 0, sizeof(pf));
    pf.ch = 1;
    pf.twopi = static_cast<float>(2.0 * M_PI);
    const int64_t n3[3] = {nx, ny, nz};
    for (int a = 0; a < 3; ++a) pf.inv_nd[a] = static_cast<float>(1.0 / ((double)n3[a] * spacing[3 * i + a]));
    pf.nx = b.nx;
    pf.ny = b.ny;
    pf.nz = b.nz;
    pf.scale = 1.0 / ((double)nx * (double)ny * (double)nz);
  }
  p->state_elems = u_off;
  p->ws_bytes = align_up(sizeof(T) * u_off) + align_up(sizeof(C) * s_off);
  if (hipMalloc(&p->d_inst, sizeof(BInst) * n_inst) != hipSuccess ||
      hipMalloc(&p->d_tw, sizeof(C) * tw.size()) != hipSuccess) {
    free_plan(p);
    return fail(EVX_ERR_NOMEM, "hipMalloc of batched plan tables failed");
  }
  if (hipMemcpy(p->d_tw, tw.data(), sizeof(C) * tw.size(), hipMemcpyHostToDevice) != hipSuccess) {
    free_plan(p);
    return fail(EVX_ERR_HIP, "batched plan upload failed");
  }
  const int st = build_items(p);
  if (st != EVX_OK) {
    free_plan(p);
    return st;
  }
  *out = p;
  return EVX_OK;
}

extern "C" int evx_batch_plan_destroy(evx_batch_plan* p) {
  free_plan(p);
  return EVX_OK;
}

extern "C" int evx_batch_state_layout(const evx_batch_plan* p, int64_t* offsets, int64_t* total) {
  if (!p || !total) return fail(EVX_ERR_INVALID, "null argument");
  if (offsets)
    for (int i = 0; i < p->n; ++i) offsets[i] = p->inst[i].u_off;
  *total = p->state_elems;
  return EVX_OK;
}

extern "C" int evx_batch_workspace_bytes(const evx_batch_plan* p, size_t* bytes) {
  if (!p || !bytes) return fail(EVX_ERR_INVALID, "null argument");
  *bytes = p->ws_bytes;
  return EVX_OK;
}

extern "C" int evx_batch_set_params(evx_batch_plan* p, const double* eps_d, double dt, double A) {
  if (!p || !eps_d) return fail(EVX_ERR_INVALID, "null argument");
  for (int i = 0; i < p->n; ++i) {
    const double eps = eps_d[2 * i], D = eps_d[2 * i + 1];
    if (!(eps != 0.0)) return fail(EVX_ERR_INVALID, "grid %d: eps must be nonzero", i);
    BInst& b = p->inst[i];
    b.k = ChCoef<T>{static_cast<T>(18.0 / eps), static_cast<T>(2.0 * eps), static_cast<T>(D)};
    b.pf.sc = static_cast<float>(((-2.0 * eps) * D) * A);  // problem_definition.py:303
    b.pf.dt = static_cast<float>(dt);
    b.pf.chan = nullptr;
    b.pf.A = A;
  }
  if (hipMemcpy(p->d_inst, p->inst.data(), sizeof(BInst) * p->n, hipMemcpyHostToDevice) != hipSuccess)
    return fail(EVX_ERR_HIP, "upload of the batched grid table failed");
  p->params_set = true;
  return EVX_OK;
}

extern "C" int evx_batch_set_tiles(evx_batch_plan* p, int xchunk) {
  if (!p) return fail(EVX_ERR_INVALID, "null plan");
  if (xchunk < 1 || xchunk > 4096) return fail(EVX_ERR_INVALID, "xchunk must be in [1, 4096]");
  // The previous launch of this plan must be complete before the item list
  // changes (the caller synchronises; set-up time only).
  if (xchunk != p->xchunk) {
    p->xchunk = xchunk;
    EVX_TRY(build_items(p));
  }
  return EVX_OK;
}

extern "C" int evx_batch_set_fft_group(evx_batch_plan* p, int tiles_per_item) {
  if (!p) return fail(EVX_ERR_INVALID, "null plan");
  if (tiles_per_item < 1 || tiles_per_item > 64) return fail(EVX_ERR_INVALID, "tiles_per_item must be in [1, 64]");
  // As for evx_batch_set_tiles: the previous launch must be complete.
  if (tiles_per_item != p->fft_group) {
    p->fft_group = tiles_per_item;
    EVX_TRY(build_items(p));
  }
  return EVX_OK;
}

extern "C" int evx_batch_get_fft_group(const evx_batch_plan* p, int* tiles_per_item) {
  if (!p || !tiles_per_item) return fail(EVX_ERR_INVALID, "null argument");
  *tiles_per_item = p->fft_group;
  return EVX_OK;
}

extern "C" int evx_batch_set_one_launch(evx_batch_plan* p, int one_launch) {
  if (!p) return fail(EVX_ERR_INVALID, "null plan");
  if (one_launch && (p->fft_group > 1 ? coop_resident<true>() : coop_resident<false>()) < 1)
    return fail(EVX_ERR_UNSUPPORTED, "cooperative launches are not available on this device");
  p->one_launch = one_launch != 0;
  return EVX_OK;
}

extern "C" int evx_batch_get_one_launch(const evx_batch_plan* p, int* one_launch) {
  if (!p || !one_launch) return fail(EVX_ERR_INVALID, "null argument");
  *one_launch = p->one_launch ? 1 : 0;
  return EVX_OK;
}

extern "C" int evx_batch_get_tiles(const evx_batch_plan* p, int* xchunk, int* n_items) {
  if (!p || !xchunk || !n_items) return fail(EVX_ERR_INVALID, "null argument");
  *xchunk = p->xchunk;
  *n_items = p->n_items;
  return EVX_OK;
}

extern "C" int evx_batch_ch_imex_step(evx_batch_plan* p, void* state, void* workspace, void* stream) {
  if (!p || !state || !workspace) return fail(EVX_ERR_INVALID, "null argument");
  if (!p->params_set) return fail(EVX_ERR_INVALID, "evx_batch_set_params has not been called");
  return launch_step(p, state, workspace, static_cast<hipStream_t>(stream));
}

// Mean time per step of `steps` steps on the current schedule, after one
// warm-up step (HIP events; synchronises the stream).
static int time_schedule(evx_batch_plan* p, void* state, void* ws, int steps, hipStream_t s, hipEvent_t e0,
                         hipEvent_t e1, double* ms_per_step) {
  if (hipStreamSynchronize(s) != hipSuccess) return fail(EVX_ERR_HIP, "stream synchronisation failed");
  EVX_TRY(launch_step(p, state, ws, s));
  if (hipEventRecord(e0, s) != hipSuccess) return fail(EVX_ERR_HIP, "hipEventRecord failed");
  for (int i = 0; i < steps; ++i) EVX_TRY(launch_step(p, state, ws, s));
  if (hipEventRecord(e1, s) != hipSuccess) return fail(EVX_ERR_HIP, "hipEventRecord failed");
  if (hipEventSynchronize(e1) != hipSuccess) return fail(EVX_ERR_HIP, "event synchronisation failed");
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return fail(EVX_ERR_HIP, "hipEventElapsedTime failed");
  *ms_per_step = ms / steps;
  return EVX_OK;
}

// Autotuner: times the candidate schedules over `steps` steps each and
// keeps the fastest: first the x-chunk of the CH tiles (x planes per tile,
// which sets the number of CH items and the halo overhead of each) at the
// current FFT group, then the FFT group (tiles per FFT item) at that
// x-chunk, then one cooperative launch per step against one launch per
// sweep.  The candidates run on the caller's scratch state, which is
// overwritten.  Synchronises the stream (set-up time only).
extern "C" int evx_batch_autotune(evx_batch_plan* p, void* scratch_state, void* workspace, int steps, void* stream,
                                  double* best_ms) {
  if (!p || !scratch_state || !workspace) return fail(EVX_ERR_INVALID, "null argument");
  if (!p->params_set) return fail(EVX_ERR_INVALID, "evx_batch_set_params has not been called");
  if (steps < 1 || steps > 1000) return fail(EVX_ERR_INVALID, "steps must be in [1, 1000]");
  hipStream_t s = static_cast<hipStream_t>(stream);
  // the item list is rewritten between candidates: nothing may be in flight
  if (hipStreamSynchronize(s) != hipSuccess) return fail(EVX_ERR_HIP, "stream synchronisation failed");
  const int xcands[] = {4, 6, 8, 12, 16, 24, 32, 48, 64};
  const int gcands[] = {1, 2, 4, 8};
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return fail(EVX_ERR_HIP, "hipEventCreate failed");
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return fail(EVX_ERR_HIP, "hipEventCreate failed");
  }
  double best = 1e300;
  int xpick = p->xchunk, gpick = p->fft_group;
  const bool one_before = p->one_launch;
  p->one_launch = false;  // the tile shapes are timed with one launch per sweep
  int st = EVX_OK;
  for (const int c : xcands) {
    double ms = 0.0;
    if ((st = evx_batch_set_tiles(p, c)) != EVX_OK) break;
    if ((st = time_schedule(p, scratch_state, workspace, steps, s, e0, e1, &ms)) != EVX_OK) break;
    if (ms < best) {
      best = ms;
      xpick = c;
    }
  }
  if (st == EVX_OK) st = evx_batch_set_tiles(p, xpick);
  for (const int g : gcands) {
    if (st != EVX_OK) break;
    double ms = 0.0;
    if ((st = evx_batch_set_fft_group(p, g)) != EVX_OK) break;
    if ((st = time_schedule(p, scratch_state, workspace, steps, s, e0, e1, &ms)) != EVX_OK) break;
    if (ms < best) {
      best = ms;
      gpick = g;
    }
  }
  // then one cooperative launch per step against one launch per sweep
  bool opick = false;
  if (st == EVX_OK) st = evx_batch_set_fft_group(p, gpick);
  const bool coop = (gpick > 1 ? coop_resident<true>() : coop_resident<false>()) > 0;
  for (const bool one : {false, true}) {
    if (st != EVX_OK || (one && !coop)) break;
    double ms = 0.0;
    p->one_launch = one;
    if ((st = time_schedule(p, scratch_state, workspace, steps, s, e0, e1, &ms)) != EVX_OK) break;
    if (ms < best) {
      best = ms;
      opick = one;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st != EVX_OK) {
    p->one_launch = one_before;
    return st;
  }
  if (hipStreamSynchronize(s) != hipSuccess) return fail(EVX_ERR_HIP, "stream synchronisation failed");
  EVX_TRY(evx_batch_set_tiles(p, xpick));
  EVX_TRY(evx_batch_set_fft_group(p, gpick));
  p->one_launch = opick;
  if (best_ms) *best_ms = best;
  return EVX_OK;
}
