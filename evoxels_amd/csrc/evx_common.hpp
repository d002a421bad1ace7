// Shared device/host helpers of libevoxels_hip.so (gfx950).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>

#include "evoxels_hip.h"

namespace evx {

// ---------------------------------------------------------------- errors --
void set_error(const char* fmt, ...);
int fail(int status, const char* fmt, ...);
// Checks hipGetLastError() after a launch; returns EVX_OK or EVX_ERR_HIP.
int check_launch(const char* what);

#define EVX_TRY(expr)                 \
  do {                                \
    int _st = (expr);                 \
    if (_st != EVX_OK) return _st;    \
  } while (0)

// ------------------------------------------------------------ geometry ----
struct Dims {
  int64_t c, nx, ny, nz;
  __host__ __device__ int64_t plane() const { return ny * nz; }
  __host__ __device__ int64_t vol() const { return nx * ny * nz; }
};

// Boundary description in kernel-argument form.  two_lo/two_hi hold 2*v as
// the reference computes it in Python double (boundary_conditions.py:54-55);
// kernels round it to the working precision before the subtraction, exactly
// like torch's scalar-minus-tensor.
struct Bc {
  int kind[3];
  double two_lo[3];
  double two_hi[3];
  double node_lo, node_hi;  // staggered_x boundary node values of the x axis
};

__host__ __device__ inline bool staggered_x(const Bc& bc) { return bc.kind[0] >= EVX_BC_NODE_VALUE; }

int make_dims(const evx_grid* g, Dims* d);
int make_bc(const evx_bc* bc, Bc* out);

__host__ __device__ inline int64_t wrap(int64_t i, int64_t n) {
  int64_t r = i % n;
  return r < 0 ? r + n : r;
}

// Ghost evaluation for a point that may lie outside the grid.  Periodic
// coordinates wrap; non-periodic ones clamp to the adjacent interior index;
// then the affine ghost rules are applied in the order x, y, z.  This equals
// the padded array of CellCenteredBCs.pad_bc (boundary_conditions.py:33-59)
// at every face, edge and corner ghost: the periodic pad runs first and each
// later axis rule reads the already overwritten planes of the earlier axes.
// `xform` maps the raw stored value (e.g. clip) before the rules apply.
template <typename T>
struct Resolved {
  int64_t off;   // element offset of the interior source point
  int8_t o[3];   // -1 low ghost, +1 high ghost, 0 inside (non-periodic axes)
};

// Periodic wrap for points at most one period outside the grid (the stencil
// halos); falls back to a division only for degenerate tiny axes.
__host__ __device__ inline int64_t wrap_near(int64_t i, int64_t n) {
  if (i < 0) {
    i += n;
    if (i < 0) i = wrap(i, n);
  } else if (i >= n) {
    i -= n;
    if (i >= n) i = wrap(i, n);
  }
  return i;
}

template <typename T>
__device__ __forceinline__ Resolved<T> resolve(int64_t x, int64_t y, int64_t z,
                                               const Dims& d, const Bc& bc) {
  Resolved<T> r;
  int64_t p[3] = {x, y, z};
  const int64_t n[3] = {d.nx, d.ny, d.nz};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    r.o[a] = 0;
    if (p[a] < 0 || p[a] >= n[a]) {
      if (bc.kind[a] == EVX_BC_PERIODIC) {
        p[a] = wrap_near(p[a], n[a]);
      } else {
        r.o[a] = p[a] < 0 ? -1 : 1;
        p[a] = p[a] < 0 ? 0 : n[a] - 1;
      }
    }
  }
  r.off = (p[0] * d.ny + p[1]) * d.nz + p[2];
  return r;
}

template <typename T>
__device__ __forceinline__ T apply_rules(T v, const int8_t o[3], const Bc& bc) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (o[a] != 0 && bc.kind[a] == EVX_BC_DIRICHLET) {
      const T two_v = static_cast<T>(o[a] < 0 ? bc.two_lo[a] : bc.two_hi[a]);
      v = two_v - v;
    }
  }
  return v;
}

template <typename T>
__device__ __forceinline__ T clip01(T v) {
  // torch.clip(c, 0, 1): min(max(v, 0), 1), NaN propagates (both
  // comparisons are false for NaN, so v passes through).  Two compares and
  // two selects; the form with an explicit v != v test compiled to branches.
  const T lo = v < T(0) ? T(0) : v;
  return lo > T(1) ? T(1) : lo;
}

// ------------------------------------------------------------- complex ----
template <typename T>
struct cpx {
  T x, y;
};

template <typename T>
__device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> csub(cpx<T> a, cpx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cpx<T> cmulc(cpx<T> a, cpx<T> b) {  // a * conj(b)
  return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
template <typename T>
__device__ __forceinline__ cpx<T> cscale(cpx<T> a, T s) { return {a.x * s, a.y * s}; }

}  // namespace evx
