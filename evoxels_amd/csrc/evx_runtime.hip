// Library-level state: error reporting and descriptor validation.
#include <cstring>

#include "evx_common.hpp"

namespace evx {
namespace {
thread_local char g_err[512] = "";
}

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int fail(int status, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return status;
}

int check_launch(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(EVX_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
  return EVX_OK;
}

int make_dims(const evx_grid* g, Dims* d) {
  if (!g) return fail(EVX_ERR_INVALID, "null grid descriptor");
  if (g->channels < 1 || g->nx < 1 || g->ny < 1 || g->nz < 1)
    return fail(EVX_ERR_INVALID, "grid extents must be positive (got %lld,%lld,%lld,%lld)", (long long)g->channels,
                (long long)g->nx, (long long)g->ny, (long long)g->nz);
  *d = Dims{g->channels, g->nx, g->ny, g->nz};
  return EVX_OK;
}

int make_bc(const evx_bc* bc, Bc* out) {
  for (int a = 0; a < 3; ++a) {
    const int k = bc->kind[a];
    if (k == EVX_BC_NODE_VALUE || k == EVX_BC_NODE_EXTRAP) {
      // staggered_x rules: x only, y and z periodic (boundary_conditions.py:85-131)
      if (a != 0 || bc->kind[1] != EVX_BC_PERIODIC || bc->kind[2] != EVX_BC_PERIODIC)
        return fail(EVX_ERR_INVALID, "staggered_x boundary rules need a non-periodic x axis and periodic y, z");
    } else if (k != EVX_BC_PERIODIC && k != EVX_BC_NEUMANN && k != EVX_BC_DIRICHLET) {
      return fail(EVX_ERR_INVALID, "Unsupported BC type code %d on axis %d", k, a);
    }
    out->kind[a] = k;
    // 2.0 * value in Python double (boundary_conditions.py:54-55)
    out->two_lo[a] = 2.0 * bc->lo[a];
    out->two_hi[a] = 2.0 * bc->hi[a];
  }
  out->node_lo = bc->lo[0];
  out->node_hi = bc->hi[0];
  return EVX_OK;
}

}  // namespace evx

extern "C" int evx_version(void) { return EVX_ABI_VERSION; }

extern "C" const char* evx_last_error(void) { return evx::g_err; }
