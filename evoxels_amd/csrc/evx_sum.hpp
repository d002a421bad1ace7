// Compensated float64 sums shared by the reductions and the step epilogues:
// (sum, compensation) pairs merged with the error-free TwoSum, in a fixed
// order, so results are reproducible run to run.
#pragma once
#include "evx_common.hpp"

namespace evx {

struct SumPair {
  double s, c;
};

__host__ __device__ inline SumPair sum_merge(SumPair a, SumPair b) {
#pragma clang fp contract(off)
  const double s = a.s + b.s;
  const double bb = s - a.s;
  const double e = (a.s - (s - bb)) + (b.s - bb);
  return SumPair{s, (a.c + b.c) + e};
}

// Per-block record a step's last sweep leaves for the state it wrote: the
// compensated sum of the block's outputs and whether any was NaN / +-Inf.
struct SumRec {
  double s, c;
  int32_t nan, inf;
  int32_t pad[2];
};
static_assert(sizeof(SumRec) == 32, "records are 32 bytes");

#if defined(__HIP_DEVICE_COMPILE__) || defined(__HIPCC__)
// Merge of a pair across the 64 lanes of a wavefront (xor butterflies); every
// lane ends with the wavefront's pair.
__device__ inline SumPair wave_sum(SumPair p) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const SumPair q{__shfl_xor(p.s, o, 64), __shfl_xor(p.c, o, 64)};
    p = sum_merge(p, q);
  }
  return p;
}
#endif

}  // namespace evx
