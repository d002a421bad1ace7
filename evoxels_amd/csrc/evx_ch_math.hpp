// Point formulas of the finite-difference operators, shared by the stencil
// kernels and the fused right-hand-side + FFT kernel.
//
// Two flavours: Ops<true> evaluates every product and sum with contraction
// off, so results round exactly where the reference's eager torch expression
// rounds (fd_stencils.py:20-75, problem_definition.py:350-371); Ops<false>
// lets the compiler fuse multiply-adds and rearranges the flux algebra.  The explicit Euler / RK4 paths use the
// strict flavour; the semi-implicit spectral path, whose FFT already rounds
// differently from pocketfft, uses the contracted one.
#pragma once
#include "evx_common.hpp"

namespace evx {

template <typename T>
struct Geo {
  T idx, idy, idz;     // div_dx
  T idx2, idy2, idz2;  // div_dx2
  T sum2;              // torch.sum(div_dx2), accumulated x, y, z
};

template <typename T>
inline Geo<T> make_geo_from(const double dd[3], const double dd2[3]) {
  Geo<T> o;
  o.idx = static_cast<T>(dd[0]);
  o.idy = static_cast<T>(dd[1]);
  o.idz = static_cast<T>(dd[2]);
  o.idx2 = static_cast<T>(dd2[0]);
  o.idy2 = static_cast<T>(dd2[1]);
  o.idz2 = static_cast<T>(dd2[2]);
  o.sum2 = (o.idx2 + o.idy2) + o.idz2;
  return o;
}

template <typename T>
struct ChCoef {
  T s18;      // 18/eps
  T two_eps;  // 2*eps
  T D;
};

#define EVX_OPS_BODY(PRAGMA)                                                                              \
  /* 7-point Laplacian in the order of fd_stencils.py:70-74. */                                           \
  template <typename T>                                                                                   \
  static __device__ __forceinline__ T lap7(T c, T xp, T xm, T yp, T ym, T zp, T zm, const Geo<T>& g) {    \
    PRAGMA                                                                                                \
    return (((xp + xm) * g.idx2 + (yp + ym) * g.idy2) + (zp + zm) * g.idz2) - (T(2) * c) * g.sum2;        \
  }                                                                                                       \
  /* mu = ((18/eps*c)*(1-c))*(1-2c) - (2eps)*lap(c)   (problem_definition.py:305,353) */                 \
  template <typename T>                                                                                   \
  static __device__ __forceinline__ T mu(T c, T xp, T xm, T yp, T ym, T zp, T zm, const Geo<T>& g,        \
                                         const ChCoef<T>& k) {                                            \
    PRAGMA                                                                                                \
    const T lap = lap7(c, xp, xm, yp, ym, zp, zm, g);                                                     \
    return ((k.s18 * c) * (T(1) - c)) * (T(1) - T(2) * c) - k.two_eps * lap;                              \
  }                                                                                                       \
  /* (cf*(1-cf)) * grad_face(mu), cf = to_face(c)   (problem_definition.py:356-369) */                   \
  template <typename T>                                                                                   \
  static __device__ __forceinline__ T flux(T c_lo, T c_hi, T m_lo, T m_hi, T id) {                        \
    PRAGMA                                                                                                \
    const T cf = T(0.5) * (c_hi + c_lo);                                                                  \
    return (cf * (T(1) - cf)) * ((m_hi - m_lo) * id);                                                     \
  }                                                                                                       \
  /* D * ((div_x + div_y) + div_z) from the six face fluxes. */                                           \
  template <typename T>                                                                                   \
  static __device__ __forceinline__ T div(T fxm, T fxp, T fym, T fyp, T fzm, T fzp, const Geo<T>& g, T D) { \
    PRAGMA                                                                                                \
    const T dv = ((fxp - fxm) * g.idx + (fyp - fym) * g.idy) + (fzp - fzm) * g.idz;                       \
    return D * dv;                                                                                        \
  }

template <bool STRICT>
struct Ops;
template <>
struct Ops<true> {
  EVX_OPS_BODY(_Pragma("clang fp contract(off)"))
};
// Contracted flavour for the semi-implicit path (multiply-adds may fuse and
// the algebra is rearranged; the results differ from the strict flavour by
// roundings only):
// * mu = (18/eps*c)*p(c) - 2eps*lap(c) with p(c) = (1-c)(1-2c) = 1 - 3c + 2c^2
//   in Horner form and 2eps folded into the Laplacian weights;
// * the face flux is returned as s(2-s)*(m_hi-m_lo) = (1 - t^2)*(m_hi-m_lo)
//   with t = 1 - c_lo - c_hi, i.e. 4*cf(1-cf)*grad_face(mu)/id: the 1/4 and
//   the face-gradient weight id are folded into the divergence weights,
//   D*id^2/4 per axis.  flux_q takes q = 1 - c of one side precomputed.
template <>
struct Ops<false> {
  template <typename T>
  static __device__ __forceinline__ T lap7(T c, T xp, T xm, T yp, T ym, T zp, T zm, const Geo<T>& g) {
    return (((xp + xm) * g.idx2 + (yp + ym) * g.idy2) + (zp + zm) * g.idz2) - (T(2) * c) * g.sum2;
  }
  template <typename T>
  static __device__ __forceinline__ T mu(T c, T xp, T xm, T yp, T ym, T zp, T zm, const Geo<T>& g,
                                         const ChCoef<T>& k) {
    const T w = k.two_eps;
    const T l = (((xp + xm) * (w * g.idx2) + (yp + ym) * (w * g.idy2)) + (zp + zm) * (w * g.idz2)) -
                c * ((T(2) * w) * g.sum2);
    const T p = (T(2) * c - T(3)) * c + T(1);
    return (k.s18 * c) * p - l;
  }
  template <typename T>
  static __device__ __forceinline__ T flux_q(T q_lo, T c_hi, T m_lo, T m_hi) {
    const T t = q_lo - c_hi;
    return (T(1) - t * t) * (m_hi - m_lo);
  }
  template <typename T>
  static __device__ __forceinline__ T flux(T c_lo, T c_hi, T m_lo, T m_hi, T /*id*/) {
    return flux_q(T(1) - c_lo, c_hi, m_lo, m_hi);
  }
  template <typename T>
  static __device__ __forceinline__ T div(T fxm, T fxp, T fym, T fyp, T fzm, T fzp, const Geo<T>& g, T D) {
    const T q = T(0.25) * D;
    return ((fxp - fxm) * (q * g.idx2) + (fyp - fym) * (q * g.idy2)) + (fzp - fzm) * (q * g.idz2);
  }
};
#undef EVX_OPS_BODY

// The contracted flavour's coefficient products folded once (on the host, or
// once per block for per-channel parameters) instead of at every point: the
// same float products Ops<false>::mu and Ops<false>::div form, so the
// results are the same.
template <typename T>
struct ChW {
  T wx, wy, wz, wc;  // (2eps)*idx2, (2eps)*idy2, (2eps)*idz2, (2*(2eps))*sum2
  T s18;             // 18/eps
  T dx, dy, dz;      // (D/4)*idx2, (D/4)*idy2, (D/4)*idz2
};

template <typename T>
__host__ __device__ inline ChW<T> ch_weights(const Geo<T>& g, const ChCoef<T>& k) {
  ChW<T> o;
  const T w = k.two_eps;
  o.wx = w * g.idx2;
  o.wy = w * g.idy2;
  o.wz = w * g.idz2;
  o.wc = (T(2) * w) * g.sum2;
  o.s18 = k.s18;
  const T q = T(0.25) * k.D;
  o.dx = q * g.idx2;
  o.dy = q * g.idy2;
  o.dz = q * g.idz2;
  return o;
}

// Ops<false>::mu and Ops<false>::div with folded weights.
template <typename T>
__device__ __forceinline__ T mu_w(T c, T xp, T xm, T yp, T ym, T zp, T zm, const ChW<T>& W) {
  const T l = (((xp + xm) * W.wx + (yp + ym) * W.wy) + (zp + zm) * W.wz) - c * W.wc;
  const T p = (T(2) * c - T(3)) * c + T(1);
  return (W.s18 * c) * p - l;
}
template <typename T>
__device__ __forceinline__ T div_w(T fxm, T fxp, T fym, T fyp, T fzm, T fzp, const ChW<T>& W) {
  return ((fxp - fxm) * W.dx + (fyp - fym) * W.dy) + (fzp - fzm) * W.dz;
}

// Strict-order shorthands used by code that must match the reference rounding.
template <typename T>
__device__ __forceinline__ T lap7(T c, T xp, T xm, T yp, T ym, T zp, T zm, const Geo<T>& g) {
  return Ops<true>::lap7(c, xp, xm, yp, ym, zp, zm, g);
}
template <typename T>
__device__ __forceinline__ T ch_mu(T c, T xp, T xm, T yp, T ym, T zp, T zm, const Geo<T>& g, const ChCoef<T>& k) {
  return Ops<true>::mu(c, xp, xm, yp, ym, zp, zm, g, k);
}
template <typename T>
__device__ __forceinline__ T face_flux(T c_lo, T c_hi, T m_lo, T m_hi, T id) {
  return Ops<true>::flux(c_lo, c_hi, m_lo, m_hi, id);
}
template <typename T>
__device__ __forceinline__ T ch_div(T fxm, T fxp, T fym, T fyp, T fzm, T fzp, const Geo<T>& g, T D) {
  return Ops<true>::div(fxm, fxp, fym, fyp, fzm, fzp, g, D);
}

}  // namespace evx
