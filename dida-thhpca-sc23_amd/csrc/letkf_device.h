// Device-side helpers shared by the LETKF kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "letkf_kernels.h"

namespace dida {

constexpr int kWave = 64;

// Single-precision literals of the reference, widened exactly to double:
// pro_info.F90:148 (a = 6371.393) and :156 (pi = atan(1.0)*4 in REAL(4)),
// letkf_yb.F90:256-261 (1.666667, 0.6666667).
constexpr double kEarthA = (double)6371.393f;
constexpr double kPiF32 = (double)3.14159274101257324f;
constexpr double kGc53 = (double)1.666667f;
constexpr double kGc23 = (double)0.6666667f;

// Gaspari-Cohn weight of letkf_yb.F90:245-269 (schr2) or the boxcar of
// :270-285 (schr), evaluated in the reference's operation order.
__device__ __forceinline__ double gc_weight(double ri, int gc_filter) {
#pragma clang fp contract(off)
  if (!gc_filter) return ri > 2.0 ? 0.0 : 1.0;
  double rci, rj, rk;
  if (ri > 2.0) {
    rci = 0.0;
  } else if (ri > 1.0) {
    rj = ri / 12.0 - 0.5;
    rj = rj * ri + 0.625;
    rj = rj * ri + kGc53;
    rci = (rj * ri - 5.0) * ri + 4.0 - kGc23 / ri;
  } else if (ri > 0.0) {
    rk = ri * ri;
    rj = (-0.25 * ri + 0.5) * ri + 0.625;
    rci = (rj * ri - kGc53) * rk + 1.0;
  } else {
    rci = 1.0;
  }
  return rci;
}

// Degrees to the reference's radians (x/180*pi with the REAL(4) pi).
__device__ __forceinline__ double deg2rad_ref(double x) {
#pragma clang fp contract(off)
  return x / 180.0 * kPiF32;
}

// Cosine of the central angle, as latlon_dist_r8 forms it
// (pro_info.F90:157): cos(x1)*cos(x2)*cos(y1-y2) + sin(x1)*sin(x2).
__device__ __forceinline__ double central_cos(double cx1, double sx1, double y1r,
                                              double cx2, double sx2, double y2r) {
#pragma clang fp contract(off)
  double v = cx1 * cx2 * cos(y1r - y2r) + sx1 * sx2;
  return v;
}

__device__ __forceinline__ double dist_from_cos(double v) {
#pragma clang fp contract(off)
  if (!(v < 1.0)) v = 1.0;
  return kEarthA * acos(v);
}

// Wave-wide helpers ---------------------------------------------------------
__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ double shfl_xor_d(double v, int mask) {
  return __shfl_xor(v, mask, kWave);
}

// FP64 MFMA 16x16x4: D(16x16) += A(16x4) * B(4x16).
// Operand fragments: lane l holds A[l%16][l/16] and B[l/16][l%16]; the result
// holds D[4*r + l/16][l%16] in element r (rows interleaved across the four
// 16-lane groups; measured on gfx950).
typedef double double4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ double4_t mfma_f64_16x16x4(double a, double b, double4_t c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// v of the lane whose byte address (4 * lane) is addr
__device__ __forceinline__ double bperm_d(int addr, double v) {
  const int lo = __builtin_amdgcn_ds_bpermute(addr, __double2loint(v));
  const int hi = __builtin_amdgcn_ds_bpermute(addr, __double2hiint(v));
  return __hiloint2double(hi, lo);
}

// 1 / sqrt(x) for positive normal x: the hardware estimate (relative error
// about 2^-23) refined by NR Newton steps, y (3 - x y^2) / 2: two reach a
// few ulp, one about 2^-44.
template <int NR = 2>
__device__ __forceinline__ double rsq_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double hx = 0.5 * x;
#pragma unroll
  for (int i = 0; i < NR; ++i) y = y * fma(-hx * y, y, 1.5);
  return y;
}

// Scaled one-sided Jacobi rotation of a column pair (p, q) stored as
// sp x_p, sq x_q, from their squared norms al, be and gs = x_p . x_q:
// t = tan(theta), gm = gamma = gs sp sq, cs = cos(theta), ic = 1/cos(theta).
// A pair is rotated while gamma^2 > tol2 al be (t = 0, cs = ic = 1 below);
// *big records gamma^2 > tolp al be (the next sweep would still rotate).
// t = sign(d e) |e| / (|d| + sqrt(d^2 + e^2)), d = be - al, e = 2 gamma
// (for d = 0 either sign zeroes gamma), from the hardware estimates of
// 1/sqrt and 1/x: an error of t (about 2^-22 relative) only leaves a rotated
// pair that far from orthogonal, which the next sweep removes, while cs keeps
// every applied rotation orthogonal to the accuracy of its NR Newton steps
// (two: a few ulp; one, for the split half-precision Grams whose own error
// is some 1e-6: about 2^-44).
template <int NR = 2>
__device__ __forceinline__ void rotation_angle(double gs, double al, double be, double sp,
                                               double sq, double tol2, double tolp,
                                               bool& rotated, bool& big, double& t, double& gm,
                                               double& cs, double& ic) {
  gm = gs * (sp * sq);
  const double gg = gm * gm, ab = al * be;
  const bool need = gg > tol2 * ab;
  big |= gg > tolp * ab;
  rotated |= need;
  const double d = be - al, e = 2.0 * gm;
  const double r2 = fma(d, d, e * e);
  const double den = fma(r2, __builtin_amdgcn_rsq(r2), fabs(d));
  t = copysign(fabs(e) * __builtin_amdgcn_rcp(den), d * e);
  t = need ? t : 0.0;
  const double u = fma(t, t, 1.0);
  cs = rsq_nr<NR>(u);
  ic = u * cs;
  cs = need ? cs : 1.0;
  ic = need ? ic : 1.0;
}

// Half-precision MFMA 16x16x32 (FP32 accumulate) of a Gram precision:
// lane l holds A[l%16][8(l/16) + e] and B[8(l/16) + e][l%16] in element e
// of its 8-vector operands; the result holds D[4(l/16) + r][l%16] in
// element r (measured on gfx950).
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef _Float16 f16x2_t __attribute__((ext_vector_type(2)));

template <int GRAM>
struct H16;
template <>
struct H16<DK_GRAM_BF16X3> {
  typedef bf16x8_t vec;
  typedef bf16x2_t vec2;  // two elements converted by one instruction
  typedef __bf16 elt;
  static __device__ __forceinline__ f32x4_t mfma(vec a, vec b, f32x4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
};
template <>
struct H16<DK_GRAM_FP16X3> {
  typedef f16x8_t vec;
  typedef f16x2_t vec2;
  typedef _Float16 elt;
  static __device__ __forceinline__ f32x4_t mfma(vec a, vec b, f32x4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
  }
};
template <>
struct H16<DK_GRAM_BF16> : H16<DK_GRAM_BF16X3> {};

// Per-phase clock accounting (wall_clock64), one thread per workgroup.
struct PhaseClock {
  unsigned long long* out;
  unsigned long long t;
  __device__ explicit PhaseClock(unsigned long long* o) : out(threadIdx.x == 0 ? o : nullptr) {
    if (out) t = wall_clock64();
  }
  __device__ void mark(int phase) {
    if (out) {
      const unsigned long long n = wall_clock64();
      atomicAdd(out + phase, n - t);
      t = n;
    }
  }
};

}  // namespace dida
