// Shared device helpers for the MI355X (gfx950) QR / polar library.
//
// Layout conventions (the same as Julia's Matrix and LAPACK):
//   * matrices are column-major with an explicit leading dimension;
//   * ComplexF64 is stored interleaved (re, im), 16 bytes per element, which is
//     the layout of Julia's ComplexF64 and numpy's complex128.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mak {

constexpr int WAVE = 64;

// ---------------------------------------------------------------------------
// Scalar types: double and an interleaved complex double.
// ---------------------------------------------------------------------------
struct zd {
  double x, y;
};

__host__ __device__ inline zd mkz(double a, double b) { zd r; r.x = a; r.y = b; return r; }

template <typename T> struct ScalarTraits;
struct zd2 {
  zd a, b;
};
template <> struct ScalarTraits<double> {
  static constexpr bool is_complex = false;
  using vec2 = double2;  // two consecutive elements (16-byte loads)
  __host__ __device__ static double lo(const double2& v) { return v.x; }
  __host__ __device__ static double hi(const double2& v) { return v.y; }
};
template <> struct ScalarTraits<zd> {
  static constexpr bool is_complex = true;
  using vec2 = zd2;
  __host__ __device__ static zd lo(const zd2& v) { return v.a; }
  __host__ __device__ static zd hi(const zd2& v) { return v.b; }
};

// real arithmetic
__device__ inline double s_zero(double) { return 0.0; }
__device__ inline double s_one(double) { return 1.0; }
__device__ inline double s_conj(double a) { return a; }
__device__ inline double s_re(double a) { return a; }
__device__ inline double s_im(double) { return 0.0; }
__device__ inline double s_abs2(double a) { return a * a; }
__device__ inline double s_absv(double a) { return fabs(a); }
__device__ inline double s_add(double a, double b) { return a + b; }
__device__ inline double s_sub(double a, double b) { return a - b; }
__device__ inline double s_mul(double a, double b) { return a * b; }
__device__ inline double s_scale(double a, double r) { return a * r; }
// a*b + c
__device__ inline double s_fma(double a, double b, double c) { return fma(a, b, c); }
// conj(a)*b + c
__device__ inline double s_cfma(double a, double b, double c) { return fma(a, b, c); }
__device__ inline double s_div(double a, double b) { return a / b; }
__device__ inline double s_mkreal(double r) { return r; }
__device__ inline double s_neg(double a) { return -a; }

// complex arithmetic
__device__ inline zd s_zero(zd) { return mkz(0.0, 0.0); }
__device__ inline zd s_one(zd) { return mkz(1.0, 0.0); }
__device__ inline zd s_conj(zd a) { return mkz(a.x, -a.y); }
__device__ inline double s_re(zd a) { return a.x; }
__device__ inline double s_im(zd a) { return a.y; }
__device__ inline double s_abs2(zd a) { return a.x * a.x + a.y * a.y; }
__device__ inline double s_absv(zd a) { return hypot(a.x, a.y); }
__device__ inline zd s_add(zd a, zd b) { return mkz(a.x + b.x, a.y + b.y); }
__device__ inline zd s_sub(zd a, zd b) { return mkz(a.x - b.x, a.y - b.y); }
__device__ inline zd s_mul(zd a, zd b) {
  return mkz(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}
__device__ inline zd s_scale(zd a, double r) { return mkz(a.x * r, a.y * r); }
__device__ inline zd s_fma(zd a, zd b, zd c) {
  return mkz(fma(a.x, b.x, fma(-a.y, b.y, c.x)), fma(a.x, b.y, fma(a.y, b.x, c.y)));
}
__device__ inline zd s_cfma(zd a, zd b, zd c) {  // conj(a)*b + c
  return mkz(fma(a.x, b.x, fma(a.y, b.y, c.x)), fma(a.x, b.y, fma(-a.y, b.x, c.y)));
}
__device__ inline zd s_div(zd a, zd b) {
  // Smith's algorithm for a / b
  if (fabs(b.x) >= fabs(b.y)) {
    double r = b.y / b.x, d = b.x + b.y * r;
    return mkz((a.x + a.y * r) / d, (a.y - a.x * r) / d);
  } else {
    double r = b.x / b.y, d = b.x * r + b.y;
    return mkz((a.x * r + a.y) / d, (a.y * r - a.x) / d);
  }
}
__device__ inline zd s_mkreal(zd, double r) { return mkz(r, 0.0); }
__device__ inline double s_mkreal(double, double r) { return r; }
__device__ inline zd s_neg(zd a) { return mkz(-a.x, -a.y); }

// 1 / sqrt(x) and 1 / x for finite x > 0 (x != 0) to about one ulp: the
// hardware estimates (v_rsq_f64, v_rcp_f64, relative error about 2^-23)
// refined by two Newton-Raphson steps each, with the residual formed by fma:
//   rsqrt: e = 1 - x y^2, y <- y + (y / 2) e;   recip: e = 1 - x y, y <- y + y e.
// The error roughly squares per step.  They replace IEEE sqrt / division on
// the serial pivot chain of the in-wave factorisations, where latency counts.
__device__ inline double rsqrt_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const double e = fma(-x * y, y, 1.0);
    y = fma(0.5 * y, e, y);
  }
  return y;
}
__device__ inline double recip_nr(double x) {
  double y = __builtin_amdgcn_rcp(x);
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const double e = fma(-x, y, 1.0);
    y = fma(y, e, y);
  }
  return y;
}
__device__ inline double s_recip(double a) { return recip_nr(a); }
__device__ inline zd s_recip(zd a) {  // conj(a) / |a|^2
  const double r = recip_nr(fma(a.x, a.x, a.y * a.y));
  return mkz(a.x * r, -a.y * r);
}

template <typename T> __device__ inline T s_from_real(double r);
template <> __device__ inline double s_from_real<double>(double r) { return r; }
template <> __device__ inline zd s_from_real<zd>(double r) { return mkz(r, 0.0); }

// branch-free value selects (keep private arrays in registers)
__device__ inline double sel(bool c, double a, double b) { return c ? a : b; }
__device__ inline zd sel(bool c, zd a, zd b) { return mkz(c ? a.x : b.x, c ? a.y : b.y); }

// sign_safe (MatrixAlgebraKit src/common/safemethods.jl:10-11): real -> +-1
// with 0 -> +1; complex -> z/|z| with 0 -> 1.
__device__ inline double sign_safe(double a) { return a < 0.0 ? -1.0 : 1.0; }
__device__ inline zd sign_safe(zd a) {
  double r = hypot(a.x, a.y);
  if (r == 0.0) return mkz(1.0, 0.0);
  return mkz(a.x / r, a.y / r);
}

// ---------------------------------------------------------------------------
// Cross-lane helpers (wave64)
// ---------------------------------------------------------------------------
__device__ inline double shfl_xor_d(double v, int mask) { return __shfl_xor(v, mask, WAVE); }
__device__ inline double shfl_d(double v, int src) { return __shfl(v, src, WAVE); }
__device__ inline zd shfl_xor_z(zd v, int mask) {
  return mkz(__shfl_xor(v.x, mask, WAVE), __shfl_xor(v.y, mask, WAVE));
}
__device__ inline double shfl_xor_s(double v, int mask) { return shfl_xor_d(v, mask); }
__device__ inline zd shfl_xor_s(zd v, int mask) { return shfl_xor_z(v, mask); }

template <typename T> __device__ inline T wave_sum(T v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = s_add(v, shfl_xor_s(v, m));
  return v;
}

__device__ inline int lane_id() { return threadIdx.x & (WAVE - 1); }

// ---------------------------------------------------------------------------
// MFMA wrappers (gfx950).  Fragment layouts for the 16x16 shapes used here
// (measured on gfx950 with a one-wave probe):
//   A (16 x K): lane l holds row (l % 16), k-group (l / 16)
//   B (K x 16): lane l holds col (l % 16), k-group (l / 16)
//   D (16 x 16), f32 / bf16: lane l, register r holds row 4 * (l / 16) + r
//   D (16 x 16), f64:        lane l, register r holds row 4 * r + (l / 16)
//   and column l % 16 in both cases.
// For f64/f32 16x16x4 each k-group is one k; for bf16 16x16x32 each k-group is
// 8 consecutive k.
// ---------------------------------------------------------------------------
typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

__device__ inline v4d mfma_f64(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ inline v4f mfma_f32(float a, float b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ inline v4f mfma_bf16(v8bf a, v8bf b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// row of result register r held by lane l
__device__ inline int mfma_f64_row(int lane, int r) { return 4 * r + (lane >> 4); }
__device__ inline int mfma_f32_row(int lane, int r) { return 4 * (lane >> 4) + r; }

// Make LDS writes of this wave visible to its own later LDS reads (LDS ops of
// one wave execute in order; this only stops the compiler from reordering).
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__host__ __device__ inline int cdiv(int a, int b) { return (a + b - 1) / b; }
__host__ __device__ inline int64_t cdiv64(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace mak
