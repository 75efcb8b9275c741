// Shared device helpers for the gfx950 phaseless-AFQMC kernels.
//
// Complex numbers are interleaved (re, im) doubles, the layout of
// torch.complex128 / numpy.complex128.  All MFMA work uses the CDNA3/CDNA4
// 16x16 fragment convention:
//   A operand (16 x 4):  lane l supplies A[l % 16][k(l / 16)]
//   B operand (4 x 16):  lane l supplies B[k(l / 16)][l % 16]
//   accumulator (16x16): lane l holds D[4 * r + l / 16][l % 16], r = 0..3
//   (register r of lane group q holds row 4r + q; measured on gfx950)
// The FP32-output instructions (v_mfma_f32_16x16x4f32, _16x16x32_bf16) hold
// D[4 * (l / 16) + r][l % 16] instead; the bf16 one takes 8 k values per lane:
// A[l % 16][8 (l / 16) + j], B[8 (l / 16) + j][l % 16], j = 0..7.
// The k index of a lane group only has to agree between A and B, so loaders
// give lane group q the contiguous k range [4q, 4q + 4) of a 16-wide K chunk
// and issue the four MFMAs of that chunk with s = 0..3 (k = 4q + s).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace afqmc {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

struct cplx {
  double re, im;
};

__device__ __forceinline__ cplx cmake(double r, double i) { return cplx{r, i}; }
__device__ __forceinline__ cplx cadd(cplx a, cplx b) { return cplx{a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cplx csub(cplx a, cplx b) { return cplx{a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ cplx cmul(cplx a, cplx b) {
  return cplx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ cplx cconj(cplx a) { return cplx{a.re, -a.im}; }
__device__ __forceinline__ cplx cscale(cplx a, double s) { return cplx{a.re * s, a.im * s}; }
// a += b * c
__device__ __forceinline__ void cfma(cplx& a, cplx b, cplx c) {
  a.re = fma(b.re, c.re, a.re);
  a.re = fma(-b.im, c.im, a.re);
  a.im = fma(b.re, c.im, a.im);
  a.im = fma(b.im, c.re, a.im);
}
__device__ __forceinline__ double cabs(cplx a) { return hypot(a.re, a.im); }

// Split-BF16 pair of an FP64 value: hi = bf16(x) and lo = bf16(x - hi), each
// rounded to nearest even from FP32 (|x - hi - lo| <= 2^-17 |x|).
__device__ __forceinline__ uint16_t bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ void split_bf16_pair(double x, uint16_t& hi, uint16_t& lo) {
  hi = bf16_rne((float)x);
  lo = bf16_rne((float)(x - (double)__uint_as_float(((uint32_t)hi) << 16)));
}
__device__ __forceinline__ cplx cdiv(cplx a, cplx b) {
  // Smith's algorithm (no overflow for moderate exponents).
  if (fabs(b.re) >= fabs(b.im)) {
    double r = b.im / b.re, d = b.re + b.im * r;
    return cplx{(a.re + a.im * r) / d, (a.im - a.re * r) / d};
  } else {
    double r = b.re / b.im, d = b.re * r + b.im;
    return cplx{(a.re * r + a.im) / d, (a.im * r - a.re) / d};
  }
}
__device__ __forceinline__ cplx cexp(cplx a) {
  double e = exp(a.re);
  double s, c;
  sincos(a.im, &s, &c);
  return cplx{e * c, e * s};
}
__device__ __forceinline__ cplx clog(cplx a) { return cplx{log(cabs(a)), atan2(a.im, a.re)}; }

__device__ __forceinline__ cplx ld(const double* p, int64_t i) {
  const double2 v = reinterpret_cast<const double2*>(p)[i];
  return cplx{v.x, v.y};
}
__device__ __forceinline__ void st(double* p, int64_t i, cplx v) {
  reinterpret_cast<double2*>(p)[i] = make_double2(v.re, v.im);
}

__device__ __forceinline__ v4d mfma64(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ v4f mfma32(float a, float b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ v4d zero4d() { return v4d{0.0, 0.0, 0.0, 0.0}; }
__device__ __forceinline__ v4f zero4f() { return v4f{0.0f, 0.0f, 0.0f, 0.0f}; }

// Wave-level sum of a double over all 64 lanes (result valid in every lane).
__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block-level sum; scratch must hold blockDim.x / 64 doubles.  Deterministic:
// the lane and wave order of the additions is fixed.
__device__ __forceinline__ double block_sum(double v, double* scratch) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if (lane == 0) scratch[wid] = v;
  __syncthreads();
  double t = 0.0;
  for (int i = 0; i < nw; ++i) t += scratch[i];
  return t;
}

__host__ __device__ __forceinline__ int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace afqmc
