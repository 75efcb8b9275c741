// Exchange as a quadratic form in the walker's half-rotated Green's function.
//
// With T_x = L_x G^T (L_x the [n, N] half-rotated Cholesky vector x, G = Ghalf
// [n, N]),
//   sum_x tr(T_x T_x) = sum_{i,j,m,m'} G[i][m'] G[j][m] sum_x L_x[i][m] L_x[j][m']
//                     = g^T Wq g,   g = vec(G),  Wq[(i,m'),(j,m)] = sum_x L_x[i][m] L_x[j][m'],
// i.e. Wq = R^T R (R = rchol as [X, n N]) with every N x N block transposed; Wq is
// real symmetric and walker-independent.  Per step the exchange of all walkers
// is one GEMM Y = G Wq (M = W, N = K = n N) and a dot product per walker, about
// X / N times fewer flops than forming every T_x (and no n x n trace epilogue).
// Same value as ipie/estimators/local_energy_sd.py:263-303 (exx_kernel_batch_real_rchol).
// Wq's symmetry halves the BF16 form: with column tiles of QUAD_TILE, the planes
// hold W' = Wq on the diagonal tiles, 2 Wq above them and 0 below, and column
// tile b of Y = G W' sums K over the tiles up to b only (g^T W' g = g^T Wq g).
//   quad_w_kernel       Wq from rchol, FP64 accumulation, written as FP64 [nN][nN]
//                       or as the split-BF16 planes of W' [2][n Kp][n Kp] (zero padding)
//   quad_reduce_kernel  E_w = sum_z sum_a g_w[a] Y_z[w][a]  (FP64 form)
//   quad_sum_kernel     E_w = sum of the per-tile dot products written by the
//                       BF16 GEMM's epilogue
#include "afqmc_internal.hpp"

namespace afqmc {

namespace {

constexpr int QW_T = 64;   // output tile edge
constexpr int QW_XC = 16;  // Cholesky vectors staged per chunk

__device__ __forceinline__ uint16_t qf2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float qbf2f(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }

// One 64 x 64 tile of Wq for the orbital pair (i, j): rows a = (i, m'), columns
// b = (j, m); out[a][b] = sum_x R[x][j][m'] R[x][i][m].  Np is the column
// stride per orbital of the output (N for FP64, Kp for the padded planes).
__global__ void __launch_bounds__(256) quad_w_kernel(const double* __restrict__ R, int X, int n, int N, int Np,
                                                    double* __restrict__ w64, uint16_t* __restrict__ wplanes) {
  __shared__ double fa[QW_XC][QW_T];  // R[x][j][m' range]
  __shared__ double fb[QW_XC][QW_T];  // R[x][i][m range]
  const int mblocks = (Np + QW_T - 1) / QW_T;
  int blk = blockIdx.x;
  const int mb = blk % mblocks;
  blk /= mblocks;
  const int ma = blk % mblocks;
  blk /= mblocks;
  const int j = blk % n, i = blk / n;
  const int m0a = ma * QW_T, m0b = mb * QW_T;
  const int tid = threadIdx.x, ty = tid >> 4, tx = tid & 15;
  double acc[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[r][c] = 0.0;
  const int64_t nN = (int64_t)n * N;
  for (int x0 = 0; x0 < X; x0 += QW_XC) {
    __syncthreads();
    for (int e = tid; e < QW_XC * QW_T; e += 256) {
      const int xx = e / QW_T, k = e % QW_T, x = x0 + xx;
      const int mpa = m0a + k, mpb = m0b + k;
      fa[xx][k] = (x < X && mpa < N) ? R[x * nN + (int64_t)j * N + mpa] : 0.0;
      fb[xx][k] = (x < X && mpb < N) ? R[x * nN + (int64_t)i * N + mpb] : 0.0;
    }
    __syncthreads();
#pragma unroll 4
    for (int xx = 0; xx < QW_XC; ++xx) {
      double av[4], bv[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) av[r] = fa[xx][ty * 4 + r];
#pragma unroll
      for (int c = 0; c < 4; ++c) bv[c] = fb[xx][tx * 4 + c];
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[r][c] += av[r] * bv[c];
    }
  }
  const int64_t ldo = (int64_t)n * Np;  // row length of the output
  const int64_t plane = ldo * ldo;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int mp = m0a + ty * 4 + r;  // m'
    if (mp >= Np) continue;
    const int64_t a = (int64_t)i * Np + mp;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int m = m0b + tx * 4 + c;
      if (m >= Np) continue;
      const int64_t b = (int64_t)j * Np + m;
      if (w64) {
        w64[a * ldo + b] = acc[r][c];
      } else {
        // plane row a is the output column, b the K index: W'_{b a}
        const int ta = (int)(a / QUAD_TILE), tb = (int)(b / QUAD_TILE);
        const double f = tb < ta ? 2.0 : (tb == ta ? 1.0 : 0.0);
        const double v = (mp < N && m < N) ? f * acc[r][c] : 0.0;
        const uint16_t hi = qf2bf((float)v);
        const uint16_t lo = qf2bf((float)(v - (double)qbf2f(hi)));
        wplanes[a * ldo + b] = hi;
        wplanes[plane + a * ldo + b] = lo;
      }
    }
  }
}

// E_w = sum_z sum_a d_w[a] Y_z[w][a]; Y index a = i Np + m (m < N used), d = g
// from Ghalf [W, n, N] (or its sampled rows: walker_of), minus gref [n, N] when
// given (the control variate's dG).  One workgroup per walker, result into
// part[w] (complex).
__global__ void __launch_bounds__(256) quad_reduce_kernel(const double* __restrict__ y, int ksplit, int nwalkers,
                                                         int n, int N, int Np, const double* __restrict__ g,
                                                         const int* __restrict__ walker_of,
                                                         double* __restrict__ part, const double* __restrict__ gref) {
  __shared__ double red[256 / 64];
  const int w = blockIdx.x;
  const int64_t ldy = (int64_t)n * Np;
  const int src = walker_of ? walker_of[w] : w;
  const double* gw = g + 2 * (int64_t)src * n * N;
  double re = 0.0, im = 0.0;
  for (int z = 0; z < ksplit; ++z) {
    const double* yw = y + 2 * ((int64_t)z * nwalkers + w) * ldy;
    for (int64_t e = threadIdx.x; e < (int64_t)n * N; e += blockDim.x) {
      const int i = (int)(e / N), m = (int)(e % N);
      cplx gv = ld(gw, e);
      if (gref) gv = csub(gv, ld(gref, e));
      const cplx yv = ld(yw, (int64_t)i * Np + m);
      re += gv.re * yv.re - gv.im * yv.im;
      im += gv.re * yv.im + gv.im * yv.re;
    }
  }
  re = block_sum(re, red);
  im = block_sum(im, red);
  if (threadIdx.x == 0) st(part, w, cmake(re, im));
}

// out[w] = sum_s part[s][w]: WL walkers x 256 / WL slot lanes per workgroup, the
// slot lanes combined in a fixed order (few walkers, many slots: more lanes
// per walker keep the serial chain of each lane short).
template <int WL>
__global__ void __launch_bounds__(256) quad_sum_kernel(const double* __restrict__ part, int nslots, int nwalkers,
                                                      double* __restrict__ out) {
  constexpr int SL = 256 / WL;
  __shared__ double sr[SL][WL + 1], si[SL][WL + 1];
  const int wl = threadIdx.x % WL, sl = threadIdx.x / WL;
  const int w = blockIdx.x * WL + wl;
  double re = 0.0, im = 0.0;
  if (w < nwalkers) {
    for (int s = sl; s < nslots; s += SL) {
      const cplx v = ld(part, (int64_t)s * nwalkers + w);
      re += v.re;
      im += v.im;
    }
  }
  sr[sl][wl] = re;
  si[sl][wl] = im;
  __syncthreads();
  if (sl == 0 && w < nwalkers) {
    double tr = 0.0, ti = 0.0;
    for (int k = 0; k < SL; ++k) {
      tr += sr[k][wl];
      ti += si[k][wl];
    }
    st(out, w, cmake(tr, ti));
  }
}

// ---------------------------------------------------------------- FP8 operands
// FP8 (E4M3, round to nearest even) of four values, packed little-endian
__device__ __forceinline__ uint32_t pack_fp8x4(float v0, float v1, float v2, float v3) {
  int p = 0;
  p = __builtin_amdgcn_cvt_pk_fp8_f32(v0, v1, p, false);
  p = __builtin_amdgcn_cvt_pk_fp8_f32(v2, v3, p, true);
  return (uint32_t)p;
}

// power-of-two scale putting the largest magnitude at most at 224 (half the
// E4M3 maximum: the lo part of a split value and rounding stay in range)
__device__ __forceinline__ double fp8_scale(double amax) {
  return amax > 0.0 ? exp2(floor(log2(224.0 / amax))) : 1.0;
}

// max |x| over n BF16 values into *absmax (FP32 bits; non-negative floats
// order like their bit patterns, so the result does not depend on timing)
__global__ void __launch_bounds__(256) quad_absmax_kernel(const uint16_t* __restrict__ x, int64_t n,
                                                        unsigned int* __restrict__ absmax) {
  float m = 0.0f;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
    m = fmaxf(m, fabsf(qbf2f(x[e])));
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(absmax, __float_as_uint(m));
}

__global__ void quad_w8_scale_kernel(const unsigned int* __restrict__ absmax, float* __restrict__ scale) {
  if (threadIdx.x == 0) scale[0] = (float)fp8_scale((double)__uint_as_float(absmax[0]));
}

// w8[r][k], 4 consecutive k per thread, zero outside ldn x ldn
__global__ void __launch_bounds__(256) quad_w8_kernel(const uint16_t* __restrict__ wp, int64_t ldn, int64_t ld8,
                                                    uint8_t* __restrict__ w8, const float* __restrict__ scale) {
  const int64_t q4 = ld8 / 4;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= ld8 * q4) return;
  const int64_t r = idx / q4, k0 = (idx % q4) * 4;
  const float s = scale[0];
  float v[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int64_t k = k0 + j;
    v[j] = (r < ldn && k < ldn) ? (qbf2f(wp[r * ldn + k]) + qbf2f(wp[ldn * ldn + r * ldn + k])) * s : 0.0f;
  }
  *reinterpret_cast<uint32_t*>(w8 + r * ld8 + k0) = pack_fp8x4(v[0], v[1], v[2], v[3]);
}

// one workgroup per walker: the largest |re|, |im| of d_w, then its scaled FP8 planes
__global__ void __launch_bounds__(256) split_fp8_rows_kernel(const double* __restrict__ g,
                                                           const double* __restrict__ gref, int nwalkers, int n,
                                                           int N, int Kp, int64_t ld8, uint8_t* __restrict__ a8,
                                                           const float* __restrict__ wscale,
                                                           double* __restrict__ row_inv) {
  __shared__ double red[256 / 64];
  const int w = blockIdx.x, tid = threadIdx.x;
  const double* gw = g + 2 * (int64_t)w * n * N;
  double m = 0.0;
  for (int e = tid; e < n * N; e += 256) {
    cplx d = ld(gw, e);
    if (gref) d = csub(d, ld(gref, e));
    m = fmax(m, fmax(fabs(d.re), fabs(d.im)));
  }
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  if ((tid & 63) == 0) red[tid >> 6] = m;
  __syncthreads();
  m = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
  const double s = fp8_scale(m);
  uint8_t* pre = a8 + (int64_t)w * ld8;
  uint8_t* pim = a8 + ((int64_t)nwalkers + w) * ld8;
  for (int64_t k0 = (int64_t)tid * 4; k0 < ld8; k0 += 4 * 256) {
    float vr[4], vi[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      // a walker row holds n Kp < 2^31 values: 32-bit row and column
      const int k = (int)(k0 + j);
      const int i = k / Kp, mm = k - i * Kp;
      vr[j] = vi[j] = 0.0f;
      if (i < n && mm < N) {
        cplx d = ld(gw, (int64_t)i * N + mm);
        if (gref) d = csub(d, ld(gref, (int64_t)i * N + mm));
        vr[j] = (float)(d.re * s);
        vi[j] = (float)(d.im * s);
      }
    }
    *reinterpret_cast<uint32_t*>(pre + k0) = pack_fp8x4(vr[0], vr[1], vr[2], vr[3]);
    *reinterpret_cast<uint32_t*>(pim + k0) = pack_fp8x4(vi[0], vi[1], vi[2], vi[3]);
  }
  if (tid == 0) row_inv[w] = 1.0 / (s * (double)wscale[0]);
}

// The same planes with the walker's dG held in registers between the maximum
// and the conversion (G read once): one walker per workgroup of 1024 threads,
// J chunks of 4 consecutive padded K positions per thread (ld8 <= 4096 J;
// Kp a multiple of 4, so a chunk never crosses a row of Kp).
template <int J>
__global__ void __launch_bounds__(1024) split_fp8_rows_reg_kernel(const double* __restrict__ g,
                                                                 const double* __restrict__ gref, int nwalkers,
                                                                 int n, int N, int Kp, int ld8,
                                                                 uint8_t* __restrict__ a8,
                                                                 const float* __restrict__ wscale,
                                                                 double* __restrict__ row_inv) {
  __shared__ double red[16];
  const int w = blockIdx.x, tid = threadIdx.x;
  const double* gw = g + 2 * (int64_t)w * n * N;
  cplx v[J][4];
  double m = 0.0;
#pragma unroll
  for (int j = 0; j < J; ++j) {
    const int k0 = 4 * (tid + 1024 * j);
    const int i = k0 / Kp, m0 = k0 - i * Kp;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      v[j][r] = cplx{0.0, 0.0};
      if (k0 < ld8 && i < n && m0 + r < N) {
        cplx d = ld(gw, i * N + m0 + r);
        if (gref) d = csub(d, ld(gref, i * N + m0 + r));
        v[j][r] = d;
        m = fmax(m, fmax(fabs(d.re), fabs(d.im)));
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  if ((tid & 63) == 0) red[tid >> 6] = m;
  __syncthreads();
#pragma unroll
  for (int t = 0; t < 16; ++t) m = fmax(m, red[t]);
  const double s = fp8_scale(m);
  uint8_t* pre = a8 + (int64_t)w * ld8;
  uint8_t* pim = a8 + ((int64_t)nwalkers + w) * ld8;
#pragma unroll
  for (int j = 0; j < J; ++j) {
    const int k0 = 4 * (tid + 1024 * j);
    if (k0 < ld8) {
      *reinterpret_cast<uint32_t*>(pre + k0) =
          pack_fp8x4((float)(v[j][0].re * s), (float)(v[j][1].re * s), (float)(v[j][2].re * s), (float)(v[j][3].re * s));
      *reinterpret_cast<uint32_t*>(pim + k0) =
          pack_fp8x4((float)(v[j][0].im * s), (float)(v[j][1].im * s), (float)(v[j][2].im * s), (float)(v[j][3].im * s));
    }
  }
  if (tid == 0) row_inv[w] = 1.0 / (s * (double)wscale[0]);
}

}  // namespace

hipError_t launch_quad_w8(const uint16_t* wplanes, int64_t ldn, int64_t ld8, uint8_t* w8, float* scale,
                          unsigned int* absmax, hipStream_t stream) {
  if (ldn < 1 || ld8 < ldn || ld8 % 64 != 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(quad_absmax_kernel, dim3(2048), dim3(256), 0, stream, wplanes, ldn * ldn, absmax);
  hipLaunchKernelGGL(quad_w8_scale_kernel, dim3(1), dim3(64), 0, stream, absmax, scale);
  const int64_t total = ld8 * (ld8 / 4);
  hipLaunchKernelGGL(quad_w8_kernel, dim3((unsigned)ceil_div(total, 256)), dim3(256), 0, stream, wplanes, ldn, ld8,
                     w8, scale);
  return hipGetLastError();
}

hipError_t launch_split_fp8_rows(const double* g, const double* gref, int nwalkers, int n, int N, int Kp,
                                 int64_t ld8, uint8_t* a8, const float* wscale, double* row_inv, hipStream_t stream) {
  if (nwalkers < 1 || n < 1 || Kp < N || ld8 < (int64_t)n * Kp || ld8 % 64 != 0 || ld8 > INT32_MAX)
    return hipErrorInvalidValue;
  if (Kp % 4 == 0 && ld8 <= 5 * 4096) {
    const int l = (int)ld8;
    if (ld8 <= 2 * 4096)
      hipLaunchKernelGGL(split_fp8_rows_reg_kernel<2>, dim3((unsigned)nwalkers), dim3(1024), 0, stream, g, gref,
                         nwalkers, n, N, Kp, l, a8, wscale, row_inv);
    else if (ld8 <= 3 * 4096)
      hipLaunchKernelGGL(split_fp8_rows_reg_kernel<3>, dim3((unsigned)nwalkers), dim3(1024), 0, stream, g, gref,
                         nwalkers, n, N, Kp, l, a8, wscale, row_inv);
    else
      hipLaunchKernelGGL(split_fp8_rows_reg_kernel<5>, dim3((unsigned)nwalkers), dim3(1024), 0, stream, g, gref,
                         nwalkers, n, N, Kp, l, a8, wscale, row_inv);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(split_fp8_rows_kernel, dim3((unsigned)nwalkers), dim3(256), 0, stream, g, gref, nwalkers, n, N,
                     Kp, ld8, a8, wscale, row_inv);
  return hipGetLastError();
}

hipError_t launch_quad_sum(const double* part, int nslots, int nwalkers, double* out, hipStream_t stream) {
  if (nslots < 1 || nwalkers < 1) return hipErrorInvalidValue;
  if (nwalkers >= 512)
    hipLaunchKernelGGL(quad_sum_kernel<16>, dim3((unsigned)ceil_div(nwalkers, 16)), dim3(256), 0, stream, part,
                       nslots, nwalkers, out);
  else
    hipLaunchKernelGGL(quad_sum_kernel<4>, dim3((unsigned)ceil_div(nwalkers, 4)), dim3(256), 0, stream, part,
                       nslots, nwalkers, out);
  return hipGetLastError();
}

hipError_t launch_quad_w(const double* rchol, int X, int n, int N, int Np, double* w64, uint16_t* wplanes,
                         hipStream_t stream) {
  if ((w64 == nullptr) == (wplanes == nullptr) || Np < N) return hipErrorInvalidValue;
  const int64_t mblocks = ceil_div(Np, QW_T);
  const int64_t grid = (int64_t)n * n * mblocks * mblocks;
  hipLaunchKernelGGL(quad_w_kernel, dim3((unsigned)grid), dim3(256), 0, stream, rchol, X, n, N, Np, w64, wplanes);
  return hipGetLastError();
}

hipError_t launch_quad_reduce(const double* y, int ksplit, int nwalkers, int n, int N, int Np, const double* g,
                              const int* walker_of, double* part, hipStream_t stream, const double* gref) {
  hipLaunchKernelGGL(quad_reduce_kernel, dim3(nwalkers), dim3(256), 0, stream, y, ksplit, nwalkers, n, N, Np, g,
                     walker_of, part, gref);
  return hipGetLastError();
}

}  // namespace afqmc
