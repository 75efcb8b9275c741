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

}  // namespace

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
