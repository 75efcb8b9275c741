// Set-up producers on the device: the static tensors of the phaseless step,
// built once per Hamiltonian and trial (ipie builds them on the host):
//
//   setup_gemm_kernel    C = alpha op(A) op(B) + beta C in FP64 for operands of
//                        any strides, 64 x 64 tiles through LDS, optional
//                        split-K into partial planes summed in a fixed order
//                        by setup_reduce_kernel; the store can remap columns
//                        (the half-rotated Cholesky layout [X, n N])
//   pack_kernel          chol_packed[p, x] = L[i_p, j_p, x] over the upper
//                        triangle in numpy.triu_indices order
//   and small elementwise kernels (fill, identity add, scale + shift, norm).
//
// Restates ipie/hamiltonians/generic.py:31-43 (h1e_mod), 64-72 with
// utils/pack_numba.py:36-45 (packing), trial_wavefunction/half_rotate.py:
// 11-116 (rchol, rH1), propagation/phaseless_base.py:153-172 (mean-field
// shift) and 27-74 (one-body propagator, whose scipy.linalg.expm is replaced
// by scaling and squaring of a degree-18 Taylor polynomial: Moler and Van
// Loan, "Nineteen dubious ways to compute the exponential of a matrix",
// method 3, with ||A / 2^s||_1 <= 1/2).
#include "afqmc_internal.hpp"

namespace afqmc {

namespace {

constexpr int SG_TILE = 64, SG_BK = 16, SG_THREADS = 256;

__device__ __forceinline__ int64_t setup_store_index(const SetupGemmArgs& g, int64_t i, int64_t j) {
  if (g.store == SETUP_STORE_HALF_ROTATED) {  // column j = m X + x -> rchol[x][i N + m]
    const int64_t m = j / g.hr_x, x = j - m * g.hr_x;
    return x * g.hr_nN + i * g.hr_N + m;
  }
  return i * g.scr + j * g.scc;
}

__device__ __forceinline__ void setup_store(const SetupGemmArgs& g, int64_t i, int64_t j, double v) {
  const int64_t idx = setup_store_index(g, i, j);
  double out = g.alpha * v;
  if (g.beta != 0.0) out += g.beta * g.c[idx];
  g.c[idx] = out;
}

__global__ void __launch_bounds__(SG_THREADS) setup_gemm_kernel(SetupGemmArgs g) {
  __shared__ double As[SG_BK][SG_TILE + 1];
  __shared__ double Bs[SG_BK][SG_TILE + 1];
  const int t = threadIdx.x, tx = t % 16, ty = t / 16;
  const int64_t i0 = (int64_t)blockIdx.y * SG_TILE, j0 = (int64_t)blockIdx.x * SG_TILE;
  const int64_t kb = (int64_t)blockIdx.z * g.kslice;
  const int64_t ke = kb + g.kslice < g.k ? kb + g.kslice : g.k;
  const bool a_kfast = g.sac == 1, b_jfast = g.sbc == 1;
  double acc[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[r][q] = 0.0;
  for (int64_t k0 = kb; k0 < ke; k0 += SG_BK) {
    // 64 x 16 tiles of A and 16 x 64 of B, 4 elements per thread; the
    // thread order follows the operand's contiguous index
    for (int e = t; e < SG_TILE * SG_BK; e += SG_THREADS) {
      const int ii = a_kfast ? e / SG_BK : e % SG_TILE;
      const int kk = a_kfast ? e % SG_BK : e / SG_TILE;
      const int64_t i = i0 + ii, k = k0 + kk;
      As[kk][ii] = (i < g.m && k < ke) ? g.a[i * g.sar + k * g.sac] : 0.0;
      const int jj = b_jfast ? e % SG_TILE : e / SG_BK;
      const int kb2 = b_jfast ? e / SG_TILE : e % SG_BK;
      const int64_t j = j0 + jj, k2 = k0 + kb2;
      Bs[kb2][jj] = (j < g.n && k2 < ke) ? g.b[k2 * g.sbr + j * g.sbc] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < SG_BK; ++kk) {
      double av[4], bv[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) av[r] = As[kk][ty + 16 * r];
#pragma unroll
      for (int q = 0; q < 4; ++q) bv[q] = Bs[kk][tx + 16 * q];
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[r][q] = fma(av[r], bv[q], acc[r][q]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int64_t i = i0 + ty + 16 * r, j = j0 + tx + 16 * q;
      if (i >= g.m || j >= g.n) continue;
      if (g.part) g.part[((int64_t)blockIdx.z * g.m + i) * g.n + j] = acc[r][q];
      else setup_store(g, i, j, acc[r][q]);
    }
}

// sum of the split-K planes in slice order, then the epilogue
__global__ void setup_reduce_kernel(SetupGemmArgs g, int nz) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= g.m * g.n) return;
  double s = 0.0;
  for (int z = 0; z < nz; ++z) s += g.part[(int64_t)z * g.m * g.n + e];
  setup_store(g, e / g.n, e % g.n, s);
}

// chol_packed[p, x] = chol[(i_p N + j_p) X + x], (i_p, j_p) the p-th entry of
// the upper triangle in row-major order
__global__ void pack_kernel(int N, int X, const double* __restrict__ chol, double* __restrict__ packed) {
  const int64_t P = (int64_t)N * (N + 1) / 2;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= P * X) return;
  const int64_t p = e / X, x = e - p * X;
  // row i starts at off(i) = i N - i (i - 1) / 2: the last row with off(i) <= p
  int lo = 0, hi = N - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    const int64_t off = (int64_t)mid * N - (int64_t)mid * (mid - 1) / 2;
    if (off <= p) lo = mid;
    else hi = mid - 1;
  }
  const int64_t i = lo, off = i * N - i * (i - 1) / 2;
  const int64_t j = i + (p - off);
  packed[e] = chol[(i * N + j) * X + x];
}

__global__ void fill_kernel(double* out, int64_t n, double v) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) out[e] = v;
}

// out = scale (h + v)
__global__ void scale_shift_kernel(int64_t n, const double* h, const double* v, double scale, double* out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) out[e] = scale * (h[e] + v[e]);
}

// out = alpha x + beta y (out may alias x or y: each element is read before it is written)
__global__ void axpby_kernel(int64_t n, double alpha, const double* x, double beta, const double* y, double* out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) out[e] = alpha * x[e] + beta * y[e];
}

// a[i][i] += 1
__global__ void add_identity_kernel(int N, double* a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) a[(int64_t)i * N + i] += 1.0;
}

// max over columns of sum_i |a[i][j]| (one workgroup, fixed order)
__global__ void __launch_bounds__(256) norm1_kernel(int N, const double* a, double* out) {
  __shared__ double red[256];
  double m = 0.0;
  for (int j = threadIdx.x; j < N; j += 256) {
    double s = 0.0;
    for (int i = 0; i < N; ++i) s += fabs(a[(int64_t)i * N + j]);
    m = fmax(m, s);
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + h]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

unsigned blocks(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

}  // namespace

int64_t setup_gemm_part_doubles(int64_t m, int64_t n, int nz) { return nz > 1 ? (int64_t)nz * m * n : 0; }

hipError_t launch_setup_gemm(SetupGemmArgs g, int nz, hipStream_t stream) {
  if (g.m < 1 || g.n < 1 || g.k < 1 || nz < 1) return hipErrorInvalidValue;
  if (nz > 1 && !g.part) return hipErrorInvalidValue;
  const int64_t gy = (g.m + SG_TILE - 1) / SG_TILE, gx = (g.n + SG_TILE - 1) / SG_TILE;
  if (gy > 65535 || gx > 0x7fffffff || nz > 65535) return hipErrorInvalidValue;
  g.kslice = (g.k + nz - 1) / nz;
  g.kslice = (g.kslice + SG_BK - 1) / SG_BK * SG_BK;
  const int nzr = (int)((g.k + g.kslice - 1) / g.kslice);  // slices that hold K
  if (nzr == 1) {
    SetupGemmArgs d = g;
    d.part = nullptr;
    hipLaunchKernelGGL(setup_gemm_kernel, dim3((unsigned)gx, (unsigned)gy, 1), dim3(SG_THREADS), 0, stream, d);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(setup_gemm_kernel, dim3((unsigned)gx, (unsigned)gy, (unsigned)nzr), dim3(SG_THREADS), 0,
                     stream, g);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(setup_reduce_kernel, dim3(blocks(g.m * g.n, 256)), dim3(256), 0, stream, g, nzr);
  return hipGetLastError();
}

hipError_t launch_setup_pack(int N, int X, const double* chol, double* packed, hipStream_t stream) {
  const int64_t tot = (int64_t)N * (N + 1) / 2 * X;
  hipLaunchKernelGGL(pack_kernel, dim3(blocks(tot, 256)), dim3(256), 0, stream, N, X, chol, packed);
  return hipGetLastError();
}

hipError_t launch_setup_fill(double* out, int64_t n, double v, hipStream_t stream) {
  hipLaunchKernelGGL(fill_kernel, dim3(blocks(n, 256)), dim3(256), 0, stream, out, n, v);
  return hipGetLastError();
}

hipError_t launch_setup_scale_shift(int64_t n, const double* h, const double* v, double scale, double* out,
                                    hipStream_t stream) {
  hipLaunchKernelGGL(scale_shift_kernel, dim3(blocks(n, 256)), dim3(256), 0, stream, n, h, v, scale, out);
  return hipGetLastError();
}

hipError_t launch_setup_axpby(int64_t n, double alpha, const double* x, double beta, const double* y, double* out,
                              hipStream_t stream) {
  hipLaunchKernelGGL(axpby_kernel, dim3(blocks(n, 256)), dim3(256), 0, stream, n, alpha, x, beta, y, out);
  return hipGetLastError();
}

hipError_t launch_setup_add_identity(int N, double* a, hipStream_t stream) {
  hipLaunchKernelGGL(add_identity_kernel, dim3(blocks(N, 256)), dim3(256), 0, stream, N, a);
  return hipGetLastError();
}

hipError_t launch_setup_norm1(int N, const double* a, double* out, hipStream_t stream) {
  hipLaunchKernelGGL(norm1_kernel, dim3(1), dim3(256), 0, stream, N, a, out);
  return hipGetLastError();
}

}  // namespace afqmc
