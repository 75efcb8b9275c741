// Mixed-precision exchange: split-BF16 products with FP32 accumulation.
//
// Every FP64 operand x is split into x_hi = bf16(x), x_lo = bf16(x - x_hi);
// T = A B^T is formed as A_hi B_hi + A_hi B_lo + A_lo B_hi with
// v_mfma_f32_16x16x32_bf16 (FP32 accumulation), which keeps about 16
// significant bits per product (relative error ~1e-5 per T element, ~1e-7 on
// the exchange energy).  The per-walker trace sum_ij T_ij T_ji is accumulated
// in FP64.  Same contraction and epilogue as the FP64 kernel
// (ipie/estimators/local_energy_sd.py:263-303, 1040-1071).
//
// Operand planes are bf16 [rows, Kp] with Kp = N rounded up to 32 and zero
// padding, so a 32-wide K chunk of a row is one aligned 64-byte segment:
//   Ghalf: 4 planes (re_hi, re_lo, im_hi, im_lo), rows = (walker, j)
//   rchol: 2 planes (hi, lo), rows = (x, i)
// Workgroup: 4 waves in a 2 x 2 grid, each wave owns (RT/2) x (CT/2) output
// tiles of 16 x 16; double-buffered LDS staging; one barrier per K chunk.
// Workgroups that run together on one XCD cover 4 walker blocks x 8 Cholesky
// ranges, so both operand tiles are shared through that XCD's L2.
#include "afqmc_internal.hpp"

namespace afqmc {

namespace {

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

constexpr int BX_THREADS = 256;
constexpr int BX_KC = 32;

__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }

__device__ __forceinline__ void split_bf16(double x, uint16_t& hi, uint16_t& lo) {
  hi = f2bf((float)x);
  lo = f2bf((float)(x - (double)bf2f(hi)));
}

__device__ __forceinline__ v8bf as_v8bf(uint4 u) {
  v8bf r;
  __builtin_memcpy(&r, &u, 16);
  return r;
}

__device__ __forceinline__ v4f mfma_bf16(v8bf a, v8bf b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Ghalf [W*n, N] complex -> 4 bf16 planes [rows, Kp]; with walker_of, plane
// row t*n + j takes Ghalf row walker_of[t]*n + j (sampled walkers, compacted).
__global__ void split_ghalf_kernel(const double* __restrict__ g, int64_t rows, int N, int Kp,
                                   uint16_t* __restrict__ planes, const int* __restrict__ walker_of, int nocc) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= rows * Kp) return;
  const int64_t r = idx / Kp;
  const int k = (int)(idx % Kp);
  uint16_t rh = 0, rl = 0, ih = 0, il = 0;
  if (k < N) {
    const int64_t src = walker_of ? (int64_t)walker_of[r / nocc] * nocc + r % nocc : r;
    const cplx v = ld(g, src * N + k);
    split_bf16(v.re, rh, rl);
    split_bf16(v.im, ih, il);
  }
  const int64_t plane = rows * Kp;
  planes[idx] = rh;
  planes[plane + idx] = rl;
  planes[2 * plane + idx] = ih;
  planes[3 * plane + idx] = il;
}

// rchol [rows, N] real -> 2 bf16 planes [rows, Kp].
__global__ void split_real_kernel(const double* __restrict__ x, int64_t rows, int N, int Kp,
                                  uint16_t* __restrict__ planes) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= rows * Kp) return;
  const int64_t r = idx / Kp;
  const int k = (int)(idx % Kp);
  uint16_t h = 0, l = 0;
  if (k < N) split_bf16(x[r * N + k], h, l);
  planes[idx] = h;
  planes[rows * Kp + idx] = l;
}

template <int RT, int CT>
__global__ void __launch_bounds__(BX_THREADS) exchange_bf16_kernel(ExchangeBf16Args a) {
  constexpr int R = RT * 16, C = CT * 16;
  constexpr int WR = RT / 2, WC = CT / 2;  // tiles per wave
  // staging (uint16 units): A planes [4][R][32], B planes [2][C][32]
  constexpr int A_ELEMS = 4 * R * BX_KC, B_ELEMS = 2 * C * BX_KC;
  constexpr int STAGE = A_ELEMS + B_ELEMS;
  constexpr int A_VEC = A_ELEMS / 8, B_VEC = B_ELEMS / 8;  // 16-byte vectors
  constexpr int A_PER = (A_VEC + BX_THREADS - 1) / BX_THREADS;
  constexpr int B_PER = (B_VEC + BX_THREADS - 1) / BX_THREADS;
  extern __shared__ uint4 lds_u4[];
  uint16_t* lds = reinterpret_cast<uint16_t*>(lds_u4);
  __shared__ double red[BX_THREADS / 64];
  // epilogue staging aliases the buffers: S[row][col] complex float, stride n + 1
  float2* S = reinterpret_cast<float2*>(lds_u4);

  const int n = a.nocc, Kp = a.kpad;
  const int rows_valid = a.wb * n, cols_valid = a.xb * n;
  const int nwb = (int)ceil_div(a.nwalkers, a.wb);
  const int lin = xcd_remap(blockIdx.x, gridDim.x);
  const int split = lin % a.xsplit, bw = lin / a.xsplit;
  const int nxb = (int)ceil_div(a.nchol, a.xb);
  const int xb_begin = (int)((int64_t)nxb * split / a.xsplit);
  const int xb_end = (int)((int64_t)nxb * (split + 1) / a.xsplit);
  if (bw >= nwb) return;

  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, l16 = lane & 15, q = lane >> 4;
  const int wr = wave & 1, wc = wave >> 1;
  const int w0 = bw * a.wb;
  const int64_t arow_base = (int64_t)w0 * n, arow_limit = (int64_t)a.nwalkers * n;
  const int64_t brow_limit = (int64_t)a.nchol * n;
  const int64_t aplane = arow_limit * Kp, bplane = brow_limit * Kp;
  const int nch = Kp / BX_KC;

  double wacc_r = 0.0, wacc_i = 0.0;
  uint4 pa[A_PER], pb[B_PER];

  for (int xblk = xb_begin; xblk < xb_end; ++xblk) {
    const int x0 = xblk * a.xb;
    const int64_t brow_base = (int64_t)x0 * n;
    auto fetch = [&](int c) {
      const int k0 = c * BX_KC;
#pragma unroll
      for (int i = 0; i < A_PER; ++i) {
        const int v = tid + i * BX_THREADS;  // vector index: plane, row, 4 vectors per row
        pa[i] = make_uint4(0, 0, 0, 0);
        if (v < A_VEC) {
          const int p = v / (R * 4), rem = v % (R * 4), row = rem / 4, seg = rem % 4;
          const int64_t ar = arow_base + row;
          if (row < rows_valid && ar < arow_limit)
            pa[i] = *reinterpret_cast<const uint4*>(a.gplanes + p * aplane + ar * Kp + k0 + seg * 8);
        }
      }
#pragma unroll
      for (int i = 0; i < B_PER; ++i) {
        const int v = tid + i * BX_THREADS;
        pb[i] = make_uint4(0, 0, 0, 0);
        if (v < B_VEC) {
          const int p = v / (C * 4), rem = v % (C * 4), col = rem / 4, seg = rem % 4;
          const int64_t br = brow_base + col;
          if (col < cols_valid && br < brow_limit)
            pb[i] = *reinterpret_cast<const uint4*>(a.lplanes + p * bplane + br * Kp + k0 + seg * 8);
        }
      }
    };
    auto stash = [&](int buf) {
      uint4* base = lds_u4 + buf * (STAGE / 8);
#pragma unroll
      for (int i = 0; i < A_PER; ++i) {
        const int v = tid + i * BX_THREADS;
        if (v < A_VEC) base[v] = pa[i];
      }
#pragma unroll
      for (int i = 0; i < B_PER; ++i) {
        const int v = tid + i * BX_THREADS;
        if (v < B_VEC) base[A_VEC + v] = pb[i];
      }
    };
    v4f re[WR][WC], im[WR][WC];
#pragma unroll
    for (int i = 0; i < WR; ++i)
#pragma unroll
      for (int j = 0; j < WC; ++j) {
        re[i][j] = zero4f();
        im[i][j] = zero4f();
      }
    fetch(0);
    __syncthreads();
    stash(0);
    __syncthreads();
    for (int c = 0; c < nch; ++c) {
      if (c + 1 < nch) fetch(c + 1);
      const uint16_t* As = lds + (c & 1) * STAGE;
      const uint16_t* Bs = As + A_ELEMS;
      // B fragments of the wave's WC column tiles (hi, lo)
      v8bf bh[WC], bl[WC];
#pragma unroll
      for (int j = 0; j < WC; ++j) {
        const int col = (wc * WC + j) * 16 + l16;
        bh[j] = as_v8bf(*reinterpret_cast<const uint4*>(Bs + (0 * C + col) * BX_KC + 8 * q));
        bl[j] = as_v8bf(*reinterpret_cast<const uint4*>(Bs + (1 * C + col) * BX_KC + 8 * q));
      }
#pragma unroll
      for (int i = 0; i < WR; ++i) {
        const int row = (wr * WR + i) * 16 + l16;
        const v8bf arh = as_v8bf(*reinterpret_cast<const uint4*>(As + (0 * R + row) * BX_KC + 8 * q));
        const v8bf arl = as_v8bf(*reinterpret_cast<const uint4*>(As + (1 * R + row) * BX_KC + 8 * q));
        const v8bf aih = as_v8bf(*reinterpret_cast<const uint4*>(As + (2 * R + row) * BX_KC + 8 * q));
        const v8bf ail = as_v8bf(*reinterpret_cast<const uint4*>(As + (3 * R + row) * BX_KC + 8 * q));
#pragma unroll
        for (int j = 0; j < WC; ++j) {
          re[i][j] = mfma_bf16(arh, bh[j], re[i][j]);
          re[i][j] = mfma_bf16(arh, bl[j], re[i][j]);
          re[i][j] = mfma_bf16(arl, bh[j], re[i][j]);
          im[i][j] = mfma_bf16(aih, bh[j], im[i][j]);
          im[i][j] = mfma_bf16(aih, bl[j], im[i][j]);
          im[i][j] = mfma_bf16(ail, bh[j], im[i][j]);
        }
      }
      if (c + 1 < nch) stash((c + 1) & 1);
      __syncthreads();
    }
    // epilogue, one Cholesky vector at a time (accumulator row = 4 q + r)
    for (int xl = 0; xl < a.xb; ++xl) {
      if (x0 + xl >= a.nchol) break;
      const int cbeg = xl * n;
#pragma unroll
      for (int j = 0; j < WC; ++j) {
        const int col = (wc * WC + j) * 16 + l16 - cbeg;
        if (col >= 0 && col < n) {
#pragma unroll
          for (int i = 0; i < WR; ++i)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int row = (wr * WR + i) * 16 + 4 * q + r;
              if (row < rows_valid) S[row * (n + 1) + col] = make_float2(re[i][j][r], im[i][j][r]);
            }
        }
      }
      __syncthreads();
      const int nn = n * n;
      for (int e = tid; e < a.wb * nn; e += BX_THREADS) {
        const int wl = e % a.wb, ij = e / a.wb, ii = ij / n, jj = ij % n;
        const float2 tij = S[(wl * n + jj) * (n + 1) + ii];
        const float2 tji = S[(wl * n + ii) * (n + 1) + jj];
        wacc_r += (double)tij.x * tji.x - (double)tij.y * tji.y;
        wacc_i += (double)tij.x * tji.y + (double)tij.y * tji.x;
      }
      __syncthreads();
    }
  }
  for (int wl = 0; wl < a.wb; ++wl) {
    const bool mine = (tid % a.wb) == wl;
    const double vr = block_sum(mine ? wacc_r : 0.0, red);
    const double vi = block_sum(mine ? wacc_i : 0.0, red);
    const int w = w0 + wl;
    if (tid == 0 && w < a.nwalkers) st(a.exx_part, (int64_t)split * a.nwalkers + w, cmake(vr, vi));
  }
}

template <int RT, int CT>
hipError_t launch_bf16(const ExchangeBf16Args& a, hipStream_t stream) {
  constexpr int STAGE_BYTES = 2 * (4 * RT * 16 * BX_KC + 2 * CT * 16 * BX_KC);
  const size_t epi = sizeof(float2) * (size_t)(a.wb * a.nocc) * (a.nocc + 1);
  const size_t lds = (size_t)2 * STAGE_BYTES > epi ? (size_t)2 * STAGE_BYTES : epi;
  hipError_t e = allow_lds((const void*)exchange_bf16_kernel<RT, CT>, lds);
  if (e != hipSuccess) return e;
  const int nwb = (int)ceil_div(a.nwalkers, a.wb);
  hipLaunchKernelGGL((exchange_bf16_kernel<RT, CT>), dim3(nwb * a.xsplit), dim3(BX_THREADS), lds, stream, a);
  return hipGetLastError();
}

}  // namespace

// Geometry for the split-BF16 exchange: wb walkers x n rows and xb vectors x n
// columns padded to an even number (<= 10) of 16-row tiles (every even pair
// is instantiated); wb a power of two.
bool bf16_exchange_plan(int nocc, int* wb, int* xb, int* rt, int* ct) {
  double best_score = -1.0, best_eff = 0.0;
  for (int w = 1; w <= 8; w *= 2) {
    const int R = (w * nocc + 15) / 16;
    if (R > 10) break;
    const int RT = R + (R & 1);
    for (int x = 1; x <= 32; ++x) {
      const int Cn = (x * nocc + 15) / 16;
      if (Cn > 10) break;
      const int CT = Cn + (Cn & 1);
      const double eff = (double)(w * nocc) * (x * nocc) / (256.0 * RT * CT);
      const double score = eff * 1000.0 + RT * CT;
      if (score > best_score) {
        best_score = score;
        best_eff = eff;
        *wb = w;
        *xb = x;
        *rt = RT;
        *ct = CT;
      }
    }
  }
  return best_score > 0.0 && best_eff > 0.7;
}

hipError_t launch_split_ghalf(const double* g, int64_t rows, int N, int Kp, uint16_t* planes, hipStream_t stream,
                              const int* walker_of, int nocc) {
  const int64_t total = rows * Kp;
  hipLaunchKernelGGL(split_ghalf_kernel, dim3((unsigned)ceil_div(total, 256)), dim3(256), 0, stream, g, rows, N, Kp,
                     planes, walker_of, nocc);
  return hipGetLastError();
}

hipError_t launch_split_real(const double* x, int64_t rows, int N, int Kp, uint16_t* planes, hipStream_t stream) {
  const int64_t total = rows * Kp;
  hipLaunchKernelGGL(split_real_kernel, dim3((unsigned)ceil_div(total, 256)), dim3(256), 0, stream, x, rows, N, Kp,
                     planes);
  return hipGetLastError();
}

hipError_t launch_exchange_bf16(const ExchangeBf16Args& a, hipStream_t stream) {
#define AFQMC_BX(rt, ct) \
  if (a.rt_ == rt && a.ct_ == ct) return launch_bf16<rt, ct>(a, stream);
#define AFQMC_BX_ROW(rt) AFQMC_BX(rt, 2) AFQMC_BX(rt, 4) AFQMC_BX(rt, 6) AFQMC_BX(rt, 8) AFQMC_BX(rt, 10)
  AFQMC_BX_ROW(2) AFQMC_BX_ROW(4) AFQMC_BX_ROW(6) AFQMC_BX_ROW(8) AFQMC_BX_ROW(10)
#undef AFQMC_BX_ROW
#undef AFQMC_BX
  return hipErrorInvalidValue;
}

}  // namespace afqmc
