// Fused exchange energy (FP64):
//   exx[w] = sum_x sum_{ij} T[x,w][i,j] T[x,w][j,i],
//   T[x,w][i,j] = sum_m rchol[x, i, m] Ghalf[w, j, m]
// Restates ipie/estimators/local_energy_sd.py:263-303 (CPU) and the GPU path
// local_energy_sd.py:1040-1071 + estimators/kernels/gpu/exchange.py:64-96,
// without ever storing T: the [W*n, X*n] product is formed tile by tile with
// v_mfma_f64_16x16x4f64 and reduced in the epilogue.
//
// Workgroup tile: rows = (walker, j) for wb walkers, cols = (x, i) for xb
// Cholesky vectors; both are flattened and padded to multiples of 16.  K = N
// (the basis index) is streamed through LDS 16 at a time; the 16x16 output
// tiles are dealt round-robin to the 4 waves.  After the K loop the tile is
// staged one Cholesky vector at a time into LDS, where sum_ij T_ij T_ji is
// accumulated per walker.  Each workgroup walks a contiguous range of x
// blocks and writes one partial sum per walker; partials are added in a
// fixed order later, so the result is reproducible.
#include "afqmc_internal.hpp"

namespace afqmc {

namespace {

constexpr int EX_THREADS = 256;
constexpr int EX_KC = 16;

struct ExPlan {
  int wb, xb, RT, CT;
};

__host__ __device__ inline int round16(int v) { return (v + 15) / 16 * 16; }

template <int MAXT>
__global__ void __launch_bounds__(EX_THREADS) exchange_kernel(ExchangeArgs a, ExPlan pl) {
  extern __shared__ double lds[];
  const int n = a.nocc, N = a.nbasis;
  const int R = pl.RT * 16, C = pl.CT * 16;
  // k-major staging strides padded so the 4 lane groups of a fragment read
  // fall in different LDS banks
  const int RS = R + 1, CS = C + 2;
  const int rows_valid = pl.wb * n, cols_valid = pl.xb * n;
  // staging buffers for one K chunk, k-major: A_s[k][row] complex, B_s[k][col] real
  cplx* A_s = reinterpret_cast<cplx*>(lds);
  double* B_s = lds + 2 * EX_KC * RS;
  // epilogue buffer aliases the staging area: S[row][i] complex, stride n + 1
  cplx* S = reinterpret_cast<cplx*>(lds);
  __shared__ double red[EX_THREADS / 64];

  const int nwb = (int)ceil_div(a.nwalkers, pl.wb);
  const int lin = xcd_remap(blockIdx.x, gridDim.x);
  const int bw = lin % nwb, split = lin / nwb;
  const int nxb = (int)ceil_div(a.nchol, pl.xb);
  const int xb_begin = (int)((int64_t)nxb * split / a.xsplit);
  const int xb_end = (int)((int64_t)nxb * (split + 1) / a.xsplit);

  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, l16 = lane & 15, q = lane >> 4;
  const int ntiles = pl.RT * pl.CT;
  const int w0 = bw * pl.wb;
  const int64_t arow_base = (int64_t)w0 * n;             // first flattened Ghalf row
  const int64_t arow_limit = (int64_t)a.nwalkers * n;

  // wb divides EX_THREADS, so thread tid always accumulates walker tid % wb
  double wacc_r = 0.0, wacc_i = 0.0;

  const int nchunks = (int)ceil_div(N, EX_KC);
  for (int xblk = xb_begin; xblk < xb_end; ++xblk) {
    const int x0 = xblk * pl.xb;
    const int64_t brow_base = (int64_t)x0 * n;
    const int64_t brow_limit = (int64_t)a.nchol * n;
    v4d re[MAXT], im[MAXT];
#pragma unroll
    for (int t = 0; t < MAXT; ++t) {
      re[t] = zero4d();
      im[t] = zero4d();
    }
    for (int c = 0; c < nchunks; ++c) {
      const int k0 = c * EX_KC;
      __syncthreads();
      // A chunk: R rows x 16 k (complex)
      for (int e = tid; e < R * EX_KC; e += EX_THREADS) {
        const int row = e / EX_KC, k = e % EX_KC;
        const int64_t ar = arow_base + row;
        cplx v = cmake(0.0, 0.0);
        if (row < rows_valid && ar < arow_limit && k0 + k < N) v = ld(a.ghalf, ar * N + k0 + k);
        A_s[k * RS + row] = v;
      }
      // B chunk: C rows x 16 k (real)
      for (int e = tid; e < C * EX_KC; e += EX_THREADS) {
        const int col = e / EX_KC, k = e % EX_KC;
        const int64_t br = brow_base + col;
        double v = 0.0;
        if (col < cols_valid && br < brow_limit && k0 + k < N) v = a.rchol[br * N + k0 + k];
        B_s[k * CS + col] = v;
      }
      __syncthreads();
#pragma unroll 1
      for (int s = 0; s < 4; ++s) {
        const int k = 4 * q + s;
#pragma unroll
        for (int m = 0; m < MAXT; ++m) {
          const int t = wave + 4 * m;
          if (t < ntiles) {
            const int rt = t / pl.CT, ct = t % pl.CT;
            const cplx av = A_s[k * RS + rt * 16 + l16];
            const double bv = B_s[k * CS + ct * 16 + l16];
            re[m] = mfma64(av.re, bv, re[m]);
            im[m] = mfma64(av.im, bv, im[m]);
          }
        }
      }
    }
    // epilogue: one Cholesky vector of the block at a time
    for (int xl = 0; xl < pl.xb; ++xl) {
      if (x0 + xl >= a.nchol) break;
      __syncthreads();
      const int cbeg = xl * n;
#pragma unroll
      for (int m = 0; m < MAXT; ++m) {
        const int t = wave + 4 * m;
        if (t < ntiles) {
          const int rt = t / pl.CT, ct = t % pl.CT;
          const int col = ct * 16 + l16 - cbeg;
          if (col >= 0 && col < n) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int row = rt * 16 + 4 * r + q;
              if (row < rows_valid) S[row * (n + 1) + col] = cmake(re[m][r], im[m][r]);
            }
          }
        }
      }
      __syncthreads();
      const int nn = n * n;
      for (int e = tid; e < pl.wb * nn; e += EX_THREADS) {
        const int wl = e % pl.wb, ij = e / pl.wb, i = ij / n, j = ij % n;
        const cplx tij = S[(wl * n + j) * (n + 1) + i];
        const cplx tji = S[(wl * n + i) * (n + 1) + j];
        wacc_r += tij.re * tji.re - tij.im * tji.im;
        wacc_i += tij.re * tji.im + tij.im * tji.re;
      }
    }
  }
  // per-walker block reduction (fixed order)
  for (int wl = 0; wl < pl.wb; ++wl) {
    const bool mine = (tid % pl.wb) == wl;
    const double vr = block_sum(mine ? wacc_r : 0.0, red);
    const double vi = block_sum(mine ? wacc_i : 0.0, red);
    const int w = w0 + wl;
    if (tid == 0 && w < a.nwalkers) st(a.exx_part, (int64_t)split * a.nwalkers + w, cmake(vr, vi));
  }
}

// Tiled variant with compile-time geometry: RT x CT output tiles of 16x16
// dealt round-robin to NW waves, double-buffered LDS staging with the next K
// chunk fetched into registers while the current one is multiplied (one
// barrier per chunk).
template <int RT, int CT, int NW>
__global__ void __launch_bounds__(64 * NW) exchange_tiled_kernel(ExchangeArgs a, ExPlan pl) {
  constexpr int THREADS = 64 * NW;
  constexpr int R = RT * 16, C = CT * 16;
  constexpr int RS = R + 1, CS = C + 2;
  constexpr int NT = RT * CT;
  constexpr int MAXT = (NT + NW - 1) / NW;
  constexpr int STAGE = 2 * EX_KC * RS + EX_KC * CS;  // doubles per staging buffer
  constexpr int A_PER = (R * EX_KC + THREADS - 1) / THREADS;
  constexpr int B_PER = (C * EX_KC + THREADS - 1) / THREADS;
  extern __shared__ double lds[];
  __shared__ double red[NW];
  const int n = a.nocc, N = a.nbasis;
  const int rows_valid = pl.wb * n, cols_valid = pl.xb * n;
  cplx* S = reinterpret_cast<cplx*>(lds);

  const int nwb = (int)ceil_div(a.nwalkers, pl.wb);
  const int lin = xcd_remap(blockIdx.x, gridDim.x);
  const int bw = lin % nwb, split = lin / nwb;
  const int nxb = (int)ceil_div(a.nchol, pl.xb);
  const int xb_begin = (int)((int64_t)nxb * split / a.xsplit);
  const int xb_end = (int)((int64_t)nxb * (split + 1) / a.xsplit);

  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, l16 = lane & 15, q = lane >> 4;
  const int w0 = bw * pl.wb;
  const int64_t arow_base = (int64_t)w0 * n;
  const int64_t arow_limit = (int64_t)a.nwalkers * n;
  const int64_t brow_limit = (int64_t)a.nchol * n;

  int a_off[MAXT], b_off[MAXT];
#pragma unroll
  for (int m = 0; m < MAXT; ++m) {
    const int t = wave + NW * m;
    a_off[m] = (t / CT) * 16 + l16;
    b_off[m] = (t % CT) * 16 + l16;
  }

  double wacc_r = 0.0, wacc_i = 0.0;
  const int nch = (int)ceil_div(N, EX_KC);
  cplx pa[A_PER];
  double pb[B_PER];

  for (int xblk = xb_begin; xblk < xb_end; ++xblk) {
    const int x0 = xblk * pl.xb;
    const int64_t brow_base = (int64_t)x0 * n;
    auto fetch = [&](int c) {
      const int k0 = c * EX_KC;
#pragma unroll
      for (int i = 0; i < A_PER; ++i) {
        const int e = tid + i * THREADS;
        const int row = e / EX_KC, k = e % EX_KC;
        const int64_t ar = arow_base + row;
        pa[i] = cmake(0.0, 0.0);
        if (e < R * EX_KC && row < rows_valid && ar < arow_limit && k0 + k < N) pa[i] = ld(a.ghalf, ar * N + k0 + k);
      }
#pragma unroll
      for (int i = 0; i < B_PER; ++i) {
        const int e = tid + i * THREADS;
        const int col = e / EX_KC, k = e % EX_KC;
        const int64_t br = brow_base + col;
        pb[i] = 0.0;
        if (e < C * EX_KC && col < cols_valid && br < brow_limit && k0 + k < N) pb[i] = a.rchol[br * N + k0 + k];
      }
    };
    auto stash = [&](int buf) {
      cplx* As = reinterpret_cast<cplx*>(lds + buf * STAGE);
      double* Bs = lds + buf * STAGE + 2 * EX_KC * RS;
#pragma unroll
      for (int i = 0; i < A_PER; ++i) {
        const int e = tid + i * THREADS;
        if (e < R * EX_KC) As[(e % EX_KC) * RS + e / EX_KC] = pa[i];
      }
#pragma unroll
      for (int i = 0; i < B_PER; ++i) {
        const int e = tid + i * THREADS;
        if (e < C * EX_KC) Bs[(e % EX_KC) * CS + e / EX_KC] = pb[i];
      }
    };
    v4d re[MAXT], im[MAXT];
#pragma unroll
    for (int m = 0; m < MAXT; ++m) {
      re[m] = zero4d();
      im[m] = zero4d();
    }
    fetch(0);
    __syncthreads();  // previous x block's epilogue is done with the aliased area
    stash(0);
    __syncthreads();
    for (int c = 0; c < nch; ++c) {
      if (c + 1 < nch) fetch(c + 1);
      const cplx* As = reinterpret_cast<const cplx*>(lds + (c & 1) * STAGE);
      const double* Bs = lds + (c & 1) * STAGE + 2 * EX_KC * RS;
#pragma unroll 1
      for (int s = 0; s < 4; ++s) {
        const int k = 4 * q + s;
#pragma unroll
        for (int m = 0; m < MAXT; ++m) {
          if (wave + NW * m < NT) {
            const cplx av = As[k * RS + a_off[m]];
            const double bv = Bs[k * CS + b_off[m]];
            re[m] = mfma64(av.re, bv, re[m]);
            im[m] = mfma64(av.im, bv, im[m]);
          }
        }
      }
      if (c + 1 < nch) stash((c + 1) & 1);
      __syncthreads();
    }
    for (int xl = 0; xl < pl.xb; ++xl) {
      if (x0 + xl >= a.nchol) break;
      const int cbeg = xl * n;
#pragma unroll
      for (int m = 0; m < MAXT; ++m) {
        if (wave + NW * m < NT) {
          const int col = b_off[m] - cbeg;
          if (col >= 0 && col < n) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int row = a_off[m] - l16 + 4 * r + q;
              if (row < rows_valid) S[row * (n + 1) + col] = cmake(re[m][r], im[m][r]);
            }
          }
        }
      }
      __syncthreads();
      const int nn = n * n;
      for (int e = tid; e < pl.wb * nn; e += THREADS) {
        const int wl = e % pl.wb, ij = e / pl.wb, i = ij / n, j = ij % n;
        const cplx tij = S[(wl * n + j) * (n + 1) + i];
        const cplx tji = S[(wl * n + i) * (n + 1) + j];
        wacc_r += tij.re * tji.re - tij.im * tji.im;
        wacc_i += tij.re * tji.im + tij.im * tji.re;
      }
      __syncthreads();
    }
  }
  for (int wl = 0; wl < pl.wb; ++wl) {
    const bool mine = (tid % pl.wb) == wl;
    const double vr = block_sum(mine ? wacc_r : 0.0, red);
    const double vi = block_sum(mine ? wacc_i : 0.0, red);
    const int w = w0 + wl;
    if (tid == 0 && w < a.nwalkers) st(a.exx_part, (int64_t)split * a.nwalkers + w, cmake(vr, vi));
  }
}

template <int RT, int CT, int NW>
size_t tiled_lds_bytes(const ExPlan& pl, int nocc) {
  constexpr int STAGE = 2 * EX_KC * (RT * 16 + 1) + EX_KC * (CT * 16 + 2);
  const size_t stage = sizeof(double) * 2 * (size_t)STAGE;
  const size_t epi = sizeof(cplx) * (size_t)(pl.wb * nocc) * (nocc + 1);
  return stage > epi ? stage : epi;
}

template <int RT, int CT, int NW>
hipError_t launch_tiled(const ExchangeArgs& a, const ExPlan& pl, dim3 grid, hipStream_t stream) {
  const size_t lds = tiled_lds_bytes<RT, CT, NW>(pl, a.nocc);
  hipError_t e = allow_lds((const void*)exchange_tiled_kernel<RT, CT, NW>, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((exchange_tiled_kernel<RT, CT, NW>), grid, dim3(64 * NW), lds, stream, a, pl);
  return hipGetLastError();
}

ExPlan make_plan(int nocc, int max_tiles) {
  // Choose walkers-per-tile (wb, a power of two dividing the block size) and
  // vectors-per-tile (xb) minimising padding, with at most 5 row tiles, 10
  // column tiles and max_tiles tiles in total.
  ExPlan best{1, 1, 1, 1};
  double best_score = -1.0;
  for (int wb = 1; wb <= 8; wb *= 2) {
    const int R = round16(wb * nocc);
    if (R / 16 > 5) break;
    for (int xb = 1; xb <= 64; ++xb) {
      const int C = round16(xb * nocc);
      if (C / 16 > 10) break;
      const int tiles = (R / 16) * (C / 16);
      if (tiles > max_tiles) break;
      const double eff = (double)(wb * nocc) * (xb * nocc) / ((double)R * C);
      // prefer efficient tiles, then bigger tiles (more reuse)
      const double score = eff * 1000.0 + tiles;
      if (score > best_score) {
        best_score = score;
        best = ExPlan{wb, xb, R / 16, C / 16};
      }
    }
  }
  return best;
}

// Plan used by the launcher: the tiled kernel when its geometry is compiled in.
bool tiled_geometry(const ExPlan& pl) { return pl.RT >= 1 && pl.RT <= 5 && pl.CT >= 8 && pl.CT <= 10; }

ExPlan choose_plan(int nocc) {
  const ExPlan big = make_plan(nocc, 50);
  if (tiled_geometry(big)) return big;
  return make_plan(nocc, 32);
}

size_t ex_lds_bytes(const ExPlan& pl, int nocc) {
  const size_t stage = sizeof(double) * (size_t)(2 * EX_KC * (pl.RT * 16 + 1) + EX_KC * (pl.CT * 16 + 2));
  const size_t epi = sizeof(cplx) * (size_t)(pl.wb * nocc) * (nocc + 1);
  return stage > epi ? stage : epi;
}

}  // namespace

int exchange_xsplit(int nwalkers, int nchol, int nocc, int want) {
  const ExPlan pl = choose_plan(nocc);
  const int nwb = (int)ceil_div(nwalkers, pl.wb);
  const int nxb = (int)ceil_div(nchol, pl.xb);
  // want > 0: that split; otherwise aim for >= 2048 workgroups.  Never more
  // splits than x blocks.
  int xs = want > 0 ? want : (int)ceil_div(2048, nwb);
  if (xs > nxb) xs = nxb;
  if (xs < 1) xs = 1;
  return xs;
}

hipError_t launch_exchange(const ExchangeArgs& a, hipStream_t stream) {
  if (a.nocc <= 0) return hipSuccess;
  if (a.nocc > 80) return hipErrorInvalidValue;
  const ExPlan pl = choose_plan(a.nocc);
  const int nwb = (int)ceil_div(a.nwalkers, pl.wb);
  const dim3 grid(nwb * a.xsplit);
  if (tiled_geometry(pl)) {
#define AFQMC_TILED(rt, ct) \
  if (pl.RT == rt && pl.CT == ct) return launch_tiled<rt, ct, (rt * ct > 32 ? 8 : 4)>(a, pl, grid, stream);
    AFQMC_TILED(1, 8) AFQMC_TILED(1, 9) AFQMC_TILED(1, 10)
    AFQMC_TILED(2, 8) AFQMC_TILED(2, 9) AFQMC_TILED(2, 10)
    AFQMC_TILED(3, 8) AFQMC_TILED(3, 9) AFQMC_TILED(3, 10)
    AFQMC_TILED(4, 8) AFQMC_TILED(4, 9) AFQMC_TILED(4, 10)
    AFQMC_TILED(5, 8) AFQMC_TILED(5, 9) AFQMC_TILED(5, 10)
#undef AFQMC_TILED
  }
  const size_t lds = ex_lds_bytes(pl, a.nocc);
  const int ntiles = pl.RT * pl.CT;
  hipError_t e = allow_lds(ntiles <= 16 ? (const void*)exchange_kernel<4> : (const void*)exchange_kernel<8>, lds);
  if (e != hipSuccess) return e;
  if (ntiles <= 16)
    hipLaunchKernelGGL(exchange_kernel<4>, grid, dim3(EX_THREADS), lds, stream, a, pl);
  else
    hipLaunchKernelGGL(exchange_kernel<8>, grid, dim3(EX_THREADS), lds, stream, a, pl);
  return hipGetLastError();
}

}  // namespace afqmc
