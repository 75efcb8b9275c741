// Batched single-determinant Green's function and overlap (FP64), four kernels:
//   gf_obuild_kernel  O_s = phi_s^T conj(psi_s)           [n, n] per walker and spin
//                     (v_mfma_f64_16x16x4f64, one wave per 16-row tile of O)
//   gf_lu_kernel      P O_s = L U (partial pivoting on |re| + |im|, the first
//                     index on ties, LAPACK's izamax rule); log|det O_s| =
//                     sum log|U_kk|, phase = prod U_kk / |U_kk| * (-1)^swaps;
//                     optionally O_s^{-1}, in place (U^{-1}, then inv(O) L =
//                     U^{-1}, then the pivots' column interchanges: one matrix
//                     of LDS).  One wave per walker and spin, lane = row.
//   gf_combine_kernel ovlp = phase_a phase_b exp(logdet_a + logdet_b - log_shift)
//   gf_ghalf_kernel   Ghalf_s = O_s^{-1} phi_s^T            [n, N]
//                     (FP64 MFMA, one wave per 16-row tile of Ghalf)
// Restates ipie/estimators/greens_function_single_det.py:62-154 and
// ipie/propagation/overlap.py:29-59.  All reductions are in a fixed order, so
// the result is bitwise reproducible run to run.
#include "afqmc_internal.hpp"

namespace afqmc {

namespace {

constexpr int GF_WAVES = 4;  // waves per workgroup of the MFMA kernels

__device__ __forceinline__ int gf_nspin(const GfArgs& a) { return (a.nbeta > 0 && !a.rhf) ? 2 : 1; }

// O[w][s][i][j] = sum_m phi_s[w][m][i] conj(psi_s[m][j]); wave = (w, s, row tile it).
// MFMA operands: A[i][k = m] = phi[m][i], B[k = m][j] = conj(psi[m][j]).  REAL:
// psi has zero imaginary parts, whose two MFMAs per step would add only zeros.
template <int NT, bool REAL>
__global__ void __launch_bounds__(64 * GF_WAVES) gf_obuild_kernel(GfArgs a, double* __restrict__ obuf, int nmax) {
  const int gw = blockIdx.x * GF_WAVES + (threadIdx.x >> 6);
  const int nspin = gf_nspin(a);
  const int task = gw / NT, it = gw % NT;
  if (task >= a.nwalkers * nspin) return;
  const int w = task / nspin, s = task % nspin;
  const int n = s == 0 ? a.nalpha : a.nbeta;
  if (it * 16 >= n) return;
  const int N = a.nbasis;
  const double* phi = (s == 0 ? a.phia : a.phib) + 2 * (int64_t)w * N * n;
  const double* psi = s == 0 ? a.psia : a.psib;
  const int lane = threadIdx.x & 63, r16 = lane & 15, kq = lane >> 4;
  const int i = it * 16 + r16;
  v4d re[NT], im[NT];
#pragma unroll
  for (int jt = 0; jt < NT; ++jt) {
    re[jt] = zero4d();
    im[jt] = zero4d();
  }
  // four K steps per iteration, their loads issued before the first MFMA (the
  // same products in the same order as one step at a time)
  constexpr int KU = 4;
  for (int m0 = 0; m0 < N; m0 += 4 * KU) {
    cplx av[KU], bv[KU][NT];
#pragma unroll
    for (int u = 0; u < KU; ++u) {
      const int m = m0 + 4 * u + kq;
      av[u] = (m < N && i < n) ? ld(phi, (int64_t)m * n + i) : cmake(0.0, 0.0);
#pragma unroll
      for (int jt = 0; jt < NT; ++jt) {
        const int j = jt * 16 + r16;
        if (REAL)
          bv[u][jt] = cmake((m < N && j < n) ? psi[2 * ((int64_t)m * n + j)] : 0.0, 0.0);
        else
          bv[u][jt] = (m < N && j < n) ? cconj(ld(psi, (int64_t)m * n + j)) : cmake(0.0, 0.0);
      }
    }
#pragma unroll
    for (int u = 0; u < KU; ++u) {
      if (m0 + 4 * u >= N) break;  // wave-uniform: a step past the end adds nothing
#pragma unroll
      for (int jt = 0; jt < NT; ++jt) {
        re[jt] = mfma64(av[u].re, bv[u][jt].re, re[jt]);
        if (!REAL) re[jt] = mfma64(-av[u].im, bv[u][jt].im, re[jt]);
        if (!REAL) im[jt] = mfma64(av[u].re, bv[u][jt].im, im[jt]);
        im[jt] = mfma64(av[u].im, bv[u][jt].re, im[jt]);
      }
    }
  }
  double* o = obuf + 2 * ((int64_t)w * 2 + s) * nmax * nmax;
#pragma unroll
  for (int jt = 0; jt < NT; ++jt)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = it * 16 + 4 * r + kq, col = jt * 16 + r16;  // accumulator row = 4 r + lane / 16
      if (row < n && col < n) st(o, (int64_t)row * nmax + col, cmake(re[jt][r], im[jt][r]));
    }
}

// sum_{k0 <= k < k1} row[k] col[k ld] (complex): eight loads of each operand in
// flight, four independent partial sums.
__device__ __forceinline__ cplx lu_dot(const cplx* row, const cplx* col, int ld, int k0, int k1) {
  cplx s4[4] = {cmake(0.0, 0.0), cmake(0.0, 0.0), cmake(0.0, 0.0), cmake(0.0, 0.0)};
  int k = k0;
  for (; k + 8 <= k1; k += 8) {
    cplx r[8], x[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      r[u] = row[k + u];
      x[u] = col[(k + u) * ld];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) cfma(s4[u & 3], r[u], x[u]);
  }
  for (; k < k1; ++k) cfma(s4[0], row[k], col[k * ld]);
  return cadd(cadd(s4[0], s4[1]), cadd(s4[2], s4[3]));
}

// max over the 64 lanes of a wave of v (every lane active), in every lane:
// DPP swaps within rows of 16 lanes (quad_perm [1,0,3,2], quad_perm [2,3,0,1],
// row_half_mirror, row_mirror pair each lane with one of another group), then
// two shuffles across rows.
__device__ __forceinline__ double dpp_max(double v, int ctrl) {
  int lo = __double2loint(v), hi = __double2hiint(v);
  switch (ctrl) {  // the control must be a constant of the instruction
    case 0xB1:
      lo = __builtin_amdgcn_update_dpp(lo, lo, 0xB1, 0xF, 0xF, false);
      hi = __builtin_amdgcn_update_dpp(hi, hi, 0xB1, 0xF, 0xF, false);
      break;
    case 0x4E:
      lo = __builtin_amdgcn_update_dpp(lo, lo, 0x4E, 0xF, 0xF, false);
      hi = __builtin_amdgcn_update_dpp(hi, hi, 0x4E, 0xF, 0xF, false);
      break;
    case 0x141:
      lo = __builtin_amdgcn_update_dpp(lo, lo, 0x141, 0xF, 0xF, false);
      hi = __builtin_amdgcn_update_dpp(hi, hi, 0x141, 0xF, 0xF, false);
      break;
    default:
      lo = __builtin_amdgcn_update_dpp(lo, lo, 0x140, 0xF, 0xF, false);
      hi = __builtin_amdgcn_update_dpp(hi, hi, 0x140, 0xF, 0xF, false);
      break;
  }
  return fmax(v, __hiloint2double(hi, lo));
}
__device__ __forceinline__ double wave_max(double v) {
  v = dpp_max(v, 0xB1);
  v = dpp_max(v, 0x4E);
  v = dpp_max(v, 0x141);
  v = dpp_max(v, 0x140);
  v = fmax(v, __shfl_xor(v, 16, 64));
  return fmax(v, __shfl_xor(v, 32, 64));
}

// One wave per walker: LU of O_s (in LDS, row stride nmax + 1), log-det and
// phase, optionally the inverse into oinv, then the overlap of the walker.
// A single-wave workgroup: __syncthreads() orders the LDS traffic between lanes.
__global__ void __launch_bounds__(64) gf_lu_kernel(GfArgs a, const double* __restrict__ obuf,
                                                  double* __restrict__ oinv, int nmax, double* __restrict__ res) {
  extern __shared__ cplx smem[];
  const int lr = nmax + 1;        // LDS row stride
  cplx* LU = smem;                // [nmax][lr]; with oinv, then inverted in place
  cplx* work = smem + nmax * lr;  // [nmax], a column of L (oinv only)
  __shared__ int piv[64];
  const int lane = threadIdx.x;
  const int nspin = gf_nspin(a);
  const int w = blockIdx.x / nspin, s = blockIdx.x % nspin;
  double logdet = 0.0;
  cplx phase = cmake(1.0, 0.0);
  int singular = 0;
  {
    const int n = s == 0 ? a.nalpha : a.nbeta;
    const double* o = obuf + 2 * ((int64_t)w * 2 + s) * nmax * nmax;
    for (int i = 0; i < n; ++i)
      if (lane < n) LU[i * lr + lane] = ld(o, (int64_t)i * nmax + lane);
    __syncthreads();
    int swaps = 0;
    for (int k = 0; k < n; ++k) {
      // pivot: the largest |re| + |im| in rows k..n-1, the first such row on ties
      const bool cand = lane >= k && lane < n;
      double v = -1.0;
      if (cand) {
        const cplx z = LU[lane * lr + k];
        v = fabs(z.re) + fabs(z.im);
      }
      const double vmax = wave_max(v);
      const unsigned long long hit = __ballot(cand && v == vmax);
      const int p = hit ? __ffsll(hit) - 1 : k;  // the same in every lane
      if (lane == 0) piv[k] = p;
      if (p != k) {
        ++swaps;
        if (lane < n) {
          const cplx t = LU[k * lr + lane];
          LU[k * lr + lane] = LU[p * lr + lane];
          LU[p * lr + lane] = t;
        }
      }
      __syncthreads();
      const cplx pv = LU[k * lr + k];
      const bool zero = (pv.re == 0.0 && pv.im == 0.0);
      if (zero) singular = 1;
      if (!zero && lane > k && lane < n) {
        const cplx rinv = cdiv(cmake(1.0, 0.0), pv);
        cplx* row = LU + lane * lr;
        const cplx l = cmul(row[k], rinv);
        row[k] = l;
        const cplx* urow = LU + k * lr;
        int j = k + 1;
        // eight columns at a time, loads before stores (row k is not this lane's row)
        for (; j + 8 <= n; j += 8) {
          cplx u[8], x[8];
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            u[q] = urow[j + q];
            x[q] = row[j + q];
          }
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            x[q].re -= l.re * u[q].re - l.im * u[q].im;
            x[q].im -= l.re * u[q].im + l.im * u[q].re;
            row[j + q] = x[q];
          }
        }
        for (; j < n; ++j) {
          const cplx u = urow[j];
          cplx x = row[j];
          x.re -= l.re * u.re - l.im * u.im;
          x.im -= l.re * u.im + l.im * u.re;
          row[j] = x;
        }
      }
      __syncthreads();
    }
    // lane k takes log|U_kk| and U_kk / |U_kk|; a butterfly (every lane adds
    // and multiplies the same pairs, so all lanes agree) sums and multiplies them
    double lg = 0.0;
    cplx ph = cmake(1.0, 0.0);
    if (lane < n) {
      const cplx u = LU[lane * lr + lane];
      const double ab = cabs(u);
      lg = log(ab);
      ph = ab > 0.0 ? cmake(u.re / ab, u.im / ab) : cmake(0.0, 0.0);
    }
    for (int off = 1; off < 64; off <<= 1) {
      lg += __shfl_xor(lg, off, 64);
      const cplx o = cmake(__shfl_xor(ph.re, off, 64), __shfl_xor(ph.im, off, 64));
      ph = cmul(ph, o);
    }
    if (swaps & 1) ph = cmake(-ph.re, -ph.im);
    logdet = lg;
    phase = ph;
    if (oinv) {
      // O^{-1} in place from the factors, one matrix of LDS (lane = row i):
      // U^{-1} column by column, then inv(O) L = U^{-1} for the columns from
      // the right (each column of L saved to `work` first), then the column
      // interchanges of the pivots in reverse.  Every sum runs over k in a fixed
      // order (lu_dot).
      const int i = lane;
      for (int j = 0; j < n; ++j) {
        // column j of U^{-1}: x_i = sum_{i <= k < j} Uinv[i][k] U[k][j], times -1 / U[j][j]
        const cplx t = i < j ? lu_dot(LU + i * lr, LU + j, lr, i, j) : cmake(0.0, 0.0);
        const cplx djj = cdiv(cmake(1.0, 0.0), LU[j * lr + j]);
        __syncthreads();
        if (i < j) LU[i * lr + j] = cmul(t, cmake(-djj.re, -djj.im));
        if (i == j) LU[j * lr + j] = djj;
        __syncthreads();
      }
      for (int j = n - 2; j >= 0; --j) {
        if (i > j && i < n) {
          work[i] = LU[i * lr + j];
          LU[i * lr + j] = cmake(0.0, 0.0);
        }
        __syncthreads();
        if (i < n) LU[i * lr + j] = csub(LU[i * lr + j], lu_dot(LU + i * lr, work, 1, j + 1, n));
        __syncthreads();
      }
      for (int j = n - 2; j >= 0; --j) {
        const int p = piv[j];
        if (p != j && i < n) {
          const cplx t = LU[i * lr + j];
          LU[i * lr + j] = LU[i * lr + p];
          LU[i * lr + p] = t;
        }
      }
      __syncthreads();
      double* oi = oinv + 2 * ((int64_t)w * 2 + s) * nmax * nmax;
      for (int r = 0; r < n; ++r)
        if (lane < n) st(oi, (int64_t)r * nmax + lane, LU[r * lr + lane]);
    }
    __syncthreads();
  }
  if (lane == 0) {
    double* r = res + 4 * ((int64_t)w * 2 + s);
    r[0] = logdet;
    r[1] = phase.re;
    r[2] = phase.im;
    r[3] = singular ? 1.0 : 0.0;
  }
}

// ovlp[w] = phase_a phase_b exp(logdet_a + logdet_b - log_shift[w]) from the
// per-spin results of gf_lu_kernel (RHF: the alpha factors squared).
__global__ void __launch_bounds__(256) gf_combine_kernel(GfArgs a, const double* __restrict__ res) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= a.nwalkers) return;
  const double* r = res + 8 * (int64_t)w;
  const cplx pa = cmake(r[1], r[2]);
  cplx ph;
  double lg;
  bool singular = r[3] != 0.0;
  if (a.nbeta > 0 && !a.rhf) {
    ph = cmul(pa, cmake(r[5], r[6]));
    lg = r[0] + r[4];
    singular = singular || r[7] != 0.0;
  } else if (a.nbeta > 0 && a.rhf) {
    ph = cmul(pa, pa);
    lg = r[0] + r[0];
  } else {
    ph = pa;
    lg = r[0];
  }
  const double lsh = a.log_shift ? a.log_shift[w] : 0.0;
  const double e = exp(lg - lsh);
  st(a.ovlp, w, cmake(ph.re * e, ph.im * e));
  if (singular && a.flags) atomicOr(a.flags, AFQMC_FLAG_SINGULAR);
  if (singular && a.wflags) atomicOr(a.wflags + w, AFQMC_FLAG_SINGULAR);
}

// Ghalf_s[w][i][m] = sum_j Oinv[w][s][i][j] phi_s[w][m][j]; wave = (w, s, row tile it).
// MFMA operands: A[i][k = j] = Oinv[i][j] (held for all K steps), B[k = j][m] = phi[m][j].
template <int KS, bool PLANES>
__global__ void __launch_bounds__(64 * GF_WAVES) gf_ghalf_kernel(GfArgs a, const double* __restrict__ oinv, int nmax,
                                                                int ntile) {
  const int gw = blockIdx.x * GF_WAVES + (threadIdx.x >> 6);
  const int nspin = gf_nspin(a);
  const int task = gw / ntile, it = gw % ntile;
  if (task >= a.nwalkers * nspin) return;
  const int w = task / nspin, s = task % nspin;
  const int n = s == 0 ? a.nalpha : a.nbeta;
  if (it * 16 >= n) return;
  const int N = a.nbasis;
  const double* phi = (s == 0 ? a.phia : a.phib) + 2 * (int64_t)w * N * n;
  const double* oi = oinv + 2 * ((int64_t)w * 2 + s) * nmax * nmax;
  double* gh = (s == 0 ? a.ghalfa : a.ghalfb) + 2 * (int64_t)w * n * N;
  uint16_t* gp = PLANES ? a.gplanes[s] : nullptr;  // split-BF16 planes of Ghalf_s, or null
  const int64_t pstride = (int64_t)a.nwalkers * n * a.kpad;
  uint16_t* gprow = gp ? gp + (int64_t)w * n * a.kpad : nullptr;
  const int lane = threadIdx.x & 63, r16 = lane & 15, kq = lane >> 4;
  const int i = it * 16 + r16;
  cplx af[KS];
#pragma unroll
  for (int ks = 0; ks < KS; ++ks) {
    const int k = ks * 4 + kq;
    af[ks] = (i < n && k < n) ? ld(oi, (int64_t)i * nmax + k) : cmake(0.0, 0.0);
  }
  // two column tiles per iteration: the loads of both are issued before the
  // MFMAs and stores of either
  constexpr int MU = 2;
  const int nmt = (int)ceil_div(N, 16);
  for (int mt0 = 0; mt0 < nmt; mt0 += MU) {
    cplx bv[MU][KS];
#pragma unroll
    for (int u = 0; u < MU; ++u) {
      const int m = (mt0 + u) * 16 + r16;
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        const int k = ks * 4 + kq;
        bv[u][ks] = (m < N && k < n) ? ld(phi, (int64_t)m * n + k) : cmake(0.0, 0.0);
      }
    }
#pragma unroll
    for (int u = 0; u < MU; ++u) {
      if (mt0 + u >= nmt) break;  // wave-uniform
      const int m = (mt0 + u) * 16 + r16;
      v4d re = zero4d(), im = zero4d();
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        if (ks * 4 < n) {  // K steps past n have zero operands (n = 40 takes 10 of 12); a branch keeps the unroll
          re = mfma64(af[ks].re, bv[u][ks].re, re);
          re = mfma64(-af[ks].im, bv[u][ks].im, re);
          im = mfma64(af[ks].re, bv[u][ks].im, im);
          im = mfma64(af[ks].im, bv[u][ks].re, im);
        }
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = it * 16 + 4 * r + kq;  // accumulator row = 4 r + lane / 16
        if (row < n && m < N) {
          st(gh, (int64_t)row * N + m, cmake(re[r], im[r]));
          if (PLANES && gprow) {
            uint16_t rh, rl, ih, il;
            split_bf16_pair(re[r], rh, rl);
            split_bf16_pair(im[r], ih, il);
            const int64_t o = (int64_t)row * a.kpad + m;
            gprow[o] = rh;
            gprow[pstride + o] = rl;
            gprow[2 * pstride + o] = ih;
            gprow[3 * pstride + o] = il;
          }
        }
      }
    }
  }
  if (PLANES && gprow) {  // zero K padding of this wave's rows
    const int pad = a.kpad - N;
    for (int e = lane; e < 16 * pad; e += 64) {
      const int row = it * 16 + e / pad, m = N + e % pad;
      if (row < n) {
        const int64_t o = (int64_t)row * a.kpad + m;
        gprow[o] = 0;
        gprow[pstride + o] = 0;
        gprow[2 * pstride + o] = 0;
        gprow[3 * pstride + o] = 0;
      }
    }
  }
}

__global__ void __launch_bounds__(256) imag_any_kernel(const double* __restrict__ z, int64_t n, int* flag) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n && z[2 * i + 1] != 0.0) *flag = 1;
}

}  // namespace

hipError_t launch_imag_any(const double* z, int64_t n, int* flag, hipStream_t stream) {
  if (n < 1) return hipSuccess;
  hipLaunchKernelGGL(imag_any_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, stream, z, n, flag);
  return hipGetLastError();
}

size_t gf_lds_bytes(int nalpha, int nbeta) {
  const int nmax = nalpha > nbeta ? nalpha : nbeta;
  return sizeof(cplx) * ((size_t)nmax * (nmax + 1) + nmax);
}

// O matrices [W, 2, nmax, nmax] complex, then the per-(walker, spin) LU results
// [W, 2, 4] (log|det|, phase re, im, singular)
size_t gf_work_bytes(int nwalkers, int nalpha, int nbeta) {
  const int nmax = nalpha > nbeta ? nalpha : nbeta;
  return sizeof(cplx) * (size_t)nwalkers * 2 * nmax * nmax + sizeof(double) * 8 * (size_t)nwalkers;
}

hipError_t launch_gf(const GfArgs& a, bool want_ghalf, double* obuf, double* oinv, hipStream_t stream) {
  const int nmax = a.nalpha > a.nbeta ? a.nalpha : a.nbeta;
  if (nmax < 1 || nmax > 64 || a.nwalkers < 1) return hipErrorInvalidValue;
  const int nspin = (a.nbeta > 0 && !a.rhf) ? 2 : 1;
  const int nt = (nmax + 15) / 16;
  const int64_t waves = (int64_t)a.nwalkers * nspin * nt;
  const unsigned grid = (unsigned)ceil_div(waves, GF_WAVES);
#define AFQMC_OBUILD(NT)                                                                                  \
  do {                                                                                                   \
    if (a.psi_real)                                                                                      \
      hipLaunchKernelGGL((gf_obuild_kernel<NT, true>), dim3(grid), dim3(64 * GF_WAVES), 0, stream, a, obuf, \
                         nmax);                                                                          \
    else                                                                                                 \
      hipLaunchKernelGGL((gf_obuild_kernel<NT, false>), dim3(grid), dim3(64 * GF_WAVES), 0, stream, a, obuf, \
                         nmax);                                                                          \
  } while (0)
  switch (nt) {
    case 1: AFQMC_OBUILD(1); break;
    case 2: AFQMC_OBUILD(2); break;
    case 3: AFQMC_OBUILD(3); break;
    default: AFQMC_OBUILD(4); break;
  }
#undef AFQMC_OBUILD
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  // the inverse needs a column of L besides the factors; the overlap alone does not
  const size_t lds = gf_lds_bytes(a.nalpha, a.nbeta) - (want_ghalf ? 0 : sizeof(cplx) * (size_t)nmax);
  e = allow_lds((const void*)gf_lu_kernel, lds);
  if (e != hipSuccess) return e;
  double* res = obuf + 2 * (int64_t)a.nwalkers * 2 * nmax * nmax;
  hipLaunchKernelGGL(gf_lu_kernel, dim3((unsigned)(a.nwalkers * nspin)), dim3(64), lds, stream, a, obuf,
                     want_ghalf ? oinv : nullptr, nmax, res);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(gf_combine_kernel, dim3((unsigned)ceil_div(a.nwalkers, 256)), dim3(256), 0, stream, a, res);
  e = hipGetLastError();
  if (e != hipSuccess || !want_ghalf) return e;
  const bool planes = a.gplanes[0] != nullptr;
#define AFQMC_GHALF(KS)                                                                                      \
  do {                                                                                                       \
    if (planes)                                                                                              \
      hipLaunchKernelGGL((gf_ghalf_kernel<KS, true>), dim3(grid), dim3(64 * GF_WAVES), 0, stream, a, oinv, nmax, nt); \
    else                                                                                                     \
      hipLaunchKernelGGL((gf_ghalf_kernel<KS, false>), dim3(grid), dim3(64 * GF_WAVES), 0, stream, a, oinv, nmax, nt); \
  } while (0)
  switch (nt) {  // K steps of 4 cover nmax <= 16 nt
    case 1: AFQMC_GHALF(4); break;
    case 2: AFQMC_GHALF(8); break;
    case 3: AFQMC_GHALF(12); break;
    default: AFQMC_GHALF(16); break;
  }
#undef AFQMC_GHALF
  return hipGetLastError();
}

}  // namespace afqmc
