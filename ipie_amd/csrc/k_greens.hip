// Batched single-determinant Green's function and overlap (FP64).
//
// One workgroup per walker handles both spins in turn:
//   O_s   = phi_s^T conj(psi_s)                         [n, n]
//   P O_s = L U      (partial pivoting on |re| + |im|, LAPACK's izamax rule)
//   log|det O_s| = sum log|U_kk|,  phase = prod U_kk / |U_kk| * (-1)^swaps
//   Ghalf_s = O_s^{-1} phi_s^T                          [n, N]
//   ovlp  = phase_a phase_b exp(logdet_a + logdet_b - log_shift)
// Restates ipie/estimators/greens_function_single_det.py:62-154 and
// ipie/propagation/overlap.py:29-59.  All reductions are in a fixed order, so
// the result is bitwise reproducible run to run.
#include "afqmc_internal.hpp"

namespace afqmc {

namespace {

constexpr int GF_THREADS = 256;
constexpr int GF_MC = 16;  // rows of phi / psi staged per chunk

struct GfSpinOut {
  double logdet;
  cplx phase;
  int singular;
};

// Computes O = phi^T conj(psi) into lds_O (row stride ldo) using staged chunks.
__device__ void build_overlap_matrix(const double* __restrict__ phi, const double* __restrict__ psi, int N,
                                     int n, cplx* lds_O, int ldo, cplx* stage) {
  cplx* sphi = stage;             // [GF_MC][n]
  cplx* spsi = stage + GF_MC * n; // [GF_MC][n]
  const int tid = threadIdx.x;
  const int nn = n * n;
  // each thread owns outputs idx = tid + t * GF_THREADS
  constexpr int MAXOUT = (64 * 64 + GF_THREADS - 1) / GF_THREADS;  // 16
  cplx acc[MAXOUT];
#pragma unroll
  for (int t = 0; t < MAXOUT; ++t) acc[t] = cmake(0.0, 0.0);
  for (int m0 = 0; m0 < N; m0 += GF_MC) {
    const int mc = min(GF_MC, N - m0);
    __syncthreads();
    for (int e = tid; e < mc * n; e += GF_THREADS) {
      const int r = e / n, c = e % n;
      sphi[r * n + c] = ld(phi, (int64_t)(m0 + r) * n + c);
      spsi[r * n + c] = cconj(ld(psi, (int64_t)(m0 + r) * n + c));
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < MAXOUT; ++t) {
      const int idx = tid + t * GF_THREADS;
      if (idx < nn) {
        const int i = idx / n, j = idx % n;
        cplx a = acc[t];
        for (int r = 0; r < mc; ++r) cfma(a, sphi[r * n + i], spsi[r * n + j]);
        acc[t] = a;
      }
    }
  }
#pragma unroll
  for (int t = 0; t < MAXOUT; ++t) {
    const int idx = tid + t * GF_THREADS;
    if (idx < nn) lds_O[(idx / n) * ldo + (idx % n)] = acc[t];
  }
  __syncthreads();
}

// In-place LU with partial pivoting of lds_O; piv[k] = row swapped with k.
__device__ GfSpinOut lu_factor(cplx* O, int ldo, int n, int* piv, double* red) {
  const int tid = threadIdx.x;
  __shared__ int s_piv;
  int swaps = 0;
  int singular = 0;
  for (int k = 0; k < n; ++k) {
    // pivot search by wave 0 (n <= 64 rows, one lane per row)
    if (tid < 64) {
      const int i = k + tid;
      double v = -1.0;
      if (i < n) {
        const cplx z = O[i * ldo + k];
        v = fabs(z.re) + fabs(z.im);
      }
      int best = i;
      for (int off = 1; off < 64; off <<= 1) {
        const double ov = __shfl_xor(v, off, 64);
        const int ob = __shfl_xor(best, off, 64);
        if (ov > v || (ov == v && ob < best)) {
          v = ov;
          best = ob;
        }
      }
      if (tid == 0) s_piv = best;
    }
    __syncthreads();
    const int p = s_piv;
    if (tid == 0) piv[k] = p;
    if (p != k) {
      ++swaps;
      for (int j = tid; j < n; j += GF_THREADS) {
        const cplx t = O[k * ldo + j];
        O[k * ldo + j] = O[p * ldo + j];
        O[p * ldo + j] = t;
      }
    }
    __syncthreads();
    const cplx pv = O[k * ldo + k];
    const bool zero = (pv.re == 0.0 && pv.im == 0.0);
    if (zero) singular = 1;
    if (!zero) {
      const cplx rinv = cdiv(cmake(1.0, 0.0), pv);
      for (int i = k + 1 + tid; i < n; i += GF_THREADS) O[i * ldo + k] = cmul(O[i * ldo + k], rinv);
    }
    __syncthreads();
    if (!zero) {
      const int m = n - k - 1;
      for (int e = tid; e < m * m; e += GF_THREADS) {
        const int i = k + 1 + e / m, j = k + 1 + e % m;
        cplx a = O[i * ldo + j];
        const cplx l = O[i * ldo + k];
        const cplx u = O[k * ldo + j];
        a.re -= l.re * u.re - l.im * u.im;
        a.im -= l.re * u.im + l.im * u.re;
        O[i * ldo + j] = a;
      }
    }
    __syncthreads();
  }
  GfSpinOut out;
  double logdet = 0.0;
  cplx phase = cmake((swaps & 1) ? -1.0 : 1.0, 0.0);
  for (int k = 0; k < n; ++k) {
    const cplx u = O[k * ldo + k];
    const double a = cabs(u);
    logdet += log(a);
    if (a > 0.0) phase = cmul(phase, cmake(u.re / a, u.im / a));
    else phase = cmake(0.0, 0.0);
  }
  (void)red;
  out.logdet = logdet;
  out.phase = phase;
  out.singular = singular;
  return out;
}

// Oinv = O^{-1} from its LU factors (one column of the identity per thread).
__device__ void lu_inverse(const cplx* LU, int ldo, int n, const int* piv, cplx* Oinv, int ldi) {
  const int tid = threadIdx.x;
  // Column j of Oinv solves (L U) x = P e_j; the work column lives in Oinv[:, j].
  for (int j = tid; j < n; j += GF_THREADS) {
    for (int i = 0; i < n; ++i) Oinv[i * ldi + j] = cmake(i == j ? 1.0 : 0.0, 0.0);
    for (int k = 0; k < n; ++k) {
      const int p = piv[k];
      if (p != k) {
        const cplx t = Oinv[k * ldi + j];
        Oinv[k * ldi + j] = Oinv[p * ldi + j];
        Oinv[p * ldi + j] = t;
      }
    }
    for (int i = 0; i < n; ++i) {
      cplx s = Oinv[i * ldi + j];
      for (int k = 0; k < i; ++k) {
        const cplx l = LU[i * ldo + k];
        const cplx x = Oinv[k * ldi + j];
        s.re -= l.re * x.re - l.im * x.im;
        s.im -= l.re * x.im + l.im * x.re;
      }
      Oinv[i * ldi + j] = s;
    }
    for (int i = n - 1; i >= 0; --i) {
      cplx s = Oinv[i * ldi + j];
      for (int k = i + 1; k < n; ++k) {
        const cplx u = LU[i * ldo + k];
        const cplx x = Oinv[k * ldi + j];
        s.re -= u.re * x.re - u.im * x.im;
        s.im -= u.re * x.im + u.im * x.re;
      }
      Oinv[i * ldi + j] = cdiv(s, LU[i * ldo + i]);
    }
  }
  __syncthreads();
}

// Ghalf[i][m] = sum_j Oinv[i][j] phi[m][j]
__device__ void build_ghalf(const cplx* Oinv, int ldi, const double* __restrict__ phi, int N, int n,
                            double* __restrict__ ghalf) {
  const int tid = threadIdx.x;
  for (int e = tid; e < n * N; e += GF_THREADS) {
    const int i = e / N, m = e % N;
    cplx a = cmake(0.0, 0.0);
    const double* prow = phi + 2 * (int64_t)m * n;
    for (int j = 0; j < n; ++j) cfma(a, Oinv[i * ldi + j], ld(prow, j));
    st(ghalf, (int64_t)i * N + m, a);
  }
}

template <bool WANT_GHALF>
__global__ void __launch_bounds__(GF_THREADS)
gf_kernel(GfArgs a) {
  extern __shared__ cplx smem[];
  const int w = blockIdx.x;
  const int N = a.nbasis;
  const int nmax = max(a.nalpha, a.nbeta);
  const int ld = nmax + 1;
  cplx* O = smem;                       // [nmax][ld]
  cplx* Oinv = O + nmax * ld;           // [nmax][ld]
  cplx* stage = Oinv;                   // [2][GF_MC][nmax], only used before Oinv is formed
  __shared__ int piv[64];
  __shared__ double red[GF_THREADS / 64];

  const int nspin = (a.nbeta > 0 && !a.rhf) ? 2 : 1;
  double logdet[2] = {0.0, 0.0};
  cplx phase[2] = {cmake(1.0, 0.0), cmake(1.0, 0.0)};
  int singular = 0;
  for (int s = 0; s < nspin; ++s) {
    const int n = s == 0 ? a.nalpha : a.nbeta;
    const double* phi = (s == 0 ? a.phia : a.phib) + 2 * (int64_t)w * N * n;
    const double* psi = s == 0 ? a.psia : a.psib;
    build_overlap_matrix(phi, psi, N, n, O, ld, stage);
    GfSpinOut r = lu_factor(O, ld, n, piv, red);
    logdet[s] = r.logdet;
    phase[s] = r.phase;
    singular |= r.singular;
    if (WANT_GHALF) {
      __syncthreads();
      lu_inverse(O, ld, n, piv, Oinv, ld);
      double* gh = (s == 0 ? a.ghalfa : a.ghalfb) + 2 * (int64_t)w * n * N;
      build_ghalf(Oinv, ld, phi, N, n, gh);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double lsh = a.log_shift ? a.log_shift[w] : 0.0;
    cplx ph;
    double lg;
    if (a.nbeta > 0 && !a.rhf) {
      ph = cmul(phase[0], phase[1]);
      lg = logdet[0] + logdet[1];
    } else if (a.nbeta > 0 && a.rhf) {
      ph = cmul(phase[0], phase[0]);
      lg = logdet[0] + logdet[0];
    } else {
      ph = phase[0];
      lg = logdet[0];
    }
    const double e = exp(lg - lsh);
    st(a.ovlp, w, cmake(ph.re * e, ph.im * e));
    if (singular && a.flags) atomicOr(a.flags, AFQMC_FLAG_SINGULAR);
  }
}

}  // namespace

size_t gf_lds_bytes(int nalpha, int nbeta) {
  const int nmax = nalpha > nbeta ? nalpha : nbeta;
  const size_t inv = (size_t)nmax * (nmax + 1), stage = (size_t)2 * GF_MC * nmax;
  return sizeof(cplx) * ((size_t)nmax * (nmax + 1) + (inv > stage ? inv : stage));
}

hipError_t launch_gf(const GfArgs& a, bool want_ghalf, hipStream_t stream) {
  const size_t lds = gf_lds_bytes(a.nalpha, a.nbeta);
  hipError_t e = allow_lds(want_ghalf ? (const void*)gf_kernel<true> : (const void*)gf_kernel<false>, lds);
  if (e != hipSuccess) return e;
  if (want_ghalf)
    hipLaunchKernelGGL(gf_kernel<true>, dim3(a.nwalkers), dim3(GF_THREADS), lds, stream, a);
  else
    hipLaunchKernelGGL(gf_kernel<false>, dim3(a.nwalkers), dim3(GF_THREADS), lds, stream, a);
  return hipGetLastError();
}

}  // namespace afqmc
