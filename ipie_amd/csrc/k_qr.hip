// Walker re-orthogonalisation (FP64): phi_s = Q_s R_s with diag(R_s) > 0.
//
// ipie's CPU path (ipie/walkers/uhf_walkers.py:108-143) takes LAPACK's
// economic QR and multiplies Q by sign(diag R), i.e. it returns the unique Q
// whose R has a positive real diagonal.  Here the same Q is produced by
// shifted Cholesky-QR in three passes (a shifted pass followed by CholeskyQR2,
// Fukaya, Kannan, Nakatsukasa, Yamamoto, Yanagisawa, SIAM J. Sci. Comput.
// 42 (2020) A477), which is as accurate as Householder QR for any numerically
// full-rank walker and maps onto per-row independent work:
//   G = X^H X (+ shift on pass 1),  G = R^H R,  X <- X R^{-1}.
// log det R = sum over passes of sum_i log R_ii, and
//   detR = exp(log det R_a + log det R_b - detR_shift),
//   log_detR += log detR,  ovlp /= detR.
#include "afqmc_internal.hpp"

namespace afqmc {

namespace {

constexpr int QR_THREADS = 256;
constexpr int QR_MC = 16;

// G = X^H X, returns trace(G) (real) in every thread.
__device__ double gram(const double* __restrict__ X, int N, int n, cplx* G, int ldg, cplx* stage) {
  const int tid = threadIdx.x;
  const int nn = n * n;
  constexpr int MAXOUT = (64 * 64 + QR_THREADS - 1) / QR_THREADS;
  cplx acc[MAXOUT];
#pragma unroll
  for (int t = 0; t < MAXOUT; ++t) acc[t] = cmake(0.0, 0.0);
  for (int m0 = 0; m0 < N; m0 += QR_MC) {
    const int mc = min(QR_MC, N - m0);
    __syncthreads();
    for (int e = tid; e < mc * n; e += QR_THREADS) stage[e] = ld(X, (int64_t)m0 * n + e);
    __syncthreads();
#pragma unroll
    for (int t = 0; t < MAXOUT; ++t) {
      const int idx = tid + t * QR_THREADS;
      if (idx < nn) {
        const int i = idx / n, j = idx % n;
        cplx a = acc[t];
        for (int r = 0; r < mc; ++r) cfma(a, cconj(stage[r * n + i]), stage[r * n + j]);
        acc[t] = a;
      }
    }
  }
#pragma unroll
  for (int t = 0; t < MAXOUT; ++t) {
    const int idx = tid + t * QR_THREADS;
    if (idx < nn) G[(idx / n) * ldg + (idx % n)] = acc[t];
  }
  __syncthreads();
  double tr = 0.0;
  for (int i = 0; i < n; ++i) tr += G[i * ldg + i].re;
  return tr;
}

// In-place upper Cholesky G = R^H R (R stored in the upper triangle of G).
// Returns sum_i log R_ii, or NaN if G is not numerically positive definite.
__device__ double cholesky_upper(cplx* G, int ldg, int n) {
  const int tid = threadIdx.x;
  __shared__ double s_d;
  for (int k = 0; k < n; ++k) {
    if (tid == 0) {
      const double d = G[k * ldg + k].re;
      s_d = (d > 0.0) ? sqrt(d) : NAN;
    }
    __syncthreads();
    const double d = s_d;
    if (!(d > 0.0)) return NAN;
    if (tid == 0) G[k * ldg + k] = cmake(d, 0.0);
    for (int j = k + 1 + tid; j < n; j += QR_THREADS) G[k * ldg + j] = cscale(G[k * ldg + j], 1.0 / d);
    __syncthreads();
    const int m = n - k - 1;
    for (int e = tid; e < m * m; e += QR_THREADS) {
      const int i = k + 1 + e / m, j = k + 1 + e % m;
      if (j >= i) {
        // G_ij -= conj(R_ki) R_kj
        const cplx rki = G[k * ldg + i], rkj = G[k * ldg + j];
        cplx g = G[i * ldg + j];
        g.re -= rki.re * rkj.re + rki.im * rkj.im;
        g.im -= rki.re * rkj.im - rki.im * rkj.re;
        G[i * ldg + j] = g;
      }
    }
    __syncthreads();
  }
  double ld = 0.0;
  for (int i = 0; i < n; ++i) ld += log(G[i * ldg + i].re);
  return ld;
}

// Rinv = R^{-1} (upper triangular), one column per thread.
__device__ void upper_inverse(const cplx* R, int ldr, int n, cplx* Ri, int ldi) {
  const int tid = threadIdx.x;
  for (int j = tid; j < n; j += QR_THREADS) {
    for (int i = n - 1; i >= 0; --i) {
      if (i > j) {
        Ri[i * ldi + j] = cmake(0.0, 0.0);
        continue;
      }
      cplx s = cmake(i == j ? 1.0 : 0.0, 0.0);
      for (int k = i + 1; k <= j; ++k) {
        const cplx r = R[i * ldr + k], x = Ri[k * ldi + j];
        s.re -= r.re * x.re - r.im * x.im;
        s.im -= r.re * x.im + r.im * x.re;
      }
      Ri[i * ldi + j] = cscale(s, 1.0 / R[i * ldr + i].re);
    }
  }
  __syncthreads();
}

// Y = X Rinv (Rinv upper triangular).
__device__ void apply_rinv(const double* __restrict__ X, int N, int n, const cplx* Ri, int ldi,
                           double* __restrict__ Y) {
  for (int e = threadIdx.x; e < N * n; e += QR_THREADS) {
    const int m = e / n, j = e % n;
    const double* xr = X + 2 * (int64_t)m * n;
    cplx a = cmake(0.0, 0.0);
    for (int k = 0; k <= j; ++k) cfma(a, ld(xr, k), Ri[k * ldi + j]);
    st(Y, (int64_t)m * n + j, a);
  }
  __syncthreads();
}

__global__ void __launch_bounds__(QR_THREADS) reortho_kernel(ReorthoArgs a) {
  extern __shared__ cplx smem[];
  const int w = blockIdx.x;
  const int N = a.nbasis;
  const int nmax = max(a.nalpha, a.nbeta);
  const int ldg = nmax + 1;
  cplx* G = smem;
  cplx* Ri = G + nmax * ldg;
  cplx* stage = Ri;  // Gram staging, only used before Ri is formed
  const int nspin = a.nbeta > 0 ? 2 : 1;
  double logdet = 0.0;
  bool bad = false;
  for (int s = 0; s < nspin; ++s) {
    const int n = s == 0 ? a.nalpha : a.nbeta;
    double* phi = (s == 0 ? a.phia : a.phib) + 2 * (int64_t)w * N * n;
    double* work = a.work + 2 * (int64_t)w * N * nmax;
    double* src = phi;
    double* dst = work;
    for (int pass = 0; pass < 3; ++pass) {
      const double tr = gram(src, N, n, G, ldg, stage);
      if (pass == 0) {
        // shift = 11 (N n + n (n + 1)) u ||X||_F^2, u = 2^-53
        const double shift = 11.0 * ((double)N * n + (double)n * (n + 1)) * 1.1102230246251565e-16 * tr;
        if (threadIdx.x == 0)
          for (int i = 0; i < n; ++i) G[i * ldg + i].re += shift;
        __syncthreads();
      }
      const double ldr = cholesky_upper(G, ldg, n);
      if (!(ldr == ldr)) {
        bad = true;
        break;
      }
      logdet += ldr;
      upper_inverse(G, ldg, n, Ri, ldg);
      apply_rinv(src, N, n, Ri, ldg, dst);
      double* t = src;
      src = dst;
      dst = t;
    }
    if (bad) break;
    // three passes leave the result in `work`; copy back into phi
    for (int e = threadIdx.x; e < N * n; e += QR_THREADS) st(phi, e, ld(src, e));
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (bad) {
      if (a.flags) atomicOr(a.flags, AFQMC_FLAG_CHOLESKY);
      a.detR[w] = NAN;
      return;
    }
    const double shift = a.detR_shift ? a.detR_shift[w] : 0.0;
    const double detR = exp(logdet - shift);
    a.detR[w] = detR;
    if (a.log_detR) a.log_detR[w] += log(detR);
    if (a.ovlp) {
      const cplx o = ld(a.ovlp, w);
      st(a.ovlp, w, cmake(o.re / detR, o.im / detR));
    }
  }
}

}  // namespace

hipError_t launch_reortho(const ReorthoArgs& a, hipStream_t stream) {
  const int nmax = a.nalpha > a.nbeta ? a.nalpha : a.nbeta;
  const size_t inv = (size_t)nmax * (nmax + 1), stage = (size_t)QR_MC * nmax;
  const size_t lds = sizeof(cplx) * ((size_t)nmax * (nmax + 1) + (inv > stage ? inv : stage));
  hipError_t e = allow_lds((const void*)reortho_kernel, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(reortho_kernel, dim3(a.nwalkers), dim3(QR_THREADS), lds, stream, a);
  return hipGetLastError();
}

}  // namespace afqmc
