// Walker re-orthogonalisation (FP64): phi_s = Q_s R_s with diag(R_s) > 0.
//
// ipie's CPU path (ipie/walkers/uhf_walkers.py:108-143) takes LAPACK's
// economic QR and multiplies Q by sign(diag R), i.e. it returns the unique Q
// whose R has a positive real diagonal.  Here the same Q is produced by
// Cholesky-QR, which maps onto per-row independent work:
//   G = X^H X (+ shift),  G = R^H R,  X <- X R^{-1},
// as sweeps over the rows of X on FP64 MFMA: the Gram matrix of pass p + 1 is
// formed from the rows of X R_p^{-1} while they are written.
//   - CholeskyQR2 (two passes, three sweeps; Yamamoto, Nakatsukasa, Yanagisawa,
//     Fukaya, ETNA 44 (2015) 306) when the unshifted Cholesky of X^H X succeeds
//     and 8 kappa sqrt(N n u + n (n + 1) u) <= 1/2 with kappa bounded by
//     ||R||_F ||R^-1||_F >= kappa_2(X): then Q is orthogonal to O(u).  Walkers
//     re-orthogonalised every few steps are far inside this bound.
//   - otherwise shifted CholeskyQR3 (a shifted pass, then CholeskyQR2; Fukaya,
//     Kannan, Nakatsukasa, Yamamoto, Yanagisawa, SIAM J. Sci. Comput. 42
//     (2020) A477), as accurate as Householder QR for any numerically
//     full-rank walker: the Gram matrix is formed again (the unshifted
//     attempt overwrote it), then three passes and four sweeps.
// log det R = sum over passes of sum_i log R_ii, and
//   detR = exp(log det R_a + log det R_b - detR_shift),
//   log_detR += log detR,  ovlp /= detR.
#include "afqmc_internal.hpp"

namespace afqmc {

namespace {

constexpr int QR_THREADS = 256;
constexpr int QR_WAVES = QR_THREADS / 64;

// Row stride of G and Ri for at most 16 NT columns.
__host__ __device__ inline int qr_ld(int nt) { return 16 * nt + 1; }

// One sweep over the rows of X [N][n] (n <= 16 NT) in 16-row tiles, tile t on
// wave t % 4, on FP64 MFMA (16 x 16 x 4).  With Ri (upper triangular, zero
// rows and columns from n to 16 NT, row stride ldr) the rows are replaced by
// Y = X Ri in place; with G the Gram matrix Y^H Y (Y = X without Ri) is formed
// in G's upper triangle (row stride ldg) and its trace returned.  The
// accumulator register r of a Y tile holds row 4 r + lane / 16, column
// lane % 16, which is the A / B operand layout of K step r over those rows, so
// the Gram product takes Y straight from the registers.  The waves' partial
// Gram matrices are added into G in wave order (reproducible sums).
template <int NT>
__device__ double qr_sweep(double* __restrict__ X, int N, int n, const cplx* Ri, int ldr, cplx* G, int ldg) {
  constexpr int NPAIR = NT * (NT + 1) / 2;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, l16 = lane & 15, q = lane >> 4;
  v4d gr[NPAIR], gi[NPAIR];
#pragma unroll
  for (int p = 0; p < NPAIR; ++p) {
    gr[p] = zero4d();
    gi[p] = zero4d();
  }
  const int ntile = (N + 15) / 16;
  for (int t = wave; t < ntile; t += QR_WAVES) {
    const int m0 = 16 * t;
    v4d yr[NT], yi[NT];
    if (Ri) {
#pragma unroll
      for (int J = 0; J < NT; ++J) {
        yr[J] = zero4d();
        yi[J] = zero4d();
      }
      const int mrow = m0 + l16;  // A operand: X[mrow][k0 + q]
#pragma unroll
      for (int k0 = 0; k0 < 16 * NT; k0 += 4) {
        const int k = k0 + q;
        const cplx x = (mrow < N && k < n) ? ld(X, (int64_t)mrow * n + k) : cmake(0.0, 0.0);
#pragma unroll
        for (int J = 0; J < NT; ++J) {
          if (k0 < 16 * (J + 1)) {  // Ri is upper triangular
            const cplx r = Ri[k * ldr + 16 * J + l16];
            yr[J] = mfma64(x.re, r.re, yr[J]);
            yr[J] = mfma64(-x.im, r.im, yr[J]);
            yi[J] = mfma64(x.re, r.im, yi[J]);
            yi[J] = mfma64(x.im, r.re, yi[J]);
          }
        }
      }
#pragma unroll
      for (int J = 0; J < NT; ++J)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = m0 + 4 * r + q, col = 16 * J + l16;
          if (row < N && col < n) st(X, (int64_t)row * n + col, cmake(yr[J][r], yi[J][r]));
        }
    } else {
#pragma unroll
      for (int J = 0; J < NT; ++J)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = m0 + 4 * r + q, col = 16 * J + l16;
          const cplx v = (row < N && col < n) ? ld(X, (int64_t)row * n + col) : cmake(0.0, 0.0);
          yr[J][r] = v.re;
          yi[J][r] = v.im;
        }
    }
    if (G) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int p = 0;
#pragma unroll
        for (int I = 0; I < NT; ++I)
#pragma unroll
          for (int J = I; J < NT; ++J, ++p) {
            // G_IJ += conj(Y_I)^T Y_J over rows 4 r .. 4 r + 3 of the tile
            gr[p] = mfma64(yr[I][r], yr[J][r], gr[p]);
            gr[p] = mfma64(yi[I][r], yi[J][r], gr[p]);
            gi[p] = mfma64(yr[I][r], yi[J][r], gi[p]);
            gi[p] = mfma64(-yi[I][r], yr[J][r], gi[p]);
          }
      }
    }
  }
  if (!G) {
    __syncthreads();
    return 0.0;
  }
  for (int w = 0; w < QR_WAVES; ++w) {
    if (wave == w) {
      int p = 0;
#pragma unroll
      for (int I = 0; I < NT; ++I)
#pragma unroll
        for (int J = I; J < NT; ++J, ++p)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int i = 16 * I + 4 * r + q, j = 16 * J + l16;
            if (i < n && j < n && j >= i) {
              const cplx v = cmake(gr[p][r], gi[p][r]);
              G[i * ldg + j] = w == 0 ? v : cadd(G[i * ldg + j], v);
            }
          }
    }
    __syncthreads();
  }
  double tr = 0.0;
  for (int i = 0; i < n; ++i) tr += G[i * ldg + i].re;
  return tr;
}

// In-place upper Cholesky G = R^H R (R in the upper triangle of G), one barrier
// per step: step k updates the trailing block with the unscaled row k
// (G_ij -= conj(G_ki) G_kj / G_kk) while row k - 1 is scaled by 1 / R_{k-1,k-1}.
// Returns sum_i log R_ii, or NaN if G is not numerically positive definite.
__device__ double qr_cholesky(cplx* G, int ldg, int n) {
  const int tid = threadIdx.x;
  double logdet = 0.0, dprev = 0.0;
  for (int k = 0; k <= n; ++k) {
    if (k > 0) {
      const double inv = 1.0 / dprev;
      for (int j = k + tid; j < n; j += QR_THREADS) G[(k - 1) * ldg + j] = cscale(G[(k - 1) * ldg + j], inv);
      if (tid == 0) G[(k - 1) * ldg + k - 1] = cmake(dprev, 0.0);
    }
    if (k == n) break;
    const double dkk = G[k * ldg + k].re;  // the same value in every thread
    if (!(dkk > 0.0)) {
      __syncthreads();
      return NAN;
    }
    const double d = sqrt(dkk), idkk = 1.0 / dkk;
    logdet += log(d);
    // trailing block: thread (tid / 16, tid % 16) takes rows k + 1 + tid / 16 + 16 a, columns
    // k + 1 + tid % 16 + 16 b (no integer division)
    for (int i = k + 1 + (tid >> 4); i < n; i += QR_THREADS / 16) {
      const cplx gki = G[k * ldg + i];
      int j = k + 1 + (tid & 15);
      if (j < i) j += (i - j + 15) / 16 * 16;  // first column >= i of this thread's residue class
      for (; j < n; j += 16) {
        const cplx gkj = G[k * ldg + j];
        cplx g = G[i * ldg + j];
        g.re -= (gki.re * gkj.re + gki.im * gkj.im) * idkk;
        g.im -= (gki.re * gkj.im - gki.im * gkj.re) * idkk;
        G[i * ldg + j] = g;
      }
    }
    dprev = d;
    __syncthreads();
  }
  __syncthreads();
  return logdet;
}

// ||R||_F ||R^{-1}||_F from the upper triangles (an upper bound of kappa_2(R)),
// the same value in every thread (fixed-order sums).
__device__ double qr_kappa_f(const cplx* R, const cplx* Ri, int ld, int n) {
  __shared__ double part[2][QR_WAVES];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double a = 0.0, b = 0.0;
  for (int e = threadIdx.x; e < n * n; e += QR_THREADS) {
    const int i = e / n, j = e % n;
    if (j >= i) {
      const cplx r = R[i * ld + j], x = Ri[i * ld + j];
      a += r.re * r.re + r.im * r.im;
      b += x.re * x.re + x.im * x.im;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_down(a, off, 64);
    b += __shfl_down(b, off, 64);
  }
  if (lane == 0) {
    part[0][wave] = a;
    part[1][wave] = b;
  }
  __syncthreads();
  double sa = 0.0, sb = 0.0;
  for (int w = 0; w < QR_WAVES; ++w) {
    sa += part[0][w];
    sb += part[1][w];
  }
  __syncthreads();  // part is rewritten by the next call
  return sqrt(sa * sb);
}

// Ri = R^{-1} (upper triangular, zero below and in the padding rows and
// columns up to np), one column per thread.
__device__ void qr_upper_inverse(const cplx* R, int ldr, int n, int np, cplx* Ri, int ldi) {
  const int j = threadIdx.x;
  if (j < np) {
    for (int i = np - 1; i > j || (j >= n && i >= 0); --i) Ri[i * ldi + j] = cmake(0.0, 0.0);
  }
  if (j < n) {
    for (int i = j; i >= 0; --i) {
      // four independent partial sums, loads issued four at a time
      cplx s4[4] = {cmake(i == j ? 1.0 : 0.0, 0.0), cmake(0.0, 0.0), cmake(0.0, 0.0), cmake(0.0, 0.0)};
      int k = i + 1;
      for (; k + 3 <= j; k += 4) {
        cplx r[4], x[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          r[u] = R[i * ldr + k + u];
          x[u] = Ri[(k + u) * ldi + j];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          s4[u].re -= r[u].re * x[u].re - r[u].im * x[u].im;
          s4[u].im -= r[u].re * x[u].im + r[u].im * x[u].re;
        }
      }
      for (; k <= j; ++k) {
        const cplx r = R[i * ldr + k], x = Ri[k * ldi + j];
        s4[0].re -= r.re * x.re - r.im * x.im;
        s4[0].im -= r.re * x.im + r.im * x.re;
      }
      const cplx t = cadd(cadd(s4[0], s4[1]), cadd(s4[2], s4[3]));
      Ri[i * ldi + j] = cscale(t, 1.0 / R[i * ldr + i].re);
    }
  }
  __syncthreads();
}

template <int NT>
__global__ void __launch_bounds__(QR_THREADS) reortho_kernel(ReorthoArgs a) {
  extern __shared__ cplx smem[];
  const int w = blockIdx.x;
  const int N = a.nbasis;
  const int ldg = qr_ld(NT);
  cplx* G = smem;                  // [16 NT][ldg]: Gram matrix, then R
  cplx* Ri = G + 16 * NT * ldg;    // [16 NT][ldg]: R^{-1}, zero padded
  const int nspin = a.nbeta > 0 ? 2 : 1;
  double logdet = 0.0;
  bool bad = false;
  constexpr double u = 1.1102230246251565e-16;  // 2^-53
  for (int s = 0; s < nspin; ++s) {
    const int n = s == 0 ? a.nalpha : a.nbeta;
    double* phi = (s == 0 ? a.phia : a.phib) + 2 * (int64_t)w * N * n;
    const double mn = (double)N * n + (double)n * (n + 1);
    // One loop, one call site per stage (the sweep is inlined once):
    //   apply: X <- X Ri in this sweep;  gram: form G = X^H X in it.
    // The first Gram pass is unshifted; if its Cholesky fails or R fails the
    // CholeskyQR2 bound, the Gram matrix is formed again and shifted.  Then
    // `left` apply + Gram passes (1 for CholeskyQR2, 2 after a shifted pass)
    // and a final apply.
    bool apply = false, gram = true, shifted = false;
    int left = 0;
    for (;;) {
      const double tr = qr_sweep<NT>(phi, N, n, apply ? Ri : nullptr, ldg, gram ? G : nullptr, ldg);
      if (!gram) break;
      if (shifted && !apply) {
        // shift = 11 (N n + n (n + 1)) u ||X||_F^2
        if (threadIdx.x == 0)
          for (int i = 0; i < n; ++i) G[i * ldg + i].re += 11.0 * mn * u * tr;
        __syncthreads();
      }
      const double ldr = qr_cholesky(G, ldg, n);
      const bool ok = ldr == ldr;  // the same in every thread
      if (ok) qr_upper_inverse(G, ldg, n, 16 * NT, Ri, ldg);
      if (!apply && !shifted && !(ok && 8.0 * qr_kappa_f(G, Ri, ldg, n) * sqrt(mn * u) <= 0.5)) {
        shifted = true;
        continue;
      }
      if (!ok) {
        bad = true;
        break;
      }
      logdet += ldr;
      left = apply ? left - 1 : (shifted ? 2 : 1);
      apply = true;
      gram = left > 0;
    }
    if (bad) break;
  }
  if (threadIdx.x == 0) {
    if (bad) {
      if (a.flags) atomicOr(a.flags, AFQMC_FLAG_CHOLESKY);
      if (a.wflags) atomicOr(a.wflags + w, AFQMC_FLAG_CHOLESKY);
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
  const int nt = (nmax + 15) / 16;
  if (nt < 1 || nt > 4) return hipErrorInvalidValue;
  const size_t lds = sizeof(cplx) * 2 * 16 * (size_t)nt * qr_ld(nt);
  const void* k = nt == 1 ? (const void*)reortho_kernel<1> : nt == 2 ? (const void*)reortho_kernel<2>
                : nt == 3 ? (const void*)reortho_kernel<3> : (const void*)reortho_kernel<4>;
  hipError_t e = allow_lds(k, lds);
  if (e != hipSuccess) return e;
  switch (nt) {
    case 1: hipLaunchKernelGGL(reortho_kernel<1>, dim3(a.nwalkers), dim3(QR_THREADS), lds, stream, a); break;
    case 2: hipLaunchKernelGGL(reortho_kernel<2>, dim3(a.nwalkers), dim3(QR_THREADS), lds, stream, a); break;
    case 3: hipLaunchKernelGGL(reortho_kernel<3>, dim3(a.nwalkers), dim3(QR_THREADS), lds, stream, a); break;
    default: hipLaunchKernelGGL(reortho_kernel<4>, dim3(a.nwalkers), dim3(QR_THREADS), lds, stream, a); break;
  }
  return hipGetLastError();
}

}  // namespace afqmc
