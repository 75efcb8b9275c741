// FP64 MFMA contractions of the propagation step.
//
//  * force bias   vbias[w, x] = sum_s sum_k Ghalf_s[w, k] rchol_s[x, k]
//                 (ipie/propagation/force_bias.py:79-121)
//  * VHS          VHS[w] = i sqrt(dt) sum_x xshifted[w, x] L_x, built from the
//                 packed upper triangle and written to both triangles
//                 (ipie/propagation/phaseless_generic.py:67-97,
//                  ipie/utils/pack_numba.py:22-33)
//  * batched NN   C[w] = alpha A[w] B[w] for the one-body propagator
//                 (ipie/propagation/operations.py:26-55) and every Taylor term
//                 (operations.py:85-117)
//
// All use v_mfma_f64_16x16x4f64.  Operands are loaded straight from global
// memory into the fragment registers (each lane reads a contiguous 4-element
// k range), and the next K chunk is fetched while the current one is
// multiplied.  A real operand times a complex operand is two real MFMAs (re,
// im); complex times complex is four.
#include "afqmc_internal.hpp"

namespace afqmc {

namespace {

// ---------------------------------------------------------------- NT mainloop
template <int RM, int RN>
struct NtFrag {
  cplx a[RM][4];
  double b[RN][4];
};

template <int RM, int RN, bool FAST>
__device__ __forceinline__ void nt_load(const NtSegment& sg, int arow0, int Mrows, int bcol0, int Ncols, int kk,
                                        NtFrag<RM, RN>& f) {
  const int l16 = threadIdx.x & 15;
#pragma unroll
  for (int i = 0; i < RM; ++i) {
    const int r = arow0 + 16 * i + l16;
    const bool okr = r < Mrows;
    const double* p = sg.A + 2 * ((int64_t)r * sg.lda + kk);
    if (FAST) {
      if (okr && kk < sg.K) {
        const double2* q = reinterpret_cast<const double2*>(p);
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const double2 v = q[s];
          f.a[i][s] = cplx{v.x, v.y};
        }
      } else {
#pragma unroll
        for (int s = 0; s < 4; ++s) f.a[i][s] = cplx{0.0, 0.0};
      }
    } else {
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        if (okr && kk + s < sg.K) f.a[i][s] = cplx{p[2 * s], p[2 * s + 1]};
        else f.a[i][s] = cplx{0.0, 0.0};
      }
    }
  }
#pragma unroll
  for (int j = 0; j < RN; ++j) {
    const int c = bcol0 + 16 * j + l16;
    const bool okc = c < Ncols;
    const double* p = sg.B + (int64_t)c * sg.ldb + kk;
    if (FAST) {
      if (okc && kk < sg.K) {
        const double2* q = reinterpret_cast<const double2*>(p);
        const double2 v0 = q[0], v1 = q[1];
        f.b[j][0] = v0.x;
        f.b[j][1] = v0.y;
        f.b[j][2] = v1.x;
        f.b[j][3] = v1.y;
      } else {
#pragma unroll
        for (int s = 0; s < 4; ++s) f.b[j][s] = 0.0;
      }
    } else {
#pragma unroll
      for (int s = 0; s < 4; ++s) f.b[j][s] = (okc && kk + s < sg.K) ? p[s] : 0.0;
    }
  }
}

template <int RM, int RN>
__device__ __forceinline__ void nt_mma(const NtFrag<RM, RN>& f, v4d (&re)[RM][RN], v4d (&im)[RM][RN]) {
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int i = 0; i < RM; ++i)
#pragma unroll
      for (int j = 0; j < RN; ++j) {
        re[i][j] = mfma64(f.a[i][s].re, f.b[j][s], re[i][j]);
        im[i][j] = mfma64(f.a[i][s].im, f.b[j][s], im[i][j]);
      }
}

// Accumulates chunks [c_begin, c_end) of a segment into (re, im).
template <int RM, int RN, bool FAST>
__device__ __forceinline__ void nt_mainloop(const NtSegment& sg, int arow0, int Mrows, int bcol0, int Ncols,
                                            int c_begin, int c_end, v4d (&re)[RM][RN], v4d (&im)[RM][RN]) {
  if (c_begin >= c_end) return;
  const int q = (threadIdx.x & 63) >> 4;
  NtFrag<RM, RN> cur, nxt;
  nt_load<RM, RN, FAST>(sg, arow0, Mrows, bcol0, Ncols, 16 * c_begin + 4 * q, cur);
  for (int c = c_begin; c < c_end; ++c) {
    if (c + 1 < c_end) nt_load<RM, RN, FAST>(sg, arow0, Mrows, bcol0, Ncols, 16 * (c + 1) + 4 * q, nxt);
    nt_mma<RM, RN>(cur, re, im);
    if (c + 1 < c_end) cur = nxt;
  }
}

// ---------------------------------------------------------------- force bias
constexpr int FB_RM = 2, FB_RN = 4, FB_WM = 2, FB_WN = 2;
constexpr int FB_TM = 16 * FB_RM * FB_WM;  // 64 walkers
constexpr int FB_TN = 16 * FB_RN * FB_WN;  // 128 fields

template <bool FAST>
__global__ void __launch_bounds__(64 * FB_WM * FB_WN) force_bias_kernel(ForceBiasArgs a) {
  const int nbm = (int)ceil_div(a.nwalkers, FB_TM);
  const int nbn = (int)ceil_div(a.nchol, FB_TN);
  const int lin = xcd_remap(blockIdx.x, gridDim.x);
  const int bm = lin % nbm, bn = lin / nbm;
  const int z = blockIdx.y;
  const int wave = threadIdx.x >> 6;
  const int wm = wave % FB_WM, wn = wave / FB_WM;
  const int arow0 = bm * FB_TM + wm * 16 * FB_RM;
  const int bcol0 = bn * FB_TN + wn * 16 * FB_RN;
  (void)nbn;
  v4d re[FB_RM][FB_RN], im[FB_RM][FB_RN];
#pragma unroll
  for (int i = 0; i < FB_RM; ++i)
#pragma unroll
    for (int j = 0; j < FB_RN; ++j) {
      re[i][j] = zero4d();
      im[i][j] = zero4d();
    }
  for (int s = 0; s < a.nseg; ++s) {
    const NtSegment& sg = a.seg[s];
    const int nch = (int)ceil_div(sg.K, 16);
    const int cb = (int)((int64_t)nch * z / a.ksplit), ce = (int)((int64_t)nch * (z + 1) / a.ksplit);
    nt_mainloop<FB_RM, FB_RN, FAST>(sg, arow0, a.nwalkers, bcol0, a.nchol, cb, ce, re, im);
  }
  const int lane = threadIdx.x & 63;
  double* out = a.vbias_part + 2 * (int64_t)z * a.nwalkers * a.nchol;
#pragma unroll
  for (int i = 0; i < FB_RM; ++i)
#pragma unroll
    for (int j = 0; j < FB_RN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int w = arow0 + 16 * i + 4 * r + (lane >> 4);
        const int x = bcol0 + 16 * j + (lane & 15);
        if (w < a.nwalkers && x < a.nchol)
          st(out, (int64_t)w * a.nchol + x, cmake(a.scale * re[i][j][r], a.scale * im[i][j][r]));
      }
}

// ---------------------------------------------------------------- VHS
constexpr int VH_RM = 2, VH_RN = 4, VH_WM = 2, VH_WN = 2;
constexpr int VH_TM = 16 * VH_RM * VH_WM;  // 64 walkers
constexpr int VH_TN = 16 * VH_RN * VH_WN;  // 128 packed indices

template <bool FAST>
__global__ void __launch_bounds__(64 * VH_WM * VH_WN) vhs_kernel(VhsArgs a) {
  const int P = a.nbasis * (a.nbasis + 1) / 2;
  const int nbm = (int)ceil_div(a.nwalkers, VH_TM);
  const int lin = xcd_remap(blockIdx.x, gridDim.x);
  const int bm = lin % nbm, bn = lin / nbm;
  const int wave = threadIdx.x >> 6;
  const int wm = wave % VH_WM, wn = wave / VH_WM;
  const int arow0 = bm * VH_TM + wm * 16 * VH_RM;
  const int bcol0 = bn * VH_TN + wn * 16 * VH_RN;
  NtSegment sg;
  sg.A = a.xshifted;
  sg.lda = a.nchol;
  sg.B = a.chol_packed;
  sg.ldb = a.nchol;
  sg.K = a.nchol;
  v4d re[VH_RM][VH_RN], im[VH_RM][VH_RN];
#pragma unroll
  for (int i = 0; i < VH_RM; ++i)
#pragma unroll
    for (int j = 0; j < VH_RN; ++j) {
      re[i][j] = zero4d();
      im[i][j] = zero4d();
    }
  nt_mainloop<VH_RM, VH_RN, FAST>(sg, arow0, a.nwalkers, bcol0, P, 0, (int)ceil_div(a.nchol, 16), re, im);
  const int lane = threadIdx.x & 63;
  const int64_t NN = (int64_t)a.nbasis * a.nbasis;
#pragma unroll
  for (int j = 0; j < VH_RN; ++j) {
    const int p = bcol0 + 16 * j + (lane & 15);
    if (p >= P) continue;
    const int ii = a.tri_i[p], jj = a.tri_j[p];
#pragma unroll
    for (int i = 0; i < VH_RM; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int w = arow0 + 16 * i + 4 * r + (lane >> 4);
        if (w >= a.nwalkers) continue;
        // i sqrt(dt) (re + i im) = sqrt(dt) (-im + i re)
        const cplx v = cmake(-a.sqrt_dt * im[i][j][r], a.sqrt_dt * re[i][j][r]);
        double* base = a.vhs + 2 * (int64_t)w * NN;
        st(base, (int64_t)ii * a.nbasis + jj, v);
        if (ii != jj) st(base, (int64_t)jj * a.nbasis + ii, v);
      }
  }
}

// ---------------------------------------------------------------- batched NN
constexpr int NN_WAVES = 4;  // 64 rows per workgroup

template <int RN, bool ACPLX, bool FAST>
__global__ void __launch_bounds__(64 * NN_WAVES) batched_nn_kernel(BatchedNNArgs a) {
  const int nrb = (int)ceil_div(a.nrows, 16 * NN_WAVES);
  const int lin = xcd_remap(blockIdx.x, gridDim.x);
  const int w = lin / nrb, rb = lin % nrb;
  if (w >= a.nwalkers) return;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int l16 = lane & 15, q = lane >> 4;
  const int row0 = rb * 16 * NN_WAVES + wave * 16;
  if (row0 >= a.nrows) return;
  const int K = a.nrows;
  const int arow = row0 + l16;
  const bool okr = arow < a.nrows;
  const double* Ab = a.A + (ACPLX ? 2 : 1) * ((int64_t)w * a.a_stride + (int64_t)arow * K);
  const double* Bb = a.B + 2 * (int64_t)w * a.b_stride;
  v4d re[RN], im[RN];
#pragma unroll
  for (int j = 0; j < RN; ++j) {
    re[j] = zero4d();
    im[j] = zero4d();
  }
  const int nch = (int)ceil_div(K, 16);
  for (int c = 0; c < nch; ++c) {
    const int kk = 16 * c + 4 * q;
    cplx af[4];
    if (FAST) {
      if (okr && kk < K) {
        if (ACPLX) {
          const double2* p = reinterpret_cast<const double2*>(Ab + 2 * kk);
#pragma unroll
          for (int s = 0; s < 4; ++s) {
            const double2 v = p[s];
            af[s] = cplx{v.x, v.y};
          }
        } else {
          const double2* p = reinterpret_cast<const double2*>(Ab + kk);
          const double2 v0 = p[0], v1 = p[1];
          af[0] = cplx{v0.x, 0.0};
          af[1] = cplx{v0.y, 0.0};
          af[2] = cplx{v1.x, 0.0};
          af[3] = cplx{v1.y, 0.0};
        }
      } else {
#pragma unroll
        for (int s = 0; s < 4; ++s) af[s] = cplx{0.0, 0.0};
      }
    } else {
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        if (okr && kk + s < K) {
          if (ACPLX) af[s] = cplx{Ab[2 * (kk + s)], Ab[2 * (kk + s) + 1]};
          else af[s] = cplx{Ab[kk + s], 0.0};
        } else {
          af[s] = cplx{0.0, 0.0};
        }
      }
    }
#pragma unroll
    for (int j = 0; j < RN; ++j) {
      const int col = 16 * j + l16;
      cplx bf[4];
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int k = kk + s;
        if (col < a.ncol && k < K) bf[s] = ld(Bb, (int64_t)k * a.ldb + col);
        else bf[s] = cplx{0.0, 0.0};
      }
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        re[j] = mfma64(af[s].re, bf[s].re, re[j]);
        im[j] = mfma64(af[s].re, bf[s].im, im[j]);
        if (ACPLX) {
          re[j] = mfma64(-af[s].im, bf[s].im, re[j]);
          im[j] = mfma64(af[s].im, bf[s].re, im[j]);
        }
      }
    }
  }
  double* Cb = a.C + 2 * (int64_t)w * a.c_stride;
  double* Sb = a.accum ? a.accum + 2 * (int64_t)w * a.c_stride : nullptr;
  const double* Ib = a.accum_init ? a.accum_init + 2 * (int64_t)w * a.c_stride : nullptr;
#pragma unroll
  for (int j = 0; j < RN; ++j) {
    const int col = 16 * j + l16;
    if (col >= a.ncol) continue;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = row0 + 4 * r + q;
      if (row >= a.nrows) continue;
      const int64_t off = (int64_t)row * a.ldc + col;
      const cplx v = cmake(a.alpha * re[j][r], a.alpha * im[j][r]);
      st(Cb, off, v);
      if (Sb) {
        const cplx base = Ib ? ld(Ib, off) : ld(Sb, off);
        st(Sb, off, cadd(base, v));
      }
    }
  }
}

template <int RN, bool ACPLX>
hipError_t launch_nn_rn(const BatchedNNArgs& a, bool fast, hipStream_t stream) {
  const int nrb = (int)ceil_div(a.nrows, 16 * NN_WAVES);
  const dim3 grid(a.nwalkers * nrb);
  if (fast)
    hipLaunchKernelGGL((batched_nn_kernel<RN, ACPLX, true>), grid, dim3(64 * NN_WAVES), 0, stream, a);
  else
    hipLaunchKernelGGL((batched_nn_kernel<RN, ACPLX, false>), grid, dim3(64 * NN_WAVES), 0, stream, a);
  return hipGetLastError();
}

template <bool ACPLX>
hipError_t launch_nn(const BatchedNNArgs& a, bool fast, hipStream_t stream) {
  const int rn = (int)ceil_div(a.ncol, 16);
  switch (rn) {
    case 1: return launch_nn_rn<1, ACPLX>(a, fast, stream);
    case 2: return launch_nn_rn<2, ACPLX>(a, fast, stream);
    case 3: return launch_nn_rn<3, ACPLX>(a, fast, stream);
    case 4: return launch_nn_rn<4, ACPLX>(a, fast, stream);
    case 5: return launch_nn_rn<5, ACPLX>(a, fast, stream);
    case 6: return launch_nn_rn<6, ACPLX>(a, fast, stream);
    case 7: return launch_nn_rn<7, ACPLX>(a, fast, stream);
    case 8: return launch_nn_rn<8, ACPLX>(a, fast, stream);
    default: return hipErrorInvalidValue;
  }
}

inline bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

}  // namespace

hipError_t launch_force_bias(const ForceBiasArgs& a, hipStream_t stream) {
  bool fast = true;
  for (int s = 0; s < a.nseg; ++s)
    fast = fast && (a.seg[s].K % 4 == 0) && (a.seg[s].ldb % 2 == 0) && aligned16(a.seg[s].A) && aligned16(a.seg[s].B);
  const int nbm = (int)ceil_div(a.nwalkers, FB_TM), nbn = (int)ceil_div(a.nchol, FB_TN);
  const dim3 grid(nbm * nbn, a.ksplit);
  if (fast)
    hipLaunchKernelGGL(force_bias_kernel<true>, grid, dim3(64 * FB_WM * FB_WN), 0, stream, a);
  else
    hipLaunchKernelGGL(force_bias_kernel<false>, grid, dim3(64 * FB_WM * FB_WN), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_vhs(const VhsArgs& a, hipStream_t stream) {
  const int P = a.nbasis * (a.nbasis + 1) / 2;
  const bool fast = (a.nchol % 4 == 0) && aligned16(a.xshifted) && aligned16(a.chol_packed);
  const int nbm = (int)ceil_div(a.nwalkers, VH_TM), nbn = (int)ceil_div(P, VH_TN);
  const dim3 grid(nbm * nbn);
  if (fast)
    hipLaunchKernelGGL(vhs_kernel<true>, grid, dim3(64 * VH_WM * VH_WN), 0, stream, a);
  else
    hipLaunchKernelGGL(vhs_kernel<false>, grid, dim3(64 * VH_WM * VH_WN), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_batched_nn(const BatchedNNArgs& a, hipStream_t stream) {
  const bool fast = (a.nrows % 4 == 0) && aligned16(a.A) && (a.a_complex || (a.nrows % 2 == 0));
  if (a.a_complex) return launch_nn<true>(a, fast, stream);
  return launch_nn<false>(a, fast, stream);
}

}  // namespace afqmc
