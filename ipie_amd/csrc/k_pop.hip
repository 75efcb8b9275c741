// Population control for the walkers of one device (one rank).
//
//   pop_plan_kernel    one workgroup: total |w|, rescaling to the target weight,
//                      then the method's decisions as a source map src[i]
//                      (slot i takes walker src[i]'s buffer) and the weights
//                      afterwards:
//                        pair_branch  stable sort of |w| (bitonic on the key
//                                     (|w|, index)), lightest and heaviest
//                                     paired while outside [min, max]; the
//                                     k-th branched walker (slot order) fills
//                                     the k-th killed slot
//                        comb         systematic comb over the cumulative
//                                     weights, the i-th clone fills the i-th
//                                     killed slot, weights reset to 1
//                        stochastic reconfiguration   searchsorted of
//                                     (i + zeta) / W of the total
//                        rescale      the rescaling only
//   pop_gather_kernel  out[i] = in[src[i]] for walker rows of any size
//   walker_sums_kernel the per-step sums of ipie's WalkerAccumulator.update
//                      (base_walkers.py:32-50): sum w, sum unscaled w,
//                      sum w E, added to a device accumulator
//   clip_weights_kernel the driver's weight clip (afqmc.py:584-588)
//
// Restates ipie/walkers/pop_controller.py:78-126 (rescale), 304-402
// (pair_branch), 199-301 (comb), 405-450 (stochastic_reconfiguration) for
// comm.size == 1.  The uniforms are passed in, consumed in the reference's
// order (one per branched pair, one per comb / reconfiguration).  Sums and
// cumulative sums run serially left to right, the order numpy.cumsum uses
// (numpy.sum of the total sums pairwise: the last bits may differ).
//
// pair_branch on one rank: every walker message carries tag 0 (the tags are
// built from rank indices, pop_controller.py:398-420), and MPI delivers
// messages of one source and tag in the order sent, so the walkers branched
// are sent in slot order and received by the killed slots in slot order: the
// k-th branched walker fills the k-th killed slot, whatever pair each came
// from, and carries its own new weight.
#include "afqmc_internal.hpp"

namespace afqmc {

namespace {

constexpr int POP_THREADS = 1024;

__global__ void __launch_bounds__(POP_THREADS) pop_plan_kernel(PopPlanArgs a) {
  __shared__ double key[POP_MAX];   // |scaled w| (sorted for pair_branch; cumulative for comb / SR)
  __shared__ int idx[POP_MAX];      // walker index of key (pair_branch); parent count (comb)
  __shared__ unsigned char stat[POP_MAX];  // pair_branch: 1 live, 2 branched, 0 killed
  __shared__ double s_total, s_scale;
  const int tid = threadIdx.x, nw = a.nwalkers;
  if (tid == 0) {
    double t = 0.0;
    for (int i = 0; i < nw; ++i) t += fabs(a.weight[i]);
    s_total = t;
    s_scale = t / a.target;
    a.total[0] = t;
  }
  __syncthreads();
  const double total = s_total, scale = s_scale;
  const bool ok = total >= 1e-8;
  for (int i = tid; i < nw; i += POP_THREADS) {
    const double ws = ok ? a.weight[i] / scale : a.weight[i];
    a.new_weight[i] = ws;
    a.src[i] = i;
  }
  if (!ok) {  // the caller raises on total < 1e-8; weights and map untouched
    if (tid == 0) a.npairs[0] = 0;
    return;
  }
  __syncthreads();
  if (a.method == POP_RESCALE) {
    if (tid == 0) a.npairs[0] = 0;
  } else if (a.method == POP_PAIR_BRANCH) {
    int np2 = 1;
    while (np2 < nw) np2 <<= 1;
    for (int i = tid; i < np2; i += POP_THREADS) {
      key[i] = i < nw ? fabs(a.new_weight[i]) : INFINITY;
      idx[i] = i;
      stat[i] = 1;
    }
    __syncthreads();
    // bitonic sort ascending by (key, index): equal keys keep the input order
    for (int size = 2; size <= np2; size <<= 1)
      for (int stride = size >> 1; stride > 0; stride >>= 1) {
        for (int i = tid; i < np2; i += POP_THREADS) {
          const int j = i ^ stride;
          if (j > i) {
            const bool up = (i & size) == 0;
            const bool gt = key[i] > key[j] || (key[i] == key[j] && idx[i] > idx[j]);
            if (gt == up) {
              const double tk = key[i];
              key[i] = key[j];
              key[j] = tk;
              const int ti = idx[i];
              idx[i] = idx[j];
              idx[j] = ti;
            }
          }
        }
        __syncthreads();
      }
    if (tid == 0) {
      int s = 0, e = nw - 1, k = 0;
      while (s < e) {
        const double ws = key[s], we = key[e];
        if (!(ws < a.min_weight || we > a.max_weight)) break;
        const double wab = ws + we;
        const double r = a.uniforms[k++];
        const int is = idx[s], ie = idx[e];
        const bool heavy = r < we / wab;  // clone the heavy walker, kill the light one
        const int ib = heavy ? ie : is;
        stat[ib] = 2;
        stat[heavy ? is : ie] = 0;
        a.new_weight[ib] = 0.5 * wab;
        ++s;
        --e;
      }
      a.npairs[0] = k;
      int kb = 0, kk = 0;
      for (int p = 0; p < k; ++p) {
        while (stat[kb] != 2) ++kb;
        while (stat[kk] != 0) ++kk;
        a.src[kk] = kb;
        a.new_weight[kk] = a.new_weight[kb];
        ++kb;
        ++kk;
      }
    }
  } else if (a.method == POP_COMB) {
    if (tid == 0) {
      // cumulative scaled weights and the comb (i + r) total / target
      double c = 0.0;
      for (int i = 0; i < nw; ++i) {
        c += fabs(a.new_weight[i]);
        key[i] = c;
        idx[i] = 0;
      }
      double tw = 0.0;
      for (int i = 0; i < nw; ++i) tw += fabs(a.new_weight[i]);
      const int target = (int)a.target;
      const double r = a.uniforms[0];
      int iw = 0, ic = 0;
      while (ic < target && iw < nw) {
        const double point = (ic + r) * (tw / a.target);
        if (point < key[iw]) {
          ++idx[iw];
          ++ic;
        } else {
          ++iw;
        }
      }
      // zip(clone, kill): the i-th walker with more than one parent point fills
      // the i-th walker with none
      int kc = 0, kk = 0;
      while (true) {
        while (kc < nw && idx[kc] <= 1) ++kc;
        while (kk < nw && idx[kk] != 0) ++kk;
        if (kc >= nw || kk >= nw) break;
        a.src[kk] = kc;
        ++kc;
        ++kk;
      }
      a.npairs[0] = 0;
    }
    __syncthreads();
    for (int i = tid; i < nw; i += POP_THREADS) a.new_weight[i] = 1.0;
  } else {  // POP_RECONFIGURATION
    if (tid == 0) {
      double c = 0.0;
      for (int i = 0; i < nw; ++i) {
        c += fabs(a.new_weight[i]);
        key[i] = c;
      }
      a.npairs[0] = 0;
    }
    __syncthreads();
    const double tw = key[nw - 1], zeta = a.uniforms[0];
    for (int i = tid; i < nw; i += POP_THREADS) {
      const double z = (i + zeta) / nw * tw;
      int lo = 0, hi = nw;  // first j with key[j] >= z (numpy.searchsorted, side="left")
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (key[mid] < z) lo = mid + 1;
        else hi = mid;
      }
      a.src[i] = lo < nw ? lo : nw - 1;
    }
    __syncthreads();
    for (int i = tid; i < nw; i += POP_THREADS) a.new_weight[i] = tw / nw;
  }
}

// out row i = in row src[i] (bytes_per_walker each); rows with src[i] == i are
// skipped when in == out, rows whose src[i] is outside [0, nw) are left alone.  16-byte vectors when rows and pointers allow.
__global__ void __launch_bounds__(256) pop_gather_kernel(int nw, const int* __restrict__ src, const char* in,
                                                        char* out, int64_t row_bytes) {
  const int i = blockIdx.y;
  if (i >= nw) return;
  const int s = src[i];
  if (s < 0 || s >= nw || (in == out && s == i)) return;  // a map entry out of range copies nothing
  const char* from = in + (int64_t)s * row_bytes;
  char* to = out + (int64_t)i * row_bytes;
  if ((row_bytes & 15) == 0 && (reinterpret_cast<uintptr_t>(in) & 15) == 0 &&
      (reinterpret_cast<uintptr_t>(out) & 15) == 0) {
    const int64_t nv = row_bytes / 16;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += (int64_t)gridDim.x * blockDim.x)
      reinterpret_cast<uint4*>(to)[v] = reinterpret_cast<const uint4*>(from)[v];
  } else {
    for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b < row_bytes; b += (int64_t)gridDim.x * blockDim.x)
      to[b] = from[b];
  }
}

constexpr int SUM_THREADS = 256;

// One workgroup: thread t sums walkers t, t + 256, ... in order, then a fixed
// tree over the threads, so the result does not depend on scheduling.
__global__ void __launch_bounds__(SUM_THREADS) walker_sums_kernel(int nw, const double* weight, const double* unscaled,
                                                                  const double* values, int64_t vstride, double* acc) {
  __shared__ double red[5][SUM_THREADS];
  const int t = threadIdx.x;
  double sw = 0.0, su = 0.0, sr = 0.0, si = 0.0;
  for (int i = t; i < nw; i += SUM_THREADS) {
    const double w = weight[i];
    sw += w;
    if (unscaled) su += unscaled[i];
    if (values) {
      sr += w * values[2 * (int64_t)i * vstride];
      si += w * values[2 * (int64_t)i * vstride + 1];
    }
  }
  red[0][t] = sw;
  red[1][t] = su;
  red[2][t] = sr;
  red[3][t] = si;
  __syncthreads();
  for (int h = SUM_THREADS / 2; h > 0; h >>= 1) {
    if (t < h)
      for (int q = 0; q < 4; ++q) red[q][t] += red[q][t + h];
    __syncthreads();
  }
  if (t == 0) {
    acc[0] += red[0][0];
    if (unscaled) acc[2] += red[1][0];
    if (values) {
      acc[4] += red[2][0];
      acc[5] += red[3][0];
    }
  }
}

__global__ void clip_weights_kernel(int nw, double* weight, double bound) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nw) return;
  double w = weight[i];
  w = w < -bound ? -bound : w;  // comparisons are false for NaN: NaN is kept, as numpy.clip does
  w = w > bound ? bound : w;
  weight[i] = w;
}

}  // namespace

hipError_t launch_walker_sums(int nw, const double* weight, const double* unscaled, const double* values,
                              int64_t vstride, double* acc, hipStream_t stream) {
  if (nw < 1 || !weight || !acc || vstride < 1) return hipErrorInvalidValue;
  hipLaunchKernelGGL(walker_sums_kernel, dim3(1), dim3(SUM_THREADS), 0, stream, nw, weight, unscaled, values, vstride,
                     acc);
  return hipGetLastError();
}

hipError_t launch_clip_weights(int nw, double* weight, double bound, hipStream_t stream) {
  if (nw < 1 || !weight) return hipErrorInvalidValue;
  hipLaunchKernelGGL(clip_weights_kernel, dim3((unsigned)ceil_div(nw, 256)), dim3(256), 0, stream, nw, weight, bound);
  return hipGetLastError();
}

hipError_t launch_pop_plan(const PopPlanArgs& a, hipStream_t stream) {
  if (a.nwalkers < 1 || a.nwalkers > POP_MAX || !(a.target > 0.0) || !a.weight || !a.src || !a.new_weight ||
      !a.total || !a.npairs || (!a.uniforms && a.method != POP_RESCALE) || a.method < POP_PAIR_BRANCH ||
      a.method > POP_RESCALE)
    return hipErrorInvalidValue;
  hipLaunchKernelGGL(pop_plan_kernel, dim3(1), dim3(POP_THREADS), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_pop_gather(int nw, const int* src, const void* in, void* out, int64_t row_bytes,
                             hipStream_t stream) {
  if (nw < 1 || row_bytes < 1 || !src || !in || !out) return hipErrorInvalidValue;
  const int64_t per = (row_bytes + 15) / 16;
  const unsigned gx = (unsigned)std::min<int64_t>(std::max<int64_t>(ceil_div(per, 256), 1), 64);
  hipLaunchKernelGGL(pop_gather_kernel, dim3(gx, (unsigned)nw), dim3(256), 0, stream, nw, src,
                     static_cast<const char*>(in), static_cast<char*>(out), row_bytes);
  return hipGetLastError();
}

}  // namespace afqmc
