// Learned control variate for the exchange energy.
//
// The exchange sum E_x(w) = sum_x tr(T_x T_x), T_x = L_x G_w^T (complex, both
// spins), is expanded about the trial's half-rotated Green's function G_T:
//     E_x(G) = E_T + 2 <K, dG> + Q(dG),   dG = G - G_T,
//     E_T = sum_x tr(T0_x T0_x),  T0_x = L_x G_T^T,  K = sum_x T0_x L_x,
// with Q(dG) = sum_x tr(dT_x dT_x), dT_x = L_x dG^T.  E_T and K are computed
// once in FP64; per step the linear term is an FP64 dot product and Q is
// evaluated with single BF16 products (1 MFMA per product instead of 3).  The
// predictor is g(w) = E_T + 2 <K, dG_w> + Q_bf16(w) + alpha m(f_w), where m is
// a small MLP, m(f) = W3 silu(W2 silu(W1 f + b1) + b2) + b3 (BF16 weights, FP32
// accumulation), over cheap per-walker features (one-body energy, Coulomb
// energy, mean and max |xbar|, log|overlap|, overlap phase), trained on the
// residual E_x - (E_T + 2 <K, dG> + Q_bf16).  The gain alpha in [0, 1] is the
// least-squares coefficient of that residual on m (both centred per step) over
// the exact samples of earlier steps (exponentially decayed), so an MLP that
// does not predict the residual is switched off.
//
// The exact exchange is evaluated on a subset S of k walkers in which every
// walker has inclusion probability k / W, and the estimate
//     E_x_hat(w) = g(w) + [w in S] (W / k) (E_x(w) - g(w))
// is unbiased for every walker whatever g is (Horvitz-Thompson).  S is the
// image of [0, k) under w = a^{-1} (t - b) mod W for a random a coprime to W
// and random b, drawn afresh every step.
//
// Training: Adam on the mean squared error of normalised targets, on exact
// samples only (all walkers during warm-up, the subset afterwards); inputs and
// targets are normalised with running statistics accumulated during warm-up.
// Every reduction has a fixed order, so a run is reproducible.
#include "afqmc_internal.hpp"

namespace afqmc {

namespace {

constexpr int SG_H = SUR_HIDDEN;
constexpr int SG_F = SUR_FEATURES;
// samples per gradient workgroup: each walks its samples serially, so few per
// workgroup and many workgroups (partial gradients are summed in fixed order)
constexpr int SG_SAMPLES_PER_WG = 4;
// row stride of the H x H matrices in the gradient kernel's LDS: thread j walks
// row j, so an odd stride puts the 64 lanes of a wave on distinct banks
constexpr int SG_LD = SG_H + 1;

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
  const float s = 1.0f / (1.0f + __expf(-z));
  return s * (1.0f + z * (1.0f - s));
}
__device__ __forceinline__ uint16_t f2bf_rne(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }

// raw features [W][F]
__global__ void sur_features_kernel(SurrogateFeatureArgs a) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= a.nwalkers) return;
  const cplx e1 = ld(a.e1, w), ej = ld(a.ecoul, w), o = ld(a.ovlp, w);
  const double ao = cabs(o);
  float* f = a.feat + (int64_t)w * SG_F;
  f[0] = (float)e1.re;
  f[1] = (float)e1.im;
  f[2] = (float)ej.re;
  f[3] = (float)ej.im;
  f[4] = (float)a.xbar_mean[w];
  f[5] = (float)a.xbar_max[w];
  f[6] = (float)(ao > 0.0 ? log(ao) : -700.0);
  f[7] = (float)(ao > 0.0 ? o.re / ao : 1.0);
  f[8] = (float)(ao > 0.0 ? o.im / ao : 0.0);
}

// Exact exchange sum of sampled walker t (both spins, summed over the partials).
__device__ __forceinline__ cplx exact_sum(const SurrogateCombineArgs& a, int t) {
  cplx ex = cmake(0.0, 0.0);
  for (int z = 0; z < a.xs_a; ++z) ex = cadd(ex, ld(a.part_a, (int64_t)z * a.k + t));
  if (a.part_b)
    for (int z = 0; z < a.xs_b; ++z) ex = cadd(ex, ld(a.part_b, (int64_t)z * a.k + t));
  return ex;
}

// Running statistics (FP64 sums) of the features and the exact targets of
// this batch; one workgroup, fixed reduction order.
__global__ void sur_stats_kernel(SurrogateStatsArgs a) {
  __shared__ double red[256 / 64];
  const int nq = SG_F + 2;
  for (int qd = 0; qd < nq; ++qd) {
    double s1 = 0.0, s2 = 0.0;
    for (int t = threadIdx.x; t < a.count; t += blockDim.x) {
      const int w = a.walker_of ? a.walker_of[t] : t;
      double v;
      if (qd < SG_F) v = a.feat[(int64_t)w * SG_F + qd];
      else v = qd == SG_F ? a.target[2 * t] : a.target[2 * t + 1];
      s1 += v;
      s2 += v * v;
    }
    s1 = block_sum(s1, red);
    s2 = block_sum(s2, red);
    if (threadIdx.x == 0) {
      a.state->sum[qd] += s1;
      a.state->sumsq[qd] += s2;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    a.state->count += a.count;
    const double n = a.state->count;
    for (int qd = 0; qd < nq; ++qd) {
      const double m = a.state->sum[qd] / n;
      const double var = fmax(a.state->sumsq[qd] / n - m * m, 0.0);
      a.state->mean[qd] = (float)m;
      a.state->inv_std[qd] = (float)(1.0 / sqrt(var + 1e-12));
    }
  }
}

// Inference: pred[w] = denormalised g(f_w), BF16 weights, FP32 accumulation.
// One workgroup of H threads per group of walkers.
__global__ void __launch_bounds__(SG_H) sur_forward_kernel(SurrogateForwardArgs a) {
  __shared__ float h1[SG_H];
  __shared__ float red[2][SG_H / 64];
  const int j = threadIdx.x;
  const SurrogateState* sst = a.state;
  const uint16_t* w1 = a.wbf;                    // [H][F]
  const uint16_t* w2 = w1 + SG_H * SG_F;         // [H][H]
  const uint16_t* w3 = w2 + SG_H * SG_H;         // [2][H]
  const float* bias = a.bias;                    // b1 [H], b2 [H], b3 [2] (FP32)
  for (int w = blockIdx.x; w < a.nwalkers; w += gridDim.x) {
    const float* f = a.feat + (int64_t)w * SG_F;
    float z1 = bias[j];
#pragma unroll
    for (int q = 0; q < SG_F; ++q) z1 += bf2f(w1[j * SG_F + q]) * ((f[q] - sst->mean[q]) * sst->inv_std[q]);
    __syncthreads();
    h1[j] = silu(z1);
    __syncthreads();
    float z2 = bias[SG_H + j];
    for (int k = 0; k < SG_H; ++k) z2 += bf2f(w2[j * SG_H + k]) * h1[k];
    const float h2 = silu(z2);
    float o0 = bf2f(w3[j]) * h2, o1 = bf2f(w3[SG_H + j]) * h2;
    for (int off = 32; off > 0; off >>= 1) {
      o0 += __shfl_xor(o0, off, 64);
      o1 += __shfl_xor(o1, off, 64);
    }
    if ((j & 63) == 0) {
      red[0][j >> 6] = o0;
      red[1][j >> 6] = o1;
    }
    __syncthreads();
    if (j == 0) {
      float s0 = bias[2 * SG_H], s1 = bias[2 * SG_H + 1];
      for (int i = 0; i < SG_H / 64; ++i) {
        s0 += red[0][i];
        s1 += red[1][i];
      }
      const double re = sst->mean[SG_F] + (double)s0 / sst->inv_std[SG_F];
      const double im = sst->mean[SG_F + 1] + (double)s1 / sst->inv_std[SG_F + 1];
      st(a.pred, w, cmake(re, im));
    }
  }
}

// Gradient of the batch loss, partial per workgroup: grad[wg][P].
// Loss = mean_t sum_o (g_o(f_t) - y_t,o)^2 on normalised targets.
__global__ void __launch_bounds__(SG_H) sur_grad_kernel(SurrogateTrainArgs a) {
  extern __shared__ float sm[];
  float* W2s = sm;                    // [H][SG_LD] master weights
  float* dW2 = W2s + SG_H * SG_LD;    // [H][SG_LD] gradient accumulator
  float* h1 = dW2 + SG_H * SG_LD;     // [H]
  float* dz2 = h1 + SG_H;             // [H]
  float* xin = dz2 + SG_H;            // [F]
  __shared__ float red[2][SG_H / 64];
  __shared__ float dout[2];
  const int j = threadIdx.x;
  float loss_acc = 0.0f;
  const float* P = a.master;
  const float* W1 = P;                       // [H][F]
  const float* W2 = W1 + SG_H * SG_F;        // [H][H]
  const float* W3 = W2 + SG_H * SG_H;        // [2][H]
  const float* b1 = W3 + 2 * SG_H;           // [H]
  const float* b2 = b1 + SG_H;               // [H]
  const float* b3 = b2 + SG_H;               // [2]
  const SurrogateState* sst = a.state;
  for (int e = j; e < SG_H * SG_H; e += SG_H) {
    W2s[(e / SG_H) * SG_LD + e % SG_H] = W2[e];
    dW2[(e / SG_H) * SG_LD + e % SG_H] = 0.0f;
  }
  float gW1[SG_F];
#pragma unroll
  for (int q = 0; q < SG_F; ++q) gW1[q] = 0.0f;
  float gb1 = 0.0f, gb2 = 0.0f, gW3a = 0.0f, gW3b = 0.0f, gb3a = 0.0f, gb3b = 0.0f;
  float w1row[SG_F];
#pragma unroll
  for (int q = 0; q < SG_F; ++q) w1row[q] = W1[j * SG_F + q];
  const float w3a = W3[j], w3b = W3[SG_H + j];
  __syncthreads();
  const int t0 = blockIdx.x * SG_SAMPLES_PER_WG;
  const int t1 = min(t0 + SG_SAMPLES_PER_WG, a.count);
  const float inv_count = 1.0f / (float)a.count;
  for (int t = t0; t < t1; ++t) {
    const int w = a.walker_of ? a.walker_of[t] : t;
    if (j < SG_F) xin[j] = (a.feat[(int64_t)w * SG_F + j] - sst->mean[j]) * sst->inv_std[j];
    __syncthreads();
    float z1 = b1[j];
#pragma unroll
    for (int q = 0; q < SG_F; ++q) z1 += w1row[q] * xin[q];
    h1[j] = silu(z1);
    __syncthreads();
    float z2 = b2[j];
    for (int k = 0; k < SG_H; ++k) z2 += W2s[j * SG_LD + k] * h1[k];
    const float h2 = silu(z2);
    float o0 = w3a * h2, o1 = w3b * h2;
    for (int off = 32; off > 0; off >>= 1) {
      o0 += __shfl_xor(o0, off, 64);
      o1 += __shfl_xor(o1, off, 64);
    }
    if ((j & 63) == 0) {
      red[0][j >> 6] = o0;
      red[1][j >> 6] = o1;
    }
    __syncthreads();
    if (j == 0) {
      float s0 = b3[0], s1 = b3[1];
      for (int i = 0; i < SG_H / 64; ++i) {
        s0 += red[0][i];
        s1 += red[1][i];
      }
      const float y0 = (float)((a.target[2 * t] - sst->mean[SG_F]) * sst->inv_std[SG_F]);
      const float y1 = (float)((a.target[2 * t + 1] - sst->mean[SG_F + 1]) * sst->inv_std[SG_F + 1]);
      dout[0] = 2.0f * (s0 - y0) * inv_count;
      dout[1] = 2.0f * (s1 - y1) * inv_count;
      loss_acc += (s0 - y0) * (s0 - y0) + (s1 - y1) * (s1 - y1);
    }
    __syncthreads();
    const float d0 = dout[0], d1 = dout[1];
    gW3a += d0 * h2;
    gW3b += d1 * h2;
    if (j == 0) {
      gb3a += d0;
      gb3b += d1;
    }
    const float g2 = (w3a * d0 + w3b * d1) * dsilu(z2);
    dz2[j] = g2;
    gb2 += g2;
    for (int k = 0; k < SG_H; ++k) dW2[j * SG_LD + k] += g2 * h1[k];
    __syncthreads();
    float dh1 = 0.0f;
    for (int r = 0; r < SG_H; ++r) dh1 += W2s[r * SG_LD + j] * dz2[r];
    const float g1 = dh1 * dsilu(z1);
    gb1 += g1;
#pragma unroll
    for (int q = 0; q < SG_F; ++q) gW1[q] += g1 * xin[q];
    __syncthreads();
  }
  float* G = a.grad + (int64_t)blockIdx.x * SUR_NPARAM;
#pragma unroll
  for (int q = 0; q < SG_F; ++q) G[j * SG_F + q] = gW1[q];
  __syncthreads();  // rows of dW2 were accumulated by their own threads; written out by columns
  for (int r = 0; r < SG_H; ++r) G[SG_H * SG_F + r * SG_H + j] = dW2[r * SG_LD + j];
  G[SG_H * SG_F + SG_H * SG_H + j] = gW3a;
  G[SG_H * SG_F + SG_H * SG_H + SG_H + j] = gW3b;
  G[SG_H * SG_F + SG_H * SG_H + 2 * SG_H + j] = gb1;
  G[SG_H * SG_F + SG_H * SG_H + 3 * SG_H + j] = gb2;
  if (j == 0) {
    G[SUR_NPARAM - 2] = gb3a;
    G[SUR_NPARAM - 1] = gb3b;
    a.loss[blockIdx.x] = loss_acc;
  }
}

// Sum the per-workgroup gradients (fixed order), Adam step, refresh the BF16
// inference copy of the weights and the FP32 biases.
__global__ void sur_adam_kernel(SurrogateAdamArgs a) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= SUR_NPARAM) return;
  float g = 0.0f;
  for (int i = 0; i < a.ngroups; ++i) g += a.grad[(int64_t)i * SUR_NPARAM + p];
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float m = b1 * a.m[p] + (1.0f - b1) * g;
  const float v = b2 * a.v[p] + (1.0f - b2) * g * g;
  a.m[p] = m;
  a.v[p] = v;
  const float step = (float)(a.state->adam_t + a.it + 1);
  const float mh = m / (1.0f - powf(b1, step));
  const float vh = v / (1.0f - powf(b2, step));
  const float w = a.master[p] - a.lr * mh / (sqrtf(vh) + eps);
  a.master[p] = w;
  constexpr int NW = SG_H * SG_F + SG_H * SG_H + 2 * SG_H;  // weights, then biases
  if (p < NW) a.wbf[p] = f2bf_rne(w);
  else a.bias[p - NW] = w;
}

// Horvitz-Thompson combination: est[w] = g(w) + [w in S] (W / k) (exact - g(w)).
__global__ void sur_combine_kernel(SurrogateCombineArgs a) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= a.nwalkers) return;
  const cplx g = ld(a.pred, w);
  cplx est = g;
  const int64_t pos = ((int64_t)a.perm->perm_a * w + a.perm->perm_b) % a.nwalkers;
  if (pos < a.k) {
    const cplx ex = exact_sum(a, (int)pos);
    const double s = (double)a.nwalkers / (double)a.k;
    est = cadd(g, cscale(csub(ex, g), s));
    if (a.target) {  // the MLP's target: exact minus the physical part of g
      const cplx base = a.cv ? ld(a.cv, w) : cmake(0.0, 0.0);
      a.target[2 * pos] = ex.re - base.re;
      a.target[2 * pos + 1] = ex.im - base.im;
    }
    if (a.resid) {
      const cplx r = csub(ex, g);
      a.resid[pos] = r.re * r.re + r.im * r.im;
    }
  }
  st(a.est, w, est);
}

// walker index of sample t: w = a^{-1} (t - b) mod W
__global__ void sur_select_kernel(int nwalkers, int k, const StepState* perm, int* walker_of) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= k) return;
  int64_t v = ((int64_t)t - perm->perm_b) % nwalkers;
  if (v < 0) v += nwalkers;
  walker_of[t] = (int)((perm->inv_a * v) % nwalkers);
}

__device__ int64_t gcd_dev(int64_t a, int64_t b) {
  while (b) {
    const int64_t t = a % b;
    a = b;
    b = t;
  }
  return a;
}

// a^{-1} mod m for gcd(a, m) = 1 (extended Euclid)
__device__ int64_t inv_mod_dev(int64_t a, int64_t m) {
  int64_t g = m, x = 0, x1 = 1, a1 = a % m;
  while (a1) {
    const int64_t q = g / a1;
    int64_t t = g - q * a1;
    g = a1;
    a1 = t;
    t = x - q * x1;
    x = x1;
    x1 = t;
  }
  x %= m;
  return x < 0 ? x + m : x;
}

// The step's subset permutation.  The inclusion probability k / W of every
// walker needs only perm_b uniform on [0, W); perm_a (coprime with W) varies
// which walkers are sampled together.  Up to 64 Philox draws for perm_a, then
// perm_a = 1.
__global__ void sur_plan_kernel(StepState* s, uint64_t seed, int nwalkers, int identity) {
  if (threadIdx.x != 0) return;
  int64_t pa = 1, pb = 0;
  if (!identity && nwalkers > 1) {
    const unsigned long long d = s->sur_draws;
    const uint2 key = make_uint2((uint32_t)seed, (uint32_t)(seed >> 32));
    const uint4 rb = philox4x32(make_uint4(0u, 0x5EEDu, (uint32_t)d, (uint32_t)(d >> 32)), key);
    pb = (int64_t)((((uint64_t)rb.x << 32) | rb.y) % (uint64_t)nwalkers);
    for (uint32_t j = 1; j <= 64; ++j) {
      const uint4 r = philox4x32(make_uint4(j, 0x5EEDu, (uint32_t)d, (uint32_t)(d >> 32)), key);
      const int64_t cand = 1 + (int64_t)((((uint64_t)r.x << 32) | r.y) % (uint64_t)(nwalkers - 1));
      if (gcd_dev(cand, nwalkers) == 1) {
        pa = cand;
        break;
      }
    }
    s->sur_draws = d + 1;
  }
  s->perm_a = pa;
  s->perm_b = pb;
  s->inv_a = nwalkers > 1 ? inv_mod_dev(pa, nwalkers) : 0;
}

// Ghalf rows of the sampled walkers, compacted: out[t] = g[walker_of[t]] ([n, N] complex each).
__global__ void sur_gather_kernel(const double* __restrict__ g, const int* walker_of, int k, int64_t per_walker,
                                  double* __restrict__ out) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)k * per_walker) return;
  const int t = (int)(idx / per_walker);
  const int64_t e = idx % per_walker;
  const int64_t src = (int64_t)walker_of[t] * per_walker + e;
  reinterpret_cast<double2*>(out)[idx] = reinterpret_cast<const double2*>(g)[src];
}

// ---------------------------------------------------------------- control variate
// T0[x][i][j] = sum_m L[x][i][m] gT[j][m]   (FP64, once per Hamiltonian / trial)
__global__ void cv_t0_kernel(const double* __restrict__ L, const double* __restrict__ gT, int X, int n, int N,
                             double* __restrict__ T0) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)X * n * n) return;
  const int64_t x = idx / ((int64_t)n * n);
  const int i = (int)((idx / n) % n), j = (int)(idx % n);
  const double* l = L + (x * n + i) * N;
  const double* g = gT + 2 * (int64_t)j * N;
  double re = 0.0, im = 0.0;
  for (int m = 0; m < N; ++m) {
    re += l[m] * g[2 * m];
    im += l[m] * g[2 * m + 1];
  }
  T0[2 * idx] = re;
  T0[2 * idx + 1] = im;
}

// K[i][m] = sum_x sum_j T0[x][i][j] L[x][j][m]
__global__ void cv_k_kernel(const double* __restrict__ T0, const double* __restrict__ L, int X, int n, int N,
                            double* __restrict__ K) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n * N) return;
  const int i = (int)(idx / N), m = (int)(idx % N);
  double re = 0.0, im = 0.0;
  for (int64_t x = 0; x < X; ++x) {
    const double* t = T0 + 2 * (x * n + i) * n;
    const double* l = L + x * n * N + m;
    for (int j = 0; j < n; ++j) {
      const double lv = l[(int64_t)j * N];
      re += t[2 * j] * lv;
      im += t[2 * j + 1] * lv;
    }
  }
  K[2 * idx] = re;
  K[2 * idx + 1] = im;
}

// part[x] = sum_ij T0[x][i][j] T0[x][j][i]
__global__ void cv_et_part_kernel(const double* __restrict__ T0, int X, int n, double* __restrict__ part) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= X) return;
  const double* t = T0 + 2 * (int64_t)x * n * n;
  double re = 0.0, im = 0.0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      const cplx a = ld(t, (int64_t)i * n + j), b = ld(t, (int64_t)j * n + i);
      re += a.re * b.re - a.im * b.im;
      im += a.re * b.im + a.im * b.re;
    }
  part[2 * x] = re;
  part[2 * x + 1] = im;
}

// et = sum_x part[x], in order (one thread)
__global__ void cv_et_sum_kernel(const double* __restrict__ part, int X, double* __restrict__ et) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double re = 0.0, im = 0.0;
  for (int x = 0; x < X; ++x) {
    re += part[2 * x];
    im += part[2 * x + 1];
  }
  et[0] = re;
  et[1] = im;
}

// cv[w] = sum_s (E_T,s + 2 <K_s, G_s,w - gT_s> + Q_s,w), Q_s,w = sum_z part_s[z][w];
// pred[w] = cv[w] + gain * m[w] with m the MLP output.  One workgroup per walker.
__global__ void __launch_bounds__(256) sur_cv_kernel(SurrogateCvArgs a) {
  __shared__ double red[256 / 64];
  const int w = blockIdx.x;
  cplx cv = cmake(0.0, 0.0);
  double hr = 0.0, hi = 0.0;  // <rH1, Ghalf> over the spins
  for (int s = 0; s < a.nspin; ++s) {
    const int64_t len = (int64_t)a.nocc[s] * a.nbasis;
    const double* g = a.g[s] + 2 * (int64_t)w * len;
    double re = 0.0, im = 0.0;
    for (int64_t e = threadIdx.x; e < len; e += blockDim.x) {
      const cplx gv = ld(g, e);
      const cplx k = ld(a.K[s], e), d = csub(gv, ld(a.gT[s], e));
      re += k.re * d.re - k.im * d.im;
      im += k.re * d.im + k.im * d.re;
      if (a.e1) {
        const cplx h = ld(a.rH1[s], e);
        hr += gv.re * h.re - gv.im * h.im;
        hi += gv.re * h.im + gv.im * h.re;
      }
    }
    re = block_sum(re, red);
    im = block_sum(im, red);
    cv = cadd(cv, cadd(ld(a.et, s), cmake(2.0 * re, 2.0 * im)));
    for (int z = 0; z < a.xs[s]; ++z) cv = cadd(cv, ld(a.part[s], (int64_t)z * a.nwalkers + w));
  }
  if (a.e1) {
    hr = block_sum(hr, red);
    hi = block_sum(hi, red);
  }
  if (threadIdx.x == 0) {
    st(a.cv, w, cv);
    if (a.e1) st(a.e1, w, cmake(a.e1_factor * hr + a.ecore, a.e1_factor * hi));
    if (a.fold) {
      const cplx m = ld(a.pred, w);
      st(a.mlp, w, m);
      st(a.pred, w, cadd(cv, cscale(m, a.state->gain)));
    }
  }
}

__global__ void __launch_bounds__(256) sur_fold_kernel(int nwalkers, const double* __restrict__ cv,
                                                       double* __restrict__ pred, double* __restrict__ mlp,
                                                       const SurrogateState* __restrict__ state) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nwalkers) return;
  const cplx m = ld(pred, w);
  st(mlp, w, m);
  st(pred, w, cadd(ld(cv, w), cscale(m, state->gain)));
}

// Least-squares gain of the MLP term from this step's exact samples, for the
// next step (so each step's predictor is independent of its own sample).  The
// residual r and the MLP output m are centred on their sample means first: a
// constant offset cancels in the estimate, only the within-step variation of
// r - alpha m matters.
__global__ void __launch_bounds__(256) sur_gain_kernel(SurrogateGainArgs a) {
  __shared__ double red[256 / 64];
  double s_mr = 0.0, s_mi = 0.0, s_rr = 0.0, s_ri = 0.0;
  for (int t = threadIdx.x; t < a.k; t += blockDim.x) {
    const cplx m = ld(a.mlp, a.walker_of[t]);
    s_mr += m.re;
    s_mi += m.im;
    s_rr += a.target[2 * t];
    s_ri += a.target[2 * t + 1];
  }
  const double inv_k = 1.0 / a.k;
  const double m_re = block_sum(s_mr, red) * inv_k, m_im = block_sum(s_mi, red) * inv_k;
  const double r_re = block_sum(s_rr, red) * inv_k, r_im = block_sum(s_ri, red) * inv_k;
  double rm = 0.0, mm = 0.0;
  for (int t = threadIdx.x; t < a.k; t += blockDim.x) {
    const cplx m = ld(a.mlp, a.walker_of[t]);
    const double dmr = m.re - m_re, dmi = m.im - m_im;
    const double drr = a.target[2 * t] - r_re, dri = a.target[2 * t + 1] - r_im;
    rm += dmr * drr + dmi * dri;
    mm += dmr * dmr + dmi * dmi;
  }
  rm = block_sum(rm, red);
  mm = block_sum(mm, red);
  if (threadIdx.x == 0) {
    SurrogateState* s = a.state;
    s->gain_rm = a.decay * s->gain_rm + rm;
    s->gain_mm = a.decay * s->gain_mm + mm;
    const double g = s->gain_mm > 0.0 ? s->gain_rm / s->gain_mm : 0.0;
    s->gain = fmin(fmax(g, 0.0), a.gain_max);
  }
}

}  // namespace

hipError_t launch_sur_features(const SurrogateFeatureArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(sur_features_kernel, dim3((unsigned)ceil_div(a.nwalkers, 256)), dim3(256), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_sur_stats(const SurrogateStatsArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(sur_stats_kernel, dim3(1), dim3(256), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_sur_forward(const SurrogateForwardArgs& a, hipStream_t s) {
  const int groups = a.nwalkers < 1024 ? a.nwalkers : 1024;
  hipLaunchKernelGGL(sur_forward_kernel, dim3(groups), dim3(SG_H), 0, s, a);
  return hipGetLastError();
}
int sur_grad_groups(int count) { return (int)ceil_div(count, SG_SAMPLES_PER_WG); }
hipError_t launch_sur_grad(const SurrogateTrainArgs& a, hipStream_t s) {
  const size_t lds = sizeof(float) * (2 * SG_H * SG_LD + 2 * SG_H + SG_F);
  hipError_t e = allow_lds((const void*)sur_grad_kernel, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(sur_grad_kernel, dim3(sur_grad_groups(a.count)), dim3(SG_H), lds, s, a);
  return hipGetLastError();
}
hipError_t launch_sur_adam(const SurrogateAdamArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(sur_adam_kernel, dim3((unsigned)ceil_div(SUR_NPARAM, 256)), dim3(256), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_sur_combine(const SurrogateCombineArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(sur_combine_kernel, dim3((unsigned)ceil_div(a.nwalkers, 256)), dim3(256), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_sur_select(int nwalkers, int k, const StepState* perm, int* walker_of, hipStream_t s) {
  hipLaunchKernelGGL(sur_select_kernel, dim3((unsigned)ceil_div(k, 256)), dim3(256), 0, s, nwalkers, k, perm,
                     walker_of);
  return hipGetLastError();
}
hipError_t launch_sur_plan(StepState* state, uint64_t seed, int nwalkers, int identity, hipStream_t s) {
  hipLaunchKernelGGL(sur_plan_kernel, dim3(1), dim3(64), 0, s, state, seed, nwalkers, identity);
  return hipGetLastError();
}
hipError_t launch_sur_gather(const double* g, const int* walker_of, int k, int64_t per_walker, double* out,
                             hipStream_t s) {
  const int64_t total = (int64_t)k * per_walker;
  hipLaunchKernelGGL(sur_gather_kernel, dim3((unsigned)ceil_div(total, 256)), dim3(256), 0, s, g, walker_of, k,
                     per_walker, out);
  return hipGetLastError();
}

hipError_t launch_cv_static(const double* L, const double* gT, int X, int n, int N, double* T0, double* K,
                            double* part, double* et, hipStream_t s) {
  const int64_t nt = (int64_t)X * n * n;
  hipLaunchKernelGGL(cv_t0_kernel, dim3((unsigned)ceil_div(nt, 256)), dim3(256), 0, s, L, gT, X, n, N, T0);
  hipLaunchKernelGGL(cv_k_kernel, dim3((unsigned)ceil_div((int64_t)n * N, 64)), dim3(64), 0, s, T0, L, X, n, N, K);
  hipLaunchKernelGGL(cv_et_part_kernel, dim3((unsigned)ceil_div(X, 64)), dim3(64), 0, s, T0, X, n, part);
  hipLaunchKernelGGL(cv_et_sum_kernel, dim3(1), dim3(64), 0, s, part, X, et);
  return hipGetLastError();
}

hipError_t launch_sur_cv(const SurrogateCvArgs& a, hipStream_t s) {
  if (a.e1 && (a.rH1[0] == nullptr || (a.nspin > 1 && a.rH1[1] == nullptr))) return hipErrorInvalidValue;
  hipLaunchKernelGGL(sur_cv_kernel, dim3(a.nwalkers), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_sur_fold(int nwalkers, const double* cv, double* pred, double* mlp, const SurrogateState* state,
                           hipStream_t s) {
  if (nwalkers < 1) return hipSuccess;
  hipLaunchKernelGGL(sur_fold_kernel, dim3((unsigned)ceil_div(nwalkers, 256)), dim3(256), 0, s, nwalkers, cv, pred,
                     mlp, state);
  return hipGetLastError();
}

hipError_t launch_sur_gain(const SurrogateGainArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(sur_gain_kernel, dim3(1), dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace afqmc
