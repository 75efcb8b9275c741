// Per-walker elementwise / reduction kernels of the step:
//   fields      xbar, force-bias cap, xi, xshifted, cmf, cfb (+ Coulomb)
//               (ipie/propagation/phaseless_base.py:282-310, 360-368;
//                Coulomb: ipie/estimators/local_energy_sd.py:158-201)
//   weights     hybrid energy, phaseless importance weight
//               (phaseless_base.py:340-358, 370-380)
//   energy      e1b + Coulomb - exchange assembly
//               (ipie/estimators/local_energy_sd.py:358-403, 106-155)
//   normals     counter-based N(0, 1) fields (Philox4x32-10 + Box-Muller)
#include "afqmc_internal.hpp"

namespace afqmc {

namespace {

constexpr int FLD_THREADS = 256;

// Two standard normals (Philox key = seed, counter = (pair index, stream)) for pair index p of stream `counter`.
__device__ __forceinline__ double2 normal_pair(uint64_t seed, uint64_t counter, uint64_t p) {
  const uint4 c = make_uint4((uint32_t)p, (uint32_t)(p >> 32), (uint32_t)counter, (uint32_t)(counter >> 32));
  const uint4 r = philox4x32(c, make_uint2((uint32_t)seed, (uint32_t)(seed >> 32)));
  // 53-bit uniforms in (0, 1]
  const uint64_t a = ((uint64_t)r.x << 21) ^ ((uint64_t)r.y >> 11);
  const uint64_t b = ((uint64_t)r.z << 21) ^ ((uint64_t)r.w >> 11);
  const double u1 = ((double)(a & ((1ull << 53) - 1)) + 1.0) * (1.0 / 9007199254740992.0);
  const double u2 = (double)(b & ((1ull << 53) - 1)) * (1.0 / 9007199254740992.0);
  const double rad = sqrt(-2.0 * log(u1));
  double s, co;
  sincospi(2.0 * u2, &s, &co);
  return make_double2(rad * co, rad * s);
}

__device__ __forceinline__ double draw_normal(uint64_t seed, uint64_t counter, uint64_t idx) {
  const double2 v = normal_pair(seed, counter, idx >> 1);
  return (idx & 1) ? v.y : v.x;
}

__global__ void __launch_bounds__(FLD_THREADS) fields_kernel(FieldsArgs a) {
  __shared__ double red[FLD_THREADS / 64];
  const int w = blockIdx.x;
  const int X = a.nchol;
  const int64_t plane = (int64_t)a.nwalkers * X;
  double cmf_r = 0, cmf_i = 0, cfb_r = 0, cfb_i = 0, ec_r = 0, ec_i = 0;
  double xb_sum = 0.0, xb_max = 0.0;
  unsigned long long ncap = 0;
  const uint64_t counter = a.state ? a.state->counter : a.counter;
  for (int x = threadIdx.x; x < X; x += FLD_THREADS) {
    const int64_t idx = (int64_t)w * X + x;
    cplx vb;
    if (a.part32) {  // split-BF16 partials, FP32 (their accumulation precision), summed in FP64
      const float2* p32 = reinterpret_cast<const float2*>(a.vbias_part);
      vb = cmake(0.0, 0.0);
      for (int z = 0; z < a.ksplit; ++z) {
        const float2 v = p32[z * plane + idx];
        vb = cadd(vb, cmake((double)v.x, (double)v.y));
      }
    } else {
      vb = ld(a.vbias_part, idx);
      for (int z = 1; z < a.ksplit; ++z) vb = cadd(vb, ld(a.vbias_part, z * plane + idx));
    }
    if (a.vbias_out) st(a.vbias_out, idx, vb);
    ec_r += vb.re * vb.re - vb.im * vb.im;
    ec_i += 2.0 * vb.re * vb.im;
    const cplx mf = ld(a.mf_shift, x);
    // xbar = -sqrt(dt) (i vbias - mf_shift)
    cplx xb = cmake(a.sqrt_dt * (mf.re + vb.im), a.sqrt_dt * (mf.im - vb.re));
    const double axb = cabs(xb);
    if (axb > a.fbbound) {
      ++ncap;
      if (axb > 1e-13) xb = cmake(xb.re / axb, xb.im / axb);
    }
    const double axc = axb > a.fbbound ? 1.0 : axb;
    xb_sum += axc;
    xb_max = fmax(xb_max, axc);
    if (!a.want_fields) continue;
    const double xi = a.xi ? a.xi[idx] : draw_normal(a.seed, counter, (uint64_t)idx);
    if (a.xi_out) a.xi_out[idx] = xi;
    const cplx xs = cmake(xi - xb.re, -xb.im);
    st(a.xshifted, idx, xs);
    cmf_r += xs.re * mf.re - xs.im * mf.im;
    cmf_i += xs.re * mf.im + xs.im * mf.re;
    cfb_r += xi * xb.re - 0.5 * (xb.re * xb.re - xb.im * xb.im);
    cfb_i += xi * xb.im - xb.re * xb.im;
  }
  ec_r = block_sum(ec_r, red);
  ec_i = block_sum(ec_i, red);
  if (a.xbar_mean) {
    xb_sum = block_sum(xb_sum, red);
    // block max: wave shuffles then the per-wave maxima in order
    for (int off = 32; off > 0; off >>= 1) xb_max = fmax(xb_max, __shfl_xor(xb_max, off, 64));
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = xb_max;
    __syncthreads();
    double m = 0.0;
    for (int i = 0; i < FLD_THREADS / 64; ++i) m = fmax(m, red[i]);
    if (threadIdx.x == 0) {
      a.xbar_mean[w] = xb_sum / X;
      a.xbar_max[w] = m;
    }
  }
  if (a.want_fields) {
    cmf_r = block_sum(cmf_r, red);
    cmf_i = block_sum(cmf_i, red);
    cfb_r = block_sum(cfb_r, red);
    cfb_i = block_sum(cfb_i, red);
    if (ncap && a.nfb) atomicAdd(a.nfb, ncap);
  }
  if (threadIdx.x == 0) {
    if (a.ecoul) st(a.ecoul, w, cmake(0.5 * ec_r, 0.5 * ec_i));
    if (a.want_fields) {
      st(a.cmf, w, cmake(-a.sqrt_dt * cmf_r, -a.sqrt_dt * cmf_i));
      st(a.cfb, w, cmake(cfb_r, cfb_i));
    }
  }
}

__global__ void update_weight_kernel(WeightArgs a) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= a.nwalkers) return;
  const cplx o_old = ld(a.ovlp_old, w), o_new = ld(a.ovlp_new, w);
  const cplx ratio = cdiv(o_new, o_old);
  const cplx lr = clog(ratio);
  const cplx cfb = ld(a.cfb, w), cmf = ld(a.cmf, w);
  const double eshift = a.state ? a.state->eshift : a.eshift;
  cplx eh = cmake(-(lr.re + cfb.re + cmf.re) / a.dt, -(lr.im + cfb.im + cmf.im) / a.dt);
  if (fabs(eshift) >= 1e-10) {
    const double emax = eshift + a.ebound, emin = eshift - a.ebound;
    eh.re = fmin(fmax(eh.re, emin), emax);
  }
  const cplx prev = ld(a.hybrid_energy, w);
  const cplx arg = cmake(-a.dt * (0.5 * (eh.re + prev.re) - eshift), -a.dt * 0.5 * (eh.im + prev.im));
  const double magn = exp(arg.re);  // |exp(arg)|
  const double dtheta = -a.dt * eh.im - cfb.im;
  double cosf = cos(dtheta);
  if (cosf < 0.0) cosf = 0.0;
  const double wnew = a.weight[w] * magn * cosf;
  a.weight[w] = wnew;
  st(a.hybrid_energy, w, eh);
  st(a.walker_ovlp, w, o_new);
  if (!isfinite(wnew) && a.flags) atomicOr(a.flags, AFQMC_FLAG_NONFINITE);
  if (!isfinite(wnew) && a.wflags) atomicOr(a.wflags + w, AFQMC_FLAG_NONFINITE);
}

__global__ void __launch_bounds__(FLD_THREADS) energy_kernel(EnergyArgs a) {
  __shared__ double red[FLD_THREADS / 64];
  const int w = blockIdx.x;
  const int N = a.nbasis;
  double er = 0.0, ei = 0.0;
  const int nspin = a.e1_in ? 0 : (a.rhf || a.nbeta == 0) ? 1 : 2;
  for (int s = 0; s < nspin; ++s) {
    const int n = s == 0 ? a.nalpha : a.nbeta;
    const double* g = (s == 0 ? a.ghalfa : a.ghalfb) + 2 * (int64_t)w * n * N;
    const double* h = s == 0 ? a.rH1a : a.rH1b;
    for (int e = threadIdx.x; e < n * N; e += FLD_THREADS) {
      const cplx gv = ld(g, e), hv = ld(h, e);
      er += gv.re * hv.re - gv.im * hv.im;
      ei += gv.re * hv.im + gv.im * hv.re;
    }
  }
  er = block_sum(er, red);
  ei = block_sum(ei, red);
  if (threadIdx.x == 0) {
    const double f = a.rhf ? 2.0 : 1.0;
    const cplx e1 = a.e1_in ? ld(a.e1_in, w) : cmake(f * er + a.ecore, f * ei);
    if (a.e1_out) st(a.e1_out, w, e1);
    cplx ex = cmake(0.0, 0.0);
    if (a.exx_est) {
      ex = ld(a.exx_est, w);
    } else {
      for (int z = 0; z < a.xs_a; ++z) ex = cadd(ex, ld(a.exx_part_a, (int64_t)z * a.nwalkers + w));
      if (a.exx_part_b)
        for (int z = 0; z < a.xs_b; ++z) ex = cadd(ex, ld(a.exx_part_b, (int64_t)z * a.nwalkers + w));
    }
    // UHF: exx = 1/2 (sum_a + sum_b); RHF: exx = 2 * 1/2 * sum_a
    const double fx = a.rhf ? 1.0 : 0.5;
    ex = cscale(ex, fx);
    const cplx e2 = csub(ld(a.ecoul, w), ex);
    st(a.energy, 3 * (int64_t)w + 0, cadd(e1, e2));
    st(a.energy, 3 * (int64_t)w + 1, e1);
    st(a.energy, 3 * (int64_t)w + 2, e2);
  }
}

__global__ void step_set_kernel(StepState* s, unsigned long long counter, double eshift) {
  if (threadIdx.x == 0) {
    s->counter = counter;
    s->eshift = eshift;
  }
}

__global__ void step_advance_kernel(StepState* s, long long dcounter, long long dadam) {
  if (threadIdx.x == 0) {
    s->counter += (unsigned long long)dcounter;
    s->adam_t += dadam;
  }
}

__global__ void normals_kernel(double* out, int64_t n, uint64_t seed, uint64_t counter) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (2 * p >= n) return;
  const double2 v = normal_pair(seed, counter, (uint64_t)p);
  out[2 * p] = v.x;
  if (2 * p + 1 < n) out[2 * p + 1] = v.y;
}

}  // namespace

hipError_t launch_fields(const FieldsArgs& a, hipStream_t stream) {
  hipLaunchKernelGGL(fields_kernel, dim3(a.nwalkers), dim3(FLD_THREADS), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_update_weight(const WeightArgs& a, hipStream_t stream) {
  hipLaunchKernelGGL(update_weight_kernel, dim3((unsigned)ceil_div(a.nwalkers, 256)), dim3(256), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_energy(const EnergyArgs& a, hipStream_t stream) {
  hipLaunchKernelGGL(energy_kernel, dim3(a.nwalkers), dim3(FLD_THREADS), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_step_set(StepState* state, unsigned long long counter, double eshift, hipStream_t stream) {
  hipLaunchKernelGGL(step_set_kernel, dim3(1), dim3(64), 0, stream, state, counter, eshift);
  return hipGetLastError();
}

hipError_t launch_step_advance(StepState* state, long long dcounter, long long dadam, hipStream_t stream) {
  hipLaunchKernelGGL(step_advance_kernel, dim3(1), dim3(64), 0, stream, state, dcounter, dadam);
  return hipGetLastError();
}

// dst[i] = src[i] for i < n, then src[i] = 0 when clear (a kernel, not a
// memset, so that a captured graph replays it correctly).
__global__ void take_flags_kernel(int* src, int* dst, int n, int clear) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int v = src[i];
  if (dst) dst[i] = v;
  if (clear) src[i] = 0;
}

hipError_t launch_take_flags(int* src, int* dst, int n, int clear, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(take_flags_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, stream, src, dst, n, clear);
  return hipGetLastError();
}

hipError_t launch_draw_normals(double* out, int64_t n, uint64_t seed, uint64_t counter, hipStream_t stream) {
  const int64_t pairs = (n + 1) / 2;
  hipLaunchKernelGGL(normals_kernel, dim3((unsigned)ceil_div(pairs, 256)), dim3(256), 0, stream, out, n, seed,
                     counter);
  return hipGetLastError();
}

}  // namespace afqmc
