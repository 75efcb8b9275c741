// Split-BF16 kernels of the mixed-precision step (AFQMC_PREC_MIXED) and of the
// exchange control variate.
//
// Every FP64 operand x is split into x_hi = bf16(x), x_lo = bf16(x - x_hi);
// a product A B is formed as A_hi B_hi + A_hi B_lo + A_lo B_hi with
// v_mfma_f32_16x16x32_bf16 (FP32 accumulation), which keeps about 16
// significant bits per product (relative error ~1e-5 per element).
//   exchange_bf16_kernel  sum_x tr(T_x T_x), T_x = L_x G^T, trace in FP64
//                         (ipie/estimators/local_energy_sd.py:263-303, 1040-1071);
//                         also the single-product form (hi * hi only) used by the
//                         control variate
//   bf16_gemm_kernel      force bias (ipie/propagation/force_bias.py:79-121) and
//                         VHS (ipie/propagation/phaseless_generic.py:67-97)
//   taylor_bf16_kernel    one term of the Taylor series of exp(VHS) phi, VHS read as
//                         16 x 16 atoms of its upper triangle (VHS epilogue of
//                         bf16_gemm_kernel), mirrors by transposing LDS loads
//                         (ipie/propagation/operations.py:58-117)
//
// Operand planes are bf16 [rows, Kp] with Kp = K rounded up to 32 and zero
// padding, so a 32-wide K chunk of a row is one aligned 64-byte segment:
//   Ghalf: 4 planes (re_hi, re_lo, im_hi, im_lo), rows = (walker, j)
//   rchol: 2 planes (hi, lo), rows = (x, i)
// Exchange / GEMM workgroup: 8 waves in a 4 x 2 grid, each wave owns 2 x 5
// output tiles of 16 x 16, so one workgroup covers 128 rows (wb walkers) by
// 160 columns (xb Cholesky vectors); two waves share each SIMD, so one wave's
// LDS and barrier waits are covered by the other's MFMAs.  Double-buffered LDS
// staging, one barrier per K chunk.  Workgroups that run together on one XCD
// cover a few walker blocks x 8 Cholesky ranges, so both operand tiles are
// shared through that XCD's L2.
#include "afqmc_internal.hpp"

namespace afqmc {

namespace {

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
// native 16-byte vector: plain loads and stores, kept in registers (a uint4
// struct copy lowers to a memcpy that pins the staging arrays to scratch)
typedef unsigned int v4u __attribute__((ext_vector_type(4)));

// Workgroup tile of the BF16 GEMM: GR x GC waves, each owning MR x MC output
// tiles of 16 x 16, so R = 16 GR MR rows by C = 16 GC MC columns, with ST LDS
// stages (ST - 1 in flight while one is computed) of KS1 sub-chunks of 32 K
// values each in the 1-product form (one in the 3-product form).
template <int GR_, int GC_, int MR_, int MC_, int ST_ = 3, int KS1_ = 2>
struct Tile {
  static constexpr int GR = GR_, GC = GC_, MR = MR_, MC = MC_, ST = ST_, KS1 = KS1_;
  static constexpr int R = 16 * GR * MR, C = 16 * GC * MC, THREADS = 64 * GR * GC;
};
// The shapes bf16_gemm_kernel is built for (afqmc_tune_set picks one per
// contraction; the exchange and the quadratic form use shape 0).  All have 8
// waves; shapes 0-3 three LDS stages of the 3-product form within 160 KiB,
// shape 4 (64 x 80 per wave: 29% fewer LDS bytes per MFMA than 32 x 80) two,
// shape 5 two stages of 36 KiB, so that two workgroups share a CU (one can
// run its epilogue while the other computes: short K, heavy epilogues).
using Tile0 = Tile<4, 2, 2, 5>;     // 128 x 160
using Tile1 = Tile<4, 2, 2, 4>;     // 128 x 128
using Tile2 = Tile<8, 1, 1, 5>;     // 128 x 80
using Tile3 = Tile<4, 2, 1, 5>;     // 64 x 160
using Tile4 = Tile<2, 4, 4, 5, 2>;  // 128 x 320
using Tile5 = Tile<4, 2, 1, 5, 2>;  // 64 x 160

// 8 waves in a BX_GR x BX_GC grid, each owning BX_MR x BX_MC tiles of 16 x 16 (shape 0)
constexpr int BX_GR = Tile0::GR, BX_GC = Tile0::GC, BX_MR = Tile0::MR, BX_MC = Tile0::MC;
constexpr int BX_R = 16 * BX_GR * BX_MR;  // rows of a workgroup tile (walker orbitals)
constexpr int BX_C = 16 * BX_GC * BX_MC;  // columns (Cholesky-vector orbitals)
constexpr int BX_THREADS = 64 * BX_GR * BX_GC;
constexpr int BX_KC = 32;
static_assert(BX_C == QUAD_TILE && BX_GC == QUAD_SLOTS, "quadratic-form tiles are the GEMM's column tiles");
constexpr size_t BX_LDS_MAX = 106496;  // LDS of a workgroup: two stages of the 3-product form

// Staging, in 16-byte vectors.  SPLIT (3 products): A planes re_hi, re_lo,
// im_hi, im_lo and B planes hi, lo.  Otherwise (1 product, hi * hi): A planes
// re_hi, im_hi (physical planes 0 and 2) and B plane hi.  An LDS stage holds
// KS sub-chunks of 32 K values (the 1-product form uses the tile's KS1, two by
// default, halving the barriers per unit of MFMA work; the 3-product stage
// already fills the LDS).
template <bool SPLIT, class TL = Tile0>
struct Bx {
  static constexpr int NA = SPLIT ? 4 : 2, NB = SPLIT ? 2 : 1;
  static constexpr int KS = SPLIT ? 1 : TL::KS1;
  static constexpr int A_VEC = NA * TL::R * 4, B_VEC = NB * TL::C * 4;  // per sub-chunk
  static constexpr int SUB_VEC = A_VEC + B_VEC;
  static constexpr int STAGE_VEC = KS * SUB_VEC;
  static constexpr int A_PER = A_VEC / TL::THREADS;
  static constexpr int B_PER = (B_VEC + TL::THREADS - 1) / TL::THREADS;
  static constexpr size_t STAGES = 2 * 16 * (size_t)STAGE_VEC;  // double-buffered, bytes (exchange_bf16_kernel)
  static constexpr size_t STAGES_ALL = TL::ST * 16 * (size_t)STAGE_VEC;  // the tile's stages (bf16_gemm_kernel)
  static_assert(A_VEC % TL::THREADS == 0, "A staging is split evenly over the threads");
  static_assert((B_VEC % TL::THREADS) % 64 == 0, "the partial B vector round is wave-uniform");
  static_assert(STAGES_ALL <= 163840, "the stages fit the 160 KiB of LDS");
  __device__ static int a_plane(int p) { return SPLIT ? p : 2 * p; }
  __device__ static int b_plane(int p) { return SPLIT ? p : 0; }
};

static_assert(Bx<true>::STAGES <= BX_LDS_MAX, "two stages of the exchange kernel fit its LDS budget");

__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }

__device__ __forceinline__ void split_bf16(double x, uint16_t& hi, uint16_t& lo) { split_bf16_pair(x, hi, lo); }

__device__ __forceinline__ v8bf as_v8bf(v4u u) { return __builtin_bit_cast(v8bf, u); }

__device__ __forceinline__ v4f mfma_bf16(v8bf a, v8bf b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// LDS slot of 16-byte vector `seg` (0..3) of staged row `row`.  A fragment
// read takes one segment of 16 consecutive rows (64-byte rows: 4 slots), and a
// 16-byte LDS read is served a group of lanes at a time.  The segment is
// XOR-ed with ((row >> 1) ^ (row >> 3)) & 3 so that every aligned group of 8
// rows covers 8 distinct slots modulo 8 (32 banks of 4 bytes) and every
// aligned group of 16 rows 16 distinct slots modulo 16 (64 banks): no bank
// conflicts whichever the group size.  (XOR with bits 2..3 of the row alone
// leaves 2-way conflicts in 8-lane groups, which the bank-conflict counter
// showed on every GEMM.)
__device__ __forceinline__ int lds_swz(int row) { return ((row >> 1) ^ (row >> 3)) & 3; }
__device__ __forceinline__ int lds_vec(int row, int seg) { return row * 4 + (seg ^ lds_swz(row)); }

__device__ __forceinline__ v8bf lds_frag(const uint16_t* base, int row, int seg) {
  return as_v8bf(reinterpret_cast<const v4u*>(base)[lds_vec(row, seg)]);
}

// The Taylor kernel's staging keeps the XOR with bits 2..3 of the row: with
// its 80 x 80 tiles and 5 waves it measured 12% faster than the swizzle above
// (with register or direct staging alike).
__device__ __forceinline__ int lds_swz_ty(int row) { return (row >> 2) & 3; }
__device__ __forceinline__ int lds_vec_ty(int row, int seg) { return row * 4 + (seg ^ lds_swz_ty(row)); }
__device__ __forceinline__ v8bf lds_frag_ty(const uint16_t* base, int row, int seg) {
  return as_v8bf(reinterpret_cast<const v4u*>(base)[lds_vec_ty(row, seg)]);
}

// Ghalf [W*n, N] complex -> 4 bf16 planes [rows, Kp]; with walker_of, plane
// row t*n + j takes Ghalf row walker_of[t]*n + j (sampled walkers, compacted);
// with gref [n, N], row j of gref is subtracted first (dG = G - G_T).
__global__ void split_ghalf_kernel(const double* __restrict__ g, int64_t rows, int N, int Kp,
                                   uint16_t* __restrict__ planes, const int* __restrict__ walker_of, int nocc,
                                   const double* __restrict__ gref) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= rows * Kp) return;
  const int64_t r = idx / Kp;
  const int k = (int)(idx % Kp);
  uint16_t rh = 0, rl = 0, ih = 0, il = 0;
  if (k < N) {
    const int64_t src = walker_of ? (int64_t)walker_of[r / nocc] * nocc + r % nocc : r;
    cplx v = ld(g, src * N + k);
    if (gref) v = csub(v, ld(gref, (r % nocc) * N + k));
    split_bf16(v.re, rh, rl);
    split_bf16(v.im, ih, il);
  }
  const int64_t plane = rows * Kp;
  planes[idx] = rh;
  planes[plane + idx] = rl;
  planes[2 * plane + idx] = ih;
  planes[3 * plane + idx] = il;
}


// D = m - I (m real [N][N]) -> 2 bf16 planes [N][Kp].
__global__ void split_eye_kernel(const double* __restrict__ m, int N, int Kp, uint16_t* __restrict__ planes) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)N * Kp) return;
  const int64_t r = idx / Kp;
  const int k = (int)(idx % Kp);
  uint16_t h = 0, l = 0;
  if (k < N) split_bf16(m[r * N + k] - (r == k ? 1.0 : 0.0), h, l);
  planes[idx] = h;
  planes[(int64_t)N * Kp + idx] = l;
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

// Packed Cholesky rows [P][X] -> 2 bf16 planes [NAT 256][Xp] in VHS atom order:
// row t 256 + r 16 + c of atom t = (a, b) holds L[(16 a + r, 16 b + c)] (L is
// symmetric, so the diagonal atoms' lower halves repeat the packed rows of their
// mirror), zero outside N x N and for x >= X.
__global__ void split_chol_atoms_kernel(const double* __restrict__ chol, int N, int X, int Xp, int NA,
                                        uint16_t* __restrict__ planes) {
  const int64_t rows = (int64_t)NA * (NA + 1) / 2 * 256;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= rows * Xp) return;
  const int64_t row = idx / Xp;
  const int x = (int)(idx % Xp);
  int t = (int)(row >> 8), a = 0;
  while (t >= NA - a) {  // atoms of row a: (a, a), ..., (a, NA - 1)
    t -= NA - a;
    ++a;
  }
  const int b = a + t, rc = (int)(row & 255);
  const int i = 16 * a + (rc >> 4), j = 16 * b + (rc & 15);
  uint16_t h = 0, l = 0;
  if (i < N && j < N && x < X) {
    const int lo = min(i, j), hi = max(i, j);
    const int64_t p = (int64_t)lo * N - (int64_t)lo * (lo - 1) / 2 + (hi - lo);
    split_bf16(chol[p * X + x], h, l);
  }
  planes[idx] = h;
  planes[rows * Xp + idx] = l;
}


// Source offsets (bf16 units, K chunk 0) of the B vectors thread tid stages
// for Cholesky block xblk; rows past the end are clamped to the last row.
template <bool SPLIT, class TL = Tile0>
__device__ __forceinline__ void bx_set_b(int64_t (&b_src)[(Bx<SPLIT, TL>::B_PER)], int tid, int xblk, int xb, int n,
                                         int64_t brows, int64_t bplane, int Kp) {
#pragma unroll
  for (int i = 0; i < Bx<SPLIT, TL>::B_PER; ++i) {
    const int v = (tid + i * TL::THREADS) % Bx<SPLIT, TL>::B_VEC;  // threads past the end load a duplicate, never stored
    const int p = v / (TL::C * 4), rem = v % (TL::C * 4);
    int64_t r = (int64_t)xblk * xb * n + rem / 4;
    if (r > brows - 1) r = brows - 1;
    b_src[i] = Bx<SPLIT, TL>::b_plane(p) * bplane + r * Kp + (rem % 4) * 8;
  }
}

// Sub-chunks c32 .. c32 + nsub - 1 (32 K values each) into registers.
template <bool SPLIT, class TL = Tile0>
__device__ __forceinline__ void bx_fetch(v4u (&pa)[(Bx<SPLIT, TL>::KS)][(Bx<SPLIT, TL>::A_PER)],
                                         v4u (&pb)[(Bx<SPLIT, TL>::KS)][(Bx<SPLIT, TL>::B_PER)], const uint16_t* gp,
                                         const uint16_t* lp, const int64_t (&a_src)[(Bx<SPLIT, TL>::A_PER)],
                                         const int64_t (&b_src)[(Bx<SPLIT, TL>::B_PER)], int c32, int nsub) {
#pragma unroll
  for (int sc = 0; sc < Bx<SPLIT, TL>::KS; ++sc) {
    if (sc >= nsub) break;
    const int64_t k0 = (int64_t)(c32 + sc) * BX_KC;
#pragma unroll
    for (int i = 0; i < Bx<SPLIT, TL>::A_PER; ++i) pa[sc][i] = *reinterpret_cast<const v4u*>(gp + a_src[i] + k0);
#pragma unroll
    for (int i = 0; i < Bx<SPLIT, TL>::B_PER; ++i) pb[sc][i] = *reinterpret_cast<const v4u*>(lp + b_src[i] + k0);
  }
}

template <bool SPLIT, class TL = Tile0>
__device__ __forceinline__ void bx_stash(uint4* lds_u4, const v4u (&pa)[(Bx<SPLIT, TL>::KS)][(Bx<SPLIT, TL>::A_PER)],
                                         const v4u (&pb)[(Bx<SPLIT, TL>::KS)][(Bx<SPLIT, TL>::B_PER)], int tid, int buf,
                                         int nsub) {
#pragma unroll
  for (int sc = 0; sc < Bx<SPLIT, TL>::KS; ++sc) {
    if (sc >= nsub) break;
    v4u* base = reinterpret_cast<v4u*>(lds_u4) + buf * Bx<SPLIT, TL>::STAGE_VEC + sc * Bx<SPLIT, TL>::SUB_VEC;
#pragma unroll
    for (int i = 0; i < Bx<SPLIT, TL>::A_PER; ++i) {
      const int v = tid + i * TL::THREADS;
      base[lds_vec(v >> 2, v & 3)] = pa[sc][i];
    }
#pragma unroll
    for (int i = 0; i < Bx<SPLIT, TL>::B_PER; ++i) {
      const int v = tid + i * TL::THREADS;
      if (v < Bx<SPLIT, TL>::B_VEC) base[Bx<SPLIT, TL>::A_VEC + lds_vec(v >> 2, v & 3)] = pb[sc][i];
    }
  }
}


// Direct global -> LDS staging (global_load_lds_dwordx4, no VGPR round
// trip).  Slot s of a sub-chunk region holds staged vector (row s / 4,
// segment (s % 4) ^ ((row / 4) % 4)): the XOR swizzle of lds_vec moves from
// the LDS address to the source address, so the 64 lanes of one load fill 64
// consecutive slots, as the instruction requires.  Each wave stages an equal
// share of the A and B vectors with the same instruction count (the last
// instruction of a share may have idle lanes), so a wait on vmcnt can name
// the loads of a later stage that may stay in flight.
template <bool SPLIT, class TL = Tile0>
struct Dl {
  static constexpr int WAVES = TL::THREADS / 64;
  static constexpr int A_WAVE = Bx<SPLIT, TL>::A_VEC / WAVES, B_WAVE = Bx<SPLIT, TL>::B_VEC / WAVES;
  static constexpr int A_INS = (A_WAVE + 63) / 64, B_INS = (B_WAVE + 63) / 64;
  static constexpr int INS_SUB = A_INS + B_INS;  // load instructions per sub-chunk
  static_assert(Bx<SPLIT, TL>::A_VEC % WAVES == 0 && Bx<SPLIT, TL>::B_VEC % WAVES == 0, "equal shares per wave");
  static_assert(A_WAVE % 64 == 0, "A shares are whole instructions");
  static_assert(Bx<SPLIT, TL>::KS * INS_SUB <= 63, "vmcnt counts the loads of one stage");
};

template <int N>
__device__ __forceinline__ void wait_vm() {
  asm volatile("s_waitcnt vmcnt(%0)" ::"n"(N) : "memory");
}

// source offsets (in bf16 elements, K offset 0) of this lane's slots
template <bool SPLIT, class TL = Tile0>
__device__ __forceinline__ void dl_sources(int64_t (&a_src)[(Dl<SPLIT, TL>::A_INS)], int64_t (&b_src)[(Dl<SPLIT, TL>::B_INS)],
                                           const Bf16GemmArgs& a, int rb, int cb, int wave, int lane) {
  using C = Bx<SPLIT, TL>;
#pragma unroll
  for (int j = 0; j < Dl<SPLIT, TL>::A_INS; ++j) {
    const int s = wave * Dl<SPLIT, TL>::A_WAVE + j * 64 + lane;
    const int row = s >> 2, seg = (s & 3) ^ lds_swz(row);
    const int p = row / TL::R;
    int64_t r = (int64_t)rb * TL::R + row % TL::R;
    if (r > a.m - 1) r = a.m - 1;
    if (a.a_gstride > 0) r = (r / a.group) * a.a_gstride + a.a_goff + r % a.group;
    a_src[j] = C::a_plane(p) * a.a_plane + r * a.lda + seg * 8;
  }
#pragma unroll
  for (int j = 0; j < Dl<SPLIT, TL>::B_INS; ++j) {
    int s = wave * Dl<SPLIT, TL>::B_WAVE + j * 64 + lane;
    if (j * 64 + lane >= Dl<SPLIT, TL>::B_WAVE) s = wave * Dl<SPLIT, TL>::B_WAVE;  // idle lane: any valid address
    const int row = s >> 2, seg = (s & 3) ^ lds_swz(row);
    const int p = row / TL::C;
    int64_t r = (int64_t)cb * TL::C + row % TL::C;
    if (r > a.nn - 1) r = a.nn - 1;
    b_src[j] = C::b_plane(p) * a.b_plane + r * a.ldb + seg * 8;
  }
}

// issue the loads of sub-chunks c32 .. c32 + nsub - 1 into LDS stage `buf`
template <bool SPLIT, class TL = Tile0>
__device__ __forceinline__ void dl_issue(uint4* lds_u4, int buf, int nsub, const uint16_t* ga, const uint16_t* gb,
                                         const int64_t (&a_src)[(Dl<SPLIT, TL>::A_INS)],
                                         const int64_t (&b_src)[(Dl<SPLIT, TL>::B_INS)], int c32, int wave, int lane) {
  using C = Bx<SPLIT, TL>;
  using D = Dl<SPLIT, TL>;
#pragma unroll
  for (int sc = 0; sc < C::KS; ++sc) {
    if (sc >= nsub) break;
    const int64_t k0 = (int64_t)(c32 + sc) * BX_KC;
    uint4* base = lds_u4 + buf * C::STAGE_VEC + sc * C::SUB_VEC;
#pragma unroll
    for (int j = 0; j < D::A_INS; ++j)
      __builtin_amdgcn_global_load_lds(
          (const void*)(ga + a_src[j] + k0),
          (__attribute__((address_space(3))) void*)(base + wave * D::A_WAVE + j * 64), 16, 0, 0);
#pragma unroll
    for (int j = 0; j < D::B_INS; ++j)
      if (j * 64 + lane < D::B_WAVE)
        __builtin_amdgcn_global_load_lds(
            (const void*)(gb + b_src[j] + k0),
            (__attribute__((address_space(3))) void*)(base + C::A_VEC + wave * D::B_WAVE + j * 64), 16, 0, 0);
  }
}

// workgroup barrier for the direct-load pipeline: this wave's LDS reads are
// complete and its loads of the next stage landed (dl_wait); unlike
// __syncthreads it does not wait for the loads of the stage after, which stay
// in flight across it
__device__ __forceinline__ void dl_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// wait until at most `pend` sub-chunks of direct loads are outstanding
template <bool SPLIT, class TL = Tile0>
__device__ __forceinline__ void dl_wait(int pend) {
  constexpr int I = Dl<SPLIT, TL>::INS_SUB;
  static_assert(3 * I <= 63, "vmcnt counts three sub-chunks of loads");
  if (pend >= 3) wait_vm<3 * I>();
  else if (pend == 2) wait_vm<2 * I>();
  else if (pend == 1) wait_vm<I>();
  else wait_vm<0>();
}

// One K chunk from LDS stage `stage`: the wave's TL::MR x TL::MC tiles,
// re += Re(A) B and im += Im(A) B, each as hi*hi + hi*lo + lo*hi (SPLIT) or
// hi*hi.  TR swaps the MFMA operands, so the accumulators hold the transposed
// tiles: element r of tile (i, j) is C[i 16 + l16][j 16 + 4 q + r] instead of
// C[i 16 + 4 q + r][j 16 + l16].
__device__ __forceinline__ v4f mfma_ab(v8bf a, v8bf b, v4f c, bool tr) {
  return tr ? mfma_bf16(b, a, c) : mfma_bf16(a, b, c);
}

// FP8 (F8, single product only): each 16-byte fragment holds 16 FP8 values of
// K, fed to two 16x16x32 FP8 MFMAs (bytes 0-7 with 0-7, 8-15 with 8-15: A and
// B take the same K positions, so each K is summed once).
__device__ __forceinline__ v4f mfma_fp8_pair(v8bf a, v8bf b, v4f c) {
  const v4u ua = __builtin_bit_cast(v4u, a), ub = __builtin_bit_cast(v4u, b);
  const long a0 = (long)(((unsigned long)ua[1] << 32) | ua[0]), a1 = (long)(((unsigned long)ua[3] << 32) | ua[2]);
  const long b0 = (long)(((unsigned long)ub[1] << 32) | ub[0]), b1 = (long)(((unsigned long)ub[3] << 32) | ub[2]);
  c = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(a0, b0, c, 0, 0, 0);
  return __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(a1, b1, c, 0, 0, 0);
}

// FP8 over two sub-chunks at once: lane (l16, q) gives 16 bytes of each
// sub-chunk (128 K values of a row in all) to one 16x16x128 MFMA, twice the
// rate of the 16x16x32 FP8 form.  A and B give the same bytes of both
// sub-chunks, so each K is summed once whatever order the instruction takes
// the 32 bytes of a lane in.  Block scales 127 (E8M0 for 2^0): the operands
// carry their power-of-two scales already.
__device__ __forceinline__ v4f mfma_fp8_x128(v4u a0, v4u a1, v4u b0, v4u b1, v4f c) {
  typedef int v8i __attribute__((ext_vector_type(8)));
  const v8i a = {(int)a0[0], (int)a0[1], (int)a0[2], (int)a0[3], (int)a1[0], (int)a1[1], (int)a1[2], (int)a1[3]};
  const v8i b = {(int)b0[0], (int)b0[1], (int)b0[2], (int)b0[3], (int)b1[0], (int)b1[1], (int)b1[2], (int)b1[3]};
  return __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(a, b, c, 0, 0, 0, 127, 0, 127);
}

template <class TL>
__device__ __forceinline__ void bx_sub_f8x2(const uint4* sub, int wr, int wc, int l16, int q,
                                            v4f (&re)[(TL::MR)][(TL::MC)], v4f (&im)[(TL::MR)][(TL::MC)]) {
  using B = Bx<false, TL>;
  const uint16_t* A0 = reinterpret_cast<const uint16_t*>(sub);
  const uint16_t* A1 = A0 + 8 * B::SUB_VEC;
  const uint16_t* B0 = A0 + 8 * B::A_VEC;
  const uint16_t* B1 = A1 + 8 * B::A_VEC;
  auto frag = [](const uint16_t* base, int row, int seg) {
    return reinterpret_cast<const v4u*>(base)[lds_vec(row, seg)];
  };
  v4u b0[(TL::MC)], b1[(TL::MC)];
#pragma unroll
  for (int j = 0; j < TL::MC; ++j) {
    const int col = (wc * TL::MC + j) * 16 + l16;
    b0[j] = frag(B0, col, q);
    b1[j] = frag(B1, col, q);
  }
#pragma unroll
  for (int i = 0; i < TL::MR; ++i) {
    const int row = (wr * TL::MR + i) * 16 + l16;
    const v4u ar0 = frag(A0, row, q), ar1 = frag(A1, row, q);
    const v4u ai0 = frag(A0, TL::R + row, q), ai1 = frag(A1, TL::R + row, q);
#pragma unroll
    for (int j = 0; j < TL::MC; ++j) {
      re[i][j] = mfma_fp8_x128(ar0, ar1, b0[j], b1[j], re[i][j]);
      im[i][j] = mfma_fp8_x128(ai0, ai1, b0[j], b1[j], im[i][j]);
    }
  }
}

template <bool SPLIT, bool TR = false, class TL = Tile0, bool F8 = false>
__device__ __forceinline__ void bx_sub(const uint4* sub, int wr, int wc, int l16, int q,
                                       v4f (&re)[(TL::MR)][(TL::MC)], v4f (&im)[(TL::MR)][(TL::MC)]) {
  static_assert(!F8 || (!SPLIT && !TR), "FP8 operands: single product, plain accumulators");
  const uint16_t* As = reinterpret_cast<const uint16_t*>(sub);
  const uint16_t* Bs = As + 8 * Bx<SPLIT, TL>::A_VEC;
  v8bf bh[(TL::MC)], bl[(TL::MC)];
#pragma unroll
  for (int j = 0; j < TL::MC; ++j) {
    const int col = (wc * TL::MC + j) * 16 + l16;
    bh[j] = lds_frag(Bs, col, q);
    if (SPLIT) bl[j] = lds_frag(Bs, TL::C + col, q);
  }
#pragma unroll
  for (int i = 0; i < TL::MR; ++i) {
    const int row = (wr * TL::MR + i) * 16 + l16;
    if (SPLIT) {
      const v8bf arh = lds_frag(As, row, q);
      const v8bf arl = lds_frag(As, TL::R + row, q);
      const v8bf aih = lds_frag(As, 2 * TL::R + row, q);
      const v8bf ail = lds_frag(As, 3 * TL::R + row, q);
#pragma unroll
      for (int j = 0; j < TL::MC; ++j) {
        re[i][j] = mfma_ab(arh, bh[j], re[i][j], TR);
        re[i][j] = mfma_ab(arh, bl[j], re[i][j], TR);
        re[i][j] = mfma_ab(arl, bh[j], re[i][j], TR);
        im[i][j] = mfma_ab(aih, bh[j], im[i][j], TR);
        im[i][j] = mfma_ab(aih, bl[j], im[i][j], TR);
        im[i][j] = mfma_ab(ail, bh[j], im[i][j], TR);
      }
    } else {
      const v8bf arh = lds_frag(As, row, q);
      const v8bf aih = lds_frag(As, TL::R + row, q);
#pragma unroll
      for (int j = 0; j < TL::MC; ++j) {
        if constexpr (F8) {
          re[i][j] = mfma_fp8_pair(arh, bh[j], re[i][j]);
          im[i][j] = mfma_fp8_pair(aih, bh[j], im[i][j]);
        } else {
          re[i][j] = mfma_ab(arh, bh[j], re[i][j], TR);
          im[i][j] = mfma_ab(aih, bh[j], im[i][j], TR);
        }
      }
    }
  }
}

// The nsub sub-chunks of LDS stage `stage`.  A full stage is one straight
// block, so that the LDS reads of a sub-chunk can be scheduled under the
// MFMAs of the one before.
template <bool SPLIT, bool TR = false, class TL = Tile0, bool F8 = false>
__device__ __forceinline__ void bx_chunk(const uint4* stage, int wr, int wc, int l16, int q,
                                         v4f (&re)[(TL::MR)][(TL::MC)], v4f (&im)[(TL::MR)][(TL::MC)], int nsub) {
  constexpr int KS = Bx<SPLIT, TL>::KS;
  if constexpr (F8 && KS % 2 == 0) {
    if (nsub == KS) {
#pragma unroll
      for (int sc = 0; sc < KS; sc += 2) bx_sub_f8x2<TL>(stage + sc * Bx<SPLIT, TL>::SUB_VEC, wr, wc, l16, q, re, im);
      return;
    }
  }
  if (nsub == KS) {
#pragma unroll
    for (int sc = 0; sc < KS; ++sc)
      bx_sub<SPLIT, TR, TL, F8>(stage + sc * Bx<SPLIT, TL>::SUB_VEC, wr, wc, l16, q, re, im);
    return;
  }
#pragma unroll
  for (int sc = 0; sc < KS; ++sc) {
    if (sc >= nsub) break;
    bx_sub<SPLIT, TR, TL, F8>(stage + sc * Bx<SPLIT, TL>::SUB_VEC, wr, wc, l16, q, re, im);
  }
}

template <bool SPLIT>
__global__ void __launch_bounds__(BX_THREADS) exchange_bf16_kernel(ExchangeBf16Args a) {
  using C = Bx<SPLIT>;
  extern __shared__ uint4 lds_u4[];
  __shared__ double red[BX_THREADS / 64];
  // epilogue staging aliases the stage buffers: S[row][col] complex float, stride n + 1
  float2* S = reinterpret_cast<float2*>(lds_u4);

  const int n = a.nocc, Kp = a.kpad, wb = a.wb, xb = a.xb;
  const int rows_valid = wb * n;
  const int nwb = (int)ceil_div(a.nwalkers, wb);
  const int lin = xcd_remap(blockIdx.x, gridDim.x);
  const int split = lin % a.xsplit, bw = lin / a.xsplit;
  const int nxb = (int)ceil_div(a.nchol, xb);
  const int xb_begin = (int)((int64_t)nxb * split / a.xsplit);
  const int xb_end = (int)((int64_t)nxb * (split + 1) / a.xsplit);
  if (bw >= nwb) return;

  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, l16 = lane & 15, q = lane >> 4;
  const int wr = wave % BX_GR, wc = wave / BX_GR;
  const int64_t arows = (int64_t)a.nwalkers * n, brows = (int64_t)a.nchol * n;
  const int64_t aplane = arows * Kp, bplane = brows * Kp;
  const int nch32 = Kp / BX_KC, nch = (int)ceil_div(nch32, C::KS);  // 32-wide K chunks, LDS stages

  // Source offsets (bf16 units, K chunk 0) of the vectors this thread stages.
  // Rows past the end of an array are clamped to its last row: padding rows
  // and columns of the tile only feed outputs that the epilogue never reads.
  int64_t a_src[C::A_PER], b_src[C::B_PER];
#pragma unroll
  for (int i = 0; i < C::A_PER; ++i) {
    const int v = tid + i * BX_THREADS;
    const int p = v / (BX_R * 4), rem = v % (BX_R * 4);
    int64_t r = (int64_t)bw * wb * n + rem / 4;
    if (r > arows - 1) r = arows - 1;
    a_src[i] = C::a_plane(p) * aplane + r * Kp + (rem % 4) * 8;
  }
  v4u pa[C::KS][C::A_PER], pb[C::KS][C::B_PER];
  const int nsub0 = min(C::KS, nch32);

  // trace epilogue: walker wl of the block is summed by threads [wl G, (wl + 1) G)
  const int G = BX_THREADS / wb;
  const int my_wl = tid / G, my_idx = tid % G;
  // vectors staged per epilogue round: as many as the stage buffers hold
  const int xg = max(1, min(xb, (int)((int64_t)a.lds_bytes / ((int64_t)rows_valid * (n + 1) * sizeof(float2)))));
  double wacc_r = 0.0, wacc_i = 0.0;

  if (xb_begin < xb_end) {
    bx_set_b<SPLIT>(b_src, tid, xb_begin, xb, n, brows, bplane, Kp);
    bx_fetch<SPLIT>(pa, pb, a.gplanes, a.lplanes, a_src, b_src, 0, nsub0);
  }
  for (int xblk = xb_begin; xblk < xb_end; ++xblk) {
    const int x0 = xblk * xb;
    bx_stash<SPLIT>(lds_u4, pa, pb, tid, 0, nsub0);
    __syncthreads();
    v4f re[BX_MR][BX_MC], im[BX_MR][BX_MC];
#pragma unroll
    for (int i = 0; i < BX_MR; ++i)
#pragma unroll
      for (int j = 0; j < BX_MC; ++j) {
        re[i][j] = zero4f();
        im[i][j] = zero4f();
      }
    for (int c = 0; c < nch; ++c) {
      // next chunk, or after the last one the first chunk of the next x block
      // (stashed after the epilogue)
      const bool more = c + 1 < nch, next_blk = !more && xblk + 1 < xb_end;
      const int nsub = min(C::KS, nch32 - c * C::KS);
      const int nsub_next = more ? min(C::KS, nch32 - (c + 1) * C::KS) : nsub0;
      if (next_blk) bx_set_b<SPLIT>(b_src, tid, xblk + 1, xb, n, brows, bplane, Kp);
      if (more || next_blk)
        bx_fetch<SPLIT>(pa, pb, a.gplanes, a.lplanes, a_src, b_src, more ? (c + 1) * C::KS : 0, nsub_next);
      bx_chunk<SPLIT>(lds_u4 + (c & 1) * C::STAGE_VEC, wr, wc, l16, q, re, im, nsub);
      if (more) bx_stash<SPLIT>(lds_u4, pa, pb, tid, (c + 1) & 1, nsub_next);
      __syncthreads();
    }
    // epilogue: the block's vectors in groups of xg, staged as S[v][row][col]
    // (row stride n + 1), then sum_ij T_ij T_ji per walker
    for (int g0 = 0; g0 < xb; g0 += xg) {
      const int gn = min(min(xg, xb - g0), a.nchol - (x0 + g0));
      if (gn <= 0) break;
#pragma unroll
      for (int j = 0; j < BX_MC; ++j) {
        const int colg = (wc * BX_MC + j) * 16 + l16;
        const int v = colg / n - g0, col = colg % n;
        if (v >= 0 && v < gn) {
          float2* Sv = S + (int64_t)v * rows_valid * (n + 1) + col;
#pragma unroll
          for (int i = 0; i < BX_MR; ++i)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int row = (wr * BX_MR + i) * 16 + 4 * q + r;
              if (row < rows_valid) Sv[row * (n + 1)] = make_float2(re[i][j][r], im[i][j][r]);
            }
        }
      }
      __syncthreads();
      if (my_wl < wb) {
        // sum_ij S_ij S_ji = sum_r S_rr^2 + 2 sum_{r<c} S_rc S_cr.  Each unordered
        // pair {r, c} is visited once from r at distance d = c - r mod n,
        // d = 1 .. (n-1)/2, plus d = n/2 from r < n/2 when n is even.  Task
        // (v, r, h) of a thread walks half h of row r's distances.
        const int ntask = gn * n * 2;
        for (int task = my_idx; task < ntask; task += G) {
          const int h = task & 1, vr = task >> 1, v = vr / n, r = vr % n;
          const int cnt = (n - 1) / 2 + ((n % 2 == 0 && r < n / 2) ? 1 : 0);
          const int half = (cnt + 1) / 2;
          const int d0 = h * half, d1 = h ? cnt : half;
          const float2* Sw = S + ((int64_t)v * rows_valid + my_wl * n) * (n + 1);
          if (h == 0) {
            const float2 t = Sw[r * (n + 1) + r];
            wacc_r += (double)t.x * t.x - (double)t.y * t.y;
            wacc_i += 2.0 * (double)t.x * t.y;
          }
          int c = r + 1 + d0;
          if (c >= n) c -= n;
          const float2* p1 = Sw + r * (n + 1) + c;  // S[r][c]
          const float2* p2 = Sw + c * (n + 1) + r;  // S[c][r]
          double orr = 0.0, oii = 0.0;
          for (int d = d0; d < d1; ++d) {
            const float2 a = *p1, b = *p2;
            orr += (double)a.x * b.x - (double)a.y * b.y;
            oii += (double)a.x * b.y + (double)a.y * b.x;
            ++c;
            ++p1;
            p2 += n + 1;
            if (c == n) {
              c = 0;
              p1 -= n;
              p2 -= n * (n + 1);
            }
          }
          wacc_r += 2.0 * orr;
          wacc_i += 2.0 * oii;
        }
      }
      __syncthreads();
    }
  }
  for (int wl = 0; wl < wb; ++wl) {
    const bool mine = my_wl == wl;
    const double vr = block_sum(mine ? wacc_r : 0.0, red);
    const double vi = block_sum(mine ? wacc_i : 0.0, red);
    const int w = bw * wb + wl;
    if (tid == 0 && w < a.nwalkers) st(a.exx_part, (int64_t)split * a.nwalkers + w, cmake(vr, vi));
  }
}

// C = A B^T for split-BF16 operand planes (A: 4 planes, complex; B: 2 planes,
// real), K chunk range of split z, 128 x 160 tile per workgroup; SPLIT = false
// uses the hi planes only (1 product).  Epilogues:
//   BF16_EPI_PARTIAL  out[z_offset + z][row][col] = scale C     (force bias; complex FP32
//                     into out32 when set)
//   BF16_EPI_VHS      out[w][i][j] = out[w][j][i] = i scale C[w][p],
//                     (i, j) = (tri_i[p], tri_j[p])              (VHS); with
//                     out32, complex FP32 and the upper triangle i <= j only
//   BF16_EPI_VATOM    the split planes of i scale C[w][col], rounded to FP32, at out16[w atom_stride
//                     + atom 1024 + plane 256 + r 16 + c] for col = atom 256 + r 16 + c (VHS atoms)
//   BF16_EPI_ADD      out[w][col][c] = base[w][col][c] + scale C[r][col], r = w group + c,
//                     optionally with the split planes of the sum, transposed (one body)
//   BF16_EPI_QDOT     out[((cb ksplit + z) BX_GC + wc) m + row] = sum over the
//                     wave's columns col = (i, m') (m' < N) of d[row][i][m'] C[row][col],
//                     d = g[src(row)] - gref                     (exchange g^T Wq g)
// With a unit table (block-triangular B), the workgroup's column unit u gives
// (column tile, K chunk range) = (units[u].x, [units[u].y, units[u].z)) and the
// QDOT slot is u BX_GC + wc; otherwise column tile cb and split z of ksplit.
// TR (transposed accumulators, see bx_sub) is used with the ADD and VATOM epilogues only.
// Workgroups that run together share the B tile (row blocks vary fastest), or
// with cols_fast the A tile (column blocks vary fastest).
template <bool SPLIT, bool TR, bool DL, class TL, bool F8 = false>
__global__ void __launch_bounds__(TL::THREADS) bf16_gemm_kernel(Bf16GemmArgs a) {
  using C = Bx<SPLIT, TL>;
  extern __shared__ uint4 lds_u4[];
  const int mb = (int)ceil_div(a.m, TL::R), nb = (int)ceil_div(a.nn, TL::C);
  const int ncu = a.units ? a.nunits : nb;  // column units of the grid
  const int lin = xcd_remap(blockIdx.x, gridDim.x);
  int rb = lin % mb, cu = (lin / mb) % ncu, z = lin / mb / ncu;
  if (a.cols_fast && !a.units) {
    cu = lin % ncu;
    rb = (lin / ncu) % mb;
    z = lin / ncu / mb;
  }
  if (a.units && a.rgroup > 0 && a.rgroup < mb) {
    // groups of rgroup row blocks, each over every unit (the workgroups that run
    // together on one XCD: rgroup row blocks x several units of one K block)
    const int R = a.rgroup, rem = lin % (ncu * R);
    rb = (lin / (ncu * R)) * R + rem % R;
    cu = rem / R;
    z = 0;
    if (rb >= mb) return;
  }
  if (z >= (a.units ? 1 : a.ksplit)) return;
  int cb = cu, c0, c1;
  if (a.units) {
    const int4 e = a.units[cu];
    cb = e.x;
    c0 = e.y;
    c1 = e.z;
  } else {
    c0 = (int)((int64_t)a.kchunks * z / a.ksplit);
    c1 = (int)((int64_t)a.kchunks * (z + 1) / a.ksplit);
  }

  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, l16 = lane & 15, q = lane >> 4;
  const int wr = wave % TL::GR, wc = wave / TL::GR;
  // staged rows past the end are clamped to the last row (their outputs are not written)
  int64_t a_src[C::A_PER], b_src[C::B_PER];
#pragma unroll
  for (int i = 0; i < C::A_PER; ++i) {
    const int v = tid + i * TL::THREADS;
    const int p = v / (TL::R * 4), rem = v % (TL::R * 4);
    int64_t r = (int64_t)rb * TL::R + rem / 4;
    if (r > a.m - 1) r = a.m - 1;
    if (a.a_gstride > 0) r = (r / a.group) * a.a_gstride + a.a_goff + r % a.group;
    a_src[i] = C::a_plane(p) * a.a_plane + r * a.lda + (rem % 4) * 8;
  }
  {
#pragma unroll
    for (int i = 0; i < C::B_PER; ++i) {
      const int v = (tid + i * TL::THREADS) % C::B_VEC;
      const int p = v / (TL::C * 4), rem = v % (TL::C * 4);
      int64_t r = (int64_t)cb * TL::C + rem / 4;
      if (r > a.nn - 1) r = a.nn - 1;
      b_src[i] = C::b_plane(p) * a.b_plane + r * a.ldb + (rem % 4) * 8;
    }
    v4f re[(TL::MR)][(TL::MC)], im[(TL::MR)][(TL::MC)];
#pragma unroll
    for (int i = 0; i < TL::MR; ++i)
#pragma unroll
      for (int j = 0; j < TL::MC; ++j) {
        re[i][j] = zero4f();
        im[i][j] = zero4f();
      }
    // LDS stages of up to KS sub-chunks of 32 K values over this split's range
    // [c0, c1), TL::ST buffers: stage t + ST - 1 is staged while t is computed
    // (three stages: from registers loaded during stage t - 1 without DL)
    const int nst = (int)ceil_div(c1 - c0, C::KS);
    auto nsub_of = [&](int t) { return min(C::KS, c1 - (c0 + t * C::KS)); };
    if constexpr (DL) {
      // direct loads: stage t + ST - 1 is issued before stage t is computed; the
      // wait after it leaves that stage in flight and lands stage t + 1
      constexpr int ST = TL::ST;
      static_assert(ST >= 2 && ST <= 4 && (ST - 2) * C::KS <= 3, "two to four LDS stages, at most 3 sub-chunks pending");
      int64_t da[(Dl<SPLIT, TL>::A_INS)], db[(Dl<SPLIT, TL>::B_INS)];
      dl_sources<SPLIT, TL>(da, db, a, rb, cb, wave, lane);
      // prologue: stages 0 .. ST - 2 issued; stage 0 landed, the rest may stay in flight
      int pend0 = 0;
#pragma unroll
      for (int s = 0; s < ST - 1; ++s)
        if (s < nst) {
          dl_issue<SPLIT, TL>(lds_u4, s, nsub_of(s), a.a, a.b, da, db, c0 + s * C::KS, wave, lane);
          if (s > 0) pend0 += nsub_of(s);
        }
      dl_wait<SPLIT, TL>(pend0);
      dl_barrier();
      int buf = 0;
      for (int t = 0; t < nst; ++t) {
        if (t + ST - 1 < nst)
          dl_issue<SPLIT, TL>(lds_u4, buf == 0 ? ST - 1 : buf - 1, nsub_of(t + ST - 1), a.a, a.b, da, db,
                              c0 + (t + ST - 1) * C::KS, wave, lane);
        bx_chunk<SPLIT, TR, TL, F8>(lds_u4 + buf * C::STAGE_VEC, wr, wc, l16, q, re, im, nsub_of(t));
        // stage t + 1 must land; the ST - 2 stages after it may stay in flight
        int pend = 0;
#pragma unroll
        for (int s = 2; s < ST; ++s)
          if (t + s < nst) pend += nsub_of(t + s);
        dl_wait<SPLIT, TL>(pend);
        dl_barrier();
        buf = buf == ST - 1 ? 0 : buf + 1;
      }
    } else {
      static_assert(TL::ST == 3, "register staging uses three stages");
    v4u pa[C::KS][C::A_PER], pb[C::KS][C::B_PER];
    if (nst > 0) {
      bx_fetch<SPLIT, TL>(pa, pb, a.a, a.b, a_src, b_src, c0, nsub_of(0));
      bx_stash<SPLIT, TL>(lds_u4, pa, pb, tid, 0, nsub_of(0));
    }
    if (nst > 1) {
      bx_fetch<SPLIT, TL>(pa, pb, a.a, a.b, a_src, b_src, c0 + C::KS, nsub_of(1));
      bx_stash<SPLIT, TL>(lds_u4, pa, pb, tid, 1, nsub_of(1));
    }
    if (nst > 2) bx_fetch<SPLIT, TL>(pa, pb, a.a, a.b, a_src, b_src, c0 + 2 * C::KS, nsub_of(2));
    __syncthreads();
    int buf = 0;
    for (int t = 0; t < nst; ++t) {
      bx_chunk<SPLIT, TR, TL, F8>(lds_u4 + buf * C::STAGE_VEC, wr, wc, l16, q, re, im, nsub_of(t));
      const int nbuf = buf == 0 ? 2 : buf - 1;  // (t + 2) % 3: the buffer computed at t - 1
      if (t + 2 < nst) bx_stash<SPLIT, TL>(lds_u4, pa, pb, tid, nbuf, nsub_of(t + 2));
      if (t + 3 < nst) bx_fetch<SPLIT, TL>(pa, pb, a.a, a.b, a_src, b_src, c0 + (t + 3) * C::KS, nsub_of(t + 3));
      __syncthreads();
      buf = buf == 2 ? 0 : buf + 1;
    }
    }
    if (a.epilogue == BF16_EPI_QDOT) {
      // (the quadratic form runs with tile shape 0 and plain accumulators only:
      // launch_bf16_gemm; each epilogue is compiled only where it can run)
      if constexpr (!TR && std::is_same<TL, Tile0>::value) {
      // per-row dot with d over this wave's 80 columns, then across the 16 lanes
      // of a row; lane l16 = 4 i + r keeps the sum of row (i, r).  The lane's
      // columns and d_ref are set up once, and the next row's d values are
      // loaded before this row's dot and cross-lane sum (one row of loads in
      // flight under the arithmetic; same sums in the same order as row by row).
      int64_t offj[(TL::MC)];
      bool okj[(TL::MC)];
      cplx refj[(TL::MC)];
#pragma unroll
      for (int j = 0; j < TL::MC; ++j) {
        const int64_t col = (int64_t)cb * TL::C + (wc * TL::MC + j) * 16 + l16;
        const int io = (int)(col / a.dkp), mm = (int)(col % a.dkp);
        okj[j] = col < a.nn && mm < a.dn;
        offj[j] = okj[j] ? (int64_t)io * a.dn + mm : 0;  // offset 0 of the row outside the matrix
        refj[j] = (a.dref && okj[j]) ? ld(a.dref, offj[j]) : cmake(0.0, 0.0);
      }
      auto load_row = [&](int i, int r, cplx (&d)[(TL::MC)]) __attribute__((always_inline)) {
        int64_t row = (int64_t)rb * TL::R + (wr * TL::MR + i) * 16 + 4 * q + r;
        if (row > a.m - 1) row = a.m - 1;  // clamped rows are computed, never stored
        const double* drow = a.dg + 2 * (a.walker_of ? (int64_t)a.walker_of[row] : row) * a.dstride;
#pragma unroll
        for (int j = 0; j < TL::MC; ++j) d[j] = ld(drow, offj[j]);
      };
      double myr = 0.0, myi = 0.0;
      cplx dnext[(TL::MC)];
      load_row(0, 0, dnext);
#pragma unroll
      for (int i = 0; i < TL::MR; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          cplx d[(TL::MC)];
#pragma unroll
          for (int j = 0; j < TL::MC; ++j) d[j] = dnext[j];
          if (i * 4 + r + 1 < 4 * TL::MR) load_row((i * 4 + r + 1) / 4, (i * 4 + r + 1) % 4, dnext);
          double sr = 0.0, si = 0.0;
#pragma unroll
          for (int j = 0; j < TL::MC; ++j) {
            if (okj[j]) {
              const cplx dd = csub(d[j], refj[j]);
              const double yr = re[i][j][r], yi = im[i][j][r];
              sr += dd.re * yr - dd.im * yi;
              si += dd.re * yi + dd.im * yr;
            }
          }
#pragma unroll
          for (int o = 8; o >= 1; o >>= 1) {
            sr += __shfl_xor(sr, o, 16);
            si += __shfl_xor(si, o, 16);
          }
          if (l16 == 4 * i + r) {
            myr = sr;
            myi = si;
          }
        }
      if (l16 < 4 * TL::MR) {
        const int i = l16 >> 2, r = l16 & 3;
        const int64_t row = (int64_t)rb * TL::R + (wr * TL::MR + i) * 16 + 4 * q + r;
        const int64_t slot = (a.units ? (int64_t)cu : (int64_t)cb * a.ksplit + z) * TL::GC + wc;
        if (row < a.m) {
          const double f = a.scale * (a.row_scale ? a.row_scale[row] : 1.0);
          st(a.out, slot * a.m + row, cmake(f * myr, f * myi));
        }
      }
      }
    } else if (a.epilogue == BF16_EPI_ADD) {
      if constexpr (TR) {
      // transposed accumulators (TR): lane l16 walks the rows, so the FP64
      // reads and writes of 16 lanes are consecutive columns c of one walker
      const double* __restrict__ bsrc = a.base;
      double* __restrict__ odst = a.out;
      uint16_t* __restrict__ tpd = a.tp;
#pragma unroll
      for (int i = 0; i < TL::MR; ++i) {
        const int64_t row = (int64_t)rb * TL::R + (wr * TL::MR + i) * 16 + l16;
        if (row >= a.m) continue;
        const int64_t w = row / a.group;
        const int c = (int)(row % a.group);
#pragma unroll
        for (int j = 0; j < TL::MC; ++j) {
          const int64_t col0 = (int64_t)cb * TL::C + (wc * TL::MC + j) * 16 + 4 * q;
          // the 4 base values are loaded before any store (columns past the end
          // read the last column), so the loads are in flight together; base
          // and out are distinct buffers or the same one indexed alike
          cplx base[4];
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int64_t col = col0 + r > a.nn - 1 ? a.nn - 1 : col0 + r;
            base[r] = ld(bsrc, w * a.b_ws + col * a.b_ld + a.b_off + c);
          }
          uint16_t h[4][4];  // [plane][r]: re_hi, re_lo, im_hi, im_lo
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int64_t col = col0 + r;
            h[0][r] = h[1][r] = h[2][r] = h[3][r] = 0;
            if (col >= a.nn) continue;
            const cplx v = cmake(base[r].re + a.scale * re[i][j][r], base[r].im + a.scale * im[i][j][r]);
            st(odst, w * a.o_ws + col * a.o_ld + a.o_off + c, v);
            split_bf16(v.re, h[0][r], h[1][r]);
            split_bf16(v.im, h[2][r], h[3][r]);
          }
          if (tpd && col0 < a.nn) {
            const int64_t o = (w * a.tp_rows + a.tp_off + c) * a.tp_ld + col0;
            if (col0 + 3 < a.nn && (o & 3) == 0) {
#pragma unroll
              for (int p = 0; p < 4; ++p)
                *reinterpret_cast<ushort4*>(tpd + p * a.tp_plane + o) =
                    make_ushort4(h[p][0], h[p][1], h[p][2], h[p][3]);
            } else {
#pragma unroll
              for (int r = 0; r < 4; ++r)
                if (col0 + r < a.nn)
#pragma unroll
                  for (int p = 0; p < 4; ++p) tpd[p * a.tp_plane + o + r] = h[p][r];
            }
          }
        }
      }
      }
    } else if (a.epilogue == BF16_EPI_VATOM) {
      if constexpr (TR) {
      // transposed accumulators (TR): each lane holds 4 consecutive columns of one
      // row, so each plane takes one 8-byte store (4 entries of an atom row)
#pragma unroll
      for (int i = 0; i < (TL::MR); ++i) {
        const int64_t row = (int64_t)rb * TL::R + (wr * TL::MR + i) * 16 + l16;
        if (row >= a.m) continue;
        uint16_t* drow = a.out16 + row * a.atom_stride;
#pragma unroll
        for (int j = 0; j < (TL::MC); ++j) {
          const int64_t col0 = (int64_t)cb * TL::C + (wc * TL::MC + j) * 16 + 4 * q;  // nn % 256 == 0: all in range
          if (col0 >= a.nn) continue;
          uint16_t h[4][4];  // [plane][r]: re_hi, re_lo, im_hi, im_lo
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            // i scale (re + i im) = scale (-im + i re), rounded to FP32 as the accumulators hold it
            const float vre = (float)(-(a.scale * im[i][j][r])), vim = (float)(a.scale * re[i][j][r]);
            split_bf16((double)vre, h[0][r], h[1][r]);
            split_bf16((double)vim, h[2][r], h[3][r]);
          }
          uint16_t* dst = drow + (col0 >> 8) * 1024 + (col0 & 255);
#pragma unroll
          for (int p = 0; p < 4; ++p)
            *reinterpret_cast<ushort4*>(dst + 256 * p) = make_ushort4(h[p][0], h[p][1], h[p][2], h[p][3]);
        }
      }
      }
    } else if constexpr (!TR) {
      const int64_t NN = (int64_t)a.nbasis * a.nbasis;
#pragma unroll
      for (int j = 0; j < TL::MC; ++j) {
        const int64_t col = (int64_t)cb * TL::C + (wc * TL::MC + j) * 16 + l16;
        if (col >= a.nn) continue;
        int ti = 0, tj = 0;
        if (a.epilogue == BF16_EPI_VHS) {
          ti = a.tri_i[col];
          tj = a.tri_j[col];
        }
#pragma unroll
        for (int i = 0; i < TL::MR; ++i)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int64_t row = (int64_t)rb * TL::R + (wr * TL::MR + i) * 16 + 4 * q + r;
            if (row >= a.m) continue;
            const double vr = a.scale * re[i][j][r], vi = a.scale * im[i][j][r];
            if (a.epilogue == BF16_EPI_VHS) {
              // i scale (re + i im) = scale (-im + i re)
              if (a.out32) {
                // upper triangle only (ti <= tj); the reader mirrors it
                reinterpret_cast<float2*>(a.out32)[row * NN + (int64_t)ti * a.nbasis + tj] =
                    make_float2((float)-vi, (float)vr);
              } else {
                const cplx v = cmake(-vi, vr);
                double* base = a.out + 2 * row * NN;
                st(base, (int64_t)ti * a.nbasis + tj, v);
                if (ti != tj) st(base, (int64_t)tj * a.nbasis + ti, v);
              }
            } else if (a.out32) {
              reinterpret_cast<float2*>(a.out32)[((int64_t)(a.z_offset + z) * a.m + row) * a.nn + col] =
                  make_float2((float)vr, (float)vi);
            } else {
              st(a.out, ((int64_t)(a.z_offset + z) * a.m + row) * a.nn + col, cmake(vr, vi));
            }
          }
      }
    }
  }
}

// LDS of the exchange: the stages, or more so that the epilogue stages two
// vectors per barrier pair (capped at the 3-product stages), or one vector.
template <bool SPLIT>
size_t bf16_lds_bytes(int wb, int xb, int nocc) {
  const size_t per_vec = sizeof(float2) * (size_t)(wb * nocc) * (nocc + 1);
  size_t lds = Bx<SPLIT>::STAGES;
  const size_t two = per_vec * (size_t)(xb < 2 ? xb : 2);
  if (two > lds) lds = two < BX_LDS_MAX ? two : BX_LDS_MAX;
  return lds > per_vec ? lds : per_vec;
}

// ---------------------------------------------------------------- Taylor series
// out_k^T = out_{k-1}^T VHS / k for each walker (VHS is symmetric), complex x
// complex in split-BF16: A = out_{k-1}^T planes [4][W][rpad][kpad] (row = column
// of the walker matrix, K = basis index), B = the VHS atoms (row = output basis
// index).  re = Ar Br - Ai Bi, im = Ar Bi + Ai Br, each real product as
// hi*hi + hi*lo + lo*hi (SPLIT: 12 MFMAs per tile and K chunk) or hi*hi (4
// MFMAs; the hi planes only); -Ai comes from sign-flipped fragments.
// Workgroup: one walker, 80 rows x 80 output basis indices, 5 waves, wave v
// owns output column tile v and all 5 row tiles.
// B of a K chunk is the 5 x 2 atoms (aj, ak), aj = 5 cb + v, ak = 2 c + h.  Only
// the upper atoms are stored: an atom with aj <= ak is read as stored (row =
// output index, 16 K values per row), one with aj > ak is its mirror (ak, aj),
// whose rows are K values; the MFMA operand of a mirror atom (8 consecutive K
// of one output index per lane) is read with the transposing LDS load
// ds_read_b64_tr_b16, two per plane.  Each stored atom is thus read by the
// workgroups of both its row and its column block, which run together on one
// XCD, so the second read is served by L2.
// Epilogue: every term but the last writes the split planes of out_k^T to a
// buffer of its own (the next term's operand); the last term writes
// q = init + (the earlier terms, read back from their planes) + out_k in FP64,
// with the split planes of q^T (qout).  No running sum is kept between terms.
// The out^T planes have a walker stride fixed by the buffer capacity wcap;
// their padding rows and K entries are never written and must be zero.
constexpr int TY_WAVES = 5, TY_R = 80, TY_C = 80, TY_THREADS = 64 * TY_WAVES;
static_assert(TY_C == 5 * 16 && BX_KC == 2 * 16, "a workgroup's B chunk is 5 x 2 atoms");

typedef __bf16 v4bf __attribute__((ext_vector_type(4)));

// LDS slot (16-byte vector) of half h of row r of an atom plane: the halves of
// rows 4..11 are swapped, so that 8 or 16 lanes reading consecutive rows with
// 16-byte loads cover distinct slots modulo 8 and modulo 16 (no bank conflicts
// with 32 or 64 banks of 4 bytes), and the 16 lanes of a transposing load (4
// aligned rows x 32 bytes) cover one 128-byte span
__device__ __forceinline__ int atom_vec(int r, int h) { return r * 2 + (h ^ (((r >> 2) ^ (r >> 3)) & 1)); }

// B fragment of an atom plane at LDS vector offset vb: lane (l16, q) gets 8
// consecutive K values (8 (q & 1) .. + 7 within the atom) of output index l16.
// Stored atom: row l16, one 16-byte load.  Mirror atom: column l16 of rows
// 8 (q & 1) .. + 7, two transposing loads of 4 rows each (lane r' of a 16-lane
// group supplies row r' / 4, quarter r' % 4; lane i receives element i of each
// of the 4 rows).
__device__ __forceinline__ v8bf ty_bfrag(const uint4* lds, int vb, int l16, int q, bool mirror) {
  if (!mirror) return as_v8bf(reinterpret_cast<const v4u*>(lds)[vb + atom_vec(l16, q & 1)]);
  v4bf t[2];
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int r = 8 * (q & 1) + 4 * u + (l16 >> 2), qt = l16 & 3;
    const char* p = reinterpret_cast<const char*>(lds + vb + atom_vec(r, qt >> 1)) + 8 * (qt & 1);
    t[u] = __builtin_amdgcn_ds_read_tr16_b64_v4bf16((__attribute__((address_space(3))) v4bf*)p);
  }
  return __builtin_shufflevector(t[0], t[1], 0, 1, 2, 3, 4, 5, 6, 7);
}

__device__ __forceinline__ v8bf neg_bf16(v8bf x) {
  return __builtin_bit_cast(v8bf, __builtin_bit_cast(v4u, x) ^ (v4u){0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u});
}

// Staging of one LDS stage, in 16-byte vectors: one K chunk of the four planes
// of each operand (SPLIT) or of the hi planes (re_hi, im_hi: physical planes 0
// and 2) for the single product.  The single product's 20 KiB stages let four
// workgroups share a CU (two stages each); two K chunks per stage (fewer
// barriers, two workgroups per CU) measured 2.5% slower.
template <bool SPLIT>
struct Ty {
  static constexpr int NP = SPLIT ? 4 : 2, KS = 1;
  static constexpr int SUB_A = NP * TY_R * 4, SUB_VEC = SUB_A + NP * TY_C * 4;
  static constexpr int STAGE_VEC = KS * SUB_VEC;
  static constexpr size_t LDS = 2 * 16 * (size_t)STAGE_VEC;  // two stages
  static_assert(SUB_A % (64 * TY_WAVES) == 0, "each wave stages whole load instructions");
  __device__ static int phys(int pi) { return SPLIT ? pi : 2 * pi; }
};

// 4 consecutive K entries (of which `valid` are < N) of the four planes of a
// transposed Taylor operand at offset o
__device__ __forceinline__ void store_planes_ty(uint16_t* dst, int64_t tplane, int64_t o, const uint16_t (&h)[4][4],
                                                int valid) {
  if (valid >= 4) {
#pragma unroll
    for (int p = 0; p < 4; ++p)
      *reinterpret_cast<ushort4*>(dst + p * tplane + o) = make_ushort4(h[p][0], h[p][1], h[p][2], h[p][3]);
  } else {
#pragma unroll
    for (int r = 0; r < 4; ++r)
      if (r < valid)
#pragma unroll
        for (int p = 0; p < 4; ++p) dst[p * tplane + o + r] = h[p][r];
  }
}

// Epilogue of a Taylor term from the accumulators (transposed tiles: element r
// of tile i is row crow = i 16 + l16, a column of the walker matrix, and basis
// index kcol = cb 80 + wave 16 + 4 q + r, so 16 lanes touch 16 consecutive
// columns of q and each lane 4 consecutive K of a plane).
__device__ __forceinline__ void ty_epilogue(const TaylorBf16Args& a, const v4f (&re)[5], const v4f (&im)[5], int w,
                                            int rb, int cb, int wave, int l16, int q) {
  const int N = a.nbasis, Kp = a.kpad;
  const int64_t tplane = (int64_t)a.wcap * a.rpad * Kp;
  const int kc0 = cb * TY_C + wave * 16 + 4 * q;
#pragma unroll
  for (int i = 0; i < 5; ++i) {
    const int crow = rb * TY_R + i * 16 + l16;
    if (crow >= a.ncol || kc0 >= N) continue;
    // a multiple of 4: Kp % 32 == 0, kc0 % 4 == 0; kc0 + 3 < Kp, and the
    // planes' K padding (kcol >= N) is zero
    const int64_t o = ((int64_t)w * a.rpad + crow) * Kp + kc0;
    // last term: the earlier terms at these 4 K entries, read back from their
    // planes (hi + lo), smallest terms first, in FP64
    double sre[4] = {0.0, 0.0, 0.0, 0.0}, sim[4] = {0.0, 0.0, 0.0, 0.0};
    if (a.last) {
      for (int t = a.nterms - 1; t >= 0; --t) {
        const uint16_t* tp = a.terms[t] + o;
        const ushort4 hr = *reinterpret_cast<const ushort4*>(tp);
        const ushort4 hi = *reinterpret_cast<const ushort4*>(tp + 2 * tplane);
        const ushort4 lr = *reinterpret_cast<const ushort4*>(tp + tplane);
        const ushort4 li = *reinterpret_cast<const ushort4*>(tp + 3 * tplane);
        const double vr[4] = {(double)bf2f(hr.x) + bf2f(lr.x), (double)bf2f(hr.y) + bf2f(lr.y),
                              (double)bf2f(hr.z) + bf2f(lr.z), (double)bf2f(hr.w) + bf2f(lr.w)};
        const double vi[4] = {(double)bf2f(hi.x) + bf2f(li.x), (double)bf2f(hi.y) + bf2f(li.y),
                              (double)bf2f(hi.z) + bf2f(li.z), (double)bf2f(hi.w) + bf2f(li.w)};
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          sre[r] += vr[r];
          sim[r] += vi[r];
        }
      }
    }
    uint16_t ht[4][4], hq[4][4];  // [plane][r] of out_k and of q
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int kcol = kc0 + r;
#pragma unroll
      for (int p = 0; p < 4; ++p) ht[p][r] = hq[p][r] = 0;
      if (kcol >= N) continue;
      const cplx v = cmake(a.alpha * re[i][r], a.alpha * im[i][r]);
      if (a.last) {
        const int64_t qi = ((int64_t)w * N + kcol) * a.ncol + crow;
        const cplx qv = cadd(ld(a.init, qi), cmake(v.re + sre[r], v.im + sim[r]));
        st(a.q, qi, qv);
        split_bf16(qv.re, hq[0][r], hq[1][r]);
        split_bf16(qv.im, hq[2][r], hq[3][r]);
      }
      split_bf16(v.re, ht[0][r], ht[1][r]);
      split_bf16(v.im, ht[2][r], ht[3][r]);
    }
    // planes of out_k^T (all four: the last term sums hi + lo, a one-product
    // next term reads the hi planes 0 and 2) and, after the last term, of q^T
    if (a.tout) store_planes_ty(a.tout, tplane, o, ht, N - kc0);
    if (a.qout) store_planes_ty(a.qout, tplane, o, hq, N - kc0);
  }
}

template <bool SPLIT>
__global__ void __launch_bounds__(TY_THREADS) taylor_bf16_kernel(TaylorBf16Args a) {
  using T = Ty<SPLIT>;
  extern __shared__ uint4 lds_u4[];
  const int N = a.nbasis, Kp = a.kpad;
  const int nrb = a.rpad / TY_R, ncb = (int)ceil_div(N, TY_C);
  const int lin = xcd_remap(blockIdx.x, gridDim.x);
  const int cb = lin % ncb, rb = (lin / ncb) % nrb, w = lin / ncb / nrb;
  if (w >= a.nwalkers) return;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, l16 = lane & 15, q = lane >> 4;
  const int64_t tplane = (int64_t)a.wcap * a.rpad * Kp;
  const int nch = Kp / BX_KC, nst = (int)ceil_div(nch, T::KS);
  // direct global -> LDS staging: slot s of an A plane region holds (row s / 4,
  // segment (s % 4) ^ ((row / 4) % 4)), so each wave's loads fill 64
  // consecutive slots; every wave issues NP loads per operand and sub-chunk.
  // B: vector v of the chunk's B region is vector e = v % 32 of plane pi of
  // atom slot (sj, sk) (v / 32 = (2 sj + sk) NP + pi), i.e. atom row e / 2,
  // half (e % 2) ^ swizzle (atom_vec); the atom is (aj, ak) or its mirror.
  const int NA = a.natom_side;
  const uint16_t* dla[T::NP];
  int baj[T::NP], bsk[T::NP], boff[T::NP];
#pragma unroll
  for (int j = 0; j < T::NP; ++j) {
    const int sl = wave * T::NP * 64 + j * 64 + lane;
    const int row = sl >> 2, seg = (sl & 3) ^ lds_swz_ty(row);
    const int pi = row / TY_R, r = row % TY_R;
    dla[j] = a.tin + T::phys(pi) * tplane + ((int64_t)w * a.rpad + rb * TY_R + r) * Kp + seg * 8;
    const int sp = sl >> 5, e = sl & 31, slot = sp / T::NP, bpi = sp % T::NP, ar = e >> 1;
    baj[j] = 5 * cb + (slot >> 1);
    bsk[j] = slot & 1;
    boff[j] = T::phys(bpi) * 256 + ar * 16 + 8 * ((e & 1) ^ (((ar >> 2) ^ (ar >> 3)) & 1));
  }
  const uint16_t* wat = a.vatoms + (int64_t)w * a.atom_stride;
  auto dl_issue_ty = [&](int t, int buf) __attribute__((always_inline)) {
    uint4* base = lds_u4 + buf * T::STAGE_VEC;
#pragma unroll
    for (int sc = 0; sc < T::KS; ++sc) {
      const int c = t * T::KS + sc;
      if (c >= nch) break;
#pragma unroll
      for (int j = 0; j < T::NP; ++j) {
        __builtin_amdgcn_global_load_lds(
            (const void*)(dla[j] + c * BX_KC),
            (__attribute__((address_space(3))) void*)(base + sc * T::SUB_VEC + wave * T::NP * 64 + j * 64), 16, 0, 0);
        // atom (aj, ak) as stored, or its mirror; outside the matrix the zero atom
        const int aj = baj[j], ak = 2 * c + bsk[j];
        const int lo = min(aj, ak), hi = max(aj, ak);
        const uint16_t* src = (hi >= NA) ? a.zatom + boff[j]
                                         : wat + (int64_t)(lo * NA - lo * (lo - 1) / 2 + hi - lo) * 1024 + boff[j];
        __builtin_amdgcn_global_load_lds(
            (const void*)src,
            (__attribute__((address_space(3))) void*)(base + sc * T::SUB_VEC + T::SUB_A + wave * T::NP * 64 + j * 64),
            16, 0, 0);
      }
    }
  };
  v4f re[5], im[5];
#pragma unroll
  for (int i = 0; i < 5; ++i) {
    re[i] = zero4f();
    im[i] = zero4f();
  }
  dl_issue_ty(0, 0);
  wait_vm<0>();
  dl_barrier();
  for (int t = 0; t < nst; ++t) {
    if (t + 1 < nst) dl_issue_ty(t + 1, (t + 1) & 1);
#pragma unroll
    for (int sc = 0; sc < T::KS; ++sc) {
      if (t * T::KS + sc >= nch) break;
      const uint4* stage = lds_u4 + (t & 1) * T::STAGE_VEC + sc * T::SUB_VEC;
      const uint16_t* As = reinterpret_cast<const uint16_t*>(stage);
      const uint4* Bq = stage + T::SUB_A;
      // this wave's output atom aj against K atom ak of its lane group: stored or mirror
      const int c = t * T::KS + sc, bslot = wave * 2 + (q >> 1);
      const int aj = 5 * cb + wave, ak = 2 * c + (q >> 1);
      const bool mirror = aj > ak && aj < NA;
      // operands swapped in every MFMA: the accumulators hold the transposed tiles
      if (SPLIT) {
        const v8bf brh = ty_bfrag(Bq, (bslot * T::NP + 0) * 32, l16, q, mirror);
        const v8bf brl = ty_bfrag(Bq, (bslot * T::NP + 1) * 32, l16, q, mirror);
        const v8bf bih = ty_bfrag(Bq, (bslot * T::NP + 2) * 32, l16, q, mirror);
        const v8bf bil = ty_bfrag(Bq, (bslot * T::NP + 3) * 32, l16, q, mirror);
#pragma unroll
        for (int i = 0; i < 5; ++i) {
          const int arow = i * 16 + l16;
          const v8bf arh = lds_frag_ty(As, arow, q), arl = lds_frag_ty(As, TY_R + arow, q);
          const v8bf aih = lds_frag_ty(As, 2 * TY_R + arow, q), ail = lds_frag_ty(As, 3 * TY_R + arow, q);
          const v8bf naih = neg_bf16(aih), nail = neg_bf16(ail);
          re[i] = mfma_bf16(brh, arh, re[i]);
          re[i] = mfma_bf16(brl, arh, re[i]);
          re[i] = mfma_bf16(brh, arl, re[i]);
          re[i] = mfma_bf16(bih, naih, re[i]);
          re[i] = mfma_bf16(bil, naih, re[i]);
          re[i] = mfma_bf16(bih, nail, re[i]);
          im[i] = mfma_bf16(bih, arh, im[i]);
          im[i] = mfma_bf16(bil, arh, im[i]);
          im[i] = mfma_bf16(bih, arl, im[i]);
          im[i] = mfma_bf16(brh, aih, im[i]);
          im[i] = mfma_bf16(brl, aih, im[i]);
          im[i] = mfma_bf16(brh, ail, im[i]);
        }
      } else {
        const v8bf brh = ty_bfrag(Bq, (bslot * T::NP + 0) * 32, l16, q, mirror);
        const v8bf bih = ty_bfrag(Bq, (bslot * T::NP + 1) * 32, l16, q, mirror);
#pragma unroll
        for (int i = 0; i < 5; ++i) {
          const int arow = i * 16 + l16;
          const v8bf arh = lds_frag_ty(As, arow, q), aih = lds_frag_ty(As, TY_R + arow, q);
          re[i] = mfma_bf16(brh, arh, re[i]);
          re[i] = mfma_bf16(bih, neg_bf16(aih), re[i]);
          im[i] = mfma_bf16(bih, arh, im[i]);
          im[i] = mfma_bf16(brh, aih, im[i]);
        }
      }
    }
    wait_vm<0>();
    dl_barrier();
  }
  ty_epilogue(a, re, im, w, rb, cb, wave, l16, q);
}

// The same term with 16-wide K stages and four LDS stages per workgroup, three
// of them in flight while one is computed (v_mfma_f32_16x16x16_bf16: half the
// MFMA rate, twice the loads in flight of the 32-wide form).  A stage holds the
// 80 rows x 16 K of each A plane and the 5 atoms (5 cb + v, kc) of each B plane,
// 2 KiB per (plane, 16 rows); every lane reads 4 consecutive K (8 bytes) per
// fragment, a mirror atom's with one transposing load per plane.
template <int NP>
struct Ty16 {
  static constexpr int A_VEC = NP * TY_R * 2, B_VEC = NP * 5 * 32, STAGE_VEC = A_VEC + B_VEC;
  static constexpr int STAGES = 4;
  static constexpr size_t LDS = 16 * (size_t)STAGES * STAGE_VEC;
  static constexpr int A_INS = A_VEC / (64 * TY_WAVES), B_INS = B_VEC / (64 * TY_WAVES);  // per wave
  static constexpr int INS = A_INS + B_INS;
  static_assert(A_VEC % (64 * TY_WAVES) == 0 && B_VEC % (64 * TY_WAVES) == 0, "whole load instructions per wave");
  static_assert(2 * INS <= 63, "vmcnt counts two stages of loads");
};

// 16-byte slot of half h of row r (A: of the 80 rows; B: of an atom's 16): the
// halves of rows 8..15 (mod 16) are swapped, so the 8-byte reads of 32 lanes
// (16 rows x two quarters) cover all 64 banks once
__device__ __forceinline__ int ty16_vec(int r, int h) { return r * 2 + (h ^ ((r >> 3) & 1)); }

typedef short v4s __attribute__((ext_vector_type(4)));
__device__ __forceinline__ v4f mfma16_bf16(v4s a, v4s b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ v4s neg16(v4s x) {
  return __builtin_bit_cast(v4s, __builtin_bit_cast(unsigned long long, x) ^ 0x8000800080008000ull);
}
// 4 consecutive K values (quarter g = 4 g .. 4 g + 3 of the 16) of row l16 at LDS vector base vb
__device__ __forceinline__ v4s ty16_frag(const uint4* lds, int vb, int row, int g) {
  const char* p = reinterpret_cast<const char*>(lds + vb + ty16_vec(row, g >> 1)) + 8 * (g & 1);
  return *reinterpret_cast<const v4s*>(p);
}
// B fragment of an atom plane: stored atom row l16, or column l16 of the mirror
// atom's rows 4 g .. 4 g + 3 by one transposing load (lane r' of the 16-lane
// group addresses row 4 g + r' / 4, quarter r' % 4)
__device__ __forceinline__ v4s ty16_bfrag(const uint4* lds, int vb, int l16, int g, bool mirror) {
  if (!mirror) return ty16_frag(lds, vb, l16, g);
  const int r = 4 * g + (l16 >> 2), qt = l16 & 3;
  const char* p = reinterpret_cast<const char*>(lds + vb + ty16_vec(r, qt >> 1)) + 8 * (qt & 1);
  return __builtin_bit_cast(v4s, __builtin_amdgcn_ds_read_tr16_b64_v4bf16((__attribute__((address_space(3))) v4bf*)p));
}

template <bool SPLIT>
__global__ void __launch_bounds__(TY_THREADS) taylor16_kernel(TaylorBf16Args a) {
  constexpr int NP = SPLIT ? 4 : 2;
  using T = Ty16<NP>;
  extern __shared__ uint4 lds_u4[];
  const int N = a.nbasis, Kp = a.kpad;
  const int nrb = a.rpad / TY_R, ncb = (int)ceil_div(N, TY_C);
  const int lin = xcd_remap(blockIdx.x, gridDim.x);
  const int cb = lin % ncb, rb = (lin / ncb) % nrb, w = lin / ncb / nrb;
  if (w >= a.nwalkers) return;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, l16 = lane & 15, g = lane >> 4;
  const int64_t tplane = (int64_t)a.wcap * a.rpad * Kp;
  const int nst = Kp / 16, NA = a.natom_side;
  auto phys = [](int pi) { return SPLIT ? pi : 2 * pi; };
  // A: vector v of the stage's A region is plane v / 160, row (v % 160) / 2,
  // half (v % 2) ^ swizzle; B: vector v of the B region is plane v / 160 of atom
  // slot (v % 160) / 32, atom row (v % 32) / 2, half (v % 2) ^ swizzle
  const uint16_t* asrc[T::A_INS];
#pragma unroll
  for (int j = 0; j < T::A_INS; ++j) {
    const int v = wave * T::A_INS * 64 + j * 64 + lane;
    const int pi = v / (TY_R * 2), rem = v % (TY_R * 2), row = rem >> 1;
    const int h = (rem & 1) ^ ((row >> 3) & 1);
    asrc[j] = a.tin + phys(pi) * tplane + ((int64_t)w * a.rpad + rb * TY_R + row) * Kp + h * 8;
  }
  int baj[T::B_INS], boff[T::B_INS];
#pragma unroll
  for (int j = 0; j < T::B_INS; ++j) {
    const int v = wave * T::B_INS * 64 + j * 64 + lane;
    const int pi = v / 160, rem = v % 160, slot = rem >> 5, e = rem & 31, ar = e >> 1;
    baj[j] = 5 * cb + slot;
    boff[j] = phys(pi) * 256 + ar * 16 + 8 * ((e & 1) ^ ((ar >> 3) & 1));
  }
  const uint16_t* wat = a.vatoms + (int64_t)w * a.atom_stride;
  auto issue = [&](int t) __attribute__((always_inline)) {
    uint4* base = lds_u4 + (t & 3) * T::STAGE_VEC;
#pragma unroll
    for (int j = 0; j < T::A_INS; ++j)
      __builtin_amdgcn_global_load_lds((const void*)(asrc[j] + t * 16),
                                       (__attribute__((address_space(3))) void*)(base + wave * T::A_INS * 64 + j * 64),
                                       16, 0, 0);
#pragma unroll
    for (int j = 0; j < T::B_INS; ++j) {
      // atom (aj, t) as stored, or its mirror; outside the matrix the zero atom
      const int aj = baj[j], lo = min(aj, t), hi = max(aj, t);
      const uint16_t* src = (hi >= NA) ? a.zatom + boff[j]
                                       : wat + (int64_t)(lo * NA - lo * (lo - 1) / 2 + hi - lo) * 1024 + boff[j];
      __builtin_amdgcn_global_load_lds(
          (const void*)src, (__attribute__((address_space(3))) void*)(base + T::A_VEC + wave * T::B_INS * 64 + j * 64),
          16, 0, 0);
    }
  };
  v4f re[5], im[5];
#pragma unroll
  for (int i = 0; i < 5; ++i) {
    re[i] = zero4f();
    im[i] = zero4f();
  }
  const int pre = min(3, nst);
  for (int t = 0; t < pre; ++t) issue(t);
  for (int t = 0; t < nst; ++t) {
    // stage t has landed when at most the loads of the stages issued after it remain
    const int later = min(2, nst - 1 - t);
    if (later >= 2) wait_vm<2 * T::INS>();
    else if (later == 1) wait_vm<T::INS>();
    else wait_vm<0>();
    dl_barrier();  // stage t is visible, and stage t - 1 (the buffer reused next) is read by every wave
    if (t + 3 < nst) issue(t + 3);
    const uint4* stage = lds_u4 + (t & 3) * T::STAGE_VEC;
    const uint4* Bq = stage + T::A_VEC;
    const int aj = 5 * cb + wave;
    const bool mirror = aj > t && aj < NA;
    const int bslot = wave * 32;
    // operands swapped in every MFMA: the accumulators hold the transposed tiles
    if (SPLIT) {
      const v4s brh = ty16_bfrag(Bq, 0 * 160 + bslot, l16, g, mirror);
      const v4s brl = ty16_bfrag(Bq, 1 * 160 + bslot, l16, g, mirror);
      const v4s bih = ty16_bfrag(Bq, 2 * 160 + bslot, l16, g, mirror);
      const v4s bil = ty16_bfrag(Bq, 3 * 160 + bslot, l16, g, mirror);
#pragma unroll
      for (int i = 0; i < 5; ++i) {
        const int arow = i * 16 + l16;
        const v4s arh = ty16_frag(stage, 0 * 160, arow, g), arl = ty16_frag(stage, 1 * 160, arow, g);
        const v4s aih = ty16_frag(stage, 2 * 160, arow, g), ail = ty16_frag(stage, 3 * 160, arow, g);
        const v4s naih = neg16(aih), nail = neg16(ail);
        re[i] = mfma16_bf16(brh, arh, re[i]);
        re[i] = mfma16_bf16(brl, arh, re[i]);
        re[i] = mfma16_bf16(brh, arl, re[i]);
        re[i] = mfma16_bf16(bih, naih, re[i]);
        re[i] = mfma16_bf16(bil, naih, re[i]);
        re[i] = mfma16_bf16(bih, nail, re[i]);
        im[i] = mfma16_bf16(bih, arh, im[i]);
        im[i] = mfma16_bf16(bil, arh, im[i]);
        im[i] = mfma16_bf16(bih, arl, im[i]);
        im[i] = mfma16_bf16(brh, aih, im[i]);
        im[i] = mfma16_bf16(brl, aih, im[i]);
        im[i] = mfma16_bf16(brh, ail, im[i]);
      }
    } else {
      const v4s brh = ty16_bfrag(Bq, 0 * 160 + bslot, l16, g, mirror);
      const v4s bih = ty16_bfrag(Bq, 1 * 160 + bslot, l16, g, mirror);
#pragma unroll
      for (int i = 0; i < 5; ++i) {
        const int arow = i * 16 + l16;
        const v4s arh = ty16_frag(stage, 0 * 160, arow, g), aih = ty16_frag(stage, 1 * 160, arow, g);
        re[i] = mfma16_bf16(brh, arh, re[i]);
        re[i] = mfma16_bf16(bih, neg16(aih), re[i]);
        im[i] = mfma16_bf16(bih, arh, im[i]);
        im[i] = mfma16_bf16(brh, aih, im[i]);
      }
    }
  }
  ty_epilogue(a, re, im, w, rb, cb, wave, l16, g);
}

// phi [W][N][ncol] complex -> transposed split planes [4][W][rpad][kpad] at
// (w, col, k); 32 x 32 tiles through LDS so both sides are coalesced.
__global__ void __launch_bounds__(256) split_transpose_kernel(const double* __restrict__ phi, int nwalkers, int wcap,
                                                             int N, int ncol, int rpad, int Kp,
                                                             uint16_t* __restrict__ planes) {
  __shared__ double2 tile[32][33];
  const int ntk = (int)ceil_div(N, 32), ntc = (int)ceil_div(ncol, 32);
  int b = blockIdx.x;
  const int tc = b % ntc;
  b /= ntc;
  const int tk = b % ntk, w = b / ntk;
  if (w >= nwalkers) return;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int yy = ty; yy < 32; yy += 8) {  // rows k, columns col (contiguous)
    const int k = tk * 32 + yy, col = tc * 32 + tx;
    double2 v = make_double2(0.0, 0.0);
    if (k < N && col < ncol) {
      const cplx c = ld(phi, ((int64_t)w * N + k) * ncol + col);
      v = make_double2(c.re, c.im);
    }
    tile[yy][tx] = v;
  }
  __syncthreads();
  const int64_t plane = (int64_t)wcap * rpad * Kp;
  for (int yy = ty; yy < 32; yy += 8) {  // rows col, columns k (contiguous)
    const int col = tc * 32 + yy, k = tk * 32 + tx;
    if (col >= ncol || k >= N) continue;
    const double2 v = tile[tx][yy];
    uint16_t rh, rl, ih, il;
    split_bf16(v.x, rh, rl);
    split_bf16(v.y, ih, il);
    const int64_t o = ((int64_t)w * rpad + col) * Kp + k;
    planes[o] = rh;
    planes[plane + o] = rl;
    planes[2 * plane + o] = ih;
    planes[3 * plane + o] = il;
  }
}

}  // namespace

// Geometry for the split-BF16 exchange: as many walkers (wb) and Cholesky
// vectors (xb) as fit the 128 x 160 workgroup tile.
bool bf16_exchange_plan(int nocc, int* wb, int* xb) {
  if (nocc < 1 || nocc > BX_R || nocc > BX_C) return false;
  *wb = BX_R / nocc;
  *xb = BX_C / nocc;
  return bf16_lds_bytes<true>(*wb, *xb, nocc) <= 160 * 1024;
}

hipError_t launch_split_ghalf(const double* g, int64_t rows, int N, int Kp, uint16_t* planes, hipStream_t stream,
                              const int* walker_of, int nocc, const double* gref) {
  const int64_t total = rows * Kp;
  hipLaunchKernelGGL(split_ghalf_kernel, dim3((unsigned)ceil_div(total, 256)), dim3(256), 0, stream, g, rows, N, Kp,
                     planes, walker_of, nocc, gref);
  return hipGetLastError();
}

hipError_t launch_split_real(const double* x, int64_t rows, int N, int Kp, uint16_t* planes, hipStream_t stream) {
  const int64_t total = rows * Kp;
  hipLaunchKernelGGL(split_real_kernel, dim3((unsigned)ceil_div(total, 256)), dim3(256), 0, stream, x, rows, N, Kp,
                     planes);
  return hipGetLastError();
}

template <bool SPLIT>
hipError_t launch_exchange_bf16_t(ExchangeBf16Args a, hipStream_t stream) {
  a.lds_bytes = (int)bf16_lds_bytes<SPLIT>(a.wb, a.xb, a.nocc);
  hipError_t e = allow_lds((const void*)exchange_bf16_kernel<SPLIT>, a.lds_bytes);
  if (e != hipSuccess) return e;
  const int nwb = (int)ceil_div(a.nwalkers, a.wb);
  hipLaunchKernelGGL(exchange_bf16_kernel<SPLIT>, dim3(nwb * a.xsplit), dim3(BX_THREADS), a.lds_bytes, stream, a);
  return hipGetLastError();
}

hipError_t launch_exchange_bf16(const ExchangeBf16Args& a, hipStream_t stream) {
  if (a.nocc < 1 || a.wb < 1 || a.xb < 1 || a.wb * a.nocc > BX_R || a.xb * a.nocc > BX_C || a.kpad % BX_KC != 0 ||
      a.xsplit < 1)
    return hipErrorInvalidValue;
  return a.split ? launch_exchange_bf16_t<true>(a, stream) : launch_exchange_bf16_t<false>(a, stream);
}

// Direct global -> LDS staging of the GEMM operands (1 = default; 0 restores
// the register staging, 1-3% slower).
#ifndef AFQMC_BF16_DL
#define AFQMC_BF16_DL 1
#endif
constexpr bool kDl = AFQMC_BF16_DL != 0;

template <bool SPLIT, bool TR, class TL, bool F8 = false>
hipError_t launch_bf16_gemm_t(const Bf16GemmArgs& a, hipStream_t stream) {
  const int64_t nb = ceil_div(a.nn, TL::C), mb = ceil_div(a.m, TL::R);
  const int64_t mbg = (a.units && a.rgroup > 0 && a.rgroup < mb) ? ceil_div(mb, a.rgroup) * a.rgroup : mb;
  const int64_t grid = mbg * (a.units ? (int64_t)a.nunits : nb * a.ksplit);
  constexpr size_t lds = Bx<SPLIT, TL>::STAGES_ALL;
  hipError_t e = allow_lds((const void*)bf16_gemm_kernel<SPLIT, TR, kDl, TL, F8>, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((bf16_gemm_kernel<SPLIT, TR, kDl, TL, F8>), dim3((unsigned)grid), dim3(TL::THREADS), lds, stream,
                     a);
  return hipGetLastError();
}

template <bool TR>
hipError_t launch_bf16_gemm_shape(const Bf16GemmArgs& a, hipStream_t stream) {
  switch (a.tile) {
    case 1: return launch_bf16_gemm_t<true, TR, Tile1>(a, stream);
    case 2: return launch_bf16_gemm_t<true, TR, Tile2>(a, stream);
    case 3: return launch_bf16_gemm_t<true, TR, Tile3>(a, stream);
    case 4: return launch_bf16_gemm_t<true, TR, Tile4>(a, stream);
    case 5: return launch_bf16_gemm_t<true, TR, Tile5>(a, stream);
    default: return launch_bf16_gemm_t<true, TR, Tile0>(a, stream);
  }
}

hipError_t launch_bf16_gemm(const Bf16GemmArgs& a, hipStream_t stream) {
  if (a.m < 1 || a.nn < 1 || a.kchunks < 1 || a.ksplit < 1 || a.lda % BX_KC != 0 || a.ldb % BX_KC != 0 ||
      (int64_t)a.kchunks * BX_KC > a.lda || (int64_t)a.kchunks * BX_KC > a.ldb)
    return hipErrorInvalidValue;
  if (a.tile < 0 || a.tile >= BF16_TILES) return hipErrorInvalidValue;
  // the quadratic form's unit tables and dot slots, and the single product, use shape 0
  if (a.tile != 0 && (a.epilogue == BF16_EPI_QDOT || a.units || a.single)) return hipErrorInvalidValue;
  if (a.epilogue == BF16_EPI_VHS && (int64_t)a.nbasis * (a.nbasis + 1) / 2 != a.nn) return hipErrorInvalidValue;
  if (a.epilogue == BF16_EPI_VATOM && (a.out16 == nullptr || a.nn % 256 != 0 || a.atom_stride < a.nn / 256 * 1024))
    return hipErrorInvalidValue;
  if (a.epilogue == BF16_EPI_QDOT && (a.dg == nullptr || a.dkp < a.dn || a.dn < 1 || a.dstride < 1))
    return hipErrorInvalidValue;
  if (a.units && a.nunits < 1) return hipErrorInvalidValue;
  if ((a.a_gstride > 0 || a.epilogue == BF16_EPI_ADD) && a.group < 1) return hipErrorInvalidValue;
  if (a.epilogue == BF16_EPI_ADD && (a.base == nullptr || a.out == nullptr || a.single)) return hipErrorInvalidValue;
  if (a.fp8) {
    // FP8 operands: the quadratic form's single product
    if (!a.single || a.epilogue != BF16_EPI_QDOT || !kDl) return hipErrorInvalidValue;
    return launch_bf16_gemm_t<false, false, Tile0, true>(a, stream);
  }
  if (a.single) {
    if (a.epilogue == BF16_EPI_VHS || a.epilogue == BF16_EPI_VATOM) return hipErrorInvalidValue;
    return launch_bf16_gemm_t<false, false, Tile0>(a, stream);
  }
  // ADD and VATOM store rows of consecutive columns: transposed accumulators
  if (a.epilogue == BF16_EPI_ADD || a.epilogue == BF16_EPI_VATOM) return launch_bf16_gemm_shape<true>(a, stream);
  return launch_bf16_gemm_shape<false>(a, stream);
}

bool bf16_tile_dims(int tile, int* rows, int* cols) {
  switch (tile) {
    case 0: *rows = Tile0::R; *cols = Tile0::C; return true;
    case 1: *rows = Tile1::R; *cols = Tile1::C; return true;
    case 2: *rows = Tile2::R; *cols = Tile2::C; return true;
    case 3: *rows = Tile3::R; *cols = Tile3::C; return true;
    case 4: *rows = Tile4::R; *cols = Tile4::C; return true;
    case 5: *rows = Tile5::R; *cols = Tile5::C; return true;
    default: return false;
  }
}

hipError_t launch_taylor_bf16(const TaylorBf16Args& a, hipStream_t stream) {
  if (a.nwalkers < 1 || a.wcap < a.nwalkers || a.ncol < 1 || a.rpad % TY_R != 0 || a.rpad < a.ncol ||
      a.kpad % BX_KC != 0 || a.kpad < a.nbasis || a.init == nullptr || (a.last && a.q == nullptr) ||
      a.vatoms == nullptr || a.zatom == nullptr || a.natom_side != vhs_atom_side(a.nbasis) ||
      a.atom_stride < vhs_atom_count(a.nbasis) * 1024 ||
      a.nterms < 0 || a.nterms >= TAYLOR_MAX_TERMS || (a.last && a.nterms != 0 && a.terms[a.nterms - 1] == nullptr) ||
      a.variant < 0 || a.variant > 1)
    return hipErrorInvalidValue;
  const int64_t grid = (int64_t)a.nwalkers * (a.rpad / TY_R) * ceil_div(a.nbasis, TY_C);
  if (a.variant == 1) {
    const void* kern = a.split ? (const void*)taylor16_kernel<true> : (const void*)taylor16_kernel<false>;
    const size_t lds = a.split ? Ty16<4>::LDS : Ty16<2>::LDS;
    hipError_t e = allow_lds(kern, lds);
    if (e != hipSuccess) return e;
    if (a.split)
      hipLaunchKernelGGL(taylor16_kernel<true>, dim3((unsigned)grid), dim3(TY_THREADS), lds, stream, a);
    else
      hipLaunchKernelGGL(taylor16_kernel<false>, dim3((unsigned)grid), dim3(TY_THREADS), lds, stream, a);
    return hipGetLastError();
  }
  const void* kern = a.split ? (const void*)taylor_bf16_kernel<true> : (const void*)taylor_bf16_kernel<false>;
  const size_t lds = a.split ? Ty<true>::LDS : Ty<false>::LDS;
  hipError_t e = allow_lds(kern, lds);
  if (e != hipSuccess) return e;
  if (a.split)
    hipLaunchKernelGGL(taylor_bf16_kernel<true>, dim3((unsigned)grid), dim3(TY_THREADS), lds, stream, a);
  else
    hipLaunchKernelGGL(taylor_bf16_kernel<false>, dim3((unsigned)grid), dim3(TY_THREADS), lds, stream, a);
  return hipGetLastError();
}

hipError_t launch_split_transpose(const double* phi, int nwalkers, int wcap, int N, int ncol, int rpad, int Kp,
                                  uint16_t* planes, hipStream_t stream) {
  const int64_t grid = (int64_t)nwalkers * ceil_div(N, 32) * ceil_div(ncol, 32);
  hipLaunchKernelGGL(split_transpose_kernel, dim3((unsigned)grid), dim3(256), 0, stream, phi, nwalkers, wcap, N, ncol,
                     rpad, Kp, planes);
  return hipGetLastError();
}

hipError_t launch_split_chol_atoms(const double* chol_packed, int N, int X, int Xp, uint16_t* planes,
                                   hipStream_t stream) {
  if (N < 1 || X < 1 || Xp < X) return hipErrorInvalidValue;
  const int64_t total = vhs_atom_count(N) * 256 * Xp;
  hipLaunchKernelGGL(split_chol_atoms_kernel, dim3((unsigned)ceil_div(total, 256)), dim3(256), 0, stream, chol_packed,
                     N, X, Xp, vhs_atom_side(N), planes);
  return hipGetLastError();
}

hipError_t launch_split_eye(const double* m, int N, int Kp, uint16_t* planes, hipStream_t stream) {
  const int64_t total = (int64_t)N * Kp;
  hipLaunchKernelGGL(split_eye_kernel, dim3((unsigned)ceil_div(total, 256)), dim3(256), 0, stream, m, N, Kp, planes);
  return hipGetLastError();
}

int taylor_rows_padded(int ncol) { return (int)ceil_div(ncol, TY_R) * TY_R; }

}  // namespace afqmc
