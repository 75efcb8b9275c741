// Internal (non-exported) declarations shared by the kernel translation units.
#pragma once

#include "afqmc_common.hpp"

namespace afqmc {

enum : int {
  AFQMC_FLAG_SINGULAR = 1,    // an overlap matrix had an exactly zero pivot
  AFQMC_FLAG_NONFINITE = 2,   // a non-finite overlap / weight was produced
  AFQMC_FLAG_CHOLESKY = 4,    // a Gram matrix was not positive definite in reortho
};

// Values that change from step to step, kept in device memory so that a
// captured step graph reads them when it is replayed (the arguments of a
// captured kernel are fixed at capture).  Kernels read them only when handed
// a non-null pointer; eager calls pass their values as arguments instead.
struct StepState {
  unsigned long long counter;   // field-stream counter of the next step
  double eshift;
  unsigned long long sur_draws; // control-variate subset draws so far
  long long adam_t;             // Adam steps taken by the control variate's MLP
  long long perm_a, perm_b, inv_a;  // current subset: sample position of walker w is (a w + b) mod W
};
// state->counter = counter, state->eshift = eshift
hipError_t launch_step_set(StepState* state, unsigned long long counter, double eshift, hipStream_t stream);
// state->counter += dcounter, state->adam_t += dadam
hipError_t launch_step_advance(StepState* state, long long dcounter, long long dadam, hipStream_t stream);

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): 4 x 32 random bits per
// (counter, key).
__device__ __forceinline__ uint4 philox4x32(uint4 ctr, uint2 key) {
  constexpr uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t hi0 = __umulhi(M0, ctr.x), lo0 = M0 * ctr.x;
    const uint32_t hi1 = __umulhi(M1, ctr.z), lo1 = M1 * ctr.z;
    ctr = make_uint4(hi1 ^ ctr.y ^ key.x, lo1, hi0 ^ ctr.w ^ key.y, lo0);
    key.x += W0;
    key.y += W1;
  }
  return ctr;
}

// ---------------------------------------------------------------- GF / overlap
struct GfArgs {
  int nwalkers, nbasis, nalpha, nbeta, rhf;
  const double* phia;   // [W, N, na] complex
  const double* phib;   // [W, N, nb] complex
  const double* psia;   // [N, na] complex
  const double* psib;   // [N, nb] complex
  double* ghalfa;       // [W, na, N] complex (GF only)
  double* ghalfb;       // [W, nb, N] complex (GF only)
  double* ovlp;         // [W] complex
  const double* log_shift;  // [W] or null
  int* flags;
  int* wflags;   // [W] per-walker flags (OR of AFQMC_FLAG_*) or null
  int psi_real;  // the trial orbitals have zero imaginary parts (O takes two real MFMAs per K step, not four)
  // optional: Ghalf_s also as the 4 split-BF16 planes [W n_s, kpad] of spin s
  // (re_hi, re_lo, im_hi, im_lo; K padding zero), as split_ghalf writes them
  uint16_t* gplanes[2];
  int kpad;
};
size_t gf_lds_bytes(int nalpha, int nbeta);
// *flag = 1 if any of the n complex values z has a nonzero imaginary part (flag
// must be zero before; every writer stores 1)
hipError_t launch_imag_any(const double* z, int64_t n, int* flag, hipStream_t stream);
// bytes of each GF work buffer (O, O^{-1}): [W][2][nmax][nmax] complex
size_t gf_work_bytes(int nwalkers, int nalpha, int nbeta);
hipError_t launch_gf(const GfArgs& a, bool want_ghalf, double* obuf, double* oinv, hipStream_t stream);

// ---------------------------------------------------------------- GEMMs
// Rows of a complex operand A (K contiguous) times rows of a real operand B
// (K contiguous): C[r, c] = sum_k A[r, k] B[c, k], summed over up to two
// (A, B, K) segments.
struct NtSegment {
  const double* A;   // complex, row r at A + 2 * r * lda
  int64_t lda;       // in complex elements
  const double* B;   // real, row c at B + c * ldb
  int64_t ldb;       // in doubles
  int K;
};

struct ForceBiasArgs {
  int nwalkers, nchol, nseg, ksplit;
  NtSegment seg[2];
  double scale;        // 1, or 2 for RHF walkers
  double* vbias_part;  // [ksplit, W, X] complex
};
hipError_t launch_force_bias(const ForceBiasArgs& a, hipStream_t stream);

struct VhsArgs {
  int nwalkers, nbasis, nchol;
  const double* xshifted;     // [W, X] complex
  const double* chol_packed;  // [P, X] real
  const int* tri_i;           // [P] row of packed index p
  const int* tri_j;           // [P] column of packed index p
  double sqrt_dt;
  double* vhs;                // [W, N, N] complex (full, symmetric)
};
hipError_t launch_vhs(const VhsArgs& a, hipStream_t stream);

// Batched C[w] = A[w] * B[w] with A [N, N] (real shared or complex per walker)
// and B [N, ncol] complex (columns contiguous).  Epilogue:
//   out = alpha * (A B);  if accum: accum_buf += out
struct BatchedNNArgs {
  int nwalkers, nrows, ncol;
  const double* A;       // real [N, N] (shared) or complex [W, N, N]
  bool a_complex;
  int64_t a_stride;      // walker stride of A in elements of its type (0 = shared)
  const double* B;       // complex, walker w at B + 2 * w * b_stride, row k at + 2 * k * ldb
  int64_t b_stride, ldb;
  double* C;             // complex output, same addressing as B (c_stride, ldc)
  int64_t c_stride, ldc;
  double alpha;
  double* accum;         // optional complex accumulator, same addressing as C
  const double* accum_init;  // optional: accum = accum_init + out (else accum += out)
};
hipError_t launch_batched_nn(const BatchedNNArgs& a, hipStream_t stream);

// ---------------------------------------------------------------- exchange
struct ExchangeArgs {
  int nwalkers, nchol, nbasis, nocc;
  const double* ghalf;  // [W, n, N] complex
  const double* rchol;  // [X, n, N] real
  int xsplit;
  double* exx_part;     // [xsplit, W] complex, sum_x tr(T_x T_x) (no 1/2)
};
hipError_t launch_exchange(const ExchangeArgs& a, hipStream_t stream);
int exchange_xsplit(int nwalkers, int nchol, int nocc, int want = 0);

struct ExchangeBf16Args {
  int nwalkers, nchol, nocc, kpad;  // kpad = N rounded up to 32
  int wb, xb, xsplit;  // walkers and Cholesky vectors per workgroup tile, Cholesky ranges
  int split;           // 1: hi*hi + hi*lo + lo*hi products; 0: hi*hi only
  int lds_bytes;       // set by the launcher
  const uint16_t* gplanes;  // 4 x [W*n, kpad] bf16: re_hi, re_lo, im_hi, im_lo
  const uint16_t* lplanes;  // 2 x [X*n, kpad] bf16: hi, lo
  double* exx_part;         // [xsplit, W] complex
};
bool bf16_exchange_plan(int nocc, int* wb, int* xb);
hipError_t launch_split_ghalf(const double* g, int64_t rows, int N, int Kp, uint16_t* planes, hipStream_t stream,
                              const int* walker_of = nullptr, int nocc = 1, const double* gref = nullptr);
hipError_t launch_split_real(const double* x, int64_t rows, int N, int Kp, uint16_t* planes, hipStream_t stream);
hipError_t launch_exchange_bf16(const ExchangeBf16Args& a, hipStream_t stream);

// C = A B^T on split-BF16 planes: A 4 planes [m][lda] (re hi, re lo, im hi,
// im lo), B 2 planes [nn][ldb] (hi, lo), K = kchunks * 32 (zero padded).
enum { BF16_EPI_PARTIAL = 0, BF16_EPI_VHS = 1, BF16_EPI_QDOT = 2, BF16_EPI_ADD = 3, BF16_EPI_VATOM = 4 };
constexpr int BF16_FB_SPLIT = 8;       // K ranges per spin of the split-BF16 force bias (default)
constexpr int QUAD_TARGET_WGS = 1536;  // workgroups the quadratic-form column units aim for (default)
struct Bf16GemmArgs {
  const uint16_t* a;
  int64_t a_plane, lda, m;
  const uint16_t* b;
  int64_t b_plane, ldb, nn;
  int kchunks, ksplit;
  int single;        // 1: hi * hi only (A planes 0 and 2, B plane 0), partial epilogue
  int epilogue;
  double* out;       // PARTIAL: [z_offset + ksplit][m][nn] complex; VHS: [m][N][N] complex (or out32)
  int z_offset;
  double scale;      // PARTIAL: out = scale C; VHS: out = i scale C
  const int* tri_i;  // VHS: packed index -> (i, j)
  const int* tri_j;
  int nbasis;
  // column units {tile, first K chunk, end K chunk, 0} in grid order (grid = row
  // blocks x nunits, ksplit unused), or null: every tile over K splits of ksplit
  const int4* units;
  int nunits;
  // QDOT: d[row][i][m] = dg[src * dstride + i dn + m] - dref[i dn + m] (dref may be null),
  // src = walker_of[row] or row; output column col = i dkp + m (m < dn)
  const double* dg;
  const double* dref;
  const int* walker_of;
  int64_t dstride;
  int dn, dkp;
  // Rows as (walker, column): r = w group + c.  With a_gstride > 0 the A row
  // read for r is w a_gstride + a_goff + c (otherwise r).
  int group, a_gstride, a_goff;
  // ADD: out[w o_ws + col o_ld + o_off + c] = base[w b_ws + col b_ld + b_off + c] + scale C[r][col]
  // (complex), and when tp is set the split planes of that sum,
  // tp[p tp_plane + (w tp_rows + tp_off + c) tp_ld + col], p = re_hi, re_lo, im_hi, im_lo
  const double* base;
  int64_t b_ws, o_ws;
  int b_ld, b_off, o_ld, o_off;
  uint16_t* tp;
  int64_t tp_plane;
  int tp_rows, tp_off, tp_ld;
  float* out32;      // VHS: when set, the output as complex FP32 [m][N][N] (the accumulators' precision)
  // VATOM: columns in VHS atom order (vhs_atom_count); out16[row atom_stride + atom 1024
  // + plane 256 + r 16 + c] = the split planes (re_hi, re_lo, im_hi, im_lo) of
  // i scale C[row][col], col = atom 256 + r 16 + c, rounded to FP32 first
  uint16_t* out16;
  int64_t atom_stride;
  // workgroup tile shape 0 .. BF16_TILES - 1 (bf16_tile_dims); the quadratic form
  // (QDOT, unit tables) and the single product take shape 0
  int tile;
  // with a unit table: consecutive workgroups cover rgroup row blocks x units
  // (row blocks fastest), group after group of row blocks; 0 or >= the row-block
  // count: all row blocks x units
  int rgroup;
  // single product in FP8 (E4M3): the operands are byte arrays addressed in
  // 16-bit units (a K chunk of BX_KC units holds 64 FP8 values), the 16-byte
  // fragments of two chunks feeding one 16x16x128 FP8 MFMA; QDOT only.  row_scale[row]
  // multiplies each row's dot (the inverse of the operands' scales).
  int fp8;
  const double* row_scale;
  // without a unit table: column blocks vary fastest over the grid, so the
  // workgroups of one row block run together on one XCD and share its A tile
  // in L2 (for a large A and a few column blocks, as in the one body);
  // otherwise row blocks vary fastest and share the B tile
  int cols_fast;
};
hipError_t launch_bf16_gemm(const Bf16GemmArgs& a, hipStream_t stream);
constexpr int BF16_TILES = 6;
// rows x columns of a workgroup tile shape (false for an unknown shape)
bool bf16_tile_dims(int tile, int* rows, int* cols);

// ---------------------------------------------------------------- exchange as g^T Wq g
constexpr double QUAD_BUDGET_BYTES = 24e9;  // largest Wq per spin held on the device
// Column tile of the block-triangular split-BF16 Wq (the BF16 GEMM's column tile).
constexpr int QUAD_TILE = 160;
// Partial dot products written per column unit by the QDOT epilogue.
constexpr int QUAD_SLOTS = 2;
// Wq[(i,m'),(j,m)] = sum_x L_x[i][m] L_x[j][m'] from rchol [X, n, N]: FP64 [nN][nN]
// (Np = N), or split-BF16 planes [2][n Np][n Np] (Np = Kp, zero padding) of the
// block-triangular W' with plane row b, column a: W'_ab = 2 Wq_ab when tile(a) <
// tile(b), Wq_ab when tile(a) = tile(b), 0 otherwise (tile(a) = a / QUAD_TILE),
// so that g^T Wq g = sum_b g_b sum_{a < QUAD_TILE (tile(b) + 1)} g_a W'_ab.
// Exactly one of w64 / wplanes is non-null.
hipError_t launch_quad_w(const double* rchol, int X, int n, int N, int Np, double* w64, uint16_t* wplanes,
                         hipStream_t stream);
// part[w] = sum_z sum_{i, m < N} d[i][m] y[z][w][i Np + m], d = g[src] - gref (gref
// [n, N] or null), src = walker_of[w] or w
hipError_t launch_quad_reduce(const double* y, int ksplit, int nwalkers, int n, int N, int Np, const double* g,
                              const int* walker_of, double* part, hipStream_t stream,
                              const double* gref = nullptr);
// out[w] = sum_{s < nslots} part[s][w] (complex, fixed order)
hipError_t launch_quad_sum(const double* part, int nslots, int nwalkers, double* out, hipStream_t stream);
// FP8 (E4M3) copy of the block-triangular W' for the control variate's single
// product: w8[r][k] (rows and columns padded to ld8, a multiple of 64, with
// zeros) = fp8(s_w (hi + lo)[r][k]) from the split-BF16 planes [2][ldn][ldn];
// s_w = 2^e with max |W'| s_w <= 224, written to scale[0] (FP32, on the device).
// absmax (one unsigned int of device scratch) must hold 0 on entry.
hipError_t launch_quad_w8(const uint16_t* wplanes, int64_t ldn, int64_t ld8, uint8_t* w8, float* scale,
                          unsigned int* absmax, hipStream_t stream);
// FP8 planes of d = g - gref per walker for the FP8 quadratic form: a8[p][w][k]
// (p = re, im; k = i Kp + m, zero for m >= N and k >= n Kp, row stride ld8)
// = fp8(s_row d), s_row = 2^e with max |d_w| s_row <= 224 (per walker), and
// row_inv[w] = 1 / (s_row s_w) with s_w = wscale[0].
hipError_t launch_split_fp8_rows(const double* g, const double* gref, int nwalkers, int n, int N, int Kp,
                                 int64_t ld8, uint8_t* a8, const float* wscale, double* row_inv, hipStream_t stream);

// One split-BF16 Taylor term per walker: out_k^T = alpha out_{k-1}^T VHS,
// q = (init or q) + out_k, and (tout) the split planes of out_k^T.
// most Taylor terms of the split-BF16 series (afqmc_create refuses a larger
// exp_nmax in mixed precision: by then |VHS|^k / k! is far below FP64 rounding)
constexpr int TAYLOR_MAX_TERMS = 24;
struct TaylorBf16Args {
  int nwalkers, wcap, nbasis, ncol, rpad, kpad;
  const uint16_t* tin;      // [4][wcap][rpad][kpad] planes of out_{k-1}^T
  uint16_t* tout;           // same layout for out_k^T, or null (last term)
  const uint16_t* vatoms;   // split planes of the upper VHS atoms of every walker ([W][NAT][4][16][16])
  const uint16_t* zatom;    // one all-zero atom (4 planes), read for atoms outside the matrix
  int natom_side;           // NA = ceil(N / 16) atoms per side
  int64_t atom_stride;      // bf16 values per walker in vatoms (NAT 1024)
  const double* init;       // [W][N][ncol] complex: the series' input (read by the last term)
  double* q;                // [W][N][ncol] complex: init + sum of the terms (written by the last term)
  // last term: the planes of the earlier terms out_1 .. out_nterms (layout of tin), summed into q
  const uint16_t* terms[TAYLOR_MAX_TERMS];
  int nterms;
  int first, last;          // this term is the first / the last of the series
  int split;                // 1: three BF16 products (hi hi + hi lo + lo hi), 0: hi hi only
  double alpha;
  uint16_t* qout;           // last term: planes of q^T (layout of tin), or null
  int variant;              // 0: 32-wide K stages, two in LDS (taylor_bf16_kernel); 1: 16-wide, four
                            // (taylor16_kernel)
};
hipError_t launch_taylor_bf16(const TaylorBf16Args& a, hipStream_t stream);
// Taylor terms k <= TAYLOR_SPLIT_TERMS use three BF16 products, later ones one.
// Each term but the last writes its four split planes to a buffer of its own
// (the next term's operand, hi planes alone for a one-product term) and the
// last term adds them all (hi + lo: 16 significant bits) to P.
// Term k is about |VHS|^k / k! of phi (|VHS|_2 ~ 0.35 at dt = 0.005 for the
// normalised N = 400 Hamiltonian: terms 1e-1, 2e-2, 1e-3, 1e-4, 6e-6, 4e-7), so
// one product's 2^-8 relative error on terms 4.. stays below the three-product
// error (~1e-5 relative) of term 1.
constexpr int TAYLOR_SPLIT_TERMS = 3;
hipError_t launch_split_transpose(const double* phi, int nwalkers, int wcap, int N, int ncol, int rpad, int Kp,
                                  uint16_t* planes, hipStream_t stream);
// VHS of the split-BF16 path in 16 x 16 atoms: atom (a, b), a <= b < NA = ceil(N / 16),
// holds VHS[16 a + r][16 b + c] (the diagonal atoms both triangles, zero outside
// N x N), listed by a, then b: index a NA - a (a - 1) / 2 + b - a.  VHS is symmetric,
// so the atoms below the diagonal are these, transposed.
inline int vhs_atom_side(int N) { return (N + 15) / 16; }
inline int64_t vhs_atom_count(int N) {
  const int64_t na = vhs_atom_side(N);
  return na * (na + 1) / 2;
}
// packed Cholesky [N(N+1)/2, X] -> 2 bf16 planes [vhs_atom_count(N) 256, Xp] in atom order
// (row atom 256 + r 16 + c holds L[(16 a + r, 16 b + c)], zero outside N x N)
hipError_t launch_split_chol_atoms(const double* chol_packed, int N, int X, int Xp, uint16_t* planes,
                                   hipStream_t stream);
// D = M - I (M real [N][N]) -> 2 bf16 planes [N][Kp] (zero padding)
hipError_t launch_split_eye(const double* m, int N, int Kp, uint16_t* planes, hipStream_t stream);
int taylor_rows_padded(int ncol);

// ---------------------------------------------------------------- elementwise
struct FieldsArgs {
  int nwalkers, nchol, ksplit;
  const double* vbias_part;   // [ksplit, W, X] complex (complex FP32 when part32)
  int part32;
  const double* mf_shift;     // [X] complex
  const double* xi;           // [W, X] real, or null -> Philox draw
  uint64_t seed, counter;
  const StepState* state;     // non-null: the counter comes from state->counter
  double sqrt_dt, fbbound;
  double* xi_out;             // optional: the fields used, [W, X]
  double* xshifted;           // [W, X] complex
  double* cmf;                // [W] complex
  double* cfb;                // [W] complex
  double* ecoul;              // [W] complex: 1/2 sum_x vbias^2
  double* vbias_out;          // optional [W, X] complex (summed vbias)
  unsigned long long* nfb;    // number of capped force-bias entries (accumulated)
  int want_fields;            // 0: only vbias sum + Coulomb
  double* xbar_mean;          // optional [W]: mean |xbar| over x (after the cap)
  double* xbar_max;           // optional [W]: max |xbar|
};
hipError_t launch_fields(const FieldsArgs& a, hipStream_t stream);

struct WeightArgs {
  int nwalkers;
  const double* ovlp_old;  // [W] complex
  const double* ovlp_new;  // [W] complex
  const double* cfb;
  const double* cmf;
  double dt, eshift, ebound;
  const StepState* state;  // non-null: eshift comes from state->eshift
  double* weight;          // [W] real
  double* hybrid_energy;   // [W] complex (in: previous, out: new)
  double* walker_ovlp;     // [W] complex (out: ovlp_new)
  int* flags;
  int* wflags;             // [W] per-walker flags or null
};
hipError_t launch_update_weight(const WeightArgs& a, hipStream_t stream);
// dst[0..n) <- src[0..n) (dst may be null), then src <- 0 when clear
hipError_t launch_take_flags(int* src, int* dst, int n, int clear, hipStream_t stream);

struct EnergyArgs {
  int nwalkers, nbasis, nalpha, nbeta, rhf;
  const double* ghalfa;
  const double* ghalfb;
  const double* rH1a;     // [na, N] complex
  const double* rH1b;
  const double* ecoul;    // [W] complex
  const double* exx_part_a;  // [xs_a, W] complex
  int xs_a;
  const double* exx_part_b;  // [xs_b, W] complex (null when rhf / nbeta == 0)
  int xs_b;
  double ecore;
  double* energy;         // [W, 3] complex
  const double* exx_est;  // optional [W] complex: exchange sum to use instead of the partials
  double* e1_out;         // optional [W] complex: one-body energy
  const double* e1_in;    // optional [W] complex: one-body energy already computed (Ghalf not read)
};
hipError_t launch_energy(const EnergyArgs& a, hipStream_t stream);

hipError_t launch_draw_normals(double* out, int64_t n, uint64_t seed, uint64_t counter, hipStream_t stream);

// ---------------------------------------------------------------- surrogate
constexpr int SUR_HIDDEN = 128;
constexpr int SUR_FEATURES = 9;
constexpr int SUR_NPARAM = SUR_HIDDEN * SUR_FEATURES + SUR_HIDDEN * SUR_HIDDEN + 4 * SUR_HIDDEN + 2;

struct SurrogateState {
  double sum[SUR_FEATURES + 2], sumsq[SUR_FEATURES + 2], count;
  float mean[SUR_FEATURES + 2], inv_std[SUR_FEATURES + 2];
  // gain of the MLP term, alpha = sum Re(conj(m) r) / sum |m|^2 over the exact
  // samples of earlier steps (decayed), clamped to [0, 1]
  double gain_rm, gain_mm, gain;
};

struct SurrogateFeatureArgs {
  int nwalkers;
  const double* e1;        // [W] complex
  const double* ecoul;     // [W] complex
  const double* xbar_mean; // [W]
  const double* xbar_max;  // [W]
  const double* ovlp;      // [W] complex
  float* feat;             // [W][F]
};
struct SurrogateStatsArgs {
  int count;
  const float* feat;
  const int* walker_of;    // sample -> walker (null: identity)
  const double* target;    // [count][2]
  SurrogateState* state;
};
struct SurrogateForwardArgs {
  int nwalkers;
  const float* feat;
  const uint16_t* wbf;     // bf16 weights W1 | W2 | W3
  const float* bias;       // b1 | b2 | b3
  const SurrogateState* state;
  double* pred;            // [W] complex
};
struct SurrogateTrainArgs {
  int count;
  const float* feat;
  const int* walker_of;
  const double* target;    // [count][2]
  const float* master;     // [NPARAM]
  const SurrogateState* state;
  float* grad;             // [groups][NPARAM]
  float* loss;             // [groups]
};
struct SurrogateAdamArgs {
  int ngroups, it;         // Adam step number: state->adam_t + it + 1
  const StepState* state;
  float lr;
  const float* grad;
  float* master;
  float* m;
  float* v;
  uint16_t* wbf;
  float* bias;
};
struct SurrogateCombineArgs {
  int nwalkers, k;
  const StepState* perm;   // sample position of walker w: (perm_a w + perm_b) mod W
  const double* pred;      // [W] complex
  const double* part_a;    // [xs_a, k] complex
  int xs_a;
  const double* part_b;    // [xs_b, k] complex or null
  int xs_b;
  double* est;             // [W] complex
  double* target;          // [k][2] exact - cv of the samples (the MLP's target) or null
  double* resid;           // [k] |exact - g|^2 or null
  const double* cv;        // [W] complex physical part of g, or null
};
struct SurrogateCvArgs {
  int nwalkers, nbasis, nspin;
  int nocc[2];
  const double* g[2];      // Ghalf per spin [W, n_s, N] complex
  const double* gT[2];     // trial Ghalf per spin [n_s, N] complex
  const double* K[2];      // [n_s, N] complex
  const double* et;        // [2] complex E_T per spin
  const double* part[2];   // [xs_s, W] complex partial sums of Q
  int xs[2];
  double* cv;              // [W] complex
  double* pred;            // [W] complex: MLP output in, cv + gain * MLP out
  double* mlp;             // [W] complex copy of the MLP output
  const SurrogateState* state;
  // fold = 0: cv only, pred and mlp untouched (launch_sur_fold folds the MLP
  // term in once it is known)
  int fold;
  // optional, in the same pass over Ghalf: e1[w] = e1_factor sum_s <rH1_s, Ghalf_s> + ecore
  // (the one-body energy, as energy_kernel forms it)
  const double* rH1[2];
  double* e1;
  double e1_factor, ecore;
};
// pred = cv + gain * pred, mlp = the incoming pred (the MLP output), per walker
hipError_t launch_sur_fold(int nwalkers, const double* cv, double* pred, double* mlp, const SurrogateState* state,
                           hipStream_t s);
struct SurrogateGainArgs {
  int k;
  const int* walker_of;    // [k]
  const double* target;    // [k][2] exact - cv
  const double* mlp;       // [W] complex
  SurrogateState* state;
  double decay;
  double gain_max;  // upper clamp of the gain (0: the MLP term is off)
};
hipError_t launch_cv_static(const double* L, const double* gT, int X, int n, int N, double* T0, double* K,
                            double* part, double* et, hipStream_t s);
hipError_t launch_sur_cv(const SurrogateCvArgs& a, hipStream_t s);
hipError_t launch_sur_gain(const SurrogateGainArgs& a, hipStream_t s);
hipError_t launch_sur_features(const SurrogateFeatureArgs& a, hipStream_t s);
hipError_t launch_sur_stats(const SurrogateStatsArgs& a, hipStream_t s);
hipError_t launch_sur_forward(const SurrogateForwardArgs& a, hipStream_t s);
int sur_grad_groups(int count);
hipError_t launch_sur_grad(const SurrogateTrainArgs& a, hipStream_t s);
hipError_t launch_sur_adam(const SurrogateAdamArgs& a, hipStream_t s);
hipError_t launch_sur_combine(const SurrogateCombineArgs& a, hipStream_t s);
hipError_t launch_sur_select(int nwalkers, int k, const StepState* perm, int* walker_of, hipStream_t s);
// draws the step's subset permutation into state (perm_a coprime with W, perm_b
// uniform in [0, W)) from Philox(seed, state->sur_draws), then sur_draws += 1;
// identity != 0 sets the identity (a = 1, b = 0) without a draw
hipError_t launch_sur_plan(StepState* state, uint64_t seed, int nwalkers, int identity, hipStream_t s);
hipError_t launch_sur_gather(const double* g, const int* walker_of, int k, int64_t per_walker, double* out,
                             hipStream_t s);

// ---------------------------------------------------------------- reortho
struct ReorthoArgs {
  int nwalkers, nbasis, nalpha, nbeta;
  double* phia;
  double* phib;
  double* ovlp;          // [W] complex, divided by detR
  double* detR;          // [W] real out
  double* log_detR;      // [W] real, += log detR
  const double* detR_shift;  // [W] or null
  int* flags;
  int* wflags;               // [W] per-walker flags or null
};
hipError_t launch_reortho(const ReorthoArgs& a, hipStream_t stream);

// ---------------------------------------------------------------- population control
constexpr int POP_MAX = 4096;  // walkers per device handled by one plan workgroup
enum { POP_PAIR_BRANCH = 0, POP_COMB = 1, POP_RECONFIGURATION = 2, POP_RESCALE = 3 };
struct PopPlanArgs {
  int nwalkers, method;
  const double* weight;     // [W] current weights
  double target;            // target total weight (the walker count)
  double min_weight, max_weight;
  const double* uniforms;   // pair_branch: >= W/2 uniforms; comb, SR: 1 (null for POP_RESCALE)
  int* src;                 // [W] out: slot i takes walker src[i]'s buffer
  double* new_weight;       // [W] out: weights after the control
  double* total;            // [1] out: sum |w| before rescaling
  int* npairs;              // [1] out: branched pairs (pair_branch)
};
hipError_t launch_pop_plan(const PopPlanArgs& a, hipStream_t stream);
// ---------------------------------------------------------------- set-up producers
enum { SETUP_STORE_STRIDED = 0, SETUP_STORE_HALF_ROTATED = 1 };
struct SetupGemmArgs {
  int64_t m, n, k;          // C [m, n] = alpha sum_k A[i, k] B[k, j] + beta C
  const double* a;
  int64_t sar, sac;         // A[i, k] = a[i sar + k sac]
  const double* b;
  int64_t sbr, sbc;         // B[k, j] = b[k sbr + j sbc]
  double* c;
  int64_t scr, scc;         // C[i, j] = c[i scr + j scc] (SETUP_STORE_STRIDED)
  double alpha, beta;
  int64_t kslice;           // set by the launcher
  double* part;             // split-K planes [slices, m, n] (launcher's nz > 1)
  int store;
  int64_t hr_x, hr_nN, hr_N;  // SETUP_STORE_HALF_ROTATED: column j = m X + x -> c[x nN + i N + m]
};
int64_t setup_gemm_part_doubles(int64_t m, int64_t n, int nz);
hipError_t launch_setup_gemm(SetupGemmArgs g, int nz, hipStream_t stream);
hipError_t launch_setup_pack(int N, int X, const double* chol, double* packed, hipStream_t stream);
hipError_t launch_setup_fill(double* out, int64_t n, double v, hipStream_t stream);
hipError_t launch_setup_scale_shift(int64_t n, const double* h, const double* v, double scale, double* out,
                                    hipStream_t stream);
hipError_t launch_setup_axpby(int64_t n, double alpha, const double* x, double beta, const double* y, double* out,
                              hipStream_t stream);
hipError_t launch_setup_add_identity(int N, double* a, hipStream_t stream);
hipError_t launch_setup_norm1(int N, const double* a, double* out, hipStream_t stream);

// acc[0] += sum w, acc[1] += sum uw (uw non-null), acc[2] += sum w v[i * vstride]
// (v non-null); acc is [3] complex, sums in a fixed order (one workgroup)
hipError_t launch_walker_sums(int nw, const double* weight, const double* unscaled, const double* values,
                              int64_t vstride, double* acc, hipStream_t stream);
// weight <- min(max(weight, -bound), bound), NaN kept
hipError_t launch_clip_weights(int nw, double* weight, double bound, hipStream_t stream);
hipError_t launch_pop_gather(int nw, const int* src, const void* in, void* out, int64_t row_bytes,
                             hipStream_t stream);

// Raise a kernel's dynamic-LDS limit when it needs more than the 64 KiB default
// (gfx950 has 160 KiB per CU).
inline hipError_t allow_lds(const void* kernel, size_t bytes) {
  if (bytes <= 65536) return hipSuccess;
  if (bytes > 160 * 1024) return hipErrorInvalidValue;
  return hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

// XCD-aware remap: consecutive logical ids land on the same XCD (round-robin
// dispatch puts hardware id b on XCD b % 8).
__device__ __forceinline__ int xcd_remap(int bid, int nblocks) {
  constexpr int NX = 8;
  const int per = (nblocks + NX - 1) / NX;
  int tall = nblocks % NX;
  if (tall == 0) tall = NX;
  const int xcd = bid % NX, local = bid / NX;
  if (xcd < tall) return xcd * per + local;
  return tall * per + (xcd - tall) * (per - 1) + local;
}

}  // namespace afqmc
