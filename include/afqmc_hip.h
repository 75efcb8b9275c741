/*
 * afqmc_hip.h -- C ABI of the gfx950 phaseless-AFQMC step (libafqmc_hip.so).
 *
 * Plain C: device pointers, sizes and scalars only.  Complex arrays are
 * interleaved (re, im) float64 pairs, the memory layout of complex128; every
 * array is dense row-major (C-contiguous) with the shape given beside it: no
 * entry point takes a leading dimension or stride, so a caller holding a
 * strided view copies it to a dense buffer first (the Python binding,
 * ipie_amd/_lib.py, refuses non-contiguous tensors).  A context borrows
 * the static tensors passed to afqmc_set_static for its whole lifetime and
 * borrows walker arrays only for the duration of a call.  All work is
 * enqueued on the context's stream; nothing blocks the host except
 * afqmc_sync / afqmc_get_flags.  A context is not re-entrant; distinct
 * contexts may be used concurrently.
 *
 * Reference interfaces each entry point replaces (ipie v0.7.1 source tree):
 *   afqmc_greens_function  SingleDet.calc_greens_function
 *                          ipie/trial_wavefunction/single_det.py:180-184 ->
 *                          ipie/estimators/greens_function_single_det.py:62-154
 *   afqmc_overlap          SingleDet.calc_overlap, ipie/propagation/overlap.py:29-59
 *   afqmc_force_bias       SingleDet.calc_force_bias, ipie/propagation/force_bias.py:79-121
 *   afqmc_one_body         PhaselessBase.propagate_walkers_one_body,
 *                          ipie/propagation/phaseless_base.py:274-280, operations.py:26-55
 *   afqmc_two_body         PhaselessBase.propagate_walkers_two_body,
 *                          phaseless_base.py:282-310 (+ apply_VHS, phaseless_generic.py:35-97)
 *   afqmc_update_weight    PhaselessBase.update_weight, phaseless_base.py:340-380
 *   afqmc_propagate        PhaselessBase.propagate_walkers, phaseless_base.py:312-338
 *   afqmc_local_energy     local_energy_batch, ipie/estimators/local_energy_batch.py:36-75
 *                          (local_energy_sd.py:106-155, 358-403, 987-1080)
 *   afqmc_reortho          UHFWalkers.reortho, ipie/walkers/uhf_walkers.py:108-161
 *   afqmc_draw_normals     the xi draw, phaseless_base.py:296
 */
#ifndef AFQMC_HIP_H
#define AFQMC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define AFQMC_OK 0
#define AFQMC_ERR_INVALID_ARG 1  /* null pointer, bad size, call before set_static */
#define AFQMC_ERR_SHAPE 2        /* sizes outside what the kernels support */
#define AFQMC_ERR_UNSUPPORTED 3  /* precision mode / feature not available */
#define AFQMC_ERR_HIP 4          /* a HIP runtime call or launch failed */
#define AFQMC_ERR_ALLOC 5        /* device allocation failed */

/* device-side flags (afqmc_get_flags), OR-ed since the last clear */
#define AFQMC_FLAG_SINGULAR 1    /* an overlap matrix was exactly singular */
#define AFQMC_FLAG_NONFINITE 2   /* a non-finite weight was produced */
#define AFQMC_FLAG_CHOLESKY 4    /* reortho met a numerically rank-deficient walker */

/* phases timed by afqmc_profile (HIP events on the context's stream) */
#define AFQMC_PHASE_GF 0          /* Green's function (PropagatorTimer.tgf) */
#define AFQMC_PHASE_ONE_BODY 1    /* exp(-dt/2 H1) products and walker copies (tgemm) */
#define AFQMC_PHASE_FORCE_BIAS 2  /* force-bias contraction (tfbias) */
#define AFQMC_PHASE_FIELDS 3      /* xbar, fields, cmf, cfb, Coulomb (tfbias) */
#define AFQMC_PHASE_VHS 4         /* VHS contraction + symmetric unpack (tvhs) */
#define AFQMC_PHASE_TAYLOR 5      /* Taylor series of exp(VHS) (tgemm) */
#define AFQMC_PHASE_OVERLAP 6     /* overlap of the propagated walkers (tovlp) */
#define AFQMC_PHASE_WEIGHT 7      /* hybrid energy and weights (tupdate) */
#define AFQMC_PHASE_EXCHANGE 8    /* fused exchange contraction (local energy) */
#define AFQMC_PHASE_ENERGY 9      /* one-body energy and energy assembly */
#define AFQMC_PHASE_REORTHO 10    /* QR re-orthogonalisation */
#define AFQMC_PHASE_EXCHANGE_CV 11 /* single-product exchange of the control variate (surrogate) */
#define AFQMC_PHASE_SPLIT 12       /* Ghalf / dG into split-BF16 planes for the quadratic-form exchange */
#define AFQMC_NPHASES 13

/* precision modes */
#define AFQMC_PREC_FP64 0   /* every contraction in FP64 */
#define AFQMC_PREC_MIXED 1  /* force bias, VHS, Taylor series and exchange as split-BF16
                               products (hi*hi + hi*lo + lo*hi) with FP32 accumulation (FP64
                               trace and Taylor sums); Green's function (O, LU, inverse,
                               Ghalf), overlaps, weights, one-body products and QR stay FP64 */

typedef struct afqmc_ctx afqmc_ctx;

typedef struct afqmc_dims {
  int32_t nbasis;    /* N */
  int32_t nchol;     /* X (= nfields for real Cholesky vectors) */
  int32_t nalpha;    /* n_a, 1..64 */
  int32_t nbeta;     /* n_b, 0..64 */
  int32_t nwalkers;  /* maximum W per call */
  int32_t exp_nmax;  /* Taylor order of exp(VHS), ipie default 6 (at most 24 in
                        AFQMC_PREC_MIXED, which keeps every term's planes) */
  int32_t rhf;       /* walkers.rhf: propagate alpha only, beta == alpha */
  int32_t precision; /* AFQMC_PREC_* */
} afqmc_dims;

typedef struct afqmc_static {
  const double* chol_packed; /* [N(N+1)/2, X] real, numpy.triu_indices order */
  const double* rchola;      /* [X, n_a N] real, half-rotated Cholesky */
  const double* rcholb;      /* [X, n_b N] real */
  const double* rH1a;        /* [n_a, N] complex */
  const double* rH1b;        /* [n_b, N] complex */
  const double* expH1;       /* [2, N, N] real, exp(-dt/2 H1_shifted) per spin */
  const double* mf_shift;    /* [X] complex */
  const double* psia;        /* [N, n_a] complex, trial orbitals */
  const double* psib;        /* [N, n_b] complex */
  double dt;
  double ecore;
} afqmc_static;

int afqmc_create(afqmc_ctx** out, int device, const afqmc_dims* dims);
void afqmc_destroy(afqmc_ctx* ctx);
/* hipStream_t passed as void*; NULL = the default stream */
int afqmc_set_stream(afqmc_ctx* ctx, void* stream);
/* Borrows the static tensors.  Reads the trial orbitals once on the context's
 * stream (and waits for it): when their imaginary parts are all zero, the
 * overlap matrices take two real products per step instead of four. */
int afqmc_set_static(afqmc_ctx* ctx, const afqmc_static* s);
const char* afqmc_last_error(const afqmc_ctx* ctx);
int afqmc_sync(afqmc_ctx* ctx);
/* synchronises, returns the OR of device flags since the last clear */
int afqmc_get_flags(afqmc_ctx* ctx, int32_t* flags, int clear);
/* Per-walker status (the optional per-walker flag array of the error
 * conventions; ipie raises numpy.linalg.LinAlgError on a singular overlap,
 * greens_function_single_det.py:79-97, and lets NaN weights propagate):
 * flags [nw] int32 device array receives, for walkers 0..nw-1, the OR of the
 * AFQMC_FLAG_* conditions each walker met since the last take (singular
 * overlap matrix in the Green's function / overlap, non-finite weight in the
 * weight update, rank-deficient walker in reortho); clear = 1 resets them.
 * flags may be NULL with clear = 1.  Enqueued on the context's stream with a
 * kernel (no memset, so it can be recorded into a step graph); no
 * synchronisation. */
int afqmc_walker_flags(afqmc_ctx* ctx, int32_t nwalkers, int32_t* flags, int clear);
/* total force-bias caps accumulated on the device (synchronises) */
int afqmc_get_nfb(afqmc_ctx* ctx, uint64_t* nfb);
/* How this context evaluates the exchange sum_x tr(T_x T_x) (no counterpart in
 * ipie; for work accounting): AFQMC_EXCHANGE_T forms every T_x = L_x Ghalf^T
 * (ipie's contraction); AFQMC_EXCHANGE_QUAD evaluates g^T Wq g with g = vec(Ghalf)
 * and Wq[(i,m'),(j,m)] = sum_x L_x[i][m] L_x[j][m'] built once (used when
 * N < X and Wq fits the device budget). */
#define AFQMC_EXCHANGE_T 0
#define AFQMC_EXCHANGE_QUAD 1
int afqmc_exchange_form(const afqmc_ctx* ctx);

/* Population control of the context's walkers on its stream, one rank
 * (ipie PopController.pop_control, ipie/walkers/pop_controller.py:78-126, with
 * pair_branch :304-402, comb :199-301, stochastic_reconfiguration :405-450 for
 * comm.size == 1).  afqmc_pop_plan rescales weight [nw] to target_weight
 * (total_weight <- sum |w|; nothing else happens when that is below 1e-8, the
 * reference's ValueError, which the caller raises) and writes the method's
 * source map src [nw] (slot i takes the buffer of walker src[i]) and the
 * weights afterwards new_weight [nw] (may alias weight); npairs [1] counts the
 * branched pairs.  uniforms: pair_branch >= nw / 2 of them, consumed one per
 * branched pair in order; comb and reconfiguration one; null for rescale only.
 * nw <= 4096.  All arrays are device memory; nothing synchronises. */
#define AFQMC_POP_PAIR_BRANCH 0
#define AFQMC_POP_COMB 1
#define AFQMC_POP_RECONFIGURATION 2
#define AFQMC_POP_RESCALE 3
int afqmc_pop_plan(afqmc_ctx* ctx, int32_t method, int32_t nw, const double* weight, double target_weight,
                   double min_weight, double max_weight, const double* uniforms, int32_t* src, double* new_weight,
                   double* total_weight, int32_t* npairs);
/* out row i <- in row src[i] for nw rows of bytes_per_walker bytes (a walker
 * buffer field; ipie get_buffer / set_buffer, pop_controller.py:128-196).
 * in == out is allowed when no source row is also a destination (pair_branch,
 * comb); reconfiguration maps need a separate output.  nw <= 65535; a row
 * whose src[i] lies outside [0, nw) is left untouched. */
int afqmc_pop_gather(afqmc_ctx* ctx, int32_t nw, const int32_t* src, const void* in, void* out,
                     int64_t bytes_per_walker);

/* Set-up producers on the device (once per Hamiltonian and trial; ipie runs
 * them on the host), for a real Hamiltonian and a real trial.  Sizes come
 * from the context's dims (N = nbasis, X = nchol, n_s = nalpha / nbeta);
 * every array is device memory; each call returns when its results are
 * complete.  chol [N*N, X] f64 is ipie's GenericRealChol.chol.
 *   afqmc_setup_pack: chol_packed [N(N+1)/2, X] in numpy.triu_indices order
 *     (ipie/hamiltonians/generic.py:64-72, utils/pack_numba.py:36-45).
 *   afqmc_setup_h1e_mod: h1e_mod[s] = h1e[s] - 1/2 sum_{k,x} L[i,k,x] L[j,k,x],
 *     h1e and h1e_mod [2, N, N] (generic.py:31-43).
 *   afqmc_setup_half_rotate: rchol [X, n_s N] f64 and rh1 [n_s, N] complex for
 *     spin s from psi [N, n_s] f64 (row stride ld_psi) and h1 [N, N]
 *     (trial_wavefunction/half_rotate.py:11-116).
 *   afqmc_setup_mf_shift: mf [X] complex = i L^T Re(g) - L^T Im(g) for the
 *     charge density g [N*N] complex, the sum of both spins' G
 *     (propagation/phaseless_base.py:153-172).
 *   afqmc_setup_expH1: expH1 [2, N, N] f64 = exp(-dt/2 (h1e_mod[s] -
 *     i sum_x L_x mf_x)) for a purely imaginary mf (its real part is not read),
 *     by scaling and squaring a degree-18 Taylor polynomial in place of
 *     scipy.linalg.expm (phaseless_base.py:27-74). */
int afqmc_setup_pack(afqmc_ctx* ctx, const double* chol, double* chol_packed);
int afqmc_setup_h1e_mod(afqmc_ctx* ctx, const double* chol, const double* h1e, double* h1e_mod);
int afqmc_setup_half_rotate(afqmc_ctx* ctx, int32_t spin, const double* psi, int32_t ld_psi, const double* chol,
                            const double* h1, double* rchol, double* rh1);
int afqmc_setup_mf_shift(afqmc_ctx* ctx, const double* chol, const double* gcharge, double* mf);
int afqmc_setup_expH1(afqmc_ctx* ctx, const double* chol, const double* h1e_mod, const double* mf, double dt,
                      double* expH1);

/* The per-step walker sums of ipie's WalkerAccumulator.update
 * (ipie/walkers/base_walkers.py:43-50): acc [3] complex += (sum weight,
 * sum unscaled_weight, sum weight * values[i * value_stride]) over nw walkers;
 * unscaled_weight or values may be NULL (their entry is left alone).  values is
 * complex; value_stride counts complex entries (1 for hybrid_energy [W], 3 for
 * column 0 of an energy [W, 3]).  Sums run in a fixed order on the device; no
 * synchronisation. */
int afqmc_walker_sums(afqmc_ctx* ctx, int32_t nw, const double* weight, const double* unscaled_weight,
                      const double* values, int64_t value_stride, double* acc);
/* weight <- clip(weight, -bound, bound) in place, NaN kept: the driver's clip
 * after each step (ipie/qmc/afqmc.py:584-588, bound = 0.1 x the total weight). */
int afqmc_clip_weights(afqmc_ctx* ctx, int32_t nw, double* weight, double bound);

/* Step graphs (no counterpart in the reference, whose steps are host-driven
 * CuPy calls; for callers that replay a block of steps with one launch).
 * afqmc_graph_begin opens a hipStreamBeginCapture on the context's stream
 * (which must not be the null stream); every entry point called on this
 * context until afqmc_graph_end is recorded, not run.  While recording, the
 * field-stream counter and eshift are read from device memory, each recorded
 * step advances the counter by one, and the control variate's subset draws
 * and Adam steps also live in device memory, so a replay continues the
 * trajectory.  afqmc_graph_end instantiates the graph (replacing the previous
 * one) and returns the number of recorded steps.  afqmc_graph_launch sets the
 * counter of the first replayed step and eshift, then replays the graph
 * `replays` times on the context's stream: replay r, recorded step j uses
 * counter + r * steps + j.  Results equal the eager calls with those counters.
 * Calls that allocate workspace, synchronise or read back (afqmc_sync,
 * afqmc_get_flags, afqmc_get_nfb, afqmc_profile_read, the surrogate's
 * configure / stats and warm-up steps, afqmc_set_static) fail while a capture
 * is open: run the recorded sequence once eagerly first. */
int afqmc_graph_begin(afqmc_ctx* ctx);
int afqmc_graph_end(afqmc_ctx* ctx, int32_t* steps);
int afqmc_graph_launch(afqmc_ctx* ctx, int32_t replays, uint64_t counter, double eshift);

/* Multi-system step graphs (no counterpart in the reference, which runs one
 * system per process): independent contexts (systems of different sizes, one
 * device) recorded into one graph, each on its own stream, so their kernels
 * overlap.  afqmc_batch_begin starts the capture on the first context's stream
 * and forks every other context's stream from it (each context needs its own
 * non-null stream, set with afqmc_set_stream before begin and kept until end);
 * the calls made on each context until afqmc_batch_end are recorded as for
 * afqmc_graph_begin; afqmc_batch_end joins the streams and instantiates.
 * afqmc_batch_steps returns the steps recorded per context ([n]).
 * afqmc_batch_launch sets each context's counter and eshift ([n] each) and
 * replays the graph on the first context's stream.  Errors are reported in
 * afqmc_last_error of the first context. */
typedef struct afqmc_batch afqmc_batch;
int afqmc_batch_create(afqmc_batch** out, afqmc_ctx* const* ctxs, int32_t n);
void afqmc_batch_destroy(afqmc_batch* batch);
int afqmc_batch_begin(afqmc_batch* batch);
int afqmc_batch_end(afqmc_batch* batch);
int afqmc_batch_steps(afqmc_batch* batch, int32_t* steps);
int afqmc_batch_launch(afqmc_batch* batch, int32_t replays, const uint64_t* counters, const double* eshifts);

/* Ghalf_s = (phi_s^T psi_s^*)^{-1} phi_s^T and ovlp = prod_s det(phi_s^T psi_s^*) e^{-log_shift}.
 * phia [W, N, n_a], phib [W, N, n_b], Ghalfa [W, n_a, N], Ghalfb [W, n_b, N],
 * ovlp [W] complex; log_shift [W] real or NULL. */
int afqmc_greens_function(afqmc_ctx* ctx, int32_t nwalkers, const double* phia, const double* phib,
                          double* ghalfa, double* ghalfb, double* ovlp, const double* log_shift);
int afqmc_overlap(afqmc_ctx* ctx, int32_t nwalkers, const double* phia, const double* phib, double* ovlp,
                  const double* log_shift);
/* vbias [W, X] complex */
int afqmc_force_bias(afqmc_ctx* ctx, int32_t nwalkers, const double* ghalfa, const double* ghalfb,
                     double* vbias);
/* phi_s <- expH1[s] phi_s, in place */
int afqmc_one_body(afqmc_ctx* ctx, int32_t nwalkers, double* phia, double* phib);
/* Two-body step from the current Ghalf: force bias, xbar (capped), fields,
 * cmf / cfb [W] complex, VHS and the Taylor exponential applied in place.
 * xi [W, X] real injects the fields; NULL draws them from the counter-based
 * generator (seed, counter).  xi_out [W, X] (optional) receives the fields
 * used; vbias_out [W, X] complex (optional) the force bias. */
int afqmc_two_body(afqmc_ctx* ctx, int32_t nwalkers, double* phia, double* phib, const double* ghalfa,
                   const double* ghalfb, const double* xi, uint64_t seed, uint64_t counter, double* cmf,
                   double* cfb, double* xi_out, double* vbias_out);
/* weight [W] real, hybrid_energy [W] complex (in: previous), walker_ovlp [W] complex (out) */
int afqmc_update_weight(afqmc_ctx* ctx, int32_t nwalkers, const double* ovlp, const double* ovlp_new,
                        const double* cfb, const double* cmf, double eshift, double* weight,
                        double* hybrid_energy, double* walker_ovlp);
/* One full phaseless step (GF, one body, two body, one body, overlap, weights).
 * Ghalf receives the Green's function of the walkers at the start of the step;
 * ovlp [W] complex receives the new overlap.  energy_prev [W, 3] complex
 * (optional) receives the local energy of the incoming walkers (the state the
 * step starts from), computed from the same Green's function. */
int afqmc_propagate(afqmc_ctx* ctx, int32_t nwalkers, double* phia, double* phib, double* ghalfa,
                    double* ghalfb, const double* xi, uint64_t seed, uint64_t counter, double eshift,
                    double* weight, double* ovlp, double* hybrid_energy, const double* log_shift,
                    double* energy_prev);
/* energy [W, 3] complex = (E1 + E2, E1, E2) from Ghalf. */
int afqmc_local_energy(afqmc_ctx* ctx, int32_t nwalkers, const double* ghalfa, const double* ghalfb,
                       double* energy);
/* QR re-orthogonalisation in place; ovlp [W] complex is divided by detR [W]
 * real; log_detR [W] real is incremented; detR_shift [W] real or NULL. */
int afqmc_reortho(afqmc_ctx* ctx, int32_t nwalkers, double* phia, double* phib, double* ovlp, double* detR,
                  double* log_detR, const double* detR_shift);
/* Per-phase timing: afqmc_profile(ctx, 1) clears the totals and starts
 * recording HIP events around every phase; afqmc_profile_read synchronises
 * and returns accumulated milliseconds and launch counts per phase
 * (arrays of AFQMC_NPHASES). */
int afqmc_profile(afqmc_ctx* ctx, int enable);
int afqmc_profile_read(afqmc_ctx* ctx, double* ms, int64_t* launches);
/* Launch shapes, for the autotuner (ipie_amd/autotune.py).  No counterpart in
 * ipie, whose GPU kernels take CuPy's / numba's fixed shapes.  Knobs:
 *   AFQMC_TUNE_FB_SPLIT  K ranges of the force-bias contraction (FP64: in
 *                        total; mixed precision: per spin)
 *   AFQMC_TUNE_QUAD_WGS  workgroups the column units of the split-BF16
 *                        quadratic-form exchange aim for (sets the K-block
 *                        length of a unit; mixed precision, quadratic form)
 *   AFQMC_TUNE_EXX_SPLIT Cholesky-range split of the T-form exchange
 *                        (0 restores the heuristic)
 *   AFQMC_TUNE_QUAD_RGROUP
 *                        walker row blocks (128 walkers each) whose workgroups
 *                        are grouped over the column units of the split-BF16
 *                        quadratic form (mixed precision, quadratic form; the
 *                        default, all of them, puts every row block of a unit
 *                        on one XCD together)
 *   AFQMC_TUNE_TAYLOR    kernel of the split-BF16 Taylor series: 0 = 32-wide
 *                        K chunks, two in LDS; 1 = 16-wide chunks, four in LDS
 *                        (three loads in flight, half-rate MFMAs)
 *   AFQMC_TUNE_TILE_FB, AFQMC_TUNE_TILE_VHS, AFQMC_TUNE_TILE_ONE_BODY
 *                        workgroup tile of the split-BF16 force-bias, VHS and
 *                        one-body GEMMs (mixed precision): 0 = 128 x 160,
 *                        1 = 128 x 128, 2 = 128 x 80, 3 = 64 x 160 (three LDS
 *                        stages), 4 = 128 x 320, 5 = 64 x 160 (two LDS
 *                        stages; walker rows x output columns, 8 waves each)
 * afqmc_tune_get returns the current value and the admissible range [lo, hi];
 * hi = 0 when the knob does not apply to this context.  afqmc_tune_set
 * changes only launch shapes and summation grouping (results agree to
 * rounding); it is refused during graph capture, may allocate a larger
 * partial buffer, and keeps the buffers it replaces until afqmc_destroy, so
 * a graph captured before stays valid (with the old shapes). */
#define AFQMC_TUNE_FB_SPLIT 1
#define AFQMC_TUNE_QUAD_WGS 2
#define AFQMC_TUNE_EXX_SPLIT 3
#define AFQMC_TUNE_TILE_FB 4
#define AFQMC_TUNE_TILE_VHS 5
#define AFQMC_TUNE_TILE_ONE_BODY 6
#define AFQMC_TUNE_QUAD_RGROUP 7
#define AFQMC_TUNE_TAYLOR 8
int afqmc_tune_get(afqmc_ctx* ctx, int32_t knob, int32_t* value, int32_t* lo, int32_t* hi);
int afqmc_tune_set(afqmc_ctx* ctx, int32_t knob, int32_t value);
/* Learned control variate for the exchange energy of afqmc_propagate's
 * energy_prev output.  The predictor of each walker's exchange sum is
 *   g = E_T + 2 <K, G - G_T> + Q(G - G_T) + alpha m(f):
 * the exchange expanded about the trial's Green's function G_T (E_T and K in
 * FP64, computed once), its quadratic remainder Q in single FP8 or BF16
 * products (afqmc_surrogate_quadratic), and
 * an MLP m (2 x 128 SiLU, BF16 weights, FP32 accumulation) over the one-body
 * and Coulomb energies, mean / max |xbar|, log|overlap| and the overlap phase,
 * trained on the residual, with a gain alpha in [0, 1] fitted on earlier
 * steps' exact samples.  The exact exchange is computed for a random subset of
 * round(fraction * W) walkers (equal inclusion probability, redrawn every
 * step) and combined as
 *   E_x = g + [w in S] W / k (E_x_exact - g),
 * which is unbiased per walker.  The first warmup_steps energy evaluations are
 * all exact and train the model (train_iters Adam steps each, learning rate
 * lr); afterwards it keeps learning from the sampled walkers.  train_iters = 0
 * turns the MLP term off (gain held at 0, no training): the predictor is then
 * the physical control variate E_T + 2 <K, G - G_T> + Q alone.  enable = 0
 * restores the plain exact estimator.  afqmc_local_energy is always exact. */
int afqmc_surrogate_configure(afqmc_ctx* ctx, int enable, double fraction, int warmup_steps, int train_iters,
                              double lr, uint64_t seed);
/* The quadratic term Q(G - G_T) of the control variate:
 *   AFQMC_SURROGATE_Q_FP8 (default): one FP8 (E4M3) product, d and W' scaled
 *     by powers of two (per walker for d), FP32 accumulation, where the
 *     quadratic-form planes exist (mixed precision, N < X); elsewhere BF16;
 *   AFQMC_SURROGATE_Q_BF16: one BF16 product;
 *   AFQMC_SURROGATE_Q_NONE: no Q; the predictor is E_T + 2 <K, G - G_T>
 *     (+ alpha m(f)), the exact sample and the MLP term making up the rest.
 * The estimator stays unbiased in every mode (the rounding of Q changes only
 * how well g predicts, i.e. the variance).  Applies to steps enqueued or
 * captured after the call; the first FP8 evaluation (which builds the FP8 copy
 * of W') must not be inside a capture. */
#define AFQMC_SURROGATE_Q_NONE 0
#define AFQMC_SURROGATE_Q_BF16 1
#define AFQMC_SURROGATE_Q_FP8 2
int afqmc_surrogate_quadratic(afqmc_ctx* ctx, int mode);
/* out[0..n-1], n >= 6: energy evaluations so far, warm-up length, last subset
 * size, last training loss per sample (normalised units), mean
 * |exact - predicted|^2 on the last subset (Ha^2), variance over the warm-up of
 * the MLP's target Re(exact - control variate), and (n >= 7) the current gain of
 * the MLP term.  Synchronises. */
int afqmc_surrogate_stats(afqmc_ctx* ctx, double* out, int n);
/* cv [W] complex = E_T + 2 <K, G - G_T> + Q(G - G_T) from Ghalf, Q in the mode
 * of afqmc_surrogate_quadratic (none after AFQMC_SURROGATE_Q_NONE): the physical control variate of
 * each walker's exchange sum (the predictor above without the MLP term),
 * through the kernels afqmc_propagate runs for it.
 * Needs afqmc_surrogate_configure(enable = 1); changes no walker or model
 * state.  No counterpart in ipie; it lets the autotuner time the control
 * variate's exchange. */
int afqmc_control_variate(afqmc_ctx* ctx, int32_t nwalkers, const double* ghalfa, const double* ghalfb,
                          double* cv);
/* n standard normals (Philox4x32-10, Box-Muller) for stream `counter`. */
int afqmc_draw_normals(afqmc_ctx* ctx, double* out, int64_t n, uint64_t seed, uint64_t counter);

/* Green's function that allocates nothing.  No counterpart in ipie: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same results as
 * afqmc_greens_function (afqmc_overlap when ghalfa is NULL), without a
 * context: psia [N, n_a] / psib [N, n_b] complex are the trial orbitals, the
 * other arrays as in afqmc_greens_function, walker_flags [W] int32 (optional)
 * receives AFQMC_FLAG_SINGULAR OR-ed per walker, and the work is enqueued on
 * `stream` (a hipStream_t, NULL for the null stream).  The LU and inverse work
 * buffers live in a static device array of AFQMC_NOALLOC_SCRATCH_BYTES per
 * GPU (held until the process ends); afqmc_noalloc_scratch_bytes gives what a
 * call needs, and a call that needs more returns AFQMC_ERR_SHAPE.  Every call
 * on one GPU shares that GPU's array: calls that overlap in time on one GPU
 * race.  afqmc_noalloc_last_error returns the calling thread's message of its
 * last failed call. */
#define AFQMC_NOALLOC_SCRATCH_BYTES 134217728ULL /* 128 MiB */
size_t afqmc_noalloc_scratch_bytes(int32_t nwalkers, int32_t nalpha, int32_t nbeta);
const char* afqmc_noalloc_last_error(void);
int afqmc_greens_function_noalloc(int32_t nwalkers, int32_t nbasis, int32_t nalpha, int32_t nbeta, int32_t rhf,
                                  const double* psia, const double* psib, const double* phia, const double* phib,
                                  double* ghalfa, double* ghalfb, double* ovlp, const double* log_shift,
                                  int32_t* walker_flags, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* AFQMC_HIP_H */
