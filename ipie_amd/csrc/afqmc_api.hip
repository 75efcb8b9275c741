// C ABI of libafqmc_hip.so: context, workspace and step orchestration.
// Declared in include/afqmc_hip.h.  Every entry point validates its inputs on
// the host, enqueues kernels on the context's stream and returns a status;
// nothing here blocks the host except afqmc_sync / afqmc_get_flags /
// afqmc_get_nfb.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <cstring>

#include <algorithm>
#include <cmath>
#include <random>
#include <vector>

#include "afqmc_internal.hpp"
#include "../../include/afqmc_hip.h"

using namespace afqmc;

struct PendingEvent {
  int phase;
  hipEvent_t start, stop;
};

struct afqmc_ctx {
  afqmc_dims d;
  int device = 0;
  hipStream_t stream = nullptr;
  afqmc_static s{};
  bool have_static = false;
  char err[512] = {0};
  int ncol = 0;  // columns of the concatenated walker buffers (n_a + n_b, or n_a)
  int fb_ksplit = 1;
  int xs_a = 1, xs_b = 1;
  // workspace (device)
  double* vbias_part = nullptr;  // [ksplit, W, X] complex
  double* xshifted = nullptr;    // [W, X] complex
  double* wvec = nullptr;        // 6 x [W] complex: cmf, cfb, ecoul, ovlp_old, ovlp_new, spare
  double* exx_a = nullptr;       // [xs_a, W] complex
  double* exx_b = nullptr;       // [xs_b, W] complex
  double* vhs = nullptr;         // [W, N, N] complex
  double* buf[4] = {nullptr, nullptr, nullptr, nullptr};  // P, Ta, Tb, Q: [W, N, ncol] complex
  int* tri = nullptr;            // [2, P] packed-index rows / cols
  int* flags = nullptr;
  int* wflags = nullptr;    // [W] per-walker flags (OR of AFQMC_FLAG_*) since the last take
  int* psi_flag = nullptr;  // scratch of afqmc_set_static's check of the trial orbitals
  int g_planes_nw = 0;      // > 0: gplanes hold Ghalf of these nw walkers, written by the GF of this step
  bool psi_real = false;    // the trial orbitals are real (imaginary parts exactly zero)
  unsigned long long* nfb = nullptr;
  bool ws_ready = false;
  char describe_dims[128] = {0};
  // split-BF16 operand planes (AFQMC_PREC_MIXED)
  int kpad = 0;  // N rounded up to 32
  int xpad = 0;  // X rounded up to 32
  bool bx_ok[2] = {false, false};
  int bx_wb[2] = {1, 1}, bx_xb[2] = {1, 1};
  bool bf16_fb = false;                       // force bias from split-BF16 planes
  uint16_t* gplanes[2] = {nullptr, nullptr};  // per spin: 4 x [W * n_s, kpad] (Ghalf)
  uint16_t* lplanes[2] = {nullptr, nullptr};  // per spin: 2 x [X * n_s, kpad] (half-rotated Cholesky)
  uint16_t* xplanes = nullptr;                // 4 x [W, xpad] (shifted fields)
  uint16_t* cplanes = nullptr;                // 2 x [NAT 256, xpad] (Cholesky rows in VHS atom order)
  uint16_t* vatoms = nullptr;                 // VHS split planes, upper atoms per walker ([W][NAT][4][256])
  uint16_t* zatom = nullptr;                  // one all-zero atom (4 x 256)
  // 4 x [W, rpad, kpad] each, transposed: [0] the planes of P^T (input of the
  // Taylor series) and, after it, of Q^T; [k] those of term k (max(2, exp_nmax) buffers)
  uint16_t* tplanes[TAYLOR_MAX_TERMS] = {};
  uint16_t* dplanes = nullptr;                // split-BF16 expH1_s - I per spin, 2 x [N, kpad] each
  uint16_t* oplanes = nullptr;                // 4 x [W n_s, kpad]: transposed planes of phi_s (one body)
  bool ptp_ready = false;                     // tplanes[0] holds the planes of P^T from the one-body GEMM
  bool planes_ready = false;                  // lplanes / cplanes hold the current static tensors
  // exchange as g^T Wq g (when Wq fits the budget and N < X)
  bool quad_on = false;
  bool quad_ready = false;                    // Wq holds the current static tensors
  double* qw64[2] = {nullptr, nullptr};       // FP64 Wq per spin [n_s N, n_s N]
  uint16_t* qwp[2] = {nullptr, nullptr};      // split-BF16 Wq per spin 2 x [n_s Kp, n_s Kp]
  // FP8 quadratic form of the control variate (made on first use): W' per spin
  // [q8_ld][q8_ld] bytes, q8_ld = n_s Kp rounded up to 64; the scales s_w (FP32,
  // two) and their maximum scratch (two uints) on the device; the FP8 planes of dG
  // [2][W][max q8_ld] and each walker's inverse scale; column units in 64-value chunks
  uint8_t* qw8[2] = {nullptr, nullptr};
  int64_t q8_ld[2] = {0, 0};
  float* q8_scale = nullptr;
  uint8_t* g8 = nullptr;
  double* q8_rowinv = nullptr;
  bool q8_ready = false;
  int4* qunits8[2] = {nullptr, nullptr};
  std::vector<int> qunit8_off[2], qunit8_n[2];
  double* qy = nullptr;                       // FP64: Y = G Wq [W, n N]; BF16: dot partials [slots, W] (complex)
  int4* qunits[2] = {nullptr, nullptr};       // BF16 column units per spin, one table per row-block count
  std::vector<int> qunit_off[2], qunit_n[2];  // [mb - 1]: offset and length of the table for mb row blocks
  int xs_used[2] = {1, 1};  // partial-sum count of the last exchange per spin
  // launch shapes that afqmc_tune_set may change (heuristic defaults)
  int fb_split_spin = BF16_FB_SPLIT;  // split-BF16 force bias: K ranges per spin
  int fb_cap = 0;                     // K ranges vbias_part holds (0: not allocated yet)
  int quad_target = QUAD_TARGET_WGS;  // split-BF16 quadratic form: workgroups the column units aim for
  int tile_fb = 0, tile_vhs = 0, tile_ob = 0;  // split-BF16 GEMM tile shapes (force bias, VHS, one body)
  int quad_rgroup = 0;  // split-BF16 quadratic form: row blocks per workgroup group (0: all)
  int taylor_variant = 0;  // split-BF16 Taylor kernel (TaylorBf16Args::variant)
  int64_t qy_cap = 0;                 // complex entries qy holds (split-BF16 quadratic form)
  int exx_split = 0;                  // T-form exchange: Cholesky-range split (0: heuristic)
  std::vector<void*> retired;         // buffers a retune replaced: graphs captured before still use them
  int64_t exx_cap = 0;      // capacity (complex entries) of each exchange partial buffer
  double* gcomp = nullptr;  // [W, nmax, N] complex: Ghalf of sampled walkers (FP64 exchange)
  double* gf_o = nullptr;     // [W, 2, nmax, nmax] complex: overlap matrices
  double* gf_oinv = nullptr;  // [W, 2, nmax, nmax] complex: their inverses
  // learned exchange control variate
  struct {
    bool enabled = false;
    double fraction = 0.125;
    int warmup = 100, train_iters = 4, online_iters = 1;
    float lr = 1e-3f;
    int64_t steps = 0;
    int last_k = 0;
    uint64_t seed = 0;       // Philox key of the subset draws
    std::mt19937_64 rng;     // initial MLP weights
    float* feat = nullptr;
    double* pred = nullptr;
    double* est = nullptr;
    double* target = nullptr;
    double* resid = nullptr;
    int* walker_of = nullptr;
    double* xbar = nullptr;  // [2, W]: mean, max
    // physical control variate E_T + 2 <K, dG> + Q(dG) with Q in FP8 (qmode 2, where the
    // quadratic-form planes exist; BF16 otherwise) or BF16 (1), or E_T + 2 <K, dG> (0)
    int qmode = AFQMC_SURROGATE_Q_FP8;
    bool cv_ready = false;   // gT, K, E_T hold the current trial / Hamiltonian
    double* gT = nullptr;    // [2][nmax, N] complex trial Ghalf per spin
    double* K = nullptr;     // [2][nmax, N] complex
    double* et = nullptr;    // [2] complex E_T per spin
    double* cv = nullptr;    // [W] complex
    double* mlp = nullptr;   // [W] complex MLP output
    double* gtmp = nullptr;  // [2] complex overlap scratch of the trial GF
    double* e1 = nullptr;    // [W] complex
    float* master = nullptr;
    float* m = nullptr;
    float* v = nullptr;
    uint16_t* wbf = nullptr;
    float* bias = nullptr;
    float* grad = nullptr;
    float* loss = nullptr;
    SurrogateState* state = nullptr;
  } sur;
  // step graphs: values that change per step live in dstate; while a capture is
  // open the kernels read them from there
  StepState* dstate = nullptr;
  bool capturing = false;
  int capture_steps = 0, capture_sur = 0;  // field draws / control-variate energies recorded
  hipGraphExec_t graph = nullptr;
  int graph_steps = 0, graph_sur = 0;
  // per-phase HIP-event timing (afqmc_profile / afqmc_profile_read)
  bool prof = false;
  std::vector<hipEvent_t> ev_free;
  std::vector<PendingEvent> ev_pending;
  double prof_ms[AFQMC_NPHASES] = {0};
  int64_t prof_count[AFQMC_NPHASES] = {0};
};

namespace {

int fail(afqmc_ctx* c, int code, const char* fmt, ...) {
  if (c) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(c->err, sizeof(c->err), fmt, ap);
    va_end(ap);
  }
  return code;
}

#define HIP_TRY(c, expr)                                                                     \
  do {                                                                                       \
    hipError_t _e = (expr);                                                                  \
    if (_e != hipSuccess) return fail((c), AFQMC_ERR_HIP, "%s: %s", #expr, hipGetErrorString(_e)); \
  } while (0)

hipEvent_t take_event(afqmc_ctx* c) {
  if (!c->ev_free.empty()) {
    hipEvent_t e = c->ev_free.back();
    c->ev_free.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Adds the phases whose stop event has completed, oldest first, to the totals
// and returns their events to the pool (no wait): keeps the pending list short
// when profiling stays on and the totals are read rarely.
void drain_completed(afqmc_ctx* c) {
  size_t k = 0;
  for (; k < c->ev_pending.size(); ++k) {
    const PendingEvent& pe = c->ev_pending[k];
    if (hipEventQuery(pe.stop) != hipSuccess) break;
    float t = 0.0f;
    if (hipEventElapsedTime(&t, pe.start, pe.stop) == hipSuccess) {
      c->prof_ms[pe.phase] += t;
      c->prof_count[pe.phase] += 1;
    }
    c->ev_free.push_back(pe.start);
    c->ev_free.push_back(pe.stop);
  }
  c->ev_pending.erase(c->ev_pending.begin(), c->ev_pending.begin() + (std::ptrdiff_t)k);
}

// Records a start/stop event pair around one phase when profiling is on.
struct PhaseScope {
  afqmc_ctx* c;
  int phase;
  hipEvent_t start = nullptr;
  PhaseScope(afqmc_ctx* ctx, int ph) : c(ctx), phase(ph) {
    if (c->prof && !c->capturing) {
      if (c->ev_pending.size() >= 1024) drain_completed(c);
      start = take_event(c);
      if (start) (void)hipEventRecord(start, c->stream);
    }
  }
  ~PhaseScope() {
    if (!start) return;
    hipEvent_t stop = take_event(c);
    if (!stop) return;
    (void)hipEventRecord(stop, c->stream);
    c->ev_pending.push_back(PendingEvent{phase, start, stop});
  }
};

int check_ready(afqmc_ctx* c, int32_t nw) {
  if (!c) return AFQMC_ERR_INVALID_ARG;
  if (!c->have_static) return fail(c, AFQMC_ERR_INVALID_ARG, "afqmc_set_static has not been called");
  if (nw < 1 || nw > c->d.nwalkers)
    return fail(c, AFQMC_ERR_SHAPE, "nwalkers=%d outside [1, %d]", nw, c->d.nwalkers);
  c->err[0] = 0;
  return AFQMC_OK;
}

bool two_spins(const afqmc_ctx* c) { return c->d.nbeta > 0 && !c->d.rhf; }

double* wv(afqmc_ctx* c, int i) { return c->wvec + 2 * (int64_t)i * c->d.nwalkers; }
enum { WV_CMF = 0, WV_CFB = 1, WV_ECOUL = 2, WV_OVLP_OLD = 3, WV_OVLP_NEW = 4 };

// set while this thread holds an open step-graph capture: workspace cannot be
// allocated then (run the captured calls once eagerly first)
thread_local bool t_capturing = false;

hipError_t dmalloc(void** p, size_t bytes) {
  if (t_capturing) return hipErrorStreamCaptureUnsupported;
  if (bytes == 0) bytes = 256;
  return hipMalloc(p, bytes);
}

int refuse_in_capture(afqmc_ctx* c, const char* what) {
  return fail(c, AFQMC_ERR_INVALID_ARG, "%s is not allowed while a step graph is being captured", what);
}

// Column units of the block-triangular quadratic-form GEMM with mb row blocks
// and K chunks of cw values: column tile t sums the chunks
// [0, min(ceil(QUAD_TILE (t + 1) / cw), kchunks)), cut at the multiples of T
// chunks (T chosen for about target_wgs workgroups, at least 1024 / cw, and
// at least 2048 / cw for a single row block that still makes 256 units or
// more: the control variate's exact sample at n400 ran 10% faster in 64-chunk
// units than in 32, while small systems need the shorter units to fill the
// GPU), and
// the units are listed by K block, then tile (chunks past a tile's own rows
// cover zeros of the block-triangular W').  The grid takes the row blocks
// fastest, so the workgroups that run together on one XCD (consecutive logical
// indices) are a few tiles of the same K block: they read the same d rows and
// the same K range, which L2 then serves once.
std::vector<int4> quad_units(int64_t nb, int kchunks, int64_t mb, int target_wgs, int cw = 32) {
  auto extent = [&](int64_t t) { return std::min<int64_t>(ceil_div((int64_t)QUAD_TILE * (t + 1), cw), kchunks); };
  int64_t total = 0;
  for (int64_t t = 0; t < nb; ++t) total += extent(t);
  const int64_t target = std::max<int64_t>(ceil_div(target_wgs, mb), 1);
  int64_t T = std::max<int64_t>(ceil_div(total, target), 1024 / cw);
  if (mb == 1 && ceil_div(total, 2048 / cw) >= 256) T = std::max<int64_t>(T, 2048 / cw);
  std::vector<int4> u;
  for (int64_t z = 0; z * T < kchunks; ++z)
    for (int64_t t = 0; t < nb; ++t) {
      const int64_t kc = extent(t);
      if (z * T >= kc) continue;
      int4 e;
      e.x = (int)t;
      e.y = (int)(z * T);
      e.z = (int)std::min<int64_t>((z + 1) * T, kc);
      e.w = 0;
      u.push_back(e);
    }
  return u;
}

// Split-BF16 planes of Ghalf and rchol per spin (mixed precision, surrogate).
int alloc_exchange_planes(afqmc_ctx* c) {
  const int64_t W = c->d.nwalkers, X = c->d.nchol;
  for (int s = 0; s < 2; ++s) {
    if (!c->bx_ok[s] || c->gplanes[s]) continue;
    const int n = s == 0 ? c->d.nalpha : c->d.nbeta;
    if (dmalloc((void**)&c->gplanes[s], 2 * (size_t)4 * W * n * c->kpad) != hipSuccess ||
        dmalloc((void**)&c->lplanes[s], 2 * (size_t)2 * X * n * c->kpad) != hipSuccess)
      return fail(c, AFQMC_ERR_ALLOC, "bf16 plane allocation failed");
    c->planes_ready = false;
  }
  return AFQMC_OK;
}

// Column-unit tables of the split-BF16 quadratic form for every row-block
// count (quad_units with c->quad_target), and the dot-partial buffer qy sized
// for the largest table.  A rebuild (afqmc_tune_set) keeps the tables and qy it
// replaces alive until afqmc_destroy: a graph captured before still uses them.
int build_quad_units(afqmc_ctx* c) {
  const int64_t W = c->d.nwalkers;
  int64_t max_units = 0;
  // [0, 1]: BF16 tables (32-value chunks) per spin; [2, 3]: FP8 (64-value chunks)
  std::vector<int4> all[4];
  std::vector<int> off[4], cnt[4];
  for (int s = 0; s < 2; ++s) {
    if (!c->qwp[s]) continue;
    const int64_t ldn = (int64_t)(s == 0 ? c->d.nalpha : c->d.nbeta) * c->kpad;
    for (int f = 0; f < 2; ++f) {
      const int t = 2 * f + s, cw = f ? 64 : 32;
      const int kch = (int)(f ? ceil_div(ldn, 64) : ldn / 32);
      for (int64_t mb = 1; mb <= ceil_div(W, 128); ++mb) {
        std::vector<int4> u = quad_units(ceil_div(ldn, QUAD_TILE), kch, mb, c->quad_target, cw);
        off[t].push_back((int)all[t].size());
        cnt[t].push_back((int)u.size());
        max_units = std::max<int64_t>(max_units, (int64_t)u.size());
        all[t].insert(all[t].end(), u.begin(), u.end());
      }
    }
  }
  if (max_units == 0) return AFQMC_OK;
  const int64_t need = max_units * QUAD_SLOTS * W;
  if (need > c->qy_cap) {
    double* qy = nullptr;
    if (dmalloc((void**)&qy, 16 * (size_t)need) != hipSuccess)
      return fail(c, AFQMC_ERR_ALLOC, "quadratic-form partial buffer allocation failed (%s)", c->describe_dims);
    if (c->qy) c->retired.push_back(c->qy);
    c->qy = qy;
    c->qy_cap = need;
  }
  for (int k = 0; k < 4; ++k) {
    if (all[k].empty()) continue;
    int4* t = nullptr;
    if (dmalloc((void**)&t, sizeof(int4) * all[k].size()) != hipSuccess)
      return fail(c, AFQMC_ERR_ALLOC, "quadratic-form unit table allocation failed (%s)", c->describe_dims);
    HIP_TRY(c, hipMemcpy(t, all[k].data(), sizeof(int4) * all[k].size(), hipMemcpyHostToDevice));
    int4*& slot = k < 2 ? c->qunits[k] : c->qunits8[k - 2];
    if (slot) c->retired.push_back(slot);
    slot = t;
    (k < 2 ? c->qunit_off[k] : c->qunit8_off[k - 2]) = off[k];
    (k < 2 ? c->qunit_n[k] : c->qunit8_n[k - 2]) = cnt[k];
  }
  return AFQMC_OK;
}

// Workspace is allocated on first use by an entry point that needs it (the
// Green's function and overlap use only their own O / O^-1 buffers, see do_gf).
int ensure_ws(afqmc_ctx* c) {
  if (c->ws_ready) return AFQMC_OK;
  const afqmc_dims& d = c->d;
  if (d.nchol < 1) return fail(c, AFQMC_ERR_SHAPE, "this entry point needs nchol >= 1");
  const int64_t W = d.nwalkers, N = d.nbasis, X = d.nchol, P = N * (N + 1) / 2;
  const int nmax = d.nalpha > d.nbeta ? d.nalpha : d.nbeta;
  bool ok = true;
  ok = ok && dmalloc((void**)&c->vbias_part, 16 * (size_t)c->fb_ksplit * W * X) == hipSuccess;
  if (ok) c->fb_cap = c->fb_ksplit;
  ok = ok && dmalloc((void**)&c->xshifted, 16 * (size_t)W * X) == hipSuccess;
  ok = ok && dmalloc((void**)&c->wvec, 16 * (size_t)6 * W) == hipSuccess;
  // partial sums: at most one per Cholesky vector per walker, whatever the split
  c->exx_cap = (int64_t)X * W;
  ok = ok && dmalloc((void**)&c->exx_a, 16 * (size_t)c->exx_cap) == hipSuccess;
  ok = ok && dmalloc((void**)&c->exx_b, 16 * (size_t)c->exx_cap) == hipSuccess;
  ok = ok && dmalloc((void**)&c->gcomp, 16 * (size_t)W * nmax * N) == hipSuccess;
  if (d.precision == AFQMC_PREC_MIXED) ok = ok && alloc_exchange_planes(c) == AFQMC_OK;
  if (d.precision == AFQMC_PREC_MIXED) {
    ok = ok && dmalloc((void**)&c->xplanes, 2 * (size_t)4 * W * c->xpad) == hipSuccess;
    const int64_t NAT = vhs_atom_count((int)N);
    ok = ok && dmalloc((void**)&c->cplanes, 2 * (size_t)2 * NAT * 256 * c->xpad) == hipSuccess;
    const size_t tbytes = 2 * (size_t)4 * W * taylor_rows_padded(c->ncol) * c->kpad;
    ok = ok && dmalloc((void**)&c->vatoms, 2 * (size_t)W * NAT * 1024) == hipSuccess;
    ok = ok && dmalloc((void**)&c->zatom, 2 * (size_t)1024) == hipSuccess;
    ok = ok && hipMemset(c->zatom, 0, 2 * (size_t)1024) == hipSuccess;
    for (int i = 0; i < std::max(2, d.exp_nmax); ++i) {
      ok = ok && dmalloc((void**)&c->tplanes[i], tbytes) == hipSuccess;
      ok = ok && hipMemset(c->tplanes[i], 0, tbytes) == hipSuccess;  // padding stays zero
    }
    ok = ok && dmalloc((void**)&c->dplanes, 2 * (size_t)2 * 2 * N * c->kpad) == hipSuccess;
    const size_t obytes = 2 * (size_t)4 * W * nmax * c->kpad;
    ok = ok && dmalloc((void**)&c->oplanes, obytes) == hipSuccess;
    ok = ok && hipMemset(c->oplanes, 0, obytes) == hipSuccess;  // K padding stays zero
  }
  if (c->quad_on) {
    const bool mixed = d.precision == AFQMC_PREC_MIXED;
    const int64_t Np = mixed ? c->kpad : N;
    for (int s = 0; s < (two_spins(c) ? 2 : 1); ++s) {
      const int64_t rows = (int64_t)(s == 0 ? d.nalpha : d.nbeta) * Np;
      if (mixed) ok = ok && dmalloc((void**)&c->qwp[s], 2 * (size_t)2 * rows * rows) == hipSuccess;
      else ok = ok && dmalloc((void**)&c->qw64[s], 8 * (size_t)rows * rows) == hipSuccess;
    }
    // FP64: Y = G Wq; split-BF16: QUAD_SLOTS dot partials per column unit, sized
    // with the unit tables (build_quad_units)
    if (!mixed) ok = ok && dmalloc((void**)&c->qy, 16 * (size_t)W * nmax * Np) == hipSuccess;
  }
  // FP64 VHS [W, N, N] complex (2.6 GB at N=400); the split-BF16 path keeps VHS as atoms only
  if (d.precision != AFQMC_PREC_MIXED) ok = ok && dmalloc((void**)&c->vhs, 16 * (size_t)W * N * N) == hipSuccess;
  for (int i = 0; i < 4; ++i) ok = ok && dmalloc((void**)&c->buf[i], 16 * (size_t)W * N * c->ncol) == hipSuccess;
  ok = ok && dmalloc((void**)&c->tri, sizeof(int) * 2 * (size_t)P) == hipSuccess;
  if (!ok) return fail(c, AFQMC_ERR_ALLOC, "workspace allocation failed (%s)", c->describe_dims);
  std::vector<int> tri(2 * P);
  int64_t p = 0;
  for (int i = 0; i < N; ++i)
    for (int j = i; j < N; ++j, ++p) {
      tri[p] = i;
      tri[P + p] = j;
    }
  HIP_TRY(c, hipMemcpy(c->tri, tri.data(), sizeof(int) * 2 * P, hipMemcpyHostToDevice));
  const int r = build_quad_units(c);
  if (r) return r;
  HIP_TRY(c, hipMemset(c->wvec, 0, 16 * (size_t)6 * W));
  c->ws_ready = true;
  return AFQMC_OK;
}

// Static tensors needed by the propagation / energy entry points.
int need_static(afqmc_ctx* c) {
  const bool two = c->d.nbeta > 0 && !c->d.rhf;
  if (!c->s.chol_packed || !c->s.rchola || !c->s.rH1a || !c->s.expH1 || !c->s.mf_shift)
    return fail(c, AFQMC_ERR_INVALID_ARG, "static Hamiltonian tensors were not set");
  if (two && (!c->s.rcholb || !c->s.rH1b)) return fail(c, AFQMC_ERR_INVALID_ARG, "beta static tensors were not set");
  if (!(c->s.dt > 0.0)) return fail(c, AFQMC_ERR_INVALID_ARG, "dt must be positive");
  return ensure_ws(c);
}

// Split-BF16 planes of the static tensors (half-rotated and packed Cholesky),
// made on first use after afqmc_set_static.
int ensure_planes(afqmc_ctx* c) {
  if (c->planes_ready) return AFQMC_OK;
  const int N = c->d.nbasis;
  for (int s = 0; s < 2; ++s)
    if (c->lplanes[s]) {
      const int n = s == 0 ? c->d.nalpha : c->d.nbeta;
      HIP_TRY(c, launch_split_real(s == 0 ? c->s.rchola : c->s.rcholb, (int64_t)c->d.nchol * n, N, c->kpad,
                                   c->lplanes[s], c->stream));
    }
  if (c->cplanes)
    HIP_TRY(c, launch_split_chol_atoms(c->s.chol_packed, N, c->d.nchol, c->xpad, c->cplanes, c->stream));
  if (c->dplanes) {
    HIP_TRY(c, launch_split_eye(c->s.expH1, N, c->kpad, c->dplanes, c->stream));
    if (two_spins(c))
      HIP_TRY(c, launch_split_eye(c->s.expH1 + (int64_t)N * N, N, c->kpad, c->dplanes + 2 * (int64_t)N * c->kpad,
                                  c->stream));
  }
  c->planes_ready = true;
  return AFQMC_OK;
}

// ---------------------------------------------------------------- phases
int do_gf(afqmc_ctx* c, int nw, const double* phia, const double* phib, double* ga, double* gb, double* ovlp,
          const double* log_shift, bool want_ghalf, bool fb_planes = false) {
  GfArgs a{};
  a.nwalkers = nw;
  a.nbasis = c->d.nbasis;
  a.nalpha = c->d.nalpha;
  a.nbeta = c->d.nbeta;
  a.rhf = c->d.rhf;
  a.phia = phia;
  a.phib = phib;
  a.psia = c->s.psia;
  a.psib = c->s.psib;
  a.ghalfa = ga;
  a.ghalfb = gb;
  a.ovlp = ovlp;
  a.log_shift = log_shift;
  a.flags = c->flags;
  a.wflags = c->wflags;
  a.psi_real = c->psi_real ? 1 : 0;
  // the split-BF16 force bias that follows in the same step reads Ghalf's planes
  // from this kernel instead of splitting Ghalf again
  c->g_planes_nw = 0;
  if (fb_planes && want_ghalf && c->bf16_fb && c->gplanes[0] && (!two_spins(c) || c->gplanes[1])) {
    a.gplanes[0] = c->gplanes[0];
    a.gplanes[1] = two_spins(c) ? c->gplanes[1] : nullptr;
    a.kpad = c->kpad;
    c->g_planes_nw = nw;
  }
  if (!c->gf_o) {
    const size_t bytes = gf_work_bytes((int)c->d.nwalkers, c->d.nalpha, c->d.nbeta);
    if (dmalloc((void**)&c->gf_o, bytes) != hipSuccess || dmalloc((void**)&c->gf_oinv, bytes) != hipSuccess)
      return fail(c, AFQMC_ERR_ALLOC, "Green's function workspace allocation failed");
  }
  PhaseScope ps(c, want_ghalf ? AFQMC_PHASE_GF : AFQMC_PHASE_OVERLAP);
  HIP_TRY(c, launch_gf(a, want_ghalf, c->gf_o, c->gf_oinv, c->stream));
  return AFQMC_OK;
}

// Force bias from split-BF16 planes: one GEMM per spin, K split in
// BF16_FB_SPLIT ranges, partial sums [spin * BF16_FB_SPLIT + z][W][X] summed
// by the fields kernel.
int do_force_bias_bf16(afqmc_ctx* c, int nw, const double* ga, const double* gb) {
  const int N = c->d.nbasis;
  int r = ensure_planes(c);
  if (r) return r;
  PhaseScope ps(c, AFQMC_PHASE_FORCE_BIAS);
  const int nspin = two_spins(c) ? 2 : 1;
  for (int s = 0; s < nspin; ++s) {
    const int n = s == 0 ? c->d.nalpha : c->d.nbeta;
    if (c->g_planes_nw != nw)
      HIP_TRY(c, launch_split_ghalf(s == 0 ? ga : gb, (int64_t)nw * n, N, c->kpad, c->gplanes[s], c->stream));
    Bf16GemmArgs b{};
    b.a = c->gplanes[s];
    b.a_plane = (int64_t)nw * n * c->kpad;
    b.lda = (int64_t)n * c->kpad;
    b.m = nw;
    b.b = c->lplanes[s];
    b.b_plane = (int64_t)c->d.nchol * n * c->kpad;
    b.ldb = (int64_t)n * c->kpad;
    b.nn = c->d.nchol;
    b.kchunks = (int)(b.lda / 32);
    b.ksplit = c->fb_split_spin;
    b.epilogue = BF16_EPI_PARTIAL;
    b.tile = c->tile_fb;
    b.out = c->vbias_part;
    b.out32 = reinterpret_cast<float*>(c->vbias_part);  // read back by fields_kernel as FP32 (part32)
    b.z_offset = s * c->fb_split_spin;
    b.scale = c->d.rhf ? 2.0 : 1.0;  // force_bias.py:101-104 doubles for rhf whatever ndown
    HIP_TRY(c, launch_bf16_gemm(b, c->stream));
  }
  c->g_planes_nw = 0;  // the exchange reuses the planes
  return AFQMC_OK;
}

int do_force_bias(afqmc_ctx* c, int nw, const double* ga, const double* gb) {
  if (c->bf16_fb) return do_force_bias_bf16(c, nw, ga, gb);
  const int N = c->d.nbasis, na = c->d.nalpha, nb = c->d.nbeta;
  ForceBiasArgs a;
  a.nwalkers = nw;
  a.nchol = c->d.nchol;
  a.ksplit = c->fb_ksplit;
  a.nseg = 1;
  a.seg[0] = NtSegment{ga, (int64_t)na * N, c->s.rchola, (int64_t)na * N, na * N};
  a.scale = 1.0;
  if (two_spins(c)) {
    a.nseg = 2;
    a.seg[1] = NtSegment{gb, (int64_t)nb * N, c->s.rcholb, (int64_t)nb * N, nb * N};
  } else if (c->d.rhf) {  // force_bias.py:101-104 doubles for rhf whatever ndown
    a.scale = 2.0;
  }
  a.vbias_part = c->vbias_part;
  PhaseScope ps(c, AFQMC_PHASE_FORCE_BIAS);
  HIP_TRY(c, launch_force_bias(a, c->stream));
  return AFQMC_OK;
}

int do_fields(afqmc_ctx* c, int nw, bool want_fields, const double* xi, uint64_t seed, uint64_t counter,
              double* xi_out, double* vbias_out) {
  FieldsArgs a;
  a.nwalkers = nw;
  a.nchol = c->d.nchol;
  a.ksplit = c->fb_ksplit;
  a.vbias_part = c->vbias_part;
  a.part32 = c->bf16_fb ? 1 : 0;
  a.mf_shift = c->s.mf_shift;
  a.xi = xi;
  a.seed = seed;
  a.counter = counter;
  a.state = c->capturing ? c->dstate : nullptr;
  a.sqrt_dt = sqrt(c->s.dt);
  a.fbbound = 1.0;
  a.xi_out = xi_out;
  a.xshifted = c->xshifted;
  a.cmf = wv(c, WV_CMF);
  a.cfb = wv(c, WV_CFB);
  a.ecoul = wv(c, WV_ECOUL);
  a.vbias_out = vbias_out;
  a.nfb = c->nfb;
  a.want_fields = want_fields ? 1 : 0;
  a.xbar_mean = c->sur.enabled ? c->sur.xbar : nullptr;
  a.xbar_max = c->sur.enabled ? c->sur.xbar + c->d.nwalkers : nullptr;
  PhaseScope ps(c, AFQMC_PHASE_FIELDS);
  HIP_TRY(c, launch_fields(a, c->stream));
  return AFQMC_OK;
}

int do_vhs(afqmc_ctx* c, int nw) {
  if (c->cplanes) {  // split-BF16: VHS[w] = i sqrt(dt) sum_x xshifted[w, x] L_x, as atoms
    int r = ensure_planes(c);
    if (r) return r;
    PhaseScope ps(c, AFQMC_PHASE_VHS);
    const int64_t ncols = vhs_atom_count(c->d.nbasis) * 256;
    HIP_TRY(c, launch_split_ghalf(c->xshifted, nw, c->d.nchol, c->xpad, c->xplanes, c->stream));
    Bf16GemmArgs b{};
    b.a = c->xplanes;
    b.a_plane = (int64_t)nw * c->xpad;
    b.lda = c->xpad;
    b.m = nw;
    b.b = c->cplanes;
    b.b_plane = ncols * c->xpad;
    b.ldb = c->xpad;
    b.nn = ncols;
    b.kchunks = c->xpad / 32;
    b.ksplit = 1;
    b.epilogue = BF16_EPI_VATOM;
    b.tile = c->tile_vhs;
    b.out16 = c->vatoms;
    b.atom_stride = ncols * 4;
    b.scale = sqrt(c->s.dt);
    b.nbasis = c->d.nbasis;
    HIP_TRY(c, launch_bf16_gemm(b, c->stream));
    return AFQMC_OK;
  }
  VhsArgs a;
  a.nwalkers = nw;
  a.nbasis = c->d.nbasis;
  a.nchol = c->d.nchol;
  a.xshifted = c->xshifted;
  a.chol_packed = c->s.chol_packed;
  const int P = c->d.nbasis * (c->d.nbasis + 1) / 2;
  a.tri_i = c->tri;
  a.tri_j = c->tri + P;
  a.sqrt_dt = sqrt(c->s.dt);
  a.vhs = c->vhs;
  PhaseScope ps(c, AFQMC_PHASE_VHS);
  HIP_TRY(c, launch_vhs(a, c->stream));
  return AFQMC_OK;
}

// C[w] (ldc, c_stride) = alpha * A * B[w] (ldb, b_stride), ncol columns.
int nn(afqmc_ctx* c, int nw, const double* A, bool acplx, int64_t a_stride, const double* B, int64_t b_stride,
       int64_t ldb, double* C, int64_t c_stride, int64_t ldc, int ncol, double alpha, double* accum,
       const double* accum_init) {
  BatchedNNArgs a;
  a.nwalkers = nw;
  a.nrows = c->d.nbasis;
  a.ncol = ncol;
  a.A = A;
  a.a_complex = acplx;
  a.a_stride = a_stride;
  a.B = B;
  a.b_stride = b_stride;
  a.ldb = ldb;
  a.C = C;
  a.c_stride = c_stride;
  a.ldc = ldc;
  a.alpha = alpha;
  a.accum = accum;
  a.accum_init = accum_init;
  HIP_TRY(c, launch_batched_nn(a, c->stream));
  return AFQMC_OK;
}

// Split-BF16 one body of spin s for nw walkers: out = src + D_s src, D_s =
// expH1_s - I, one GEMM with walker-column rows r = w n_s + c and the basis
// index as K.  src / out are addressed as [w][i][off + c] (walker stride, row
// length, column offset); aplanes are the transposed split planes of src
// (plane stride a_plane, A row w a_gstride + a_goff + c when a_gstride > 0,
// else r); tp, when set, receives the transposed planes of out (Taylor layout).
int one_body_bf16(afqmc_ctx* c, int nw, int s, const uint16_t* aplanes, int64_t a_plane, int a_gstride, int a_goff,
                  const double* src, int64_t s_ws, int s_ld, int s_off, double* out, int64_t o_ws, int o_ld,
                  int o_off, uint16_t* tp) {
  const int N = c->d.nbasis, n = s == 0 ? c->d.nalpha : c->d.nbeta;
  const int rpad = taylor_rows_padded(c->ncol);
  Bf16GemmArgs b{};
  b.a = aplanes;
  b.a_plane = a_plane;
  b.lda = c->kpad;
  b.m = (int64_t)nw * n;
  b.b = c->dplanes + s * 2 * (int64_t)N * c->kpad;
  b.b_plane = (int64_t)N * c->kpad;
  b.ldb = c->kpad;
  b.nn = N;
  b.kchunks = c->kpad / 32;
  b.ksplit = 1;
  b.epilogue = BF16_EPI_ADD;
  b.tile = c->tile_ob;
  b.cols_fast = 1;  // a few column blocks over the walkers' large A planes
  b.scale = 1.0;
  b.group = n;
  b.a_gstride = a_gstride;
  b.a_goff = a_goff;
  b.base = src;
  b.b_ws = s_ws;
  b.b_ld = s_ld;
  b.b_off = s_off;
  b.out = out;
  b.o_ws = o_ws;
  b.o_ld = o_ld;
  b.o_off = o_off;
  if (tp) {
    b.tp = tp;
    b.tp_plane = (int64_t)c->d.nwalkers * rpad * c->kpad;
    b.tp_rows = rpad;
    b.tp_off = o_off;
    b.tp_ld = c->kpad;
  }
  HIP_TRY(c, launch_bf16_gemm(b, c->stream));
  return AFQMC_OK;
}

// phi_s -> P (one body applied when with_expH1, plain copy otherwise).  In mixed
// precision the one body is P = phi + (expH1 - I) phi with the product in
// split-BF16, and the transposed planes of P go to tplanes[0] for the Taylor series.
int gather(afqmc_ctx* c, int nw, const double* phia, const double* phib, bool with_expH1) {
  PhaseScope ps(c, AFQMC_PHASE_ONE_BODY);
  const int N = c->d.nbasis, na = c->d.nalpha, nb = c->d.nbeta, nc = c->ncol;
  const int64_t NN = (int64_t)N * N;
  double* P = c->buf[0];
  c->ptp_ready = false;
  if (with_expH1 && c->dplanes) {
    int r = ensure_planes(c);
    if (r) return r;
    for (int s = 0; s < (two_spins(c) ? 2 : 1); ++s) {
      const int n = s == 0 ? na : nb;
      const double* phi = s == 0 ? phia : phib;
      HIP_TRY(c, launch_split_transpose(phi, nw, nw, N, n, n, c->kpad, c->oplanes, c->stream));
      r = one_body_bf16(c, nw, s, c->oplanes, (int64_t)nw * n * c->kpad, 0, 0, phi, (int64_t)N * n, n, 0, P,
                        (int64_t)N * nc, nc, s == 0 ? 0 : na, c->tplanes[0]);
      if (r) return r;
    }
    c->ptp_ready = true;
    return AFQMC_OK;
  }
  if (with_expH1) {
    int r = nn(c, nw, c->s.expH1, false, 0, phia, (int64_t)N * na, na, P, (int64_t)N * nc, nc, na, 1.0, nullptr,
               nullptr);
    if (r) return r;
    if (two_spins(c))
      return nn(c, nw, c->s.expH1 + NN, false, 0, phib, (int64_t)N * nb, nb, P + 2 * na, (int64_t)N * nc, nc, nb,
                1.0, nullptr, nullptr);
    return AFQMC_OK;
  }
  HIP_TRY(c, hipMemcpy2DAsync(P, 16 * nc, phia, 16 * na, 16 * na, (size_t)nw * N, hipMemcpyDeviceToDevice,
                              c->stream));
  if (two_spins(c))
    HIP_TRY(c, hipMemcpy2DAsync(P + 2 * na, 16 * nc, phib, 16 * nb, 16 * nb, (size_t)nw * N,
                                hipMemcpyDeviceToDevice, c->stream));
  return AFQMC_OK;
}

// src (concatenated) -> phi_s; srcT: transposed planes of src in the Taylor
// layout (mixed precision with the one body), or null.
int scatter(afqmc_ctx* c, int nw, const double* src, double* phia, double* phib, bool with_expH1,
            const uint16_t* srcT = nullptr) {
  PhaseScope ps(c, AFQMC_PHASE_ONE_BODY);
  const int N = c->d.nbasis, na = c->d.nalpha, nb = c->d.nbeta, nc = c->ncol;
  const int64_t NN = (int64_t)N * N;
  if (with_expH1 && c->dplanes) {
    int r = ensure_planes(c);
    if (r) return r;
    const int rpad = taylor_rows_padded(nc);
    if (!srcT) {
      HIP_TRY(c, launch_split_transpose(src, nw, (int)c->d.nwalkers, N, nc, rpad, c->kpad, c->tplanes[0], c->stream));
      srcT = c->tplanes[0];
    }
    for (int s = 0; s < (two_spins(c) ? 2 : 1); ++s) {
      const int n = s == 0 ? na : nb, off = s == 0 ? 0 : na;
      r = one_body_bf16(c, nw, s, srcT, (int64_t)c->d.nwalkers * rpad * c->kpad, rpad, off, src, (int64_t)N * nc,
                        nc, off, s == 0 ? phia : phib, (int64_t)N * n, n, 0, nullptr);
      if (r) return r;
    }
    return AFQMC_OK;
  }
  if (with_expH1) {
    int r = nn(c, nw, c->s.expH1, false, 0, src, (int64_t)N * nc, nc, phia, (int64_t)N * na, na, na, 1.0, nullptr,
               nullptr);
    if (r) return r;
    if (two_spins(c))
      return nn(c, nw, c->s.expH1 + NN, false, 0, src + 2 * na, (int64_t)N * nc, nc, phib, (int64_t)N * nb, nb, nb,
                1.0, nullptr, nullptr);
    return AFQMC_OK;
  }
  HIP_TRY(c, hipMemcpy2DAsync(phia, 16 * na, src, 16 * nc, 16 * na, (size_t)nw * N, hipMemcpyDeviceToDevice,
                              c->stream));
  if (two_spins(c))
    HIP_TRY(c, hipMemcpy2DAsync(phib, 16 * nb, src + 2 * na, 16 * nc, 16 * nb, (size_t)nw * N,
                                hipMemcpyDeviceToDevice, c->stream));
  return AFQMC_OK;
}

// Taylor series of exp(VHS) applied to P; returns the buffer holding the result
// and, in mixed precision, the transposed split planes of it (else null).
// Split-BF16 Taylor series: Q = P + sum_k out_k with out_k^T = out_{k-1}^T VHS / k;
// term k < n writes its planes to tplanes[k], the last term sums them into Q and
// writes the planes of Q^T to tplanes[0] (tplanes[1] when n = 1, where tplanes[0]
// is the last term's input).
int taylor_bf16(afqmc_ctx* c, int nw, const double** result, const uint16_t** resultT) {
  const int N = c->d.nbasis, nc = c->ncol, rpad = taylor_rows_padded(nc);
  double* P = c->buf[0];
  double* Q = c->buf[3];
  if (!c->ptp_ready)
    HIP_TRY(c, launch_split_transpose(P, nw, (int)c->d.nwalkers, N, nc, rpad, c->kpad, c->tplanes[0], c->stream));
  c->ptp_ready = false;
  for (int k = 1; k <= c->d.exp_nmax; ++k) {
    TaylorBf16Args t{};
    t.nwalkers = nw;
    t.wcap = (int)c->d.nwalkers;
    t.nbasis = N;
    t.ncol = nc;
    t.rpad = rpad;
    t.kpad = c->kpad;
    const int n = c->d.exp_nmax;
    t.tin = c->tplanes[k - 1];
    t.tout = k < n ? c->tplanes[k] : nullptr;
    t.vatoms = c->vatoms;
    t.zatom = c->zatom;
    t.natom_side = vhs_atom_side(N);
    t.atom_stride = vhs_atom_count(N) * 1024;
    t.init = P;
    t.q = Q;
    t.first = k == 1;
    t.last = k == c->d.exp_nmax;
    t.split = k <= TAYLOR_SPLIT_TERMS;
    t.alpha = 1.0 / k;
    t.qout = k == n ? c->tplanes[n == 1 ? 1 : 0] : nullptr;
    if (k == n) {
      t.nterms = n - 1;
      for (int j = 1; j < n; ++j) t.terms[j - 1] = c->tplanes[j];
    }
    t.variant = c->taylor_variant;
    HIP_TRY(c, launch_taylor_bf16(t, c->stream));
  }
  *result = Q;
  if (resultT) *resultT = c->tplanes[c->d.exp_nmax == 1 ? 1 : 0];
  return AFQMC_OK;
}

int taylor(afqmc_ctx* c, int nw, const double** result, const uint16_t** resultT = nullptr) {
  PhaseScope ps(c, AFQMC_PHASE_TAYLOR);
  if (resultT) *resultT = nullptr;
  if (c->vatoms && c->d.exp_nmax > 0) return taylor_bf16(c, nw, result, resultT);
  const int N = c->d.nbasis, nc = c->ncol;
  const int64_t NN = (int64_t)N * N, stride = (int64_t)N * nc;
  double* P = c->buf[0];
  double* Q = c->buf[3];
  if (c->d.exp_nmax <= 0) {
    *result = P;
    return AFQMC_OK;
  }
  const double* prev = P;
  for (int k = 1; k <= c->d.exp_nmax; ++k) {
    double* out = (k & 1) ? c->buf[1] : c->buf[2];
    int r = nn(c, nw, c->vhs, true, NN, prev, stride, nc, out, stride, nc, nc, 1.0 / k, Q, k == 1 ? P : nullptr);
    if (r) return r;
    prev = out;
  }
  *result = Q;
  return AFQMC_OK;
}

// Wq per spin from the half-rotated Cholesky vectors, once per static set.
int ensure_quad(afqmc_ctx* c) {
  if (c->quad_ready) return AFQMC_OK;
  const int N = c->d.nbasis;
  for (int s = 0; s < (two_spins(c) ? 2 : 1); ++s) {
    const int n = s == 0 ? c->d.nalpha : c->d.nbeta;
    const double* rchol = s == 0 ? c->s.rchola : c->s.rcholb;
    if (c->qwp[s]) HIP_TRY(c, launch_quad_w(rchol, c->d.nchol, n, N, c->kpad, nullptr, c->qwp[s], c->stream));
    else HIP_TRY(c, launch_quad_w(rchol, c->d.nchol, n, N, N, c->qw64[s], nullptr, c->stream));
  }
  c->quad_ready = true;
  return AFQMC_OK;
}

// FP8 copy of the split-BF16 W' per spin for the control variate's quadratic
// form (buffers allocated on first use, outside graph capture).
int ensure_quad8(afqmc_ctx* c) {
  int r = ensure_quad(c);
  if (r) return r;
  if (c->q8_ready) return AFQMC_OK;
  if (c->capturing) return refuse_in_capture(c, "the first FP8 control-variate evaluation");
  const int64_t W = c->d.nwalkers;
  if (!c->q8_scale && dmalloc((void**)&c->q8_scale, 4 * sizeof(float)) != hipSuccess)
    return fail(c, AFQMC_ERR_ALLOC, "FP8 scale allocation failed");
  int64_t ldmax = 0;
  for (int s = 0; s < 2; ++s) {
    if (!c->qwp[s]) continue;
    const int64_t ldn = (int64_t)(s == 0 ? c->d.nalpha : c->d.nbeta) * c->kpad;
    c->q8_ld[s] = ceil_div(ldn, 64) * 64;
    ldmax = std::max(ldmax, c->q8_ld[s]);
    if (!c->qw8[s] && dmalloc((void**)&c->qw8[s], (size_t)c->q8_ld[s] * c->q8_ld[s]) != hipSuccess)
      return fail(c, AFQMC_ERR_ALLOC, "FP8 quadratic-form allocation failed (%s)", c->describe_dims);
  }
  if (!c->g8 && dmalloc((void**)&c->g8, 2 * (size_t)W * ldmax) != hipSuccess)
    return fail(c, AFQMC_ERR_ALLOC, "FP8 plane allocation failed (%s)", c->describe_dims);
  if (!c->q8_rowinv && dmalloc((void**)&c->q8_rowinv, sizeof(double) * (size_t)W) != hipSuccess)
    return fail(c, AFQMC_ERR_ALLOC, "FP8 scale allocation failed");
  unsigned int* absmax = reinterpret_cast<unsigned int*>(c->q8_scale + 2);
  HIP_TRY(c, hipMemsetAsync(absmax, 0, 2 * sizeof(unsigned int), c->stream));
  for (int s = 0; s < 2; ++s) {
    if (!c->qwp[s]) continue;
    const int64_t ldn = (int64_t)(s == 0 ? c->d.nalpha : c->d.nbeta) * c->kpad;
    HIP_TRY(c, launch_quad_w8(c->qwp[s], ldn, c->q8_ld[s], c->qw8[s], c->q8_scale + s, absmax + s, c->stream));
  }
  c->q8_ready = true;
  return AFQMC_OK;
}

// FP8 quadratic form of nw walkers for one spin, d^T Wq d with d = G - gref:
// d in per-walker scaled FP8 planes, Y = d8 W8 over the FP8 column units, the
// dot products with d (FP64) in the GEMM's epilogue, rescaled per walker, their
// sum per walker into part.  The split is timed as AFQMC_PHASE_SPLIT.
int quad_fp8(afqmc_ctx* c, int nw, int spin, const double* g, const double* gref, double* part, int phase) {
  const int N = c->d.nbasis;
  const int n = spin == 0 ? c->d.nalpha : c->d.nbeta;
  const int64_t ld8 = c->q8_ld[spin];
  {
    PhaseScope ps(c, AFQMC_PHASE_SPLIT);
    HIP_TRY(c, launch_split_fp8_rows(g, gref, nw, n, N, c->kpad, ld8, c->g8, c->q8_scale + spin, c->q8_rowinv,
                                     c->stream));
  }
  PhaseScope ps(c, phase);
  const int64_t mb = ceil_div(nw, 128);
  if (mb < 1 || mb > (int64_t)c->qunit8_n[spin].size()) return fail(c, AFQMC_ERR_INVALID_ARG, "walker count");
  const int nunits = c->qunit8_n[spin][mb - 1];
  Bf16GemmArgs b{};
  // byte arrays in 16-bit units: the im plane starts nw ld8 bytes in (A plane 2)
  b.a = reinterpret_cast<const uint16_t*>(c->g8);
  b.a_plane = (int64_t)nw * ld8 / 4;
  b.lda = ld8 / 2;
  b.m = nw;
  b.b = reinterpret_cast<const uint16_t*>(c->qw8[spin]);
  b.b_plane = 0;
  b.ldb = ld8 / 2;
  b.nn = (int64_t)n * c->kpad;
  b.kchunks = (int)(ld8 / 64);
  b.ksplit = 1;
  b.single = 1;
  b.fp8 = 1;
  b.epilogue = BF16_EPI_QDOT;
  b.out = c->qy;
  b.scale = 1.0;
  b.units = c->qunits8[spin] + c->qunit8_off[spin][mb - 1];
  b.nunits = nunits;
  b.rgroup = c->quad_rgroup;
  b.dg = g;
  b.dref = gref;
  b.dstride = (int64_t)n * N;
  b.dn = N;
  b.dkp = c->kpad;
  b.row_scale = c->q8_rowinv;
  HIP_TRY(c, launch_bf16_gemm(b, c->stream));
  HIP_TRY(c, launch_quad_sum(c->qy, nunits * QUAD_SLOTS, nw, part, c->stream));
  return AFQMC_OK;
}

// Split-BF16 quadratic form of nw walkers for one spin, d^T Wq d with d = G - gref
// (gref may be null; walker_of selects the walkers): Y = d W' over the column
// units of the block-triangular planes, the per-tile dot products in the GEMM's
// epilogue, their sum per walker into part.  single: hi * hi only.  The split
// of d into planes is timed as AFQMC_PHASE_SPLIT, the rest as `phase`.
int quad_bf16(afqmc_ctx* c, int nw, int spin, const double* g, const int* walker_of, const double* gref,
              bool single, double* part, int phase) {
  const int N = c->d.nbasis;
  const int n = spin == 0 ? c->d.nalpha : c->d.nbeta;
  {
    PhaseScope ps(c, AFQMC_PHASE_SPLIT);
    HIP_TRY(c, launch_split_ghalf(g, (int64_t)nw * n, N, c->kpad, c->gplanes[spin], c->stream, walker_of, n, gref));
  }
  PhaseScope ps(c, phase);
  const int64_t ldn = (int64_t)n * c->kpad;
  const int64_t mb = ceil_div(nw, 128);
  if (mb < 1 || mb > (int64_t)c->qunit_n[spin].size()) return fail(c, AFQMC_ERR_INVALID_ARG, "walker count");
  const int nunits = c->qunit_n[spin][mb - 1];
  Bf16GemmArgs b{};
  b.a = c->gplanes[spin];
  b.a_plane = (int64_t)nw * ldn;
  b.lda = ldn;
  b.m = nw;
  b.b = c->qwp[spin];
  b.b_plane = ldn * ldn;
  b.ldb = ldn;
  b.nn = ldn;
  b.kchunks = (int)(ldn / 32);
  b.ksplit = 1;
  b.single = single ? 1 : 0;
  b.epilogue = BF16_EPI_QDOT;
  b.out = c->qy;
  b.scale = 1.0;
  b.units = c->qunits[spin] + c->qunit_off[spin][mb - 1];
  b.nunits = nunits;
  b.rgroup = c->quad_rgroup;
  b.dg = g;
  b.dref = gref;
  b.walker_of = walker_of;
  b.dstride = (int64_t)n * N;
  b.dn = N;
  b.dkp = c->kpad;
  HIP_TRY(c, launch_bf16_gemm(b, c->stream));
  HIP_TRY(c, launch_quad_sum(c->qy, nunits * QUAD_SLOTS, nw, part, c->stream));
  return AFQMC_OK;
}

// Exchange of nw walkers for one spin as g^T Wq g: Y = G Wq (FP64 NT GEMM, or
// split-BF16 GEMM with split K), then one dot product per walker.
int exchange_spin_quad(afqmc_ctx* c, int nw, int spin, const double* g, const int* walker_of) {
  const int N = c->d.nbasis;
  const int n = spin == 0 ? c->d.nalpha : c->d.nbeta;
  double* part = spin == 0 ? c->exx_a : c->exx_b;
  int r = ensure_quad(c);
  if (r) return r;
  if (c->qwp[spin]) {
    r = quad_bf16(c, nw, spin, g, walker_of, nullptr, false, part, AFQMC_PHASE_EXCHANGE);
    if (r) return r;
  } else {
    PhaseScope ps(c, AFQMC_PHASE_EXCHANGE);
    if (walker_of) {
      HIP_TRY(c, launch_sur_gather(g, walker_of, nw, (int64_t)n * N, c->gcomp, c->stream));
      g = c->gcomp;
    }
    const int64_t nN = (int64_t)n * N;
    ForceBiasArgs a;
    a.nwalkers = nw;
    a.nchol = (int)nN;
    a.ksplit = 1;
    a.nseg = 1;
    a.seg[0] = NtSegment{g, nN, c->qw64[spin], nN, (int)nN};
    a.scale = 1.0;
    a.vbias_part = c->qy;
    HIP_TRY(c, launch_force_bias(a, c->stream));
    HIP_TRY(c, launch_quad_reduce(c->qy, 1, nw, n, N, N, g, nullptr, part, c->stream));
  }
  c->xs_used[spin] = 1;
  return AFQMC_OK;
}

// Exchange partial sums of nw walkers for one spin.  With walker_of, the nw
// walkers are the sampled ones (walker_of[t], t < nw) of the Ghalf array g.
int exchange_spin(afqmc_ctx* c, int nw, int spin, const double* g, const int* walker_of) {
  const int N = c->d.nbasis;
  const int n = spin == 0 ? c->d.nalpha : c->d.nbeta;
  const double* rchol = spin == 0 ? c->s.rchola : c->s.rcholb;
  double* part = spin == 0 ? c->exx_a : c->exx_b;
  if (c->quad_on) return exchange_spin_quad(c, nw, spin, g, walker_of);
  if (c->d.precision == AFQMC_PREC_MIXED && c->bx_ok[spin]) {
    int r = ensure_planes(c);
    if (r) return r;
    PhaseScope ps(c, AFQMC_PHASE_EXCHANGE);
    HIP_TRY(c, launch_split_ghalf(g, (int64_t)nw * n, N, c->kpad, c->gplanes[spin], c->stream, walker_of, n));
    ExchangeBf16Args b{};
    b.split = 1;
    b.nwalkers = nw;
    b.nchol = c->d.nchol;
    b.nocc = n;
    b.kpad = c->kpad;
    b.wb = c->bx_wb[spin];
    b.xb = c->bx_xb[spin];
    const int64_t nwb = ceil_div(nw, b.wb), nxb = ceil_div(c->d.nchol, b.xb);
    int64_t xs = c->exx_split > 0 ? c->exx_split : ceil_div(ceil_div(2048, nwb), 8) * 8;
    if (xs > nxb) xs = nxb;
    if (xs < 1) xs = 1;
    b.xsplit = (int)xs;
    b.gplanes = c->gplanes[spin];
    b.lplanes = c->lplanes[spin];
    b.exx_part = part;
    if ((int64_t)b.xsplit * nw > c->exx_cap) return fail(c, AFQMC_ERR_SHAPE, "exchange partial buffer too small");
    HIP_TRY(c, launch_exchange_bf16(b, c->stream));
    c->xs_used[spin] = b.xsplit;
    return AFQMC_OK;
  }
  PhaseScope ps(c, AFQMC_PHASE_EXCHANGE);
  if (walker_of) {
    HIP_TRY(c, launch_sur_gather(g, walker_of, nw, (int64_t)n * N, c->gcomp, c->stream));
    g = c->gcomp;
  }
  ExchangeArgs a;
  a.nwalkers = nw;
  a.nchol = c->d.nchol;
  a.nbasis = N;
  a.nocc = n;
  a.ghalf = g;
  a.rchol = rchol;
  a.xsplit = exchange_xsplit(nw, c->d.nchol, n, c->exx_split);
  a.exx_part = part;
  if ((int64_t)a.xsplit * nw > c->exx_cap) return fail(c, AFQMC_ERR_SHAPE, "exchange partial buffer too small");
  HIP_TRY(c, launch_exchange(a, c->stream));
  c->xs_used[spin] = a.xsplit;
  return AFQMC_OK;
}

int do_exchange(afqmc_ctx* c, int nw, const double* ga, const double* gb, const int* walker_of = nullptr) {
  int r = exchange_spin(c, nw, 0, ga, walker_of);
  if (r) return r;
  if (two_spins(c)) return exchange_spin(c, nw, 1, gb, walker_of);
  return AFQMC_OK;
}

int do_energy(afqmc_ctx* c, int nw, const double* ga, const double* gb, double* energy,
              const double* exx_est = nullptr, double* e1_out = nullptr, const double* e1_in = nullptr) {
  EnergyArgs a;
  a.nwalkers = nw;
  a.nbasis = c->d.nbasis;
  a.nalpha = c->d.nalpha;
  a.nbeta = c->d.nbeta;
  a.rhf = c->d.rhf;
  a.ghalfa = ga;
  a.ghalfb = gb;
  a.rH1a = c->s.rH1a;
  a.rH1b = c->s.rH1b;
  a.ecoul = wv(c, WV_ECOUL);
  a.exx_part_a = c->exx_a;
  a.xs_a = c->xs_used[0];
  a.exx_part_b = two_spins(c) ? c->exx_b : nullptr;
  a.xs_b = c->xs_used[1];
  a.ecore = c->s.ecore;
  a.energy = energy;
  a.exx_est = exx_est;
  a.e1_out = e1_out;
  a.e1_in = e1_in;
  PhaseScope ps(c, AFQMC_PHASE_ENERGY);
  HIP_TRY(c, launch_energy(a, c->stream));
  return AFQMC_OK;
}

// Local energy of nw walkers from their Ghalf with the learned exchange control
// variate (requires the force bias + fields of this Green's function).
// Trial Ghalf, K = sum_x T0_x L_x and E_T = sum_x tr(T0_x T0_x) per spin, in
// FP64, once per trial / Hamiltonian (T0_x = L_x G_T^T).
int sur_prepare_cv(afqmc_ctx* c) {
  auto& S = c->sur;
  if (S.cv_ready) return AFQMC_OK;
  const int N = c->d.nbasis, X = c->d.nchol, nmax = std::max(c->d.nalpha, c->d.nbeta);
  const int64_t gstride = 2 * (int64_t)nmax * N;
  int r = do_gf(c, 1, c->s.psia, c->s.psib, S.gT, S.gT + gstride, S.gtmp, nullptr, true);
  if (r) return r;
  double* T0 = nullptr;
  double* part = nullptr;
  if (dmalloc((void**)&T0, 16 * (size_t)X * nmax * nmax) != hipSuccess ||
      dmalloc((void**)&part, 16 * (size_t)X) != hipSuccess) {
    if (T0) (void)hipFree(T0);
    return fail(c, AFQMC_ERR_ALLOC, "control-variate set-up allocation failed");
  }
  const int nspin = two_spins(c) ? 2 : 1;
  hipError_t e = hipSuccess;
  for (int s = 0; s < nspin && e == hipSuccess; ++s) {
    const int n = s == 0 ? c->d.nalpha : c->d.nbeta;
    e = launch_cv_static(s == 0 ? c->s.rchola : c->s.rcholb, S.gT + s * gstride, X, n, N, T0, S.K + s * gstride,
                         part, S.et + 2 * s, c->stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(T0);
  (void)hipFree(part);
  if (e != hipSuccess) return fail(c, AFQMC_ERR_HIP, "control-variate set-up: %s", hipGetErrorString(e));
  S.cv_ready = true;
  return AFQMC_OK;
}

// Q_s(w) = sum_x tr(dT_x dT_x), dT_x = L_x (G_s,w - G_T,s)^T, single BF16
// products, partial sums into the exchange partial buffer of spin s: as
// dG^T Wq dG when the split-BF16 Wq planes exist, else through every dT_x.
int exchange_q_spin(afqmc_ctx* c, int nw, int spin, const double* g) {
  const int N = c->d.nbasis;
  const int n = spin == 0 ? c->d.nalpha : c->d.nbeta;
  const int64_t gstride = 2 * (int64_t)std::max(c->d.nalpha, c->d.nbeta) * N;
  const double* gT = c->sur.gT + spin * gstride;
  double* part = spin == 0 ? c->exx_a : c->exx_b;
  if (c->qwp[spin]) {
    int r;
    if (c->sur.qmode == AFQMC_SURROGATE_Q_FP8) {
      r = ensure_quad8(c);
      if (r) return r;
      r = quad_fp8(c, nw, spin, g, gT, part, AFQMC_PHASE_EXCHANGE_CV);
    } else {
      r = ensure_quad(c);
      if (r) return r;
      r = quad_bf16(c, nw, spin, g, nullptr, gT, true, part, AFQMC_PHASE_EXCHANGE_CV);
    }
    if (r) return r;
    c->xs_used[spin] = 1;
    return AFQMC_OK;
  }
  PhaseScope ps(c, AFQMC_PHASE_EXCHANGE_CV);
  HIP_TRY(c, launch_split_ghalf(g, (int64_t)nw * n, N, c->kpad, c->gplanes[spin], c->stream, nullptr, n, gT));
  ExchangeBf16Args b{};
  b.split = 0;
  b.nwalkers = nw;
  b.nchol = c->d.nchol;
  b.nocc = n;
  b.kpad = c->kpad;
  b.wb = c->bx_wb[spin];
  b.xb = c->bx_xb[spin];
  const int64_t nwb = ceil_div(nw, b.wb), nxb = ceil_div(c->d.nchol, b.xb);
  int64_t xs = ceil_div(ceil_div(2048, nwb), 8) * 8;
  if (xs > nxb) xs = nxb;
  if (xs < 1) xs = 1;
  b.xsplit = (int)xs;
  b.gplanes = c->gplanes[spin];
  b.lplanes = c->lplanes[spin];
  b.exx_part = spin == 0 ? c->exx_a : c->exx_b;
  if ((int64_t)b.xsplit * nw > c->exx_cap) return fail(c, AFQMC_ERR_SHAPE, "exchange partial buffer too small");
  HIP_TRY(c, launch_exchange_bf16(b, c->stream));
  c->xs_used[spin] = b.xsplit;
  return AFQMC_OK;
}

int do_energy_surrogate(afqmc_ctx* c, int nw, const double* ga, const double* gb, const double* ovlp,
                        double* energy) {
  auto& S = c->sur;
  const bool warm = S.steps < S.warmup;
  if (c->capturing && (warm || S.steps == 0))
    return fail(c, AFQMC_ERR_INVALID_ARG, "the control variate's warm-up steps run eagerly, before a capture");
  // physical part of the predictor: E_T + 2 <K, dG> (+ Q(dG)), with the one-body
  // energy for the features from the same pass over Ghalf
  int r = sur_prepare_cv(c);
  if (r) return r;
  r = ensure_planes(c);
  if (r) return r;
  const int nspin = two_spins(c) ? 2 : 1;
  const int N = c->d.nbasis;
  const int64_t gstride = 2 * (int64_t)std::max(c->d.nalpha, c->d.nbeta) * N;
  for (int s = 0; s < nspin && S.qmode != AFQMC_SURROGATE_Q_NONE; ++s) {
    r = exchange_q_spin(c, nw, s, s == 0 ? ga : gb);
    if (r) return r;
  }
  SurrogateCvArgs cva{};
  cva.nwalkers = nw;
  cva.nbasis = N;
  cva.nspin = nspin;
  for (int s = 0; s < nspin; ++s) {
    cva.nocc[s] = s == 0 ? c->d.nalpha : c->d.nbeta;
    cva.g[s] = s == 0 ? ga : gb;
    cva.gT[s] = S.gT + s * gstride;
    cva.K[s] = S.K + s * gstride;
    cva.part[s] = s == 0 ? c->exx_a : c->exx_b;
    cva.xs[s] = S.qmode != AFQMC_SURROGATE_Q_NONE ? c->xs_used[s] : 0;  // partial sums of Q, when computed
  }
  cva.et = S.et;
  cva.cv = S.cv;
  cva.pred = S.pred;
  cva.mlp = S.mlp;
  cva.state = S.state;
  cva.fold = 0;
  cva.rH1[0] = c->s.rH1a;
  cva.rH1[1] = c->s.rH1b;
  cva.e1 = S.e1;
  cva.e1_factor = c->d.rhf ? 2.0 : 1.0;
  cva.ecore = c->s.ecore;
  {
    PhaseScope ps(c, AFQMC_PHASE_ENERGY);
    HIP_TRY(c, launch_sur_cv(cva, c->stream));
  }
  // the MLP term from the features, folded into the predictor
  SurrogateFeatureArgs fa{nw, S.e1, wv(c, WV_ECOUL), S.xbar, S.xbar + c->d.nwalkers, ovlp, S.feat};
  HIP_TRY(c, launch_sur_features(fa, c->stream));
  if (S.steps > 0) {
    SurrogateForwardArgs f{nw, S.feat, S.wbf, S.bias, S.state, S.pred};
    HIP_TRY(c, launch_sur_forward(f, c->stream));
  } else {
    HIP_TRY(c, hipMemsetAsync(S.pred, 0, 16 * (size_t)nw, c->stream));
  }
  HIP_TRY(c, launch_sur_fold(nw, S.cv, S.pred, S.mlp, S.state, c->stream));
  // exact exchange on the sampled subset (permutation drawn on the device)
  int k = nw;
  if (!warm) {
    k = (int)std::lround(S.fraction * nw);
    if (k < 1) k = 1;
    if (k > nw) k = nw;
  }
  S.last_k = k;
  HIP_TRY(c, launch_sur_plan(c->dstate, S.seed, nw, warm ? 1 : 0, c->stream));
  HIP_TRY(c, launch_sur_select(nw, k, c->dstate, S.walker_of, c->stream));
  r = do_exchange(c, k, ga, gb, S.walker_of);
  if (r) return r;
  SurrogateCombineArgs cb;
  cb.nwalkers = nw;
  cb.k = k;
  cb.perm = c->dstate;
  cb.pred = S.pred;
  cb.part_a = c->exx_a;
  cb.xs_a = c->xs_used[0];
  cb.part_b = two_spins(c) ? c->exx_b : nullptr;
  cb.xs_b = c->xs_used[1];
  cb.est = S.est;
  cb.target = S.target;
  cb.resid = S.resid;
  cb.cv = S.cv;
  HIP_TRY(c, launch_sur_combine(cb, c->stream));
  const SurrogateGainArgs gain{k, S.walker_of, S.target, S.mlp, S.state, 0.8, S.train_iters > 0 ? 1.0 : 0.0};
  HIP_TRY(c, launch_sur_gain(gain, c->stream));
  // train_iters = 0: the physical control variate alone (no MLP term, no training)
  const int iters = S.train_iters == 0 ? 0 : (warm ? S.train_iters : S.online_iters);
  if (warm) {
    SurrogateStatsArgs sa{k, S.feat, S.walker_of, S.target, S.state};
    HIP_TRY(c, launch_sur_stats(sa, c->stream));
  }
  for (int it = 0; it < iters; ++it) {
    SurrogateTrainArgs ta{k, S.feat, S.walker_of, S.target, S.master, S.state, S.grad, S.loss};
    HIP_TRY(c, launch_sur_grad(ta, c->stream));
    SurrogateAdamArgs ad{sur_grad_groups(k), it, c->dstate, S.lr, S.grad, S.master, S.m, S.v, S.wbf, S.bias};
    HIP_TRY(c, launch_sur_adam(ad, c->stream));
  }
  if (iters > 0) HIP_TRY(c, launch_step_advance(c->dstate, 0, iters, c->stream));
  S.steps += 1;
  if (c->capturing) c->capture_sur += 1;
  // the one-body energy of the control-variate pass (S.e1) is reused: Ghalf is not read again
  return do_energy(c, nw, ga, gb, energy, S.est, nullptr, S.e1);
}

int do_two_body_core(afqmc_ctx* c, int nw, const double* ga, const double* gb, const double* xi, uint64_t seed,
                     uint64_t counter, double* xi_out, double* vbias_out, bool energy_too, double* energy,
                     const double* ovlp = nullptr) {
  int r = do_force_bias(c, nw, ga, gb);
  if (r) return r;
  r = do_fields(c, nw, true, xi, seed, counter, xi_out, vbias_out);
  if (r) return r;
  if (c->capturing) {  // the next replayed step draws the next field stream
    HIP_TRY(c, launch_step_advance(c->dstate, 1, 0, c->stream));
    c->capture_steps += 1;
  }
  if (energy_too) {
    if (c->sur.enabled && ovlp) {
      r = do_energy_surrogate(c, nw, ga, gb, ovlp, energy);
    } else {
      r = do_exchange(c, nw, ga, gb);
      if (r) return r;
      r = do_energy(c, nw, ga, gb, energy);
    }
    if (r) return r;
  }
  return do_vhs(c, nw);
}

}  // namespace

extern "C" {

int afqmc_create(afqmc_ctx** out, int device, const afqmc_dims* dims) {
  if (!out || !dims) return AFQMC_ERR_INVALID_ARG;
  *out = nullptr;
  const afqmc_dims d = *dims;
  if (d.nbasis < 1 || d.nchol < 0 || d.nalpha < 1 || d.nbeta < 0 || d.nwalkers < 1 || d.exp_nmax < 0)
    return AFQMC_ERR_SHAPE;
  if (d.nalpha > 64 || d.nbeta > 64 || d.nalpha > d.nbasis || d.nbeta > d.nbasis) return AFQMC_ERR_SHAPE;
  if (d.precision != AFQMC_PREC_FP64 && d.precision != AFQMC_PREC_MIXED) return AFQMC_ERR_UNSUPPORTED;
  // the split-BF16 Taylor series keeps every term's planes until the last term
  if (d.precision == AFQMC_PREC_MIXED && d.exp_nmax > TAYLOR_MAX_TERMS) return AFQMC_ERR_SHAPE;
  if (hipSetDevice(device) != hipSuccess) return AFQMC_ERR_HIP;
  afqmc_ctx* c = new afqmc_ctx();
  c->d = d;
  c->device = device;
  const bool two = d.nbeta > 0 && !d.rhf;
  c->ncol = two ? d.nalpha + d.nbeta : d.nalpha;
  const int64_t W = d.nwalkers, N = d.nbasis, X = d.nchol > 0 ? d.nchol : 1;
  // force-bias split-K: enough workgroups to fill the chip, >= 64 K chunks each
  {
    const int64_t tiles = ceil_div(W, 64) * ceil_div(X, 128);
    const int64_t kch = ceil_div((int64_t)d.nalpha * N, 16) + (two ? ceil_div((int64_t)d.nbeta * N, 16) : 0);
    int64_t ks = ceil_div(1024, tiles);
    if (ks > kch / 64) ks = kch / 64;
    if (ks > 8) ks = 8;
    if (ks < 1) ks = 1;
    c->fb_ksplit = (int)ks;
  }
  c->xs_a = exchange_xsplit(d.nwalkers, (int)X, d.nalpha);
  c->xs_b = two ? exchange_xsplit(d.nwalkers, (int)X, d.nbeta) : 1;
  snprintf(c->describe_dims, sizeof(c->describe_dims), "N=%d X=%d na=%d nb=%d W=%d", d.nbasis, d.nchol, d.nalpha,
           d.nbeta, d.nwalkers);
  c->kpad = (d.nbasis + 31) / 32 * 32;
  // split-BF16 exchange geometry (mixed precision, and the surrogate's control variate)
  for (int s = 0; s < (two ? 2 : 1); ++s) {
    const int n = s == 0 ? d.nalpha : d.nbeta;
    c->bx_ok[s] = bf16_exchange_plan(n, &c->bx_wb[s], &c->bx_xb[s]);
  }
  {
    // exchange as a quadratic form: cheaper than forming every T_x when N < X;
    // Wq takes (n N)^2 values per spin (FP64, or two BF16 planes of the padded size)
    const int nmx = d.nalpha > d.nbeta ? d.nalpha : d.nbeta;
    const double per_spin = d.precision == AFQMC_PREC_MIXED ? 4.0 * std::pow((double)nmx * c->kpad, 2)
                                                            : 8.0 * std::pow((double)nmx * d.nbasis, 2);
    c->quad_on = d.nchol > d.nbasis && per_spin <= QUAD_BUDGET_BYTES;
  }
  if (d.precision == AFQMC_PREC_MIXED) {
    c->xpad = (d.nchol + 31) / 32 * 32;
    c->bf16_fb = d.nchol > 0 && c->bx_ok[0] && (!two || c->bx_ok[1]);
    if (c->bf16_fb) c->fb_ksplit = c->fb_split_spin * (two ? 2 : 1);
  }
  if (dmalloc((void**)&c->flags, sizeof(int)) != hipSuccess ||
      dmalloc((void**)&c->wflags, sizeof(int) * (size_t)d.nwalkers) != hipSuccess ||
      dmalloc((void**)&c->nfb, sizeof(unsigned long long)) != hipSuccess ||
      dmalloc((void**)&c->dstate, sizeof(StepState)) != hipSuccess) {
    afqmc_destroy(c);
    return AFQMC_ERR_ALLOC;
  }
  if (hipMemset(c->flags, 0, sizeof(int)) != hipSuccess ||
      hipMemset(c->wflags, 0, sizeof(int) * (size_t)d.nwalkers) != hipSuccess ||
      hipMemset(c->nfb, 0, sizeof(unsigned long long)) != hipSuccess ||
      hipMemset(c->dstate, 0, sizeof(StepState)) != hipSuccess) {
    afqmc_destroy(c);
    return AFQMC_ERR_HIP;
  }
  *out = c;
  return AFQMC_OK;
}

void afqmc_destroy(afqmc_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  else (void)hipDeviceSynchronize();
  void* ptrs[] = {c->vbias_part, c->xshifted, c->wvec, c->exx_a, c->exx_b, c->vhs, c->buf[0], c->buf[1],
                  c->buf[2], c->buf[3], c->tri, c->flags, c->wflags, c->psi_flag, c->nfb, c->gplanes[0], c->gplanes[1], c->lplanes[0],
                  c->lplanes[1], c->xplanes, c->cplanes, c->vatoms, c->zatom, c->dplanes, c->oplanes, c->gcomp, c->gf_o, c->gf_oinv, c->qw64[0], c->qw64[1], c->qwp[0],
                  c->qwp[1], c->qy, c->qunits[0], c->qunits[1], c->sur.feat, c->sur.pred, c->sur.est, c->sur.target,
                  c->sur.resid, c->sur.walker_of, c->sur.xbar, c->sur.e1, c->sur.master, c->sur.m,
                  c->sur.v, c->sur.wbf, c->sur.bias, c->sur.grad, c->sur.loss, c->sur.state, c->sur.gT,
                  c->sur.K, c->sur.et, c->sur.cv, c->sur.mlp, c->sur.gtmp, c->dstate};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (void* p : c->tplanes)
    if (p) (void)hipFree(p);
  void* p8[] = {c->qw8[0], c->qw8[1], c->q8_scale, c->g8, c->q8_rowinv, c->qunits8[0], c->qunits8[1]};
  for (void* p : p8)
    if (p) (void)hipFree(p);
  for (void* p : c->retired) (void)hipFree(p);
  if (c->graph) (void)hipGraphExecDestroy(c->graph);
  for (hipEvent_t e : c->ev_free) (void)hipEventDestroy(e);
  for (const PendingEvent& pe : c->ev_pending) {
    (void)hipEventDestroy(pe.start);
    (void)hipEventDestroy(pe.stop);
  }
  delete c;
}

int afqmc_set_stream(afqmc_ctx* c, void* stream) {
  if (!c) return AFQMC_ERR_INVALID_ARG;
  c->stream = (hipStream_t)stream;
  return AFQMC_OK;
}

int afqmc_set_static(afqmc_ctx* c, const afqmc_static* s) {
  if (!c || !s) return AFQMC_ERR_INVALID_ARG;
  if (c->capturing) return refuse_in_capture(c, "afqmc_set_static");
  const bool two = c->d.nbeta > 0 && !c->d.rhf;
  // The trial orbitals are always required; the Hamiltonian tensors only by the
  // entry points that use them (checked there).
  if (!s->psia) return fail(c, AFQMC_ERR_INVALID_ARG, "null psia");
  if (c->d.nbeta > 0 && !s->psib && two) return fail(c, AFQMC_ERR_INVALID_ARG, "null psib");
  c->s = *s;
  c->have_static = true;
  // real trial orbitals (the usual RHF / UHF trial of a real Hamiltonian): the
  // overlap matrices skip the products with their zero imaginary parts
  if (!c->psi_flag && dmalloc((void**)&c->psi_flag, sizeof(int)) != hipSuccess)
    return fail(c, AFQMC_ERR_ALLOC, "flag allocation failed");
  int any = 0;
  HIP_TRY(c, hipMemsetAsync(c->psi_flag, 0, sizeof(int), c->stream));
  HIP_TRY(c, launch_imag_any(s->psia, (int64_t)c->d.nbasis * c->d.nalpha, c->psi_flag, c->stream));
  if (two && c->d.nbeta > 0)
    HIP_TRY(c, launch_imag_any(s->psib, (int64_t)c->d.nbasis * c->d.nbeta, c->psi_flag, c->stream));
  HIP_TRY(c, hipMemcpyAsync(&any, c->psi_flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  c->psi_real = any == 0;
  c->planes_ready = false;
  c->quad_ready = false;
  c->q8_ready = false;
  c->sur.cv_ready = false;
  return AFQMC_OK;
}

const char* afqmc_last_error(const afqmc_ctx* c) { return c ? c->err : "null context"; }

int afqmc_exchange_form(const afqmc_ctx* c) { return (c && c->quad_on) ? AFQMC_EXCHANGE_QUAD : AFQMC_EXCHANGE_T; }

int afqmc_sync(afqmc_ctx* c) {
  if (!c) return AFQMC_ERR_INVALID_ARG;
  if (c->capturing) return refuse_in_capture(c, "afqmc_sync");
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return AFQMC_OK;
}

int afqmc_get_flags(afqmc_ctx* c, int32_t* flags, int clear) {
  if (!c || !flags) return AFQMC_ERR_INVALID_ARG;
  if (c->capturing) return refuse_in_capture(c, "afqmc_get_flags");
  int h = 0;
  HIP_TRY(c, hipMemcpyAsync(&h, c->flags, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  *flags = h;
  if (clear) HIP_TRY(c, hipMemsetAsync(c->flags, 0, sizeof(int), c->stream));
  return AFQMC_OK;
}

int afqmc_walker_flags(afqmc_ctx* c, int32_t nw, int32_t* flags, int clear) {
  if (!c) return AFQMC_ERR_INVALID_ARG;
  if (nw < 0 || nw > c->d.nwalkers)
    return fail(c, AFQMC_ERR_SHAPE, "nwalkers=%d outside [0, %d]", nw, c->d.nwalkers);
  if (!flags && !clear) return fail(c, AFQMC_ERR_INVALID_ARG, "null flags array and nothing to clear");
  HIP_TRY(c, launch_take_flags(c->wflags, flags, nw, clear ? 1 : 0, c->stream));
  return AFQMC_OK;
}

int afqmc_get_nfb(afqmc_ctx* c, uint64_t* nfb) {
  if (!c || !nfb) return AFQMC_ERR_INVALID_ARG;
  if (c->capturing) return refuse_in_capture(c, "afqmc_get_nfb");
  unsigned long long h = 0;
  HIP_TRY(c, hipMemcpyAsync(&h, c->nfb, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  *nfb = h;
  return AFQMC_OK;
}

static_assert(AFQMC_POP_PAIR_BRANCH == POP_PAIR_BRANCH && AFQMC_POP_COMB == POP_COMB &&
                  AFQMC_POP_RECONFIGURATION == POP_RECONFIGURATION && AFQMC_POP_RESCALE == POP_RESCALE,
              "population control method codes");

int afqmc_pop_plan(afqmc_ctx* c, int32_t method, int32_t nw, const double* weight, double target_weight,
                   double min_weight, double max_weight, const double* uniforms, int32_t* src, double* new_weight,
                   double* total_weight, int32_t* npairs) {
  if (!c) return AFQMC_ERR_INVALID_ARG;
  if (nw < 1 || nw > POP_MAX) return fail(c, AFQMC_ERR_SHAPE, "population control of %d walkers (1..%d)", nw, POP_MAX);
  if (method < AFQMC_POP_PAIR_BRANCH || method > AFQMC_POP_RESCALE)
    return fail(c, AFQMC_ERR_INVALID_ARG, "unknown population control method %d", method);
  if (!weight || !src || !new_weight || !total_weight || !npairs || (!uniforms && method != AFQMC_POP_RESCALE))
    return fail(c, AFQMC_ERR_INVALID_ARG, "null array");
  if (!(target_weight > 0.0)) return fail(c, AFQMC_ERR_INVALID_ARG, "target weight must be positive");
  PopPlanArgs a;
  a.nwalkers = nw;
  a.method = method;
  a.weight = weight;
  a.target = target_weight;
  a.min_weight = min_weight;
  a.max_weight = max_weight;
  a.uniforms = uniforms;
  a.src = src;
  a.new_weight = new_weight;
  a.total = total_weight;
  a.npairs = npairs;
  HIP_TRY(c, launch_pop_plan(a, c->stream));
  return AFQMC_OK;
}

// ---------------------------------------------------------------- set-up producers
namespace {

// split-K slices for >= ~1024 workgroups, K >= 256 per slice, planes <= 256 MB
int setup_splits(int64_t m, int64_t n, int64_t k) {
  const int64_t tiles = ceil_div(m, 64) * ceil_div(n, 64);
  int64_t nz = ceil_div(1024, tiles);
  nz = std::min<int64_t>(nz, std::max<int64_t>(k / 256, 1));
  nz = std::min<int64_t>(nz, std::max<int64_t>((int64_t)(256ll << 20) / (8 * m * n), 1));
  return (int)std::max<int64_t>(std::min<int64_t>(nz, 4096), 1);
}

// one set-up product on the context's stream; split-K planes are allocated
// for the call and freed after it
int setup_gemm(afqmc_ctx* c, SetupGemmArgs g) {
  const int nz = setup_splits(g.m, g.n, g.k);
  const int64_t pd = setup_gemm_part_doubles(g.m, g.n, nz);
  double* part = nullptr;
  if (pd > 0 && dmalloc((void**)&part, 8 * (size_t)pd) != hipSuccess)
    return fail(c, AFQMC_ERR_ALLOC, "set-up product: %lld split-K values", (long long)pd);
  g.part = part;
  hipError_t e = launch_setup_gemm(g, nz, c->stream);
  if (part) {
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(part);
  }
  if (e != hipSuccess) return fail(c, AFQMC_ERR_HIP, "set-up product: %s", hipGetErrorString(e));
  return AFQMC_OK;
}

SetupGemmArgs setup_args(int64_t m, int64_t n, int64_t k) {
  SetupGemmArgs g{};
  g.m = m;
  g.n = n;
  g.k = k;
  g.alpha = 1.0;
  g.beta = 0.0;
  g.store = SETUP_STORE_STRIDED;
  return g;
}

int setup_ready(afqmc_ctx* c) {
  if (!c) return AFQMC_ERR_INVALID_ARG;
  if (c->capturing) return refuse_in_capture(c, "a set-up producer");
  if (c->d.nchol < 1) return fail(c, AFQMC_ERR_SHAPE, "the set-up producers need nchol >= 1");
  c->err[0] = 0;
  return AFQMC_OK;
}

}  // namespace

int afqmc_setup_pack(afqmc_ctx* c, const double* chol, double* chol_packed) {
  int r = setup_ready(c);
  if (r) return r;
  if (!chol || !chol_packed) return fail(c, AFQMC_ERR_INVALID_ARG, "null array");
  HIP_TRY(c, launch_setup_pack(c->d.nbasis, c->d.nchol, chol, chol_packed, c->stream));
  return AFQMC_OK;
}

int afqmc_setup_h1e_mod(afqmc_ctx* c, const double* chol, const double* h1e, double* h1e_mod) {
  int r = setup_ready(c);
  if (r) return r;
  if (!chol || !h1e || !h1e_mod) return fail(c, AFQMC_ERR_INVALID_ARG, "null array");
  const int64_t N = c->d.nbasis, X = c->d.nchol, NN = N * N;
  double* v = nullptr;
  if (dmalloc((void**)&v, 8 * (size_t)NN) != hipSuccess) return fail(c, AFQMC_ERR_ALLOC, "set-up scratch");
  // v = -1/2 Lv Lv^T with Lv = chol viewed [N, N X]
  SetupGemmArgs g = setup_args(N, N, N * X);
  g.a = chol;
  g.sar = N * X;
  g.sac = 1;
  g.b = chol;
  g.sbr = 1;
  g.sbc = N * X;
  g.c = v;
  g.scr = N;
  g.scc = 1;
  g.alpha = -0.5;
  r = setup_gemm(c, g);
  hipError_t e = hipSuccess;
  for (int s = 0; s < 2 && r == AFQMC_OK && e == hipSuccess; ++s)
    e = launch_setup_scale_shift(NN, h1e + s * NN, v, 1.0, h1e_mod + s * NN, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(v);
  if (r) return r;
  if (e != hipSuccess) return fail(c, AFQMC_ERR_HIP, "h1e_mod: %s", hipGetErrorString(e));
  return AFQMC_OK;
}

int afqmc_setup_half_rotate(afqmc_ctx* c, int32_t spin, const double* psi, int32_t ld_psi, const double* chol,
                            const double* h1, double* rchol, double* rh1) {
  int r = setup_ready(c);
  if (r) return r;
  if (spin != 0 && spin != 1) return fail(c, AFQMC_ERR_INVALID_ARG, "spin must be 0 or 1");
  const int64_t N = c->d.nbasis, X = c->d.nchol;
  const int64_t n = spin == 0 ? c->d.nalpha : c->d.nbeta;
  if (n < 1) return AFQMC_OK;
  if (ld_psi < n) return fail(c, AFQMC_ERR_SHAPE, "ld_psi=%d < %lld occupied orbitals", ld_psi, (long long)n);
  if (!psi || !chol || !h1 || !rchol || !rh1) return fail(c, AFQMC_ERR_INVALID_ARG, "null array");
  // rchol[x, (i, m)] = sum_k psi[k, i] L[k, m, x]: C [n, N X] = psi^T Lv, stored as [X, n N]
  SetupGemmArgs g = setup_args(n, N * X, N);
  g.a = psi;
  g.sar = 1;
  g.sac = ld_psi;
  g.b = chol;
  g.sbr = N * X;
  g.sbc = 1;
  g.c = rchol;
  g.store = SETUP_STORE_HALF_ROTATED;
  g.hr_x = X;
  g.hr_nN = n * N;
  g.hr_N = N;
  r = setup_gemm(c, g);
  if (r) return r;
  // rH1[i, q] = sum_p psi[p, i] H1[p, q], complex with a zero imaginary part
  HIP_TRY(c, launch_setup_fill(rh1, 2 * n * N, 0.0, c->stream));
  SetupGemmArgs h = setup_args(n, N, N);
  h.a = psi;
  h.sar = 1;
  h.sac = ld_psi;
  h.b = h1;
  h.sbr = N;
  h.sbc = 1;
  h.c = rh1;
  h.scr = 2 * N;
  h.scc = 2;
  return setup_gemm(c, h);
}

int afqmc_setup_mf_shift(afqmc_ctx* c, const double* chol, const double* gcharge, double* mf) {
  int r = setup_ready(c);
  if (r) return r;
  if (!chol || !gcharge || !mf) return fail(c, AFQMC_ERR_INVALID_ARG, "null array");
  const int64_t N = c->d.nbasis, X = c->d.nchol;
  // mf = i L^T Re(g) - L^T Im(g): imaginary part from Re(g), real part from -Im(g)
  for (int part = 0; part < 2; ++part) {
    SetupGemmArgs g = setup_args(X, 1, N * N);
    g.a = chol;
    g.sar = 1;
    g.sac = X;
    g.b = gcharge + (part == 0 ? 0 : 1);
    g.sbr = 2;
    g.sbc = 1;
    g.c = mf + (part == 0 ? 1 : 0);
    g.scr = 2;
    g.scc = 1;
    g.alpha = part == 0 ? 1.0 : -1.0;
    r = setup_gemm(c, g);
    if (r) return r;
  }
  return AFQMC_OK;
}

int afqmc_setup_expH1(afqmc_ctx* c, const double* chol, const double* h1e_mod, const double* mf, double dt,
                      double* expH1) {
  int r = setup_ready(c);
  if (r) return r;
  if (!chol || !h1e_mod || !mf || !expH1) return fail(c, AFQMC_ERR_INVALID_ARG, "null array");
  if (!(dt > 0.0)) return fail(c, AFQMC_ERR_INVALID_ARG, "dt must be positive");
  const int N = c->d.nbasis;
  const int64_t X = c->d.nchol, NN = (int64_t)N * N;
  double* buf = nullptr;  // v, A, P, T, norm
  if (dmalloc((void**)&buf, 8 * (size_t)(4 * NN + 1)) != hipSuccess)
    return fail(c, AFQMC_ERR_ALLOC, "set-up scratch");
  double *v = buf, *A = buf + NN, *P = buf + 2 * NN, *T = buf + 3 * NN, *nrm = buf + 4 * NN;
  auto done = [&](int code) {
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(buf);
    return code;
  };
  auto hip_fail = [&](hipError_t e) { return done(fail(c, AFQMC_ERR_HIP, "expH1: %s", hipGetErrorString(e))); };
  // v = sum_x L_x Im(mf_x): H1' = h1e_mod - i sum_x L_x mf_x = h1e_mod + v for a purely imaginary mf
  SetupGemmArgs g = setup_args(NN, 1, X);
  g.a = chol;
  g.sar = X;
  g.sac = 1;
  g.b = mf + 1;
  g.sbr = 2;
  g.sbc = 1;
  g.c = v;
  g.scr = 1;
  g.scc = 1;
  r = setup_gemm(c, g);
  if (r) return done(r);
  constexpr int DEG = 18;  // Taylor degree on ||A||_1 <= 1/2: remainder below 1e-22
  for (int s = 0; s < 2; ++s) {
    hipError_t e = launch_setup_axpby(NN, -0.5 * dt, h1e_mod + s * NN, -0.5 * dt, v, A, c->stream);
    if (e == hipSuccess) e = launch_setup_norm1(N, A, nrm, c->stream);
    double hn = 0.0;
    if (e == hipSuccess) e = hipMemcpyAsync(&hn, nrm, sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(e);
    if (!std::isfinite(hn)) return done(fail(c, AFQMC_ERR_INVALID_ARG, "expH1: non-finite one-body operator"));
    int sq = 0;
    while (hn > 0.5 && sq < 64) {
      hn *= 0.5;
      ++sq;
    }
    const double sc = std::ldexp(1.0, -sq);
    // A <- A / 2^sq; P = I + A / DEG
    e = launch_setup_axpby(NN, sc, A, 0.0, A, A, c->stream);
    if (e == hipSuccess) e = launch_setup_axpby(NN, 1.0 / DEG, A, 0.0, A, P, c->stream);
    if (e == hipSuccess) e = launch_setup_add_identity(N, P, c->stream);
    if (e != hipSuccess) return hip_fail(e);
    // Horner: P <- I + A P / k for k = DEG - 1 .. 1
    for (int k = DEG - 1; k >= 1; --k) {
      SetupGemmArgs h = setup_args(N, N, N);
      h.a = A;
      h.sar = N;
      h.sac = 1;
      h.b = P;
      h.sbr = N;
      h.sbc = 1;
      h.c = T;
      h.scr = N;
      h.scc = 1;
      h.alpha = 1.0 / k;
      r = setup_gemm(c, h);
      if (r) return done(r);
      e = launch_setup_add_identity(N, T, c->stream);
      if (e != hipSuccess) return hip_fail(e);
      std::swap(P, T);
    }
    // squaring: exp(A 2^sq) = exp(A)^(2^sq)
    for (int q = 0; q < sq; ++q) {
      SetupGemmArgs h = setup_args(N, N, N);
      h.a = P;
      h.sar = N;
      h.sac = 1;
      h.b = P;
      h.sbr = N;
      h.sbc = 1;
      h.c = T;
      h.scr = N;
      h.scc = 1;
      r = setup_gemm(c, h);
      if (r) return done(r);
      std::swap(P, T);
    }
    e = hipMemcpyAsync(expH1 + s * NN, P, 8 * (size_t)NN, hipMemcpyDeviceToDevice, c->stream);
    if (e != hipSuccess) return hip_fail(e);
  }
  return done(AFQMC_OK);
}

int afqmc_walker_sums(afqmc_ctx* c, int32_t nw, const double* weight, const double* unscaled_weight,
                      const double* values, int64_t value_stride, double* acc) {
  if (!c) return AFQMC_ERR_INVALID_ARG;
  if (nw < 1) return fail(c, AFQMC_ERR_SHAPE, "nwalkers=%d", nw);
  if (!weight || !acc) return fail(c, AFQMC_ERR_INVALID_ARG, "null array");
  if (value_stride < 1) return fail(c, AFQMC_ERR_INVALID_ARG, "value stride must be >= 1");
  HIP_TRY(c, launch_walker_sums(nw, weight, unscaled_weight, values, value_stride, acc, c->stream));
  return AFQMC_OK;
}

int afqmc_clip_weights(afqmc_ctx* c, int32_t nw, double* weight, double bound) {
  if (!c) return AFQMC_ERR_INVALID_ARG;
  if (nw < 1) return fail(c, AFQMC_ERR_SHAPE, "nwalkers=%d", nw);
  if (!weight) return fail(c, AFQMC_ERR_INVALID_ARG, "null array");
  HIP_TRY(c, launch_clip_weights(nw, weight, bound, c->stream));
  return AFQMC_OK;
}

int afqmc_pop_gather(afqmc_ctx* c, int32_t nw, const int32_t* src, const void* in, void* out,
                     int64_t bytes_per_walker) {
  if (!c) return AFQMC_ERR_INVALID_ARG;
  if (nw < 1 || nw > 65535 || bytes_per_walker < 1)
    return fail(c, AFQMC_ERR_SHAPE, "%d walker rows of %lld bytes (1..65535 rows)", nw, (long long)bytes_per_walker);
  if (!src || !in || !out) return fail(c, AFQMC_ERR_INVALID_ARG, "null array");
  HIP_TRY(c, launch_pop_gather(nw, src, in, out, bytes_per_walker, c->stream));
  return AFQMC_OK;
}

int afqmc_greens_function(afqmc_ctx* c, int32_t nw, const double* phia, const double* phib, double* ga,
                          double* gb, double* ovlp, const double* log_shift) {
  int r = check_ready(c, nw);
  if (r) return r;
  if (!phia || !ga || !ovlp || (two_spins(c) && (!phib || !gb)))
    return fail(c, AFQMC_ERR_INVALID_ARG, "null walker array");
  return do_gf(c, nw, phia, phib, ga, gb, ovlp, log_shift, true);
}

int afqmc_overlap(afqmc_ctx* c, int32_t nw, const double* phia, const double* phib, double* ovlp,
                  const double* log_shift) {
  int r = check_ready(c, nw);
  if (r) return r;
  if (!phia || !ovlp || (two_spins(c) && !phib)) return fail(c, AFQMC_ERR_INVALID_ARG, "null walker array");
  return do_gf(c, nw, phia, phib, nullptr, nullptr, ovlp, log_shift, false);
}

int afqmc_force_bias(afqmc_ctx* c, int32_t nw, const double* ga, const double* gb, double* vbias) {
  int r = check_ready(c, nw);
  if (r) return r;
  r = need_static(c);
  if (r) return r;
  if (!ga || !vbias || (two_spins(c) && !gb)) return fail(c, AFQMC_ERR_INVALID_ARG, "null array");
  r = do_force_bias(c, nw, ga, gb);
  if (r) return r;
  return do_fields(c, nw, false, nullptr, 0, 0, nullptr, vbias);
}

int afqmc_one_body(afqmc_ctx* c, int32_t nw, double* phia, double* phib) {
  int r = check_ready(c, nw);
  if (r) return r;
  r = need_static(c);
  if (r) return r;
  if (!phia || (two_spins(c) && !phib)) return fail(c, AFQMC_ERR_INVALID_ARG, "null walker array");
  r = gather(c, nw, phia, phib, true);
  if (r) return r;
  return scatter(c, nw, c->buf[0], phia, phib, false);
}

int afqmc_two_body(afqmc_ctx* c, int32_t nw, double* phia, double* phib, const double* ga, const double* gb,
                   const double* xi, uint64_t seed, uint64_t counter, double* cmf, double* cfb, double* xi_out,
                   double* vbias_out) {
  int r = check_ready(c, nw);
  if (r) return r;
  r = need_static(c);
  if (r) return r;
  if (!phia || !ga || !cmf || !cfb || (two_spins(c) && (!phib || !gb)))
    return fail(c, AFQMC_ERR_INVALID_ARG, "null array");
  r = do_two_body_core(c, nw, ga, gb, xi, seed, counter, xi_out, vbias_out, false, nullptr);
  if (r) return r;
  HIP_TRY(c, hipMemcpyAsync(cmf, wv(c, WV_CMF), 16 * (size_t)nw, hipMemcpyDeviceToDevice, c->stream));
  HIP_TRY(c, hipMemcpyAsync(cfb, wv(c, WV_CFB), 16 * (size_t)nw, hipMemcpyDeviceToDevice, c->stream));
  r = gather(c, nw, phia, phib, false);
  if (r) return r;
  const double* res = nullptr;
  r = taylor(c, nw, &res);
  if (r) return r;
  return scatter(c, nw, res, phia, phib, false);
}

int afqmc_update_weight(afqmc_ctx* c, int32_t nw, const double* ovlp, const double* ovlp_new, const double* cfb,
                        const double* cmf, double eshift, double* weight, double* hybrid_energy,
                        double* walker_ovlp) {
  int r = check_ready(c, nw);
  if (r) return r;
  if (!(c->s.dt > 0.0)) return fail(c, AFQMC_ERR_INVALID_ARG, "dt must be positive");
  if (!ovlp || !ovlp_new || !cfb || !cmf || !weight || !hybrid_energy || !walker_ovlp)
    return fail(c, AFQMC_ERR_INVALID_ARG, "null array");
  WeightArgs a;
  a.nwalkers = nw;
  a.ovlp_old = ovlp;
  a.ovlp_new = ovlp_new;
  a.cfb = cfb;
  a.cmf = cmf;
  a.dt = c->s.dt;
  a.eshift = eshift;
  a.state = c->capturing ? c->dstate : nullptr;
  a.ebound = sqrt(2.0 / c->s.dt);
  a.weight = weight;
  a.hybrid_energy = hybrid_energy;
  a.walker_ovlp = walker_ovlp;
  a.flags = c->flags;
  a.wflags = c->wflags;
  PhaseScope ps(c, AFQMC_PHASE_WEIGHT);
  HIP_TRY(c, launch_update_weight(a, c->stream));
  return AFQMC_OK;
}

int afqmc_propagate(afqmc_ctx* c, int32_t nw, double* phia, double* phib, double* ga, double* gb,
                    const double* xi, uint64_t seed, uint64_t counter, double eshift, double* weight,
                    double* ovlp, double* hybrid_energy, const double* log_shift, double* energy_prev) {
  int r = check_ready(c, nw);
  if (r) return r;
  r = need_static(c);
  if (r) return r;
  if (!phia || !ga || !weight || !ovlp || !hybrid_energy || (two_spins(c) && (!phib || !gb)))
    return fail(c, AFQMC_ERR_INVALID_ARG, "null array");
  r = do_gf(c, nw, phia, phib, ga, gb, wv(c, WV_OVLP_OLD), log_shift, true, true);
  if (r) return r;
  r = gather(c, nw, phia, phib, true);
  if (r) return r;
  r = do_two_body_core(c, nw, ga, gb, xi, seed, counter, nullptr, nullptr, energy_prev != nullptr, energy_prev,
                       wv(c, WV_OVLP_OLD));
  c->g_planes_nw = 0;
  if (r) return r;
  const double* res = nullptr;
  const uint16_t* resT = nullptr;
  r = taylor(c, nw, &res, &resT);
  if (r) return r;
  r = scatter(c, nw, res, phia, phib, true, resT);
  if (r) return r;
  r = do_gf(c, nw, phia, phib, nullptr, nullptr, wv(c, WV_OVLP_NEW), log_shift, false);
  if (r) return r;
  return afqmc_update_weight(c, nw, wv(c, WV_OVLP_OLD), wv(c, WV_OVLP_NEW), wv(c, WV_CFB), wv(c, WV_CMF), eshift,
                             weight, hybrid_energy, ovlp);
}

int afqmc_local_energy(afqmc_ctx* c, int32_t nw, const double* ga, const double* gb, double* energy) {
  int r = check_ready(c, nw);
  if (r) return r;
  r = need_static(c);
  if (r) return r;
  if (!ga || !energy || (two_spins(c) && !gb)) return fail(c, AFQMC_ERR_INVALID_ARG, "null array");
  r = do_force_bias(c, nw, ga, gb);
  if (r) return r;
  r = do_fields(c, nw, false, nullptr, 0, 0, nullptr, nullptr);
  if (r) return r;
  r = do_exchange(c, nw, ga, gb);
  if (r) return r;
  return do_energy(c, nw, ga, gb, energy);
}

int afqmc_control_variate(afqmc_ctx* c, int32_t nw, const double* ga, const double* gb, double* cv) {
  int r = check_ready(c, nw);
  if (r) return r;
  r = need_static(c);
  if (r) return r;
  if (!ga || !cv || (two_spins(c) && !gb)) return fail(c, AFQMC_ERR_INVALID_ARG, "null array");
  if (!c->sur.enabled) return fail(c, AFQMC_ERR_INVALID_ARG, "the control variate is not configured");
  r = sur_prepare_cv(c);
  if (r) return r;
  r = ensure_planes(c);
  if (r) return r;
  auto& S = c->sur;
  const int nspin = two_spins(c) ? 2 : 1;
  const int N = c->d.nbasis;
  const int64_t gstride = 2 * (int64_t)std::max(c->d.nalpha, c->d.nbeta) * N;
  for (int s = 0; s < nspin && S.qmode != AFQMC_SURROGATE_Q_NONE; ++s) {
    r = exchange_q_spin(c, nw, s, s == 0 ? ga : gb);
    if (r) return r;
  }
  SurrogateCvArgs cva{};
  cva.nwalkers = nw;
  cva.nbasis = N;
  cva.nspin = nspin;
  for (int s = 0; s < nspin; ++s) {
    cva.nocc[s] = s == 0 ? c->d.nalpha : c->d.nbeta;
    cva.g[s] = s == 0 ? ga : gb;
    cva.gT[s] = S.gT + s * gstride;
    cva.K[s] = S.K + s * gstride;
    cva.part[s] = s == 0 ? c->exx_a : c->exx_b;
    cva.xs[s] = S.qmode != AFQMC_SURROGATE_Q_NONE ? c->xs_used[s] : 0;  // partial sums of Q, when computed
  }
  cva.et = S.et;
  cva.cv = cv;
  // the kernel also folds the MLP term into the step's predictor buffers; the
  // step rewrites both before it reads them, so its state is not changed
  cva.pred = S.pred;
  cva.mlp = S.mlp;
  cva.state = S.state;
  cva.fold = 1;
  HIP_TRY(c, launch_sur_cv(cva, c->stream));
  return AFQMC_OK;
}

int afqmc_reortho(afqmc_ctx* c, int32_t nw, double* phia, double* phib, double* ovlp, double* detR,
                  double* log_detR, const double* detR_shift) {
  int r = check_ready(c, nw);
  if (r) return r;
  r = ensure_ws(c);
  if (r) return r;
  if (!phia || !detR || (c->d.nbeta > 0 && !phib)) return fail(c, AFQMC_ERR_INVALID_ARG, "null array");
  ReorthoArgs a;
  a.nwalkers = nw;
  a.nbasis = c->d.nbasis;
  a.nalpha = c->d.nalpha;
  a.nbeta = c->d.nbeta;
  a.phia = phia;
  a.phib = phib;
  a.ovlp = ovlp;
  a.detR = detR;
  a.log_detR = log_detR;
  a.detR_shift = detR_shift;
  a.flags = c->flags;
  a.wflags = c->wflags;
  PhaseScope ps(c, AFQMC_PHASE_REORTHO);
  HIP_TRY(c, launch_reortho(a, c->stream));
  return AFQMC_OK;
}

int afqmc_graph_begin(afqmc_ctx* c) {
  if (!c) return AFQMC_ERR_INVALID_ARG;
  if (c->capturing) return fail(c, AFQMC_ERR_INVALID_ARG, "a step-graph capture is already open");
  if (!c->stream)
    return fail(c, AFQMC_ERR_INVALID_ARG, "a step graph is captured on a stream other than the null stream");
  if (t_capturing) return fail(c, AFQMC_ERR_INVALID_ARG, "this thread already captures another context");
  HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  c->capturing = true;
  t_capturing = true;
  c->capture_steps = 0;
  c->capture_sur = 0;
  return AFQMC_OK;
}

int afqmc_graph_end(afqmc_ctx* c, int32_t* steps) {
  if (!c) return AFQMC_ERR_INVALID_ARG;
  if (!c->capturing) return fail(c, AFQMC_ERR_INVALID_ARG, "no step-graph capture is open");
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (e != hipSuccess) g = nullptr;  // a failed end hands back no graph of ours to destroy
  c->capturing = false;
  t_capturing = false;
  // the recorded control-variate energies did not run
  c->sur.steps -= c->capture_sur;
  if (e != hipSuccess || !g) {
    if (g) (void)hipGraphDestroy(g);
    return fail(c, AFQMC_ERR_HIP, "step-graph capture failed: %s (run the captured calls once before capturing, so "
                "that every workspace exists)", hipGetErrorString(e));
  }
  if (c->graph) {
    (void)hipGraphExecDestroy(c->graph);
    c->graph = nullptr;
  }
  e = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    c->graph = nullptr;
    return fail(c, AFQMC_ERR_HIP, "step-graph instantiation failed: %s", hipGetErrorString(e));
  }
  c->graph_steps = c->capture_steps;
  c->graph_sur = c->capture_sur;
  if (steps) *steps = c->graph_steps;
  return AFQMC_OK;
}

int afqmc_graph_launch(afqmc_ctx* c, int32_t replays, uint64_t counter, double eshift) {
  if (!c) return AFQMC_ERR_INVALID_ARG;
  if (c->capturing) return refuse_in_capture(c, "afqmc_graph_launch");
  if (!c->graph) return fail(c, AFQMC_ERR_INVALID_ARG, "no step graph has been captured");
  if (replays < 0) return fail(c, AFQMC_ERR_INVALID_ARG, "replays must be >= 0");
  HIP_TRY(c, launch_step_set(c->dstate, counter, eshift, c->stream));
  for (int i = 0; i < replays; ++i) HIP_TRY(c, hipGraphLaunch(c->graph, c->stream));
  c->sur.steps += (int64_t)replays * c->graph_sur;
  return AFQMC_OK;
}

// ---------------------------------------------------------------- multi-system step graphs
struct afqmc_batch {
  std::vector<afqmc_ctx*> ctx;
  std::vector<hipEvent_t> ev;  // [0]: fork from the first stream; [i]: join of context i
  hipGraphExec_t graph = nullptr;
  std::vector<int> steps, sur;
  bool capturing = false;
};

int afqmc_batch_create(afqmc_batch** out, afqmc_ctx* const* ctxs, int32_t n) {
  if (!out || !ctxs || n < 1) return AFQMC_ERR_INVALID_ARG;
  *out = nullptr;
  for (int i = 0; i < n; ++i) {
    if (!ctxs[i]) return AFQMC_ERR_INVALID_ARG;
    for (int j = 0; j < i; ++j)
      if (ctxs[j] == ctxs[i]) return fail(ctxs[0], AFQMC_ERR_INVALID_ARG, "context %d appears twice", i);
    if (ctxs[i]->device != ctxs[0]->device)
      return fail(ctxs[0], AFQMC_ERR_INVALID_ARG, "the contexts of a batch share one device");
  }
  afqmc_batch* b = new afqmc_batch();
  b->ctx.assign(ctxs, ctxs + n);
  b->ev.assign(n, nullptr);
  b->steps.assign(n, 0);
  b->sur.assign(n, 0);
  for (int i = 0; i < n; ++i)
    if (hipEventCreateWithFlags(&b->ev[i], hipEventDisableTiming) != hipSuccess) {
      afqmc_batch_destroy(b);
      return fail(ctxs[0], AFQMC_ERR_HIP, "event creation failed");
    }
  *out = b;
  return AFQMC_OK;
}

void afqmc_batch_destroy(afqmc_batch* b) {
  if (!b) return;
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  for (hipEvent_t e : b->ev)
    if (e) (void)hipEventDestroy(e);
  delete b;
}

int afqmc_batch_begin(afqmc_batch* b) {
  if (!b) return AFQMC_ERR_INVALID_ARG;
  afqmc_ctx* c0 = b->ctx[0];
  const int n = (int)b->ctx.size();
  if (b->capturing || t_capturing) return fail(c0, AFQMC_ERR_INVALID_ARG, "a capture is already open");
  for (int i = 0; i < n; ++i) {
    if (!b->ctx[i]->stream)
      return fail(c0, AFQMC_ERR_INVALID_ARG, "context %d has the null stream: each context needs its own stream", i);
    if (b->ctx[i]->capturing) return fail(c0, AFQMC_ERR_INVALID_ARG, "context %d is capturing", i);
    for (int j = 0; j < i; ++j)
      if (b->ctx[j]->stream == b->ctx[i]->stream)
        return fail(c0, AFQMC_ERR_INVALID_ARG, "contexts %d and %d share a stream", j, i);
  }
  HIP_TRY(c0, hipStreamBeginCapture(c0->stream, hipStreamCaptureModeThreadLocal));
  hipError_t e = hipEventRecord(b->ev[0], c0->stream);
  for (int i = 1; i < n && e == hipSuccess; ++i) e = hipStreamWaitEvent(b->ctx[i]->stream, b->ev[0], 0);
  if (e != hipSuccess) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(c0->stream, &g);
    if (g) (void)hipGraphDestroy(g);
    return fail(c0, AFQMC_ERR_HIP, "forking the capture failed: %s", hipGetErrorString(e));
  }
  for (afqmc_ctx* c : b->ctx) {
    c->capturing = true;
    c->capture_steps = 0;
    c->capture_sur = 0;
  }
  b->capturing = true;
  t_capturing = true;
  return AFQMC_OK;
}

int afqmc_batch_end(afqmc_batch* b) {
  if (!b) return AFQMC_ERR_INVALID_ARG;
  afqmc_ctx* c0 = b->ctx[0];
  if (!b->capturing) return fail(c0, AFQMC_ERR_INVALID_ARG, "no capture is open");
  const int n = (int)b->ctx.size();
  // Join every forked stream, whatever failed before: a stream left unjoined
  // makes hipStreamEndCapture fail with hipErrorStreamCaptureUnjoined and stays
  // bound to the ended capture, so the next call on it would act on capture
  // state that no longer exists.
  hipError_t e = hipSuccess;
  for (int i = 1; i < n; ++i) {
    hipError_t ej = hipEventRecord(b->ev[i], b->ctx[i]->stream);
    if (ej == hipSuccess) ej = hipStreamWaitEvent(c0->stream, b->ev[i], 0);
    if (e == hipSuccess) e = ej;
  }
  hipGraph_t g = nullptr;
  const hipError_t e2 = hipStreamEndCapture(c0->stream, &g);
  if (e2 != hipSuccess) g = nullptr;  // a failed end hands back no graph of ours to destroy
  if (e == hipSuccess) e = e2;
  b->capturing = false;
  t_capturing = false;
  for (int i = 0; i < n; ++i) {
    afqmc_ctx* c = b->ctx[i];
    c->capturing = false;
    c->sur.steps -= c->capture_sur;  // recorded, not run
    b->steps[i] = c->capture_steps;
    b->sur[i] = c->capture_sur;
  }
  if (e != hipSuccess || !g) {
    if (g) (void)hipGraphDestroy(g);
    return fail(c0, AFQMC_ERR_HIP, "multi-system capture failed: %s (run the recorded calls once before capturing)",
                hipGetErrorString(e));
  }
  if (b->graph) {
    (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
  }
  e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    b->graph = nullptr;
    return fail(c0, AFQMC_ERR_HIP, "multi-system graph instantiation failed: %s", hipGetErrorString(e));
  }
  return AFQMC_OK;
}

int afqmc_batch_steps(afqmc_batch* b, int32_t* steps) {
  if (!b || !steps) return AFQMC_ERR_INVALID_ARG;
  for (size_t i = 0; i < b->ctx.size(); ++i) steps[i] = b->steps[i];
  return AFQMC_OK;
}

int afqmc_batch_launch(afqmc_batch* b, int32_t replays, const uint64_t* counters, const double* eshifts) {
  if (!b || !counters || !eshifts) return AFQMC_ERR_INVALID_ARG;
  afqmc_ctx* c0 = b->ctx[0];
  if (b->capturing || t_capturing) return refuse_in_capture(c0, "afqmc_batch_launch");
  if (!b->graph) return fail(c0, AFQMC_ERR_INVALID_ARG, "no multi-system graph has been captured");
  if (replays < 0) return fail(c0, AFQMC_ERR_INVALID_ARG, "replays must be >= 0");
  // every context's step values are set on the launch stream, ahead of the graph
  for (size_t i = 0; i < b->ctx.size(); ++i)
    HIP_TRY(c0, launch_step_set(b->ctx[i]->dstate, counters[i], eshifts[i], c0->stream));
  for (int r = 0; r < replays; ++r) HIP_TRY(c0, hipGraphLaunch(b->graph, c0->stream));
  for (size_t i = 0; i < b->ctx.size(); ++i) b->ctx[i]->sur.steps += (int64_t)replays * b->sur[i];
  return AFQMC_OK;
}

int afqmc_draw_normals(afqmc_ctx* c, double* out, int64_t n, uint64_t seed, uint64_t counter) {
  if (!c || !out || n < 0) return AFQMC_ERR_INVALID_ARG;
  HIP_TRY(c, launch_draw_normals(out, n, seed, counter, c->stream));
  return AFQMC_OK;
}

// Launch shapes (no counterpart in ipie: its kernels' shapes are fixed by
// CuPy / numba).  Each knob has a heuristic default; ipie_amd/autotune.py
// times the candidates on the context's own shapes and sets the fastest.
namespace {
int tune_fb_max(const afqmc_ctx* c) {
  if (c->bf16_fb) {
    int kch = c->d.nalpha * c->kpad / 32;
    if (two_spins(c) && c->d.nbeta > 0) kch = std::min(kch, c->d.nbeta * c->kpad / 32);
    return std::max(1, std::min(32, kch));  // at least one K chunk per range
  }
  const int64_t N = c->d.nbasis;
  const int64_t kch = ceil_div((int64_t)c->d.nalpha * N, 16) + (two_spins(c) ? ceil_div((int64_t)c->d.nbeta * N, 16) : 0);
  return (int)std::max<int64_t>(1, std::min<int64_t>(32, kch / 8));
}
int tune_exx_max(const afqmc_ctx* c) {
  int hi = 1;
  for (int s = 0; s < (two_spins(c) ? 2 : 1); ++s) {
    const int n = s == 0 ? c->d.nalpha : c->d.nbeta;
    if (n < 1) continue;
    const int xb = (c->d.precision == AFQMC_PREC_MIXED && c->bx_ok[s]) ? c->bx_xb[s] : 1;
    hi = std::max(hi, (int)ceil_div(c->d.nchol, xb));
  }
  return hi;
}
}  // namespace

int afqmc_tune_get(afqmc_ctx* c, int32_t knob, int32_t* value, int32_t* lo, int32_t* hi) {
  if (!c || !value || !lo || !hi) return AFQMC_ERR_INVALID_ARG;
  *value = 0;
  *lo = 0;
  *hi = 0;
  switch (knob) {
    case AFQMC_TUNE_FB_SPLIT:
      if (c->d.nchol < 1) return AFQMC_OK;
      *value = c->bf16_fb ? c->fb_split_spin : c->fb_ksplit;
      *lo = 1;
      *hi = std::max(tune_fb_max(c), *value);
      return AFQMC_OK;
    case AFQMC_TUNE_QUAD_WGS:
      if (!c->quad_on || c->d.precision != AFQMC_PREC_MIXED) return AFQMC_OK;
      *value = c->quad_target;
      *lo = 256;
      *hi = 8192;
      return AFQMC_OK;
    case AFQMC_TUNE_EXX_SPLIT:
      if (c->quad_on || c->d.nchol < 1) return AFQMC_OK;
      *value = c->exx_split;
      *lo = 1;
      *hi = tune_exx_max(c);
      return AFQMC_OK;
    case AFQMC_TUNE_TAYLOR:
      if (c->d.precision != AFQMC_PREC_MIXED || c->d.nchol < 1 || c->d.exp_nmax < 1) return AFQMC_OK;
      *value = c->taylor_variant;
      *lo = 0;
      *hi = 1;
      return AFQMC_OK;
    case AFQMC_TUNE_QUAD_RGROUP:
      if (!c->quad_on || c->d.precision != AFQMC_PREC_MIXED) return AFQMC_OK;
      *hi = (int32_t)ceil_div(c->d.nwalkers, 128);
      *lo = 1;
      *value = c->quad_rgroup > 0 ? std::min<int32_t>(c->quad_rgroup, *hi) : *hi;
      return AFQMC_OK;
    case AFQMC_TUNE_TILE_FB:
    case AFQMC_TUNE_TILE_VHS:
    case AFQMC_TUNE_TILE_ONE_BODY:
      // split-BF16 GEMM tile shapes (mixed precision; the force bias needs Cholesky vectors)
      if (c->d.precision != AFQMC_PREC_MIXED || c->d.nchol < 1) return AFQMC_OK;
      *value = knob == AFQMC_TUNE_TILE_FB ? c->tile_fb : knob == AFQMC_TUNE_TILE_VHS ? c->tile_vhs : c->tile_ob;
      *lo = 0;
      *hi = BF16_TILES - 1;
      return AFQMC_OK;
    default:
      return fail(c, AFQMC_ERR_INVALID_ARG, "unknown tuning knob %d", knob);
  }
}

int afqmc_tune_set(afqmc_ctx* c, int32_t knob, int32_t value) {
  if (!c) return AFQMC_ERR_INVALID_ARG;
  if (c->capturing || t_capturing) return refuse_in_capture(c, "afqmc_tune_set");
  int32_t cur, lo, hi;
  const int r = afqmc_tune_get(c, knob, &cur, &lo, &hi);
  if (r) return r;
  if (hi < 1) return fail(c, AFQMC_ERR_UNSUPPORTED, "tuning knob %d does not apply to this context", knob);
  const bool heuristic = knob == AFQMC_TUNE_EXX_SPLIT && value == 0;
  if (!heuristic && (value < lo || value > hi))
    return fail(c, AFQMC_ERR_INVALID_ARG, "tuning knob %d: %d outside [%d, %d]", knob, value, lo, hi);
  (void)hipSetDevice(c->device);
  if (knob == AFQMC_TUNE_FB_SPLIT) {
    if (c->bf16_fb) {
      c->fb_split_spin = value;
      c->fb_ksplit = value * (two_spins(c) ? 2 : 1);
    } else {
      c->fb_ksplit = value;
    }
    if (c->ws_ready && c->fb_ksplit > c->fb_cap) {
      double* part = nullptr;
      if (dmalloc((void**)&part, 16 * (size_t)c->fb_ksplit * c->d.nwalkers * c->d.nchol) != hipSuccess)
        return fail(c, AFQMC_ERR_ALLOC, "force-bias partial buffer allocation failed (%s)", c->describe_dims);
      c->retired.push_back(c->vbias_part);
      c->vbias_part = part;
      c->fb_cap = c->fb_ksplit;
    }
  } else if (knob == AFQMC_TUNE_QUAD_WGS) {
    c->quad_target = value;
    if (c->ws_ready) return build_quad_units(c);
  } else if (knob == AFQMC_TUNE_TAYLOR) {
    c->taylor_variant = value;
  } else if (knob == AFQMC_TUNE_QUAD_RGROUP) {
    c->quad_rgroup = value;
  } else if (knob == AFQMC_TUNE_TILE_FB) {
    c->tile_fb = value;
  } else if (knob == AFQMC_TUNE_TILE_VHS) {
    c->tile_vhs = value;
  } else if (knob == AFQMC_TUNE_TILE_ONE_BODY) {
    c->tile_ob = value;
  } else {
    c->exx_split = value;
  }
  return AFQMC_OK;
}

int afqmc_profile(afqmc_ctx* c, int enable) {
  if (!c) return AFQMC_ERR_INVALID_ARG;
  c->prof = enable != 0;
  for (int i = 0; i < AFQMC_NPHASES; ++i) {
    c->prof_ms[i] = 0.0;
    c->prof_count[i] = 0;
  }
  // phases recorded before the reset and not yet read are dropped with the
  // totals (their events go back to the pool; recording an event again
  // replaces its earlier record)
  for (const PendingEvent& pe : c->ev_pending) {
    c->ev_free.push_back(pe.start);
    c->ev_free.push_back(pe.stop);
  }
  c->ev_pending.clear();
  return AFQMC_OK;
}

int afqmc_profile_read(afqmc_ctx* c, double* ms, int64_t* launches) {
  if (!c || !ms || !launches) return AFQMC_ERR_INVALID_ARG;
  if (c->capturing) return refuse_in_capture(c, "afqmc_profile_read");
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  for (const PendingEvent& pe : c->ev_pending) {
    float t = 0.0f;
    HIP_TRY(c, hipEventElapsedTime(&t, pe.start, pe.stop));
    c->prof_ms[pe.phase] += t;
    c->prof_count[pe.phase] += 1;
    c->ev_free.push_back(pe.start);
    c->ev_free.push_back(pe.stop);
  }
  c->ev_pending.clear();
  for (int i = 0; i < AFQMC_NPHASES; ++i) {
    ms[i] = c->prof_ms[i];
    launches[i] = c->prof_count[i];
  }
  return AFQMC_OK;
}

int afqmc_surrogate_configure(afqmc_ctx* c, int enable, double fraction, int warmup_steps, int train_iters,
                              double lr, uint64_t seed) {
  if (!c) return AFQMC_ERR_INVALID_ARG;
  if (c->capturing) return refuse_in_capture(c, "afqmc_surrogate_configure");
  if (!enable) {
    c->sur.enabled = false;
    return AFQMC_OK;
  }
  if (!(fraction > 0.0 && fraction <= 1.0) || warmup_steps < 0 || train_iters < 0 || !(lr > 0.0))
    return fail(c, AFQMC_ERR_INVALID_ARG, "bad surrogate parameters");
  int r = check_ready(c, 1);
  if (r) return r;
  r = ensure_ws(c);
  if (r) return r;
  auto& S = c->sur;
  const int64_t W = c->d.nwalkers;
  const int groups = sur_grad_groups((int)W);
  if (!S.master) {
    bool ok = true;
    ok = ok && dmalloc((void**)&S.feat, sizeof(float) * W * SUR_FEATURES) == hipSuccess;
    ok = ok && dmalloc((void**)&S.pred, 16 * W) == hipSuccess;
    ok = ok && dmalloc((void**)&S.est, 16 * W) == hipSuccess;
    ok = ok && dmalloc((void**)&S.target, 16 * W) == hipSuccess;
    ok = ok && dmalloc((void**)&S.resid, 8 * W) == hipSuccess;
    ok = ok && dmalloc((void**)&S.walker_of, sizeof(int) * W) == hipSuccess;
    ok = ok && dmalloc((void**)&S.xbar, 16 * W) == hipSuccess;
    ok = ok && dmalloc((void**)&S.e1, 16 * W) == hipSuccess;
    ok = ok && dmalloc((void**)&S.master, sizeof(float) * SUR_NPARAM) == hipSuccess;
    ok = ok && dmalloc((void**)&S.m, sizeof(float) * SUR_NPARAM) == hipSuccess;
    ok = ok && dmalloc((void**)&S.v, sizeof(float) * SUR_NPARAM) == hipSuccess;
    ok = ok && dmalloc((void**)&S.wbf, sizeof(uint16_t) * SUR_NPARAM) == hipSuccess;
    ok = ok && dmalloc((void**)&S.bias, sizeof(float) * (2 * SUR_HIDDEN + 2)) == hipSuccess;
    ok = ok && dmalloc((void**)&S.grad, sizeof(float) * (size_t)groups * SUR_NPARAM) == hipSuccess;
    ok = ok && dmalloc((void**)&S.loss, sizeof(float) * (size_t)groups) == hipSuccess;
    ok = ok && dmalloc((void**)&S.state, sizeof(SurrogateState)) == hipSuccess;
    const int64_t nmax = std::max(c->d.nalpha, c->d.nbeta);
    ok = ok && dmalloc((void**)&S.gT, 2 * 16 * (size_t)nmax * c->d.nbasis) == hipSuccess;
    ok = ok && dmalloc((void**)&S.K, 2 * 16 * (size_t)nmax * c->d.nbasis) == hipSuccess;
    ok = ok && dmalloc((void**)&S.et, 2 * 16) == hipSuccess;
    ok = ok && dmalloc((void**)&S.cv, 16 * W) == hipSuccess;
    ok = ok && dmalloc((void**)&S.mlp, 16 * W) == hipSuccess;
    ok = ok && dmalloc((void**)&S.gtmp, 2 * 16) == hipSuccess;
    if (!ok) return fail(c, AFQMC_ERR_ALLOC, "surrogate allocation failed");
  }
  if (!c->bx_ok[0] || (two_spins(c) && !c->bx_ok[1]))
    return fail(c, AFQMC_ERR_UNSUPPORTED, "the surrogate needs n <= 128 occupied orbitals per spin");
  r = alloc_exchange_planes(c);
  if (r) return r;
  S.cv_ready = false;
  S.enabled = true;
  S.fraction = fraction;
  S.warmup = warmup_steps;
  S.train_iters = train_iters;
  S.lr = (float)lr;
  S.steps = 0;
  S.seed = seed;
  S.rng.seed(seed);
  // weights ~ N(0, 1/fan_in), biases 0 (host RNG, then upload)
  std::vector<float> P(SUR_NPARAM, 0.0f);
  std::normal_distribution<float> nd(0.0f, 1.0f);
  const int H = SUR_HIDDEN, F = SUR_FEATURES;
  for (int i = 0; i < H * F; ++i) P[i] = nd(S.rng) / std::sqrt((float)F);
  for (int i = 0; i < H * H; ++i) P[H * F + i] = nd(S.rng) / std::sqrt((float)H);
  for (int i = 0; i < 2 * H; ++i) P[H * F + H * H + i] = nd(S.rng) / std::sqrt((float)H);
  std::vector<uint16_t> Pb(SUR_NPARAM, 0);
  for (int i = 0; i < SUR_NPARAM; ++i) {
    uint32_t u;
    std::memcpy(&u, &P[i], 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    Pb[i] = (uint16_t)(u >> 16);
  }
  HIP_TRY(c, hipMemcpyAsync(S.master, P.data(), sizeof(float) * SUR_NPARAM, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(c, hipMemcpyAsync(S.wbf, Pb.data(), sizeof(uint16_t) * SUR_NPARAM, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(c, hipMemsetAsync(S.bias, 0, sizeof(float) * (2 * SUR_HIDDEN + 2), c->stream));
  HIP_TRY(c, hipMemsetAsync(S.m, 0, sizeof(float) * SUR_NPARAM, c->stream));
  HIP_TRY(c, hipMemsetAsync(S.v, 0, sizeof(float) * SUR_NPARAM, c->stream));
  HIP_TRY(c, hipMemsetAsync(S.state, 0, sizeof(SurrogateState), c->stream));
  HIP_TRY(c, hipMemsetAsync(c->dstate, 0, sizeof(StepState), c->stream));  // subset draws, Adam steps
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return AFQMC_OK;
}

int afqmc_surrogate_quadratic(afqmc_ctx* c, int mode) {
  if (!c || mode < AFQMC_SURROGATE_Q_NONE || mode > AFQMC_SURROGATE_Q_FP8) return AFQMC_ERR_INVALID_ARG;
  if (c->capturing) return refuse_in_capture(c, "afqmc_surrogate_quadratic");
  c->sur.qmode = mode;
  return AFQMC_OK;
}

int afqmc_surrogate_stats(afqmc_ctx* c, double* out, int n) {
  if (!c || !out || n < 6) return AFQMC_ERR_INVALID_ARG;
  if (c->capturing) return refuse_in_capture(c, "afqmc_surrogate_stats");
  auto& S = c->sur;
  out[0] = (double)S.steps;
  out[1] = (double)S.warmup;
  out[2] = (double)S.last_k;
  for (int i = 3; i < n; ++i) out[i] = 0.0;
  if (!S.enabled || S.steps == 0) return AFQMC_OK;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  const int groups = sur_grad_groups(S.last_k);
  std::vector<float> loss(groups);
  std::vector<double> resid(S.last_k);
  SurrogateState st;
  HIP_TRY(c, hipMemcpy(loss.data(), S.loss, sizeof(float) * groups, hipMemcpyDeviceToHost));
  HIP_TRY(c, hipMemcpy(resid.data(), S.resid, 8 * (size_t)S.last_k, hipMemcpyDeviceToHost));
  HIP_TRY(c, hipMemcpy(&st, S.state, sizeof(st), hipMemcpyDeviceToHost));
  double l = 0.0, rr = 0.0;
  for (float v : loss) l += v;
  for (double v : resid) rr += v;
  out[3] = l / S.last_k;                         // training loss per sample (normalised units)
  out[4] = rr / S.last_k;                        // mean |exact - prediction|^2 on the sample (Ha^2)
  const double m = st.count > 0 ? st.sum[SUR_FEATURES] / st.count : 0.0;
  out[5] = st.count > 0 ? st.sumsq[SUR_FEATURES] / st.count - m * m : 0.0;  // target variance (Re)
  if (n > 6) out[6] = st.gain;                   // gain of the MLP term
  return AFQMC_OK;
}

}  // extern "C"
