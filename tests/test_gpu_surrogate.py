"""Learned exchange control variate: exact during warm-up, unbiased afterwards.

Unbiasedness is checked directly: the same walker state is evaluated many
times (fresh random subset each time) and the mean of the per-walker estimates
must match the exact local energy within its statistical error."""

import numpy
import pytest

from tests._helpers import Pair, random_walkers, rel_err, synthetic_arrays

pytestmark = pytest.mark.gpu

N, X, D, NA, NB, W = 48, 100, 30.0, 8, 6, 64


def make(precision="fp64", warmup=3, fraction=0.125, quadratic=True):
    import torch  # noqa: F401

    print(f"[case] surrogate N={N} X={X} na={NA} nb={NB} W={W} precision={precision} quadratic={quadratic}", flush=True)
    h1e, chol, ecore, psi = synthetic_arrays(N, X, D, NA, NB, scale=(12.0 / (N * D)) ** 0.5)
    init = random_walkers(N, NA, NB, W, seed=8, amp=0.1, base=psi)
    return Pair(h1e, chol, psi, (NA, NB), W, init=init, precision=precision,
                surrogate=dict(warmup=warmup, fraction=fraction, train_iters=3, lr=2e-3, seed=5, quadratic=quadratic))


def snapshot(w):
    return {k: getattr(w, k).clone() for k in ("phia", "phib", "weight", "ovlp", "hybrid_energy")}


def restore(w, snap):
    for k, v in snap.items():
        getattr(w, k).copy_(v)


def exact_energy(p):
    from ipie_amd.estimators import local_energy_batch

    p.trial.calc_greens_function(p.walkers)
    return local_energy_batch(p.system, p.ham, p.walkers, p.trial).cpu().numpy()


def fused_energy(p):
    import torch

    e = torch.empty(W, 3, dtype=torch.complex128, device=p.walkers.device)
    p.prop.propagate_walkers(p.walkers, p.ham, p.trial, 0.0, energy_prev=e)
    return e.cpu().numpy()


def test_warmup_steps_are_exact():
    p = make(warmup=3)
    for step in range(3):
        snap = snapshot(p.walkers)
        ref = exact_energy(p)
        restore(p.walkers, snap)
        got = fused_energy(p)
        assert rel_err(got, ref) < 1e-12, step
    assert p.prop.surrogate_stats()["last_k"] == W


@pytest.mark.parametrize("precision,quadratic", [("fp64", True), ("mixed", True), ("mixed", "bf16"), ("mixed", False)])
def test_estimator_is_unbiased(precision, quadratic):
    p = make(precision=precision, warmup=4, quadratic=quadratic)
    for _ in range(4):  # warm-up (training on all walkers)
        fused_energy(p)
    snap = snapshot(p.walkers)
    ref = exact_energy(p)
    R = 384  # the per-walker estimate is a skewed two-point variable: enough draws for a z-test
    est = []
    for _ in range(R):
        restore(p.walkers, snap)
        est.append(fused_energy(p))
    est = numpy.array(est)  # [R, W, 3]
    st = p.prop.surrogate_stats()
    assert st["last_k"] == round(W / 8)
    # one-body and Coulomb parts are exact every time
    assert rel_err(est[:, :, 1], numpy.broadcast_to(ref[:, 1], est[:, :, 1].shape)) < 1e-12
    e2 = est[:, :, 2]
    mean = e2.mean(axis=0)
    sem = e2.std(axis=0) / numpy.sqrt(R) + 1e-12 * numpy.abs(ref[:, 2])
    z = numpy.abs(mean - ref[:, 2]) / sem
    tol_abs = 1e-5 * numpy.abs(ref[:, 2]).max() if precision == "mixed" else 0.0
    print(f"max z {z.max():.2f}, mean z {z.mean():.2f}, stats {st}")
    assert numpy.all((z < 5.0) | (numpy.abs(mean - ref[:, 2]) < tol_abs))
    # the pooled mean over walkers is unbiased too
    pooled = (mean - ref[:, 2]).mean() / (numpy.sqrt((sem**2).sum()) / W)
    assert abs(pooled) < 5.0


@pytest.mark.timeout(900)
def test_unbiased_at_surrogate_workload():
    """N=300, X=1800, n=30, W=2048 (the surrogate workload: 16 walker row blocks
    in the quadratic-form GEMMs, a subset of k = 256 exact walkers), with the
    physical control variate alone (no MLP, so the predictor g of a walker is the
    same in every draw): R = 64 re-evaluations of one state after a short
    warm-up.  Each walker's estimate then takes two values, g when it is not
    sampled and g + (W / k)(x - g) when it is, so the estimate is unbiased exactly
    when every sampled value implies the exact x (checked against the exact local
    energy to the mixed-precision exchange error) and walkers are sampled with
    frequency k / W (each walker's count of sampling draws against Binomial(R,
    k / W)); the pooled mean is checked too, weights finite."""
    import torch

    from ipie_amd.estimators import local_energy_batch
    from tests._big import device_state

    system, ham, trial, walkers, prop = device_state(
        "n300", precision="mixed", surrogate=dict(warmup=6, fraction=0.125, train_iters=0, lr=2e-3, seed=5))
    Wn = walkers.nwalkers
    k = round(Wn / 8)
    e = torch.empty(Wn, 3, dtype=torch.complex128, device=walkers.device)
    for _ in range(6):
        prop.propagate_walkers(walkers, ham, trial, 0.0, energy_prev=e)
    snap = snapshot(walkers)
    trial.calc_greens_function(walkers)
    ref = local_energy_batch(system, ham, walkers, trial).cpu().numpy()
    R = 64
    est = numpy.empty((R, Wn), dtype=numpy.complex128)
    for r in range(R):
        restore(walkers, snap)
        prop.propagate_walkers(walkers, ham, trial, 0.0, energy_prev=e)
        est[r] = e[:, 2].cpu().numpy()
    assert prop.surrogate_stats()["last_k"] == k
    assert bool(torch.isfinite(walkers.weight).all())
    x = ref[:, 2]
    scale = numpy.abs(x).max()
    tol_abs = 1e-5 * scale  # mixed-precision exchange error of the exact samples
    # g per walker: the median draw (7 of 8 draws leave a walker unsampled)
    g = est[numpy.argsort(est.real, axis=0)[R // 2], numpy.arange(Wn)]
    hit = numpy.abs(est - g) > 1e-9 * scale
    implied = g + (est - g) * (k / Wn)
    print(f"sampled fraction {hit.mean():.4f}, max |x - g| {numpy.abs(x - g).max():.3e}, "
          f"max implied error {numpy.abs(implied - x)[hit].max(initial=0.0):.3e}", flush=True)
    assert numpy.abs(implied - x)[hit].max(initial=0.0) <= tol_abs
    # walkers whose sampling shows: the sampled value differs from g by far more
    # than the detection threshold; each one's count of draws that sample it is
    # Binomial(R, k / W) when the subsets are uniform and independent
    res = numpy.abs(x - g) * (Wn / k) > 1e-7 * scale
    p = k / Wn
    counts = hit[:, res].sum(axis=0)
    sd = numpy.sqrt(R * p * (1 - p))
    print(f"{res.sum()} walkers show their sampling: mean count {counts.mean():.3f} (want {R * p:.3f}), "
          f"sd {counts.std():.3f} (want {sd:.3f})", flush=True)
    assert res.sum() > Wn // 2
    assert abs(counts.mean() - R * p) < 5.0 * sd / numpy.sqrt(res.sum())
    assert 0.8 < counts.std() / sd < 1.2
    # the pooled mean of the estimates against the binomial spread of the draws
    dev = (est.mean(axis=0) - x).real
    sem = numpy.abs(x - g) * (Wn / k) * numpy.sqrt(p * (1 - p) / R)
    assert abs(dev.mean()) < 5.0 * numpy.sqrt((sem**2).sum()) / Wn + tol_abs


def _bf16(x):
    """Round to BF16 (nearest even), as the planes' hi parts hold it."""
    x = numpy.asarray(x, dtype=numpy.float32)
    u = x.view(numpy.uint32).astype(numpy.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(numpy.uint32).view(numpy.float32).astype(numpy.float64)


def _fp8(x):
    """Round to FP8 E4M3 (OCP: 3 mantissa bits, exponent bias 7, subnormals below
    2^-6, |x| <= 448 assumed), nearest even, as float64 values."""
    x = numpy.asarray(x, dtype=numpy.float64)
    a = numpy.abs(x)
    e = numpy.floor(numpy.log2(numpy.where(a > 0, a, 1.0)))
    step = 2.0 ** (numpy.maximum(e, -6.0) - 3.0)
    return numpy.sign(x) * numpy.round(a / step) * step  # numpy.round: halves to even


def _pow2_scale(amax):
    return 2.0 ** numpy.floor(numpy.log2(224.0 / amax)) if amax > 0 else 1.0


def _q_fp8(rchol, G, GT, N, Kp):
    """Q per walker of one spin as the FP8 control variate forms it: the
    block-triangular W' (2 Wq left of the diagonal 160-column tiles, Wq on them,
    0 right of them, in the Kp-padded index i Kp + m) rounded to BF16 hi + lo, then
    to FP8 after a power-of-two scale; dG = G - G_T rounded to FP32 and FP8 after a
    per-walker power-of-two scale; Q = sum_b dG_b Y_b with Y = dG8 W8 unscaled."""
    from tests import _mixed_emu as emu

    n = G.shape[1]
    Wq = emu.quad_matrix(rchol, n, N)
    ldn = n * Kp
    idx = (numpy.arange(n)[:, None] * Kp + numpy.arange(N)[None, :]).ravel()
    Wp = numpy.zeros((ldn, ldn))
    Wp[numpy.ix_(idx, idx)] = Wq
    tile = numpy.arange(ldn) // 160
    # plane row b, column a: 2 Wq_ab if tile(a) < tile(b), Wq_ab on the diagonal tiles, else 0
    Wd = numpy.where(tile[None, :] < tile[:, None], 2.0 * Wp, numpy.where(tile[None, :] == tile[:, None], Wp, 0.0))
    hi = _bf16(Wd)
    lo = _bf16(Wd - hi)
    v32 = (hi + lo).astype(numpy.float32).astype(numpy.float64)
    sw = _pow2_scale(numpy.abs(hi).max())
    W8 = _fp8(v32 * sw)
    d = numpy.zeros((G.shape[0], ldn), dtype=numpy.complex128)
    d[:, idx] = (G - GT).reshape(G.shape[0], -1)
    out = numpy.empty(G.shape[0], dtype=numpy.complex128)
    for w in range(G.shape[0]):
        s = _pow2_scale(max(numpy.abs(d[w].real).max(), numpy.abs(d[w].imag).max()))
        ar = _fp8((d[w].real * s).astype(numpy.float32).astype(numpy.float64))
        ai = _fp8((d[w].imag * s).astype(numpy.float32).astype(numpy.float64))
        y = (W8 @ ar + 1j * (W8 @ ai)) / (s * sw)
        out[w] = numpy.sum(d[w] * y)
    return out


def test_fp8_rounding_known_values():
    # the encodings the device's conversion gives (1, 448, 2^-6, 2^-9 and a tie)
    x = numpy.array([1.0, 448.0, 2.0**-6, 2.0**-9, 1.0 + 2.0**-4, 1.0 + 3 * 2.0**-4, 0.3])
    want = numpy.array([1.0, 448.0, 2.0**-6, 2.0**-9, 1.0, 1.25, 0.3125])
    assert numpy.array_equal(_fp8(x), want)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_control_variate_entry_point_predicts_the_exchange_sum(precision):
    """afqmc_control_variate gives S_T + 2 <K, dG> + Q(dG), the exchange sum S of a
    walker (sum_x tr(T_x T_x) over both spins, local_energy_sd.py:263-303)
    expanded about the trial's Green's function, dG = G - G_T.  S is a quadratic
    form in G, so the expansion is exact and Q(dG) is S evaluated at dG; only
    the single product of Q rounds.  Checks: with Q in BF16, every walker within
    1% of the largest |Q| of the exact S (a dropped, doubled or sign-flipped term
    misses by |Q| or more); in FP64 contexts, whose Q forms every T_x from the
    BF16 hi parts of L and dG, within 0.05% of the largest |Q| of the same sum
    taken from BF16-rounded operands on the CPU; in mixed contexts with Q in FP8
    (the default there), within 0.05% of the largest |Q| of the same FP8
    rounding emulated on the CPU, and within 5% of the largest |Q| of S."""
    import torch

    from oracle import afqmc

    p = make(precision=precision, warmup=2)
    w = p.walkers
    ga, gb = torch.empty_like(w.Ghalfa), torch.empty_like(w.Ghalfb)
    ovlp = torch.empty_like(w.ovlp)
    ctx = w._engine.ctx
    ctx.greens_function(w.phia, w.phib, ga, gb, ovlp)
    cv = torch.empty(W, dtype=torch.complex128, device=w.device)
    ctx.surrogate_quadratic("bf16")
    ctx.control_variate(ga, gb, cv)
    torch.cuda.synchronize()
    got = cv.cpu().numpy()
    o = p.orc
    gan, gbn = ga.cpu().numpy(), gb.cpu().numpy()
    ex = afqmc.exx_kernel_batch_real_rchol  # 1/2 sum_x tr(T_x T_x) per walker
    ra, rb = o.trial._rchola, o.trial._rcholb
    exact = 2.0 * (ex(ra, gan) + ex(rb, gbn))
    # the trial's own half Green's function, (psi^T psi*)^-1 psi^T
    psa, psb = o.trial.psi0a, o.trial.psi0b
    gta = numpy.linalg.solve(psa.T @ psa.conj(), psa.T)[None]
    gtb = numpy.linalg.solve(psb.T @ psb.conj(), psb.T)[None]
    q = 2.0 * (ex(ra, gan - gta) + ex(rb, gbn - gtb))
    qmax = numpy.abs(q).max()
    err = numpy.abs(got - exact)
    print(f"control variate {precision}: max err {err.max():.3e}, max |Q| {qmax:.3e}", flush=True)
    assert qmax > 0.1
    assert err.max() <= 1e-2 * qmax, err.max() / qmax
    # without its quadratic term the control variate is S_T + 2 <K, dG> = S - Q, in FP64
    ctx.surrogate_quadratic(False)
    ctx.control_variate(ga, gb, cv)
    torch.cuda.synchronize()
    lin = cv.cpu().numpy()
    assert numpy.abs(lin - (exact - q)).max() <= 1e-10 * numpy.abs(exact).max()
    if precision == "mixed":
        ctx.surrogate_quadratic("fp8")
        ctx.control_variate(ga, gb, cv)
        torch.cuda.synchronize()
        got8 = cv.cpu().numpy()
        Kp = ((N + 31) // 32) * 32
        q8 = _q_fp8(ra, gan, gta[0], N, Kp) + _q_fp8(rb, gbn, gtb[0], N, Kp)
        emu8 = exact - q + q8
        print(f"fp8 control variate: max err vs exact {numpy.abs(got8 - exact).max() / qmax:.3e} of max |Q|, "
              f"vs emulation {numpy.abs(got8 - emu8).max() / qmax:.3e}", flush=True)
        assert numpy.abs(got8 - emu8).max() <= 5e-4 * qmax
        assert numpy.abs(got8 - exact).max() <= 5e-2 * qmax
    ctx.surrogate_quadratic("bf16")
    if precision == "fp64":
        bc = lambda z: _bf16(z.real) + 1j * _bf16(z.imag)  # noqa: E731
        q_bf = 2.0 * (ex(_bf16(ra), bc(gan - gta)) + ex(_bf16(rb), bc(gbn - gtb)))
        emu = exact - q + q_bf
        assert numpy.abs(got - emu).max() <= 5e-4 * qmax, numpy.abs(got - emu).max() / qmax


def test_fp8_split_of_long_walker_rows():
    """The FP8 split of dG keeps a walker's values in registers up to 20480
    padded values (n Kp); longer rows (here n = 64, Kp = 352: 22528) take the
    two-pass split.  Its control variate agrees with the BF16 one to the FP8
    rounding of Q (5% of the largest |Q|; the BF16 form is pinned against the
    exact exchange sum in the tests above)."""
    import torch

    Nl, Xl, nl, Wl = 336, 400, 64, 128
    print(f"[case] long-row FP8 split N={Nl} X={Xl} n={nl} W={Wl}", flush=True)
    h1e, chol, ecore, psi = synthetic_arrays(Nl, Xl, 0.75 * Nl, nl, nl, scale=(12.0 / (Nl * 0.75 * Nl)) ** 0.5)
    init = random_walkers(Nl, nl, nl, Wl, seed=4, amp=0.1, base=psi)
    p = Pair(h1e, chol, psi, (nl, nl), Wl, init=init, precision="mixed",
             surrogate=dict(warmup=1, fraction=0.125, train_iters=0, lr=1e-3, seed=5))
    w = p.walkers
    ctx = w._engine.ctx
    assert ctx.exchange_form() == "quad"
    ga, gb = torch.empty_like(w.Ghalfa), torch.empty_like(w.Ghalfb)
    ovlp = torch.empty_like(w.ovlp)
    ctx.greens_function(w.phia, w.phib, ga, gb, ovlp)
    cv = torch.empty(Wl, dtype=torch.complex128, device=w.device)
    got = {}
    for mode in ("bf16", "none", "fp8"):
        ctx.surrogate_quadratic(mode)
        ctx.control_variate(ga, gb, cv)
        torch.cuda.synchronize()
        got[mode] = cv.cpu().numpy()
    q16 = got["bf16"] - got["none"]
    q8 = got["fp8"] - got["none"]
    qmax = numpy.abs(q16).max()
    err = numpy.abs(q8 - q16).max() / qmax
    print(f"long rows: fp8 vs bf16 Q max err {err:.3e} of max |Q| = {qmax:.3e}", flush=True)
    assert numpy.all(numpy.isfinite(q8)) and qmax > 0.0
    assert err <= 5e-2
