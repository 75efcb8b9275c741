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


@pytest.mark.parametrize("precision,quadratic", [("fp64", True), ("mixed", True), ("mixed", False)])
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
    in the quadratic-form GEMMs, a subset of k = 256 exact walkers): after a
    short warm-up, R = 48 re-evaluations of one state; every walker's mean
    estimate within 6 standard errors of its exact exchange, the pooled mean
    within 5, weights finite."""
    import torch

    from ipie_amd.estimators import local_energy_batch
    from tests._big import device_state

    system, ham, trial, walkers, prop = device_state(
        "n300", precision="mixed", surrogate=dict(warmup=6, fraction=0.125, train_iters=3, lr=2e-3, seed=5))
    Wn = walkers.nwalkers
    e = torch.empty(Wn, 3, dtype=torch.complex128, device=walkers.device)
    for _ in range(6):
        prop.propagate_walkers(walkers, ham, trial, 0.0, energy_prev=e)
    snap = snapshot(walkers)
    trial.calc_greens_function(walkers)
    ref = local_energy_batch(system, ham, walkers, trial).cpu().numpy()
    R = 48
    est = numpy.empty((R, Wn), dtype=numpy.complex128)
    for r in range(R):
        restore(walkers, snap)
        prop.propagate_walkers(walkers, ham, trial, 0.0, energy_prev=e)
        est[r] = e[:, 2].cpu().numpy()
    assert prop.surrogate_stats()["last_k"] == round(Wn / 8)
    assert bool(torch.isfinite(walkers.weight).all())
    mean = est.mean(axis=0)
    sem = est.std(axis=0) / numpy.sqrt(R) + 1e-12 * numpy.abs(ref[:, 2])
    err = numpy.abs(mean - ref[:, 2])
    z = err / sem
    tol_abs = 1e-5 * numpy.abs(ref[:, 2]).max()  # mixed-precision exchange error of the exact samples
    print(f"max z {z.max():.2f} mean z {z.mean():.2f} stats {prop.surrogate_stats()}", flush=True)
    assert numpy.all((z < 6.0) | (err < tol_abs))
    pooled = (mean - ref[:, 2]).real.mean() / (numpy.sqrt((sem**2).sum()) / Wn)
    assert abs(pooled) < 5.0


def _bf16(x):
    """Round to BF16 (nearest even), as the planes' hi parts hold it."""
    x = numpy.asarray(x, dtype=numpy.float32)
    u = x.view(numpy.uint32).astype(numpy.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(numpy.uint32).view(numpy.float32).astype(numpy.float64)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_control_variate_entry_point_predicts_the_exchange_sum(precision):
    """afqmc_control_variate gives S_T + 2 <K, dG> + Q(dG), the exchange sum S of a
    walker (sum_x tr(T_x T_x) over both spins, local_energy_sd.py:263-303)
    expanded about the trial's Green's function, dG = G - G_T.  S is a quadratic
    form in G, so the expansion is exact and Q(dG) is S evaluated at dG; only
    the single BF16 product of Q rounds.  Checks: every walker within 1% of the
    largest |Q| of the exact S (a dropped, doubled or sign-flipped term misses
    by |Q| or more); in FP64 contexts, whose Q forms every T_x from the BF16
    hi parts of L and dG, within 0.05% of the largest |Q| of the same sum taken
    from BF16-rounded operands on the CPU."""
    import torch

    from oracle import afqmc

    p = make(precision=precision, warmup=2)
    w = p.walkers
    ga, gb = torch.empty_like(w.Ghalfa), torch.empty_like(w.Ghalfb)
    ovlp = torch.empty_like(w.ovlp)
    ctx = w._engine.ctx
    ctx.greens_function(w.phia, w.phib, ga, gb, ovlp)
    cv = torch.empty(W, dtype=torch.complex128, device=w.device)
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
    ctx.surrogate_quadratic(True)
    if precision == "fp64":
        bc = lambda z: _bf16(z.real) + 1j * _bf16(z.imag)  # noqa: E731
        q_bf = 2.0 * (ex(_bf16(ra), bc(gan - gta)) + ex(_bf16(rb), bc(gbn - gtb)))
        emu = exact - q + q_bf
        assert numpy.abs(got - emu).max() <= 5e-4 * qmax, numpy.abs(got - emu).max() / qmax
