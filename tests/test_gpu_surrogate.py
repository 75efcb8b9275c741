"""Learned exchange control variate: exact during warm-up, unbiased afterwards.

Unbiasedness is checked directly: the same walker state is evaluated many
times (fresh random subset each time) and the mean of the per-walker estimates
must match the exact local energy within its statistical error."""

import numpy
import pytest

from tests._helpers import Pair, random_walkers, rel_err, synthetic_arrays

pytestmark = pytest.mark.gpu

N, X, D, NA, NB, W = 48, 100, 30.0, 8, 6, 64


def make(precision="fp64", warmup=3, fraction=0.125):
    import torch  # noqa: F401

    print(f"[case] surrogate N={N} X={X} na={NA} nb={NB} W={W} precision={precision}", flush=True)
    h1e, chol, ecore, psi = synthetic_arrays(N, X, D, NA, NB, scale=(12.0 / (N * D)) ** 0.5)
    init = random_walkers(N, NA, NB, W, seed=8, amp=0.1, base=psi)
    return Pair(h1e, chol, psi, (NA, NB), W, init=init, precision=precision,
                surrogate=dict(warmup=warmup, fraction=fraction, train_iters=3, lr=2e-3, seed=5))


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


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_estimator_is_unbiased(precision):
    p = make(precision=precision, warmup=4)
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
