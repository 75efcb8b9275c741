"""Step graphs: steps replayed from one hipGraph equal the same steps run
eagerly with the same field counters.  The replay runs the same kernels with
the same arguments, except that the field counter, eshift, the control
variate's subset permutation and its Adam step number are read from device
memory; results must agree to 1e-12 (bit-exact is expected)."""

import numpy
import pytest

from tests._helpers import Pair, random_walkers, rel_err, synthetic_arrays

pytestmark = pytest.mark.gpu

TOL = 1e-12
N, X, D, NA, NB, W = 48, 100, 30.0, 8, 6, 64
CASES = {
    "fp64": dict(precision="fp64", surrogate=None),
    "mixed": dict(precision="mixed", surrogate=None),
    "mixed_surrogate": dict(precision="mixed",
                            surrogate=dict(warmup=3, fraction=0.125, train_iters=3, lr=2e-3, seed=5)),
}
FIELDS = ("phia", "phib", "weight", "ovlp", "hybrid_energy", "detR", "log_detR")


def make(case):
    import torch  # noqa: F401

    print(f"[case] graph {case} N={N} X={X} na={NA} nb={NB} W={W}", flush=True)
    h1e, chol, ecore, psi = synthetic_arrays(N, X, D, NA, NB, scale=(12.0 / (N * D)) ** 0.5)
    init = random_walkers(N, NA, NB, W, seed=8, amp=0.1, base=psi)
    return Pair(h1e, chol, psi, (NA, NB), W, init=init, **CASES[case])


def _eager_block(p, energy, eshift, n=5):
    for _ in range(n):
        p.prop.propagate_walkers(p.walkers, p.ham, p.trial, eshift, energy_prev=energy)
    p.walkers.reortho()


def _compare(a, b, ea, eb, what):
    for k in FIELDS:
        err = rel_err(getattr(b.walkers, k).cpu().numpy(), getattr(a.walkers, k).cpu().numpy())
        assert err < TOL, (what, k, err)
    assert rel_err(eb.cpu().numpy(), ea.cpu().numpy()) < TOL, (what, "energy")


@pytest.mark.parametrize("case", list(CASES))
def test_replay_equals_eager(case):
    import torch

    a, b = make(case), make(case)
    ea = torch.empty(W, 3, dtype=torch.complex128, device=a.walkers.device)
    eb = torch.empty_like(ea)
    eshift = float(a.trial.energy.real)
    # eager warm-up on both (workspace, and the control variate's warm-up)
    _eager_block(a, ea, eshift)
    _eager_block(b, eb, eshift)
    _compare(a, b, ea, eb, "warm-up")
    steps = b.prop.capture_steps(b.walkers, b.ham, b.trial, 5, reortho_every=5, energy_prev=eb)
    assert steps == 5
    for r in range(3):
        _eager_block(a, ea, eshift)
        b.prop.replay_steps(b.walkers, eshift, 1)
        _compare(a, b, ea, eb, f"replay {r}")
    # one launch of several replays continues the same trajectory
    _eager_block(a, ea, eshift)
    _eager_block(a, ea, eshift)
    b.prop.replay_steps(b.walkers, eshift, 2)
    _compare(a, b, ea, eb, "two replays")
    assert a.prop.counter == b.prop.counter
    assert numpy.all(numpy.isfinite(b.walkers.weight.cpu().numpy()))
    if CASES[case]["surrogate"]:
        sa, sb = a.prop.surrogate_stats(), b.prop.surrogate_stats()
        assert sa["steps"] == sb["steps"] and sb["last_k"] == W // 8


def test_sync_calls_refused_while_capturing():
    import torch

    from ipie_amd._lib import AfqmcError

    p = make("fp64")
    e = torch.empty(W, 3, dtype=torch.complex128, device=p.walkers.device)
    _eager_block(p, e, 0.0, n=1)
    ctx = p.walkers._engine.ctx
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        ctx.graph_begin()
        with pytest.raises(AfqmcError):
            ctx.flags()
        with pytest.raises(AfqmcError):
            ctx.sync()
        ctx.propagate(p.walkers.phia, p.walkers.phib, p.walkers.Ghalfa, p.walkers.Ghalfb, p.walkers.weight,
                      p.walkers.ovlp, p.walkers.hybrid_energy, 0.0, seed=p.prop.seed, counter=0, energy_prev=e)
        assert ctx.graph_end() == 1
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()


MULTI = [
    # N, X, D, na, nb, W, precision
    (24, 72, 18.0, 3, 3, 16, "fp64"),
    (40, 120, 30.0, 4, 3, 40, "mixed"),
    (64, 200, 48.0, 6, 6, 24, "mixed"),
]


def _make_multi(k):
    N, X, Dk, na, nb, Wk, prec = MULTI[k]
    print(f"[case] multi-system member {k}: N={N} X={X} na={na} nb={nb} W={Wk} {prec}", flush=True)
    h1e, chol, ecore, psi = synthetic_arrays(N, X, Dk, na, nb, seed=11 + k, scale=(12.0 / (N * Dk)) ** 0.5)
    init = random_walkers(N, na, nb, Wk, seed=20 + k, amp=0.1, base=psi)
    return Pair(h1e, chol, psi, (na, nb), Wk, init=init, precision=prec)


def test_multi_system_batch_equals_eager():
    """Three systems of different sizes and precisions in one graph, each on
    its own stream, against the same systems stepped eagerly."""
    import torch

    from ipie_amd.batch import MultiSystemBatch

    eager = [_make_multi(k) for k in range(len(MULTI))]
    batched = [_make_multi(k) for k in range(len(MULTI))]
    e_e = [torch.empty(p.W, 3, dtype=torch.complex128, device=p.walkers.device) for p in eager]
    e_b = [torch.empty_like(t) for t in e_e]
    shifts = [float(p.trial.energy.real) for p in eager]
    for k in range(len(MULTI)):
        _eager_block(eager[k], e_e[k], shifts[k])
        _eager_block(batched[k], e_b[k], shifts[k])
    batch = MultiSystemBatch([(p.prop, p.walkers, p.ham, p.trial) for p in batched])
    assert batch.capture(5, reortho_every=5, energies=e_b) == [5] * len(MULTI)
    batch.replay(shifts, replays=2)
    for k in range(len(MULTI)):
        _eager_block(eager[k], e_e[k], shifts[k])
        _eager_block(eager[k], e_e[k], shifts[k])
        _compare(eager[k], batched[k], e_e[k], e_b[k], f"system {k}")
        assert eager[k].prop.counter == batched[k].prop.counter
