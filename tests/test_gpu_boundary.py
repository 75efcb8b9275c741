"""The drop-in boundary on the device: per-walker status flags, the
PropagatorTimer fields filled from the fused step's phase times, and refusal
of strided views (the C ABI takes dense arrays only)."""

import numpy
import pytest

from tests._helpers import Pair, random_walkers, synthetic_arrays

pytestmark = pytest.mark.gpu


def _pair(W=6):
    N, X, D, na, nb = 24, 72, 12.0, 5, 4
    print(f"[case] boundary N={N} X={X} na={na} nb={nb} W={W}", flush=True)
    h1e, chol, ecore, psi = synthetic_arrays(N, X, D, na, nb, scale=(12.0 / (N * D)) ** 0.5)
    init = random_walkers(N, na, nb, W, seed=9, amp=0.1, base=psi)
    return Pair(h1e, chol, psi, (na, nb), W, init=init)


def test_per_walker_status_flags():
    import torch

    from ipie_amd import _lib
    from ipie_amd.engine import SingularOverlapError

    p = _pair()
    w = p.walkers
    assert torch.equal(w.status_flags(), torch.zeros(p.W, dtype=torch.int32, device=w.device))
    # walker 2: first occupied column zero -> its alpha overlap matrix is exactly singular
    w.phia[2, :, 0] = 0.0
    with pytest.raises(SingularOverlapError, match=r"walkers \[2\]"):
        p.trial.calc_greens_function(w)
    f = w.status_flags().cpu().numpy()
    assert f[2] == _lib.FLAG_SINGULAR and numpy.count_nonzero(f) == 1
    assert numpy.count_nonzero(w.status_flags().cpu().numpy()) == 0  # the take cleared them
    # walker 4: an infinite weight going in stays non-finite through the update
    w.phia.copy_(torch.from_numpy(numpy.ascontiguousarray(
        random_walkers(24, 5, 4, p.W, seed=9, amp=0.1, base=p.orc.trial.psi)[:, :, :5])))
    w.weight[4] = float("inf")
    p.prop.propagate_walkers(w, p.ham, p.trial, 0.0)
    f = w.status_flags(clear=False).cpu().numpy()
    assert f[4] & _lib.FLAG_NONFINITE and numpy.count_nonzero(f) == 1
    assert p.prop._engine.ctx.flags(clear=True) & _lib.FLAG_NONFINITE  # the OR-ed context flag agrees


def test_timer_fields_filled_by_fused_step():
    p = _pair()
    t = p.prop.timer
    assert all(getattr(t, f) == 0.0 for f in t.PHASES)
    for _ in range(3):
        p.prop.propagate_walkers(p.walkers, p.ham, p.trial, 0.0)
    vals = {f: getattr(t, f) for f in t.PHASES}
    print(vals)
    assert all(v > 0.0 for v in vals.values()), vals
    prof = p.prop._engine.ctx.profile_read()
    assert vals["tgf"] == pytest.approx(prof["gf"][0] * 1e-3)
    assert prof["gf"][1] == 3 and prof["overlap"][1] == 3 and prof["weight"][1] == 3
    t.tgf = 0.0
    assert t.tgf == 0.0
    p.prop.propagate_walkers(p.walkers, p.ham, p.trial, 0.0)
    assert 0.0 < t.tgf < vals["tgf"]
    # a reset drops phases recorded before it that were never read
    ctx = p.prop._engine.ctx
    for _ in range(2):
        p.prop.propagate_walkers(p.walkers, p.ham, p.trial, 0.0)
    ctx.profile(True)
    p.prop.propagate_walkers(p.walkers, p.ham, p.trial, 0.0)
    prof = ctx.profile_read()
    assert prof["gf"][1] == 1 and prof["overlap"][1] == 1 and prof["weight"][1] == 1, prof


def test_strided_views_are_refused():
    p = _pair()
    w = p.walkers
    ctx = p.prop._engine.ctx
    import torch

    ov = torch.empty(p.W, dtype=torch.complex128, device=w.device)
    strided = torch.empty(p.W, 5, 24, dtype=torch.complex128, device=w.device).transpose(1, 2)
    with pytest.raises(ValueError, match="contiguous"):
        ctx.greens_function(strided, w.phib, w.Ghalfa, w.Ghalfb, ov)
    with pytest.raises(ValueError, match="contiguous"):
        ctx.overlap(w.phia, w.phib, torch.empty(2 * p.W, dtype=torch.complex128, device=w.device)[::2])
