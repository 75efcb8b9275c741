"""Parity of device population control (afqmc_pop_plan / afqmc_pop_gather and
the PopController drop-in) with oracle/pop_control.py, fed the same uniforms.

Bars: the source map src and the pair count bit-exact (integer decisions);
weights within 1e-12 relative (the total |w| is summed serially on the device,
pairwise by numpy.sum: the last bits may differ); copied walker rows
bit-exact."""

import numpy
import pytest

from oracle import pop_control as pc
from tests._helpers import Pair, random_walkers, synthetic_arrays

pytestmark = pytest.mark.gpu

WTOL = 1e-12


def _torch():
    import torch

    return torch


def _ctx():
    from ipie_amd import _lib

    return _lib.Context(8, 4, 2, 2, 4)


def _weights(nw, seed, kind):
    rs = numpy.random.RandomState(seed)
    if kind == "lognormal":
        return numpy.exp(1.5 * rs.normal(size=nw))
    if kind == "ties":  # few distinct values: the sort must keep slot order among equals
        return rs.choice([0.02, 0.5, 1.0, 6.0], size=nw)
    if kind == "signed":  # the reference sorts and sums |w|
        return numpy.exp(rs.normal(size=nw)) * rs.choice([-1.0, 1.0], size=nw)
    raise ValueError(kind)


def _plan(ctx, method, w, target, uniforms, min_w=0.1, max_w=4.0):
    torch = _torch()
    nw = len(w)
    dev = ctx.device
    wt = torch.from_numpy(numpy.ascontiguousarray(w, dtype=numpy.float64)).to(dev)
    u = None if uniforms is None else torch.from_numpy(numpy.asarray(uniforms, dtype=numpy.float64)).to(dev)
    src = torch.full((nw,), -7, dtype=torch.int32, device=dev)
    nwt = torch.empty(nw, dtype=torch.float64, device=dev)
    total = torch.empty(1, dtype=torch.float64, device=dev)
    npairs = torch.full((1,), -1, dtype=torch.int32, device=dev)
    ctx.pop_plan(method, wt, target, min_w, max_w, u, src, nwt, total, npairs)
    return src.cpu().numpy(), nwt.cpu().numpy(), float(total.item()), int(npairs.item())


def _rel(a, b):
    return float(numpy.abs(a - b).max() / max(numpy.abs(b).max(), 1e-300))


@pytest.mark.parametrize("nw", [2, 7, 64, 301, 1024, 2048, 4096])
@pytest.mark.parametrize("kind", ["lognormal", "ties", "signed"])
def test_pair_branch_plan(nw, kind):
    print(f"[case] pair_branch nw={nw} kind={kind}", flush=True)
    ctx = _ctx()
    w = _weights(nw, nw + 1, kind)
    total, _, scaled = pc.rescale(w, nw)
    u = numpy.random.RandomState(nw).rand(max(nw // 2, 1))
    src_o, nwt_o, np_o = pc.pair_branch_plan(scaled, 0.1, 4.0, u)
    src, nwt, tot, npairs = _plan(ctx, "pair_branch", w, nw, u)
    assert tot == pytest.approx(total, rel=1e-13)
    assert npairs == np_o
    assert numpy.array_equal(src, src_o)
    assert _rel(nwt, nwt_o) < WTOL


@pytest.mark.parametrize("nw", [1, 5, 50, 777, 4096])
def test_comb_plan(nw):
    print(f"[case] comb nw={nw}", flush=True)
    ctx = _ctx()
    w = _weights(nw, nw + 2, "lognormal")
    total, _, scaled = pc.rescale(w, nw, pairwise=False)
    r = numpy.random.RandomState(nw).rand()
    src_o, nwt_o, _ = pc.comb_plan(scaled, nw, r)
    src, nwt, tot, _ = _plan(ctx, "comb", w, nw, [r])
    assert tot == total
    assert numpy.array_equal(src, src_o)
    assert numpy.array_equal(nwt, nwt_o)


@pytest.mark.parametrize("nw", [1, 6, 100, 2048, 4096])
def test_reconfiguration_plan(nw):
    print(f"[case] reconfiguration nw={nw}", flush=True)
    ctx = _ctx()
    w = _weights(nw, nw + 3, "signed")
    _, _, scaled = pc.rescale(w, nw)
    zeta = numpy.random.RandomState(nw).rand()
    src_o, nwt_o = pc.reconfiguration_plan(scaled, zeta)
    src, nwt, _, _ = _plan(ctx, "stochastic_reconfiguration", w, nw, [zeta])
    assert numpy.array_equal(src, src_o)
    assert _rel(nwt, nwt_o) < WTOL


def test_rescale_only_and_zero_total():
    ctx = _ctx()
    w = _weights(33, 4, "signed")
    _, _, scaled = pc.rescale(w, 33)
    src, nwt, _, npairs = _plan(ctx, "rescale", w, 33, None)
    assert numpy.array_equal(src, numpy.arange(33)) and npairs == 0
    assert _rel(nwt, scaled) < WTOL
    # total below 1e-8: the total is reported and nothing else changes
    tiny = numpy.full(9, 1e-10)
    src, nwt, tot, npairs = _plan(ctx, "pair_branch", tiny, 9, numpy.zeros(4))
    assert tot == pytest.approx(9e-10, rel=1e-14)
    assert numpy.array_equal(src, numpy.arange(9)) and npairs == 0
    assert numpy.array_equal(nwt, tiny)


def test_plan_rejects_bad_arguments():
    from ipie_amd import _lib

    ctx = _ctx()
    w = numpy.ones(4097)
    with pytest.raises(_lib.AfqmcError):
        _plan(ctx, "rescale", w, 4097, None)
    with pytest.raises(_lib.AfqmcError):
        _plan(ctx, "comb", numpy.ones(4), 4, None)  # comb needs its uniform


@pytest.mark.parametrize("row", [(3, 5), (13,), (1,)])
def test_gather_in_and_out_of_place(row):
    torch = _torch()
    ctx = _ctx()
    nw = 50
    rs = numpy.random.RandomState(1)
    if len(row) == 2:
        host = rs.normal(size=(nw,) + row) + 1j * rs.normal(size=(nw,) + row)
    else:
        host = rs.randint(0, 255, size=(nw,) + row).astype(numpy.uint8)  # rows of odd byte counts
    # out of place: an arbitrary map (reconfiguration)
    src = numpy.sort(rs.randint(0, nw, size=nw)).astype(numpy.int32)
    t = torch.from_numpy(host).to(ctx.device)
    out = torch.empty_like(t)
    ctx.pop_gather(torch.from_numpy(src).to(ctx.device), t, out)
    assert numpy.array_equal(out.cpu().numpy(), host[src])
    # in place: no source row is a destination (pair_branch, comb)
    src2 = numpy.arange(nw, dtype=numpy.int32)
    src2[[3, 10, 11]] = [20, 20, 40]
    ctx.pop_gather(torch.from_numpy(src2).to(ctx.device), t, t)
    assert numpy.array_equal(t.cpu().numpy(), host[src2])


def _pair(nw):
    N, X, D, na, nb = 10, 30, 8.0, 4, 2
    h1e, chol, ecore, psi = synthetic_arrays(N, X, D, na, nb, scale=0.3)
    init = random_walkers(N, na, nb, nw, seed=5, amp=0.2, base=psi)
    return Pair(h1e, chol, psi, (na, nb), nw, init=init)


def _fields(walkers):
    return {name: getattr(walkers, name).cpu().numpy().copy() for name in walkers.buff_names}


@pytest.mark.parametrize("method,freq", [("pair_branch", 50), ("comb", 50), ("stochastic_reconfiguration", 1),
                                         ("stochastic_reconfiguration", 2)])
def test_pop_controller_drop_in(method, freq):
    """PopController.pop_control on device walkers against the oracle with the
    same numpy.random stream; the stream afterwards is where the reference
    leaves it (one draw per branched pair / comb / reconfiguration)."""
    print(f"[case] PopController {method} freq={freq}", flush=True)
    torch = _torch()
    from ipie_amd.pop_controller import PopController

    nw = 64
    p = _pair(nw)
    wk = p.walkers
    rs = numpy.random.RandomState(11)
    wk.weight.copy_(torch.from_numpy(numpy.exp(1.4 * rs.normal(size=nw))))
    wk.hybrid_energy.copy_(torch.from_numpy(rs.normal(size=nw) + 1j * rs.normal(size=nw)))
    wk.phase.copy_(torch.from_numpy(numpy.exp(1j * rs.normal(size=nw))))
    wk.log_ovlp.copy_(torch.from_numpy(rs.normal(size=nw)))
    before = _fields(wk)

    ctl = PopController(nw, 10, pop_control_method=method, reconfiguration_freq=freq)
    numpy.random.seed(123)
    if method == "pair_branch":
        u = numpy.random.rand(nw // 2)
    elif method == "comb":
        u = numpy.array([numpy.random.random()])
    else:
        u = numpy.array([numpy.random.rand()]) if freq == 1 else numpy.zeros(0)
    follow = numpy.random.rand(nw)  # what the stream yields afterwards
    numpy.random.seed(123)
    ctl.pop_control(wk, None)
    nxt = numpy.random.rand()

    want, total = pc.pop_control(before, method, nw, u, reconfigure=(freq == 1))
    got = _fields(wk)
    assert ctl.total_weight == pytest.approx(total, rel=1e-13)
    for name in wk.buff_names:
        if name == "weight":
            assert _rel(got[name], want[name]) < WTOL
        else:
            assert numpy.array_equal(got[name], want[name]), name
    if method == "pair_branch":
        _, _, npairs = pc.pair_branch_plan(pc.rescale(before["weight"], nw)[2], 0.1, 4.0, u)
        assert 0 < npairs < nw // 2
        assert nxt == u[npairs]
    else:  # one draw (comb, reconfiguration call) or none (rescale-only call)
        assert nxt == follow[0]


def test_walker_sums_and_clip():
    torch = _torch()
    from ipie_amd.walkers import WalkerAccumulator, clip_weights

    nw = 64
    p = _pair(nw)
    wk = p.walkers
    rs = numpy.random.RandomState(2)
    w = numpy.exp(rs.normal(size=nw))
    uw = numpy.exp(rs.normal(size=nw))
    eh = rs.normal(size=nw) + 1j * rs.normal(size=nw)
    wk.weight.copy_(torch.from_numpy(w))
    wk.unscaled_weight.copy_(torch.from_numpy(uw))
    wk.hybrid_energy.copy_(torch.from_numpy(eh))
    acc = WalkerAccumulator(["Weight", "WeightFactor", "HybridEnergy"], 5)
    acc.update(wk)
    acc.update(wk)
    want = 2 * numpy.array([w.sum(), uw.sum(), (w * eh).sum()])
    assert numpy.allclose(acc.buffer, want, rtol=1e-13, atol=0)
    acc.zero()
    assert numpy.array_equal(acc.buffer, numpy.zeros(3))
    # column 0 of an energy [W, 3] (stride 3); unscaled weights left out
    e = torch.from_numpy(rs.normal(size=(nw, 3)) + 1j * rs.normal(size=(nw, 3))).to(wk.weight.device)
    out = torch.zeros(3, dtype=torch.complex128, device=wk.weight.device)
    wk._engine.ctx.walker_sums(wk.weight, out, values=e)
    got = out.cpu().numpy()
    assert got[1] == 0
    assert numpy.allclose(got[[0, 2]], [w.sum(), (w * e.cpu().numpy()[:, 0]).sum()], rtol=1e-13, atol=0)
    # clip, NaN kept as numpy.clip keeps it
    w2 = rs.normal(size=nw) * 3.0
    w2[5] = numpy.nan
    wk.weight.copy_(torch.from_numpy(w2))
    clip_weights(wk, 1.5)
    got = wk.weight.cpu().numpy()
    want = numpy.clip(w2, -1.5, 1.5)
    assert numpy.array_equal(numpy.isnan(got), numpy.isnan(want))
    assert numpy.array_equal(got[~numpy.isnan(got)], want[~numpy.isnan(want)])
