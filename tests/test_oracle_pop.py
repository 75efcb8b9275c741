"""Pins of the population-control oracle (oracle/pop_control.py) on the CPU.

The reference's own population-control tests
(ipie/walkers/tests/test_population_control.py) compare whole propagations
of a random system against golden numbers; those runs are not reproducible
from the oracle alone, so the oracle is pinned here by what the methods must
satisfy whatever the implementation:

- rescaling: sum |w| after it equals the target; ValueError below 1e-8;
- pair_branch: weight conservation; unbiasedness (the expected weight carried
  by each original walker equals its weight, by an exact sweep of the
  uniform); the slot matching of the reference's messages, replayed literally
  (status table with rank indices, tag-0 FIFO of sends and receives);
- comb: parent counts sum to the target, each count is floor or ceil of
  w_i target / total, the expected count equals w_i target / total; the
  clone/kill zip;
- stochastic reconfiguration: copy counts floor or ceil of W w_i / total with
  that expectation, new weights total / W, a non-decreasing map.
"""

from collections import defaultdict, deque

import numpy
import pytest

from oracle import pop_control as pc


def _weights(nw, seed, spread=1.0):
    rs = numpy.random.RandomState(seed)
    return numpy.exp(spread * rs.normal(size=nw))


def test_rescale_hits_target_and_rejects_zero():
    w = _weights(37, 1)
    total, unscaled, scaled = pc.rescale(w, 37)
    assert total == pytest.approx(w.sum(), rel=1e-15)
    assert numpy.array_equal(unscaled, w)
    assert numpy.abs(scaled).sum() == pytest.approx(37.0, rel=1e-14)
    with pytest.raises(ValueError):
        pc.rescale(numpy.full(4, 1e-10), 4)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_pair_branch_conserves_weight(seed):
    w = pc.rescale(_weights(64, seed, spread=1.5), 64)[2]
    u = numpy.random.RandomState(seed + 10).rand(32)
    src, nwt, npairs = pc.pair_branch_plan(w, 0.1, 4.0, u)
    assert npairs > 0
    assert nwt.sum() == pytest.approx(w.sum(), rel=1e-13)
    moved = numpy.where(src != numpy.arange(64))[0]
    assert len(moved) == npairs
    # every copied slot holds a walker that was kept, with that walker's weight
    assert numpy.all(src[src] == src)
    assert numpy.array_equal(nwt[moved], nwt[src[moved]])


def _carried_weight(src, nwt, nw):
    c = numpy.zeros(nw)
    numpy.add.at(c, src, nwt)
    return c


def test_pair_branch_unbiased_exact_sweep():
    """E[weight carried by walker j] = w_j.  Each pair's decision depends only
    on its own uniform, so sweeping one value r for all pairs over a fine grid
    gives every walker's expectation to the grid's resolution."""
    w = pc.rescale(_weights(24, 5, spread=2.0), 24)[2]
    M = 4000
    acc = numpy.zeros(24)
    for m in range(M):
        r = (m + 0.5) / M
        src, nwt, _ = pc.pair_branch_plan(w, 0.1, 4.0, numpy.full(12, r))
        acc += _carried_weight(src, nwt, 24)
    assert numpy.allclose(acc / M, w, rtol=0, atol=2.0 / M * w.max())


def _pair_branch_by_messages(weight, min_weight, max_weight, uniforms, fields):
    """The reference's pair_branch for one rank, replayed step by step: the
    [|w|, status, initial rank, final rank] table, its sort and un-sort, then
    the sends (weight updated first) and receives with their tags through a
    FIFO per (source, destination, tag)."""
    nw = len(weight)
    rank = 0
    tab = numpy.zeros((nw, 4))
    tab[:, 0] = numpy.abs(weight)
    tab[:, 1] = 1
    tab[:, 2] = rank
    tab[:, 3] = rank
    sort = numpy.argsort(tab[:, 0], kind="mergesort")
    isort = numpy.argsort(sort, kind="mergesort")
    tab = tab[sort]
    s, e, k = 0, nw - 1, 0
    while s < e:
        if tab[s][0] < min_weight or tab[e][0] > max_weight:
            wab = tab[s][0] + tab[e][0]
            r = uniforms[k]
            k += 1
            if r < tab[e][0] / wab:
                tab[e][0], tab[e][1], tab[e][3] = 0.5 * wab, 2, tab[s][2]
                tab[s][0], tab[s][1], tab[s][3] = 0.0, 0, tab[e][2]
            else:
                tab[s][0], tab[s][1], tab[s][3] = 0.5 * wab, 2, tab[e][2]
                tab[e][0], tab[e][1], tab[e][3] = 0.0, 0, tab[s][2]
            s += 1
            e -= 1
        else:
            break
    tab = tab[isort]
    f = {key: numpy.array(v, copy=True) for key, v in fields.items()}
    wires = defaultdict(deque)
    for iw in range(nw):
        if tab[iw][1] > 1:
            f["weight"][iw] = tab[iw][0]
            tag = rank * nw + tab[iw][3]
            wires[(rank, int(round(tab[iw][3])), tag)].append({key: v[iw].copy() for key, v in f.items()})
    for iw in range(nw):
        if tab[iw][1] == 0:
            tag = tab[iw][3] * nw + rank
            buff = wires[(int(round(tab[iw][3])), rank, tag)].popleft()
            for key in f:
                f[key][iw] = buff[key]
    return f


@pytest.mark.parametrize("seed", [3, 4])
def test_pair_branch_matches_reference_messages(seed):
    nw = 40
    w = pc.rescale(_weights(nw, seed, spread=1.8), nw)[2]
    u = numpy.random.RandomState(seed).rand(nw // 2)
    rs = numpy.random.RandomState(seed + 1)
    fields = {"weight": w.copy(), "unscaled_weight": rs.rand(nw), "phia": rs.normal(size=(nw, 3, 2)) + 0j}
    want = _pair_branch_by_messages(w, 0.1, 4.0, u, fields)
    src, nwt, npairs = pc.pair_branch_plan(w, 0.1, 4.0, u)
    assert npairs >= 3
    got = pc.apply_plan(fields, src, nwt)
    for key in fields:
        assert numpy.array_equal(got[key], want[key]), key


def test_pair_branch_ties_keep_slot_order():
    # equal weights sort by slot (mergesort is stable): slots 0, 1 light, 2, 3 heavy
    w = numpy.array([0.05, 0.05, 1.95, 1.95])
    src, nwt, npairs = pc.pair_branch_plan(w, 0.1, 4.0, numpy.array([0.0, 0.99]))
    assert npairs == 2
    # pair (0, 3) branches 3 (r = 0 < 0.975), pair (1, 2) branches 1 (r = 0.99);
    # branched slots 1, 3 fill killed slots 0, 2 in slot order, across pairs
    assert list(src) == [1, 1, 3, 3]
    assert list(nwt) == [1.0, 1.0, 1.0, 1.0]


@pytest.mark.parametrize("seed", [0, 7])
def test_comb_counts(seed):
    nw = 50
    w = pc.rescale(_weights(nw, seed), nw, pairwise=False)[2]
    total = float(sum(w))
    expect = w * nw / total
    M = 2000
    mean = numpy.zeros(nw)
    for m in range(M):
        r = (m + 0.5) / M
        src, nwt, parent = pc.comb_plan(w, nw, r)
        assert parent.sum() == nw
        assert numpy.all(parent >= numpy.floor(expect) - 1e-9)
        assert numpy.all(parent <= numpy.ceil(expect) + 1e-9)
        assert numpy.array_equal(nwt, numpy.ones(nw))
        clone = numpy.where(parent > 1)[0]
        kill = numpy.where(parent == 0)[0]
        n = min(len(clone), len(kill))
        assert numpy.array_equal(src[kill[:n]], clone[:n])
        mean += parent
    assert numpy.allclose(mean / M, expect, atol=2.0 / M)


@pytest.mark.parametrize("seed", [1, 9])
def test_reconfiguration_counts(seed):
    nw = 40
    w = _weights(nw, seed, spread=1.2)
    total = numpy.cumsum(w)[-1]
    expect = nw * w / total
    M = 2000
    mean = numpy.zeros(nw)
    for m in range(M):
        zeta = (m + 0.5) / M
        src, nwt = pc.reconfiguration_plan(w, zeta)
        counts = numpy.bincount(src, minlength=nw)
        assert numpy.all(counts >= numpy.floor(expect) - 1e-9)
        assert numpy.all(counts <= numpy.ceil(expect) + 1e-9)
        assert numpy.all(numpy.diff(src) >= 0)
        assert numpy.allclose(nwt, total / nw, rtol=1e-15)
        mean += counts
    assert numpy.allclose(mean / M, expect, atol=2.0 / M)


def test_pop_control_keeps_unscaled_and_fields():
    nw = 16
    w = _weights(nw, 2, spread=2.0)
    fields = {"weight": w.copy(), "unscaled_weight": numpy.zeros(nw), "ident": numpy.arange(nw, dtype=float)}
    u = numpy.random.RandomState(0).rand(nw // 2)
    out, total = pc.pop_control(fields, "pair_branch", nw, u)
    assert total == pytest.approx(w.sum(), rel=1e-15)
    src = out["ident"].astype(int)
    # unscaled weights travel with their walker
    assert numpy.array_equal(out["unscaled_weight"], w[src])
    assert out["weight"].sum() == pytest.approx(nw, rel=1e-13)
    # SR outside its reconfiguration call only rescales
    out2, _ = pc.pop_control(fields, "stochastic_reconfiguration", nw, [0.3], reconfigure=False)
    assert numpy.array_equal(out2["ident"], fields["ident"])
    assert numpy.abs(out2["weight"]).sum() == pytest.approx(nw, rel=1e-14)
