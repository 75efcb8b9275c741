"""Tolerances of the mixed-precision GPU tests, derived on the CPU.

tests/_mixed_emu.py repeats the mixed mode's split-BF16 rounding (FP32
accumulation) around the oracle's FP64 step.  These tests pin the emulation
itself (BF16 rounding against known values, the split's error bound, and the
emulated step in exact arithmetic equal to the oracle's, so it drops no term)
and then check that every tolerance constant of tests/test_gpu_mixed.py and of
the full-shape mixed test (N400_TOL) sits between 2x and 25x the largest
error the emulation makes on the same cases: loose enough for a different FP32
summation order on the device, tight enough that a lost lo-part product
(~2^-8 relative, 100x more) fails.

Mixed-mode errors are max-norm relative over walkers (a per-walker relative
error on weights is dominated by walkers whose cosine factor is near zero),
except at N=400 where every weight of the checked walkers is O(1) and the
per-walker form is used.
"""

import copy
import functools

import numpy
import pytest

from oracle import afqmc
from tests import _mixed_emu as emu
from tests._helpers import random_walkers, rel_err, rel_err_elem, synthetic_arrays

# N=400, X=2400, n=40 (the headline shape): one mixed step against the FP64 oracle
# on 8 walkers (tests/test_gpu_fullsize.py); per walker for weight, overlap,
# hybrid energy and total energy, max-norm relative over the walkers for phi.
N400_TOL = {"phi": 5e-6, "weight": 5e-6, "ovlp": 1e-5, "hybrid_energy": 5e-6, "energy": 5e-8}

LOW, HIGH = 2.0, 25.0


def test_bf16_rounding_known_values():
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties go to the even mantissa
    x = numpy.array([1.0, 1.0 + 2.0**-8, 1.0 + 3 * 2.0**-8, 1.0 + 2.0**-8 + 2.0**-20, -3.140625, 0.0, 2.0**-130])
    want = numpy.array([1.0, 1.0, 1.0 + 2.0**-6, 1.0 + 2.0**-7, -3.140625, 0.0, 2.0**-130])
    assert numpy.array_equal(emu.bf16(x).astype(numpy.float64), want)


def test_split_error_bound():
    x = numpy.random.RandomState(1).normal(size=100000) * numpy.exp(numpy.random.RandomState(2).normal(size=100000))
    hi, lo = emu.split(x)
    err = numpy.abs(x - hi.astype(numpy.float64) - lo.astype(numpy.float64)) / numpy.abs(x)
    assert err.max() <= 2.0**-16
    assert numpy.median(err) > 2.0**-22  # the lo part is rounded too (the bound is not vacuous)
    # a product with the lo parts kept is ~2^8 more accurate than hi * hi alone
    a, b = numpy.random.RandomState(3).normal(size=(64, 200)), numpy.random.RandomState(4).normal(size=(200, 64))
    e3 = numpy.abs(emu.mm(a, b, 3) - a @ b).max()
    e1 = numpy.abs(emu.mm(a, b, 1) - a @ b).max()
    assert e1 > 30 * e3


def _state(case, W=None, seed=5, amp=0.2):
    N, X, D, na, nb, Wc, scale = case
    W = Wc if W is None else W
    h1e, chol, ecore, psi = synthetic_arrays(N, X, D, na, nb, scale=scale)
    init = random_walkers(N, na, nb, W, seed=seed, amp=amp, base=psi)
    o = afqmc.build_state(h1e, chol, psi, (na, nb), W, 0.005, ecore=ecore)
    o.walkers.phia = init[:, :, :na].copy()
    o.walkers.phib = init[:, :, na:].copy()
    return o


def _errors(o, exchange, xi):
    """Emulated errors of the quantities the mixed GPU tests check."""
    m = emu.MixedEmulator(o, exchange=exchange)
    w, t = o.walkers, o.trial
    t.calc_greens_function(w)
    e_ref = afqmc.local_energy_batch(o.system, o.hamiltonian, w, t)
    exx = (afqmc.exx_kernel_batch_real_rchol(t._rchola, w.Ghalfa)
           + afqmc.exx_kernel_batch_real_rchol(t._rcholb, w.Ghalfb))
    e = m.local_energy()
    out = {"exchange": float(numpy.abs(e[:, 2] - e_ref[:, 2]).max() / numpy.abs(exx).max()),
           "energy": rel_err_elem(e[:, 0], e_ref[:, 0])}
    vb = emu.force_bias(t._rchola, t._rcholb, w.Ghalfa, w.Ghalfb)
    out["force_bias"] = rel_err(vb, t.calc_force_bias(o.hamiltonian, w))
    ob = numpy.concatenate([emu.one_body(w.phia, o.propagator.expH1[0]), emu.one_body(w.phib, o.propagator.expH1[1])],
                           axis=2)
    ob_o = numpy.concatenate([afqmc.propagate_one_body(w.phia.copy(), o.propagator.expH1[0]),
                              afqmc.propagate_one_body(w.phib.copy(), o.propagator.expH1[1])], axis=2)
    out["one_body"] = rel_err(ob, ob_o)
    ref = copy.deepcopy(o)
    ref.propagator.xi_source = lambda nw, nx: xi
    ref.propagator.propagate_walkers(ref.walkers, ref.hamiltonian, ref.trial, 0.0)
    m.propagate(xi, 0.0)
    out["phi"] = max(rel_err(w.phia, ref.walkers.phia), rel_err(w.phib, ref.walkers.phib))
    for a in ("weight", "ovlp", "hybrid_energy"):
        out[a] = rel_err(getattr(w, a), getattr(ref.walkers, a))
        out[a + "_elem"] = rel_err_elem(getattr(w, a), getattr(ref.walkers, a))
    return out


def test_emulation_in_exact_arithmetic_is_the_oracle():
    """With the rounding switched off, the emulated energy and step equal the
    oracle's (both exchange forms): the emulation contracts the same indices."""
    from tests.test_gpu_mixed import CASES

    emu.EXACT = True
    try:
        for name, exchange in (("n96_n24_n12", "quad"), ("n48_x40", "t"), ("n64_n40", "quad")):
            o = _state(CASES[name])
            xi = numpy.random.RandomState(3).normal(size=(o.walkers.nwalkers, o.hamiltonian.nchol))
            errs = _errors(o, exchange, xi)
            for k, v in errs.items():
                assert v < 1e-11, (name, k, v)
    finally:
        emu.EXACT = False


@functools.lru_cache(maxsize=1)
def _mixed_case_errors():
    from tests.test_gpu_mixed import CASES

    res = {}
    for name, case in CASES.items():
        N, X = case[0], case[1]
        o = _state(case)
        xi = numpy.random.RandomState(3).normal(size=(o.walkers.nwalkers, X))
        res[name] = _errors(o, "quad" if N < X else "t", xi)
    return res


def _within(tol, err, what):
    ratio = tol / err
    print(f"{what}: tolerance {tol:.1e}, emulated {err:.2e}, ratio {ratio:.1f}")
    assert LOW <= ratio <= HIGH, (what, tol, err, ratio)


def test_mixed_test_tolerances():
    from tests import test_gpu_mixed as g

    errs = _mixed_case_errors()

    def worst(key):
        return max(e[key] for e in errs.values())

    _within(g.EXX_TOL, worst("exchange"), "exchange")
    _within(g.FB_TOL, worst("force_bias"), "force bias")
    _within(g.ONE_BODY_TOL, worst("one_body"), "one body")
    _within(g.PHI_TOL, worst("phi"), "step phi")
    _within(g.WEIGHT_TOL, max(worst("weight"), worst("ovlp"), worst("hybrid_energy")), "step weight/overlap/hybrid")


@pytest.mark.slow
def test_headline_shape_tolerances():
    """N=400, X=2400, n=40: the 8 walkers and fields of the full-shape GPU test.
    The exchange is emulated in the T-form (the quadratic form's Wq, 2 GB per
    spin, is left to the device); both forms round the same split operands and
    differ in the length of their FP32 sums, which the 25x ceiling covers."""
    from tests._big import SHAPES, arrays, spread_subset

    N, X, D, n, W = SHAPES["n400"]
    h1e, chol, ecore, psi = arrays("n400")
    sub = spread_subset(W)
    o = afqmc.build_state(h1e, chol, psi, (n, n), len(sub), 0.005, ecore=ecore)
    init = random_walkers(N, n, n, W, seed=3, amp=0.01, base=psi)[sub]
    o.walkers.phia = init[:, :, :n].copy()
    o.walkers.phib = init[:, :, n:].copy()
    o.walkers.ovlp = o.trial.calc_greens_function(o.walkers)
    xi = numpy.random.RandomState(22).normal(0.0, 1.0, X * W).reshape(W, X)[sub]
    errs = _errors(o, "t", xi)
    _within(N400_TOL["phi"], errs["phi"], "n400 phi")
    for a in ("weight", "ovlp", "hybrid_energy"):
        _within(N400_TOL[a], errs[a + "_elem"], f"n400 {a}")
    _within(N400_TOL["energy"], errs["energy"], "n400 energy")
