"""Launch-shape knobs and the autotuner.

Every admissible value of a knob gives the force bias and local energy of
the heuristic shapes: to 1e-12 in FP64 (only the grouping of FP64 partial
sums changes) and to 2e-5 in mixed precision (FP32 partial sums regrouped).
A retune after a graph was captured leaves that graph valid: it replays with
the shapes it was captured with, bit-exact against eager steps of an untuned
twin.  The autotuner picks admissible values, caches them by shape, and a
second tuner reading the cache applies them without timing anything."""

import pytest

from tests._helpers import Pair, random_walkers, rel_err, synthetic_arrays

pytestmark = pytest.mark.gpu

# (N, X, D, na, nb, W, precision): X > N gives the quadratic-form exchange,
# X < N the T form
CASES = {
    "fp64_quad": (48, 100, 30.0, 8, 6, 64, "fp64"),
    "mixed_quad": (48, 100, 30.0, 8, 6, 200, "mixed"),
    "fp64_t": (48, 40, 30.0, 8, 6, 64, "fp64"),
    "mixed_t": (48, 40, 30.0, 8, 6, 64, "mixed"),
}
TOL = {"fp64": 1e-12, "mixed": 2e-5}


def make(case, autotune=None, surrogate=None):
    N, X, D, na, nb, W, prec = CASES[case]
    print(f"[case] autotune {case}: N={N} X={X} na={na} nb={nb} W={W} {prec}", flush=True)
    h1e, chol, ecore, psi = synthetic_arrays(N, X, D, na, nb, scale=(12.0 / (N * D)) ** 0.5)
    init = random_walkers(N, na, nb, W, seed=8, amp=0.1, base=psi)
    return Pair(h1e, chol, psi, (na, nb), W, init=init, precision=prec, autotune=autotune, surrogate=surrogate)


def _outputs(p):
    import torch

    ctx = p.walkers._engine.ctx
    w = p.walkers
    ga, gb = torch.empty_like(w.Ghalfa), torch.empty_like(w.Ghalfb)
    ovlp = torch.empty_like(w.ovlp)
    ctx.greens_function(w.phia, w.phib, ga, gb, ovlp)
    vb = torch.empty(p.W, ctx.dims.nchol, dtype=torch.complex128, device=w.device)
    en = torch.empty(p.W, 3, dtype=torch.complex128, device=w.device)
    ctx.force_bias(ga, gb, vb)
    ctx.local_energy(ga, gb, en)
    torch.cuda.synchronize()
    return vb.cpu().numpy(), en.cpu().numpy()


@pytest.mark.parametrize("case", list(CASES))
def test_every_knob_value_agrees_with_the_heuristic(case):
    p = make(case)
    ctx = p.walkers._engine.ctx
    tol = TOL[CASES[case][-1]]
    vb0, en0 = _outputs(p)
    applied = 0
    for knob in ("fb_split", "quad_wgs", "exx_split", "quad_rgroup"):
        value, lo, hi = ctx.tune_get(knob)
        if hi < 1:
            continue
        trial_values = sorted({lo, hi, (lo + hi) // 2} | ({0} if knob == "exx_split" else set()))
        for v in trial_values:
            print(f"[case] {case} {knob}={v} (range {lo}-{hi}, default {value})", flush=True)
            ctx.tune_set(knob, v)
            assert ctx.tune_get(knob)[0] == v
            vb, en = _outputs(p)
            assert rel_err(vb, vb0) < tol, (knob, v, "vbias", rel_err(vb, vb0))
            assert rel_err(en, en0) < tol, (knob, v, "energy", rel_err(en, en0))
            applied += 1
        ctx.tune_set(knob, value)
    assert applied > 0
    expect = {"fp64_quad": {"fb_split"}, "mixed_quad": {"fb_split", "quad_wgs"},
              "fp64_t": {"fb_split", "exx_split"}, "mixed_t": {"fb_split", "exx_split"}}[case]
    assert {k for k in ("fb_split", "quad_wgs", "exx_split") if ctx.tune_get(k)[2] >= 1} == expect


@pytest.mark.parametrize("case", ["mixed_quad", "mixed_t"])
def test_every_tile_shape_gives_the_same_products(case):
    """The split-BF16 force-bias, VHS (two-body step) and one-body GEMMs in every
    workgroup tile shape: the shape changes which workgroup computes an output,
    not its K order, so the results equal those of shape 0."""
    import torch

    p = make(case)
    ctx = p.walkers._engine.ctx
    w = p.walkers
    ga, gb = torch.empty_like(w.Ghalfa), torch.empty_like(w.Ghalfb)
    ovlp = torch.empty_like(w.ovlp)
    ctx.greens_function(w.phia, w.phib, ga, gb, ovlp)

    def run():
        vb = torch.empty(p.W, ctx.dims.nchol, dtype=torch.complex128, device=w.device)
        ctx.force_bias(ga, gb, vb)
        pa, pb = w.phia.clone(), w.phib.clone()
        cmf = torch.empty(p.W, dtype=torch.complex128, device=w.device)
        cfb = torch.empty_like(cmf)
        ctx.two_body(pa, pb, ga, gb, cmf, cfb, seed=5, counter=3)
        qa, qb = w.phia.clone(), w.phib.clone()
        ctx.one_body(qa, qb)
        torch.cuda.synchronize()
        return [t.cpu().numpy() for t in (vb, pa, pb, cmf, cfb, qa, qb)]

    base = run()
    for knob in ("tile_fb", "tile_vhs", "tile_one_body"):
        value, lo, hi = ctx.tune_get(knob)
        assert (value, lo, hi) == (0, 0, 5)
        for v in range(lo, hi + 1):
            print(f"[case] {case} {knob}={v}", flush=True)
            ctx.tune_set(knob, v)
            for i, (a, b) in enumerate(zip(run(), base)):
                assert rel_err(a, b) < 1e-13, (knob, v, i, rel_err(a, b))
        ctx.tune_set(knob, 0)
    # the Taylor series in 16-wide K stages: half-size MFMAs, so the FP32 sums
    # are grouped differently (agreement to FP32 rounding of the terms)
    assert ctx.tune_get("taylor") == (0, 0, 1)
    for v, tol in ((1, 2e-6),):
        ctx.tune_set("taylor", v)
        print(f"[case] {case} taylor={v}", flush=True)
        for i, (a, b) in enumerate(zip(run(), base)):
            assert rel_err(a, b) < tol, ("taylor", v, i, rel_err(a, b))
    ctx.tune_set("taylor", 0)


def test_bad_values_are_refused():
    import torch

    from ipie_amd._lib import AfqmcError

    p = make("fp64_quad")
    ctx = p.walkers._engine.ctx
    value, lo, hi = ctx.tune_get("fb_split")
    with pytest.raises(AfqmcError):
        ctx.tune_set("fb_split", hi + 1)
    with pytest.raises(AfqmcError):
        ctx.tune_set("fb_split", 0)
    with pytest.raises(AfqmcError):
        ctx.tune_set("quad_wgs", 1536)  # FP64 context: the knob does not apply
    assert ctx.tune_get("tile_vhs")[2] == 0  # FP64 context: no split-BF16 GEMM tiles
    with pytest.raises(AfqmcError):
        ctx.tune_set("tile_vhs", 1)
    assert ctx.tune_get("fb_split")[0] == value
    e = torch.empty(p.W, 3, dtype=torch.complex128, device=p.walkers.device)
    p.prop.propagate_walkers(p.walkers, p.ham, p.trial, 0.0, energy_prev=e)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        ctx.graph_begin()
        with pytest.raises(AfqmcError):
            ctx.tune_set("fb_split", lo)
        ctx.propagate(p.walkers.phia, p.walkers.phib, p.walkers.Ghalfa, p.walkers.Ghalfb, p.walkers.weight,
                      p.walkers.ovlp, p.walkers.hybrid_energy, 0.0, seed=p.prop.seed, counter=0, energy_prev=e)
        assert ctx.graph_end() == 1
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()


def _block(p, energy, eshift, n=5):
    for _ in range(n):
        p.prop.propagate_walkers(p.walkers, p.ham, p.trial, eshift, energy_prev=energy)
    p.walkers.reortho()


FIELDS = ("phia", "phib", "weight", "ovlp", "hybrid_energy")


def test_retune_keeps_a_captured_graph_valid():
    """Capture, then grow the force-bias split past its buffer and rebuild the
    quadratic-form units: the graph replays with its captured shapes (bit-exact
    against the untuned twin), and eager steps take the new shapes."""
    import torch

    a, b = make("mixed_quad"), make("mixed_quad")
    ea = torch.empty(a.W, 3, dtype=torch.complex128, device=a.walkers.device)
    eb = torch.empty_like(ea)
    eshift = float(a.trial.energy.real)
    _block(a, ea, eshift)
    _block(b, eb, eshift)
    b.prop.capture_steps(b.walkers, b.ham, b.trial, 5, reortho_every=5, energy_prev=eb)
    ctx = b.walkers._engine.ctx
    _, _, fb_hi = ctx.tune_get("fb_split")
    q0, q_lo, _ = ctx.tune_get("quad_wgs")
    ctx.tune_set("fb_split", fb_hi)
    ctx.tune_set("quad_wgs", q_lo if q0 != q_lo else q0 * 2)
    _block(a, ea, eshift)
    b.prop.replay_steps(b.walkers, eshift, 1)
    torch.cuda.synchronize()
    for k in FIELDS:
        err = rel_err(getattr(b.walkers, k).cpu().numpy(), getattr(a.walkers, k).cpu().numpy())
        assert err < 1e-12, ("replay after retune", k, err)
    assert rel_err(eb.cpu().numpy(), ea.cpu().numpy()) < 1e-12
    # eager steps with the new shapes stay within the mixed-precision tolerance
    _block(a, ea, eshift, n=1)
    _block(b, eb, eshift, n=1)
    torch.cuda.synchronize()
    assert rel_err(eb.cpu().numpy(), ea.cpu().numpy()) < 1e-4


def test_autotuner_picks_caches_and_reapplies(tmp_path):
    from ipie_amd.autotune import KNOBS, Autotuner, shape_key

    path = str(tmp_path / "tune.json")
    t1 = Autotuner(cache_path=path, repeats=2, inner=1)
    p = make("mixed_quad", autotune=t1)
    ctx = p.walkers._engine.ctx
    chosen = p.prop.tuning
    assert set(chosen) == {"fb_split", "quad_wgs", "quad_rgroup", "tile_fb", "tile_vhs", "tile_one_body", "taylor"}
    for knob, v in chosen.items():
        value, lo, hi = ctx.tune_get(knob)
        assert value == v and lo <= v <= hi
    key = shape_key(ctx)
    rep = t1.reports[key]
    assert all(str(rep[k]["chosen"]) in rep[k]["ms"] for k in rep)
    t2 = Autotuner(cache_path=path)
    assert t2.cache[key] == chosen

    def no_timing(*a, **k):
        raise AssertionError("a cached shape must not be timed again")

    t2.tune = no_timing
    q = make("mixed_quad", autotune=t2)
    assert q.prop.tuning == chosen
    assert all(q.walkers._engine.ctx.tune_get(k)[0] == v for k, v in chosen.items())
    assert set(chosen) <= set(KNOBS)
    # the tuned context steps like the untuned one (mixed-precision tolerance)
    import torch

    u = make("mixed_quad")
    e1 = torch.empty(p.W, 3, dtype=torch.complex128, device=p.walkers.device)
    e2 = torch.empty_like(e1)
    _block(p, e1, 0.0, n=2)
    _block(u, e2, 0.0, n=2)
    torch.cuda.synchronize()
    assert rel_err(e1.cpu().numpy(), e2.cpu().numpy()) < 1e-4


def test_surrogate_context_tunes_on_its_control_variate():
    """A context running the learned control variate times its exchange shapes
    on the control variate (afqmc_control_variate; no step runs, the model is
    not trained); its run matches an untuned twin to the mixed-precision
    tolerance, with the control variate trained alike."""
    import torch

    from ipie_amd.autotune import Autotuner, shape_key

    sur = dict(warmup=3, fraction=0.125, train_iters=3, lr=2e-3, seed=5)
    t = Autotuner(cache_path=None, repeats=1, inner=1)
    p = make("mixed_quad", autotune=t, surrogate=sur)
    ctx = p.walkers._engine.ctx
    assert set(p.prop.tuning) == {"fb_split", "quad_wgs", "quad_rgroup", "tile_fb", "tile_vhs", "tile_one_body",
                                  "taylor"}
    assert shape_key(ctx, True) in t.cache and shape_key(ctx, False) not in t.cache
    assert p.prop.surrogate_stats()["steps"] == 0
    u = make("mixed_quad", surrogate=sur)
    e1 = torch.empty(p.W, 3, dtype=torch.complex128, device=p.walkers.device)
    e2 = torch.empty_like(e1)
    _block(p, e1, 0.0, n=5)
    _block(u, e2, 0.0, n=5)
    torch.cuda.synchronize()
    assert p.prop.surrogate_stats()["steps"] == u.prop.surrogate_stats()["steps"]
    assert rel_err(e1.cpu().numpy(), e2.cpu().numpy()) < 1e-4


def test_shape_buckets_tune_once_and_batch():
    """Systems built with one shared tuner: a shape is timed once and later
    systems of that shape reuse its choices (a bucket); the tuned systems then
    run in one multi-system graph, equal to stepping them eagerly."""
    import torch

    from ipie_amd.autotune import Autotuner
    from ipie_amd.batch import MultiSystemBatch

    t = Autotuner(cache_path=None, repeats=1, inner=1)
    timed = []
    real_tune = t.tune

    def counting_tune(ctx, ga, gb, surrogate=False, **walkers):
        timed.append(ctx.describe())
        return real_tune(ctx, ga, gb, surrogate, **walkers)

    t.tune = counting_tune
    cases = ["mixed_quad", "mixed_quad", "fp64_t"]
    tuned = [make(c, autotune=t) for c in cases]
    assert len(timed) == 2 and len(t.cache) == 2  # two shapes, three systems
    assert tuned[0].prop.tuning == tuned[1].prop.tuning
    eager = [make(c) for c in cases]
    for q, p in zip(eager, tuned):
        q.walkers.phia.copy_(p.walkers.phia)
        q.walkers.phib.copy_(p.walkers.phib)
    e_t = [torch.empty(p.W, 3, dtype=torch.complex128, device=p.walkers.device) for p in tuned]
    e_e = [torch.empty_like(x) for x in e_t]
    for p, e in zip(tuned + eager, e_t + e_e):
        _block(p, e, 0.0)
    batch = MultiSystemBatch([(p.prop, p.walkers, p.ham, p.trial) for p in tuned])
    batch.capture(5, reortho_every=5, energies=e_t)
    batch.replay([0.0] * len(tuned), replays=1)
    for p, e in zip(eager, e_e):
        _block(p, e, 0.0)
    torch.cuda.synchronize()
    for k, (p, q) in enumerate(zip(tuned, eager)):
        tol = TOL[CASES[cases[k]][-1]] * 5
        for attr in FIELDS:
            err = rel_err(getattr(p.walkers, attr).cpu().numpy(), getattr(q.walkers, attr).cpu().numpy())
            assert err < tol, (cases[k], attr, err)
