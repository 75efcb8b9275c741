"""Host logic of the launch-shape autotuner: candidate ladders, the choice
rule (the heuristic stays unless beaten by more than min_gain), the JSON
cache keyed by shape, and re-application from the cache without timing.  A
stand-in context with a cost model replaces the device (no GPU here)."""

import json

import pytest

from ipie_amd import autotune


class FakeDims:
    nbasis, nchol, nalpha, nbeta, nwalkers, rhf = 40, 120, 4, 4, 32, 0


class FakeCtx:
    """Knob ranges like a mixed-precision quadratic-form context; the cost of a
    call is a function of the current knob values."""

    device = "cpu"
    precision = "mixed"
    dims = FakeDims()

    def __init__(self):
        self.v = {"fb_split": 8, "quad_wgs": 1536, "exx_split": 0, "quad_rgroup": 1, "tile_fb": 0, "tile_vhs": 0,
                  "tile_one_body": 0, "taylor": 0}
        self.ranges = {"fb_split": (1, 12), "quad_wgs": (256, 8192), "exx_split": (0, 0), "quad_rgroup": (1, 1),
                       "tile_fb": (0, 3), "tile_vhs": (0, 3), "tile_one_body": (0, 3), "taylor": (0, 1)}
        self.calls = []

    def tune_get(self, knob):
        lo, hi = self.ranges[knob]
        return self.v[knob], lo, hi

    def tune_set(self, knob, value):
        lo, hi = self.ranges[knob]
        assert hi >= 1 and lo <= value <= hi, (knob, value)
        self.v[knob] = value

    def cost(self, knob):
        if knob == "fb_split":  # the force bias: its split and its tile
            return abs(self.v["fb_split"] - 3) + 1.0 + 0.5 * abs(self.v["tile_fb"] - 2)
        if knob == "vhs":  # the two-body step: the VHS tile and the Taylor stages
            return 1.0 + abs(self.v["tile_vhs"] - 1) + 0.5 * (1 - self.v["taylor"])
        if knob == "one_body":
            return 1.0 + abs(self.v["tile_one_body"] - 3)
        return abs(self.v["quad_wgs"] - 3072) / 1000.0 + 1.0


def test_candidates_cover_the_range_and_the_current_value():
    assert autotune.candidates("fb_split", 8, 1, 12) == [1, 2, 3, 4, 6, 8, 10, 12]
    assert autotune.candidates("fb_split", 7, 1, 12) == [1, 2, 3, 4, 6, 7, 8, 10, 12]
    assert autotune.candidates("quad_wgs", 1536, 256, 8192)[0] == 512
    assert 0 in autotune.candidates("exx_split", 0, 1, 40)
    assert max(autotune.candidates("exx_split", 0, 1, 40)) == 32
    assert autotune.candidates("quad_wgs", 0, 0, 0) == []


def test_choice_keeps_the_heuristic_unless_clearly_beaten():
    assert autotune.choose({1: 1.0, 2: 0.9, 8: 0.99}, 8, 0.02) == 2
    assert autotune.choose({1: 1.0, 2: 0.99, 8: 1.0}, 8, 0.02) == 8
    assert autotune.choose({1: 1.0, 2: 1.0}, 5, 0.02) == 1  # default not timed: the fastest, lowest on ties


def _fake_tune(monkeypatch, tuner, ctx):
    def fake_time(c, fn):
        return ctx.cost(fn())

    monkeypatch.setattr(tuner, "_time", fake_time)
    # the knob's entry point returns which cost applies
    ctx.force_bias = lambda ga, gb, vb: "fb_split"
    ctx.local_energy = lambda ga, gb, en: "quad_wgs"
    ctx.control_variate = lambda ga, gb, cv: "quad_wgs"
    ctx.two_body = lambda *args, **kw: "vhs"
    ctx.one_body = lambda pa, pb: "one_body"


def test_tune_sets_the_fastest_and_caches_by_shape(tmp_path, monkeypatch):
    torch = pytest.importorskip("torch")
    path = tmp_path / "cache.json"
    t = autotune.Autotuner(cache_path=str(path))
    ctx = FakeCtx()
    _fake_tune(monkeypatch, t, ctx)
    g = torch.zeros(32, 4, 40, dtype=torch.complex128)
    phi = torch.zeros(32, 40, 4, dtype=torch.complex128)
    chosen = t.tune(ctx, g, g, phia=phi, phib=phi)
    assert chosen == {"fb_split": 3, "quad_wgs": 3072, "quad_rgroup": 1, "tile_fb": 2, "tile_vhs": 1,
                      "tile_one_body": 3, "taylor": 1}
    assert ctx.v["fb_split"] == 3 and ctx.v["quad_wgs"] == 3072
    assert (ctx.v["tile_fb"], ctx.v["tile_vhs"], ctx.v["tile_one_body"]) == (2, 1, 3)
    assert t.reports[autotune.shape_key(ctx)]["tile_vhs"]["ms"] == {"0": 2.5, "1": 1.5, "2": 2.5, "3": 3.5}
    key = autotune.shape_key(ctx)
    assert "N64=1" in key and "X256=1" in key and "rb=1" in key and "lib=" in key and key.endswith("mixed")
    assert autotune.shape_key(ctx, True) == key + "|surrogate"
    assert json.loads(path.read_text())[key] == chosen
    rep = t.reports[key]["fb_split"]
    assert rep["default"] == 8 and rep["chosen"] == 3 and set(rep["ms"]) == {"1", "2", "3", "4", "6", "8", "10", "12"}
    # a new tuner applies the cache to a fresh context of the same shape
    t2 = autotune.Autotuner(cache_path=str(path))
    fresh = FakeCtx()
    assert t2.apply(fresh) == chosen
    assert fresh.v["fb_split"] == 3 and fresh.v["quad_wgs"] == 3072
    other = FakeCtx()
    other.dims = type("D", (), dict(vars(FakeDims), nwalkers=300))()
    assert t2.apply(other) is None  # another bucket (three 128-walker row blocks): nothing cached


def _dims(**kw):
    return type("D", (), dict(vars(FakeDims), **kw))()


def test_shape_buckets_group_similar_systems():
    a, b = FakeCtx(), FakeCtx()
    b.dims = _dims(nbasis=60, nchol=200, nalpha=6, nbeta=6, nwalkers=100)
    assert autotune.shape_key(a) == autotune.shape_key(b)  # same classes: one bucket
    for kw in (dict(nbasis=70), dict(nalpha=9), dict(nchol=300), dict(nwalkers=129), dict(rhf=1)):
        c = FakeCtx()
        c.dims = _dims(**kw)
        assert autotune.shape_key(c) != autotune.shape_key(a), kw
    # W = 300 and W = 500 both round up to four 128-walker row blocks
    c, d = FakeCtx(), FakeCtx()
    c.dims, d.dims = _dims(nwalkers=300), _dims(nwalkers=500)
    assert autotune.shape_key(c) == autotune.shape_key(d)


def test_cached_values_outside_the_admissible_range_force_a_retune(tmp_path):
    path = tmp_path / "cache.json"
    ctx = FakeCtx()
    key = autotune.shape_key(ctx)
    tiles = {"quad_rgroup": 1, "tile_fb": 1, "tile_vhs": 0, "tile_one_body": 2, "taylor": 1}
    path.write_text(json.dumps({key: {"fb_split": 40, "quad_wgs": 3072, **tiles}}))  # 40 > this context's 12
    t = autotune.Autotuner(cache_path=str(path))
    assert t.apply(ctx) is None
    assert ctx.v == {"fb_split": 8, "quad_wgs": 1536, "exx_split": 0, "quad_rgroup": 1, "tile_fb": 0, "tile_vhs": 0,
                     "tile_one_body": 0, "taylor": 0}  # nothing was set
    path.write_text(json.dumps({key: {"fb_split": 4, "quad_wgs": 2048}}))  # tile knobs missing from the entry
    assert autotune.Autotuner(cache_path=str(path)).apply(ctx) is None
    path.write_text(json.dumps({key: {"fb_split": 4, "quad_wgs": 2048, **tiles, "tile_vhs": 4}}))  # tile 4 > 3
    assert autotune.Autotuner(cache_path=str(path)).apply(ctx) is None
    path.write_text(json.dumps({key: {"fb_split": 4, "quad_wgs": 2048, **tiles}}))
    assert autotune.Autotuner(cache_path=str(path)).apply(ctx) == {"fb_split": 4, "quad_wgs": 2048, **tiles}
    assert ctx.v["fb_split"] == 4 and ctx.v["quad_wgs"] == 2048 and ctx.v["tile_one_body"] == 2


def test_key_changes_with_the_library_build(monkeypatch):
    ctx = FakeCtx()
    k1 = autotune.shape_key(ctx)
    monkeypatch.setattr(autotune, "library_tag", lambda: "other-build")
    assert autotune.shape_key(ctx) != k1


def test_as_tuner_specs(tmp_path):
    assert autotune.as_tuner(None) is None and autotune.as_tuner(False) is None
    t = autotune.Autotuner()
    assert autotune.as_tuner(t) is t
    p = str(tmp_path / "x.json")
    assert autotune.as_tuner(p).cache_path == p
    assert autotune.as_tuner(True).cache_path == autotune.DEFAULT_CACHE


def test_surrogate_contexts_time_the_exchange_on_the_control_variate(monkeypatch):
    torch = pytest.importorskip("torch")
    t = autotune.Autotuner()
    ctx = FakeCtx()
    _fake_tune(monkeypatch, t, ctx)
    g = torch.zeros(32, 4, 40, dtype=torch.complex128)
    calls = []
    ctx.local_energy = lambda ga, gb, en: calls.append("local_energy") or "quad_wgs"
    # no walkers given: the tiles timed on walker copies are skipped
    assert t.tune(ctx, g, g, surrogate=True) == {"fb_split": 3, "quad_wgs": 3072, "quad_rgroup": 1, "tile_fb": 2}
    assert calls == []  # the exchange shapes were timed through the control variate
    assert autotune.shape_key(ctx, True) in t.cache and autotune.shape_key(ctx) not in t.cache


def test_coupled_knobs_are_tuned_jointly(monkeypatch):
    """The force-bias split and tile shape: split 2 is best with tile 0, but
    tile 3 needs split 4 (a knob-by-knob search stops at (2, 0))."""
    torch = pytest.importorskip("torch")
    t = autotune.Autotuner(cache_path=None)
    ctx = FakeCtx()
    _fake_tune(monkeypatch, t, ctx)
    cost = {(2, 0): 1.0, (4, 0): 1.02, (2, 3): 1.5, (4, 3): 0.9}

    def fb_cost(knob, base=ctx.cost):
        if knob != "fb_split":
            return base(knob)
        return cost.get((ctx.v["fb_split"], ctx.v["tile_fb"]), 2.0)

    ctx.cost = fb_cost
    g = torch.zeros(32, 4, 40, dtype=torch.complex128)
    chosen = t.tune(ctx, g, g)
    assert (chosen["fb_split"], chosen["tile_fb"]) == (4, 3)
    rep = t.reports[autotune.shape_key(ctx)]
    assert rep["fb_split"]["ms"]["2"] == 1.5 and rep["tile_fb"]["ms"]["0"] == 1.02  # lines through the choice
    assert rep["fb_split"]["grid"]["4,3"] == 0.9
