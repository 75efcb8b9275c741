"""Autotuner of a device context's launch shapes.

Eight launch shapes of the step depend on the system's dimensions:

  fb_split       K ranges of the force-bias contraction (split-K; per spin in
                 mixed precision, in total in FP64)
  quad_wgs       workgroups the column units of the split-BF16 quadratic-form
                 exchange aim for, which sets the K-block length of a unit
  exx_split      Cholesky-range split of the T-form exchange (0: heuristic)
  quad_rgroup    walker row blocks grouped over the quadratic form's column
                 units (which workgroups run together and share L2)
  tile_fb        workgroup tile of the split-BF16 force-bias GEMM
  tile_vhs       workgroup tile of the split-BF16 VHS GEMM
  tile_one_body  workgroup tile of the split-BF16 one-body GEMM
                 (tiles: 0 = 128 x 160, 1 = 128 x 128, 2 = 128 x 80,
                 3 = 64 x 160, 4 = 128 x 320, 5 = 64 x 160 with two LDS
                 stages; walker rows x output columns)
  taylor         kernel of the split-BF16 Taylor series: 32-wide K stages, two
                 in LDS (0); 16-wide stages, four in LDS (1)

A context starts from heuristics (``afqmc_create``).  ``Autotuner.tune`` times
every admissible candidate of each knob that applies on the context's own
shapes, through entry points that change no walker state
(``afqmc_force_bias`` for fb_split and tile_fb, ``afqmc_local_energy`` for the
exchange knobs, or ``afqmc_control_variate`` for them in a context that runs
the learned control variate, whose step evaluates the single-product quadratic
form of every walker) or act on scratch copies of the walkers
(``afqmc_two_body`` for tile_vhs and taylor, ``afqmc_one_body`` for
tile_one_body), and keeps the fastest unless the heuristic is within
``min_gain`` of it.  Knobs that shape one launch (fb_split with tile_fb,
quad_wgs with quad_rgroup) are timed over every pair of candidates.  The
choices are cached per device and shape: a multi-system batch whose
systems share a shape tunes that shape once, and a JSON cache file carries the
choices to the next run.  The knobs change launch shapes and the grouping of
partial sums only, so results agree with the untuned context to rounding.

A context that runs the learned control variate computes its exchange inside
``afqmc_propagate``: the single-product quadratic form of every walker, timed
here through ``afqmc_control_variate``, and the exact form of a sampled
subset, which shares the quadratic form's launch shapes (its cache entries are
keyed apart).  ``PhaselessGeneric(autotune=...)`` tunes at build time.

No counterpart in ipie, whose GPU kernels take CuPy's and numba's shapes.
"""

import itertools
import json
import os
import time

import torch

KNOBS = ("fb_split", "quad_wgs", "exx_split", "quad_rgroup", "tile_fb", "tile_vhs", "tile_one_body", "taylor")
# a context running the learned control variate: the quadratic-form knobs are
# timed on its control variate (the T-form split is not used by it)
SURROGATE_KNOBS = ("fb_split", "quad_wgs", "quad_rgroup", "tile_fb", "tile_vhs", "tile_one_body", "taylor")
# knobs timed on scratch copies of the walkers
PHI_KNOBS = ("tile_vhs", "tile_one_body", "taylor")
# knobs that shape the same launch and are timed together, every pair of
# candidates (the force-bias split that suits one tile shape can leave another
# with too few workgroups)
JOINT = (("fb_split", "tile_fb"), ("quad_wgs", "quad_rgroup"))
DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "ipie_amd", "autotune.json")

_FB = (1, 2, 3, 4, 6, 8, 10, 12, 16, 20, 24, 32)
_QUAD = (512, 768, 1024, 1536, 2048, 3072, 4096, 6144)
_EXX = (1, 2, 4, 8, 16, 24, 32, 48, 64, 96, 128, 192, 256)
_TILE = (0, 1, 2, 3, 4, 5)
_RGROUP = (1, 2, 4, 8, 16, 32)


def _ceil(a, b):
    return -(-int(a) // int(b))


def _pow2(v):
    p = 1
    while p < v:
        p *= 2
    return p


def library_tag():
    """Size and modification time of the loaded library: cached choices of an
    older build (other kernels, other admissible ranges) are not reused."""
    from ipie_amd import _lib

    try:
        st = os.stat(_lib.LIB_PATH)
        return f"{st.st_size:x}-{int(st.st_mtime):x}"
    except OSError:
        return "none"


def shape_key(ctx, surrogate=False):
    """Cache key of a shape bucket: the device, the library build, and the
    classes of the dimensions the launch shapes depend on -- the basis padded
    to 64 (the K chunks of every contraction), the occupied orbitals per spin
    rounded up to 8, the Cholesky vectors in 256-column groups (the force-bias
    and VHS column tiles), the 128-walker row blocks rounded up to a power of
    two -- the precision, rhf and whether the context runs the learned control
    variate.  Systems of one bucket share tuned values (each value is checked
    against the context's own admissible range when applied)."""
    d = ctx.dims
    name = torch.cuda.get_device_name(ctx.device) if torch.cuda.is_available() else "cpu"
    return (f"{name}|lib={library_tag()}|N64={_ceil(d.nbasis, 64)}|n8={_ceil(d.nalpha, 8)},{_ceil(d.nbeta, 8)}"
            f"|X256={_ceil(d.nchol, 256)}|rb={_pow2(_ceil(d.nwalkers, 128))}|rhf={d.rhf}|{ctx.precision}"
            + ("|surrogate" if surrogate else ""))


def candidates(knob, value, lo, hi):
    """Values of `knob` to time: a fixed ladder within [lo, hi], plus the current value."""
    if hi < 1:
        return []
    ladder = {"fb_split": _FB, "quad_wgs": _QUAD, "exx_split": _EXX, "tile_fb": _TILE, "tile_vhs": _TILE,
              "tile_one_body": _TILE, "quad_rgroup": _RGROUP, "taylor": (0, 1)}[knob]
    out = {v for v in ladder if lo <= v <= hi}
    if knob == "exx_split":
        out.add(0)  # the heuristic
    if value == 0 or lo <= value <= hi:
        out.add(value)
    return sorted(out)


def choose(times, default, min_gain):
    """The fastest value, unless `default` is within a fraction min_gain of it."""
    best = min(times, key=lambda v: (times[v], v))
    if default in times and times[default] <= times[best] * (1.0 + min_gain):
        return default
    return best


class Autotuner:
    """Times and sets the launch-shape knobs of contexts; caches by shape.

    cache_path: JSON file read at construction and rewritten after each tune
    (None: in memory only).  repeats x inner calls are timed per candidate,
    after one warm-up call; the minimum over repeats counts.
    """

    def __init__(self, cache_path=None, repeats=3, inner=2, min_gain=0.02, knobs=KNOBS):
        self.cache_path = cache_path
        self.repeats = int(repeats)
        self.inner = int(inner)
        self.min_gain = float(min_gain)
        self.knobs = tuple(knobs)
        self.cache = {}
        self.reports = {}
        if cache_path and os.path.exists(cache_path):
            with open(cache_path) as fh:
                self.cache = json.load(fh)

    # ------------------------------------------------------------------ cache
    def _save(self):
        if not self.cache_path:
            return
        os.makedirs(os.path.dirname(os.path.abspath(self.cache_path)), exist_ok=True)
        tmp = self.cache_path + f".{os.getpid()}.tmp"
        with open(tmp, "w") as fh:
            json.dump(self.cache, fh, indent=1, sort_keys=True)
        os.replace(tmp, self.cache_path)

    def apply(self, ctx, surrogate=False):
        """Set the cached choices of ctx's shape bucket; returns them, or None when
        nothing usable is cached (no entry, or a cached value outside this
        context's admissible range or for a knob the entry lacks: then the
        caller tunes afresh)."""
        entry = self.cache.get(shape_key(ctx, surrogate))
        if entry is None:
            return None
        plan = {}
        for knob in self.knobs:
            if surrogate and knob not in SURROGATE_KNOBS:
                continue
            value, lo, hi = ctx.tune_get(knob)
            if hi < 1:
                continue
            if knob not in entry:
                return None
            v = int(entry[knob])
            if not (lo <= v <= hi or (knob == "exx_split" and v == 0)):
                return None
            plan[knob] = v
        for knob, v in plan.items():
            ctx.tune_set(knob, v)
        return dict(plan)

    # ------------------------------------------------------------------ timing
    def _time(self, ctx, fn):
        fn()
        ctx.sync()
        best = float("inf")
        for _ in range(self.repeats):
            t0 = time.perf_counter()
            for _ in range(self.inner):
                fn()
            ctx.sync()
            best = min(best, (time.perf_counter() - t0) / self.inner)
        return best * 1e3

    def tune(self, ctx, ghalfa, ghalfb, surrogate=False, phia=None, phib=None):
        """Time every candidate of each applicable knob with these Green's
        functions ([W, n_s, N] complex128 on ctx's device), set the choices on
        ctx, cache and return them ({knob: value}).

        surrogate: true when the context runs the learned control variate
        (then the exchange knobs are not tuned, see the module note).
        phia / phib: walkers ([W, N, n_s] complex128); the tile_vhs and
        tile_one_body knobs are timed on copies of them (skipped without)."""
        nw = ghalfa.shape[0]
        dev = ghalfa.device
        vbias = torch.empty(nw, ctx.dims.nchol, dtype=torch.complex128, device=dev)
        energy = torch.empty(nw, 3, dtype=torch.complex128, device=dev)
        cvbuf = torch.empty(nw, dtype=torch.complex128, device=dev)
        exchange = ((lambda: ctx.control_variate(ghalfa, ghalfb, cvbuf)) if surrogate
                    else (lambda: ctx.local_energy(ghalfa, ghalfb, energy)))
        calls = {
            "fb_split": lambda: ctx.force_bias(ghalfa, ghalfb, vbias),
            "tile_fb": lambda: ctx.force_bias(ghalfa, ghalfb, vbias),
            "quad_wgs": exchange,
            "quad_rgroup": exchange,
            "exx_split": exchange,
        }
        if phia is not None:
            pa, pb = phia.clone(), phib.clone()
            cmf = torch.empty(nw, dtype=torch.complex128, device=dev)
            cfb = torch.empty_like(cmf)
            calls["tile_vhs"] = lambda: ctx.two_body(pa, pb, ghalfa, ghalfb, cmf, cfb, seed=1, counter=0)
            calls["tile_one_body"] = lambda: ctx.one_body(pa, pb)
            calls["taylor"] = calls["tile_vhs"]
        chosen, report = {}, {}

        def tunable(knob):
            if (surrogate and knob not in SURROGATE_KNOBS) or knob not in calls:
                return None
            value, lo, hi = ctx.tune_get(knob)
            cands = candidates(knob, value, lo, hi)
            return (value, cands) if cands else None

        done = set()
        for knob in self.knobs:
            if knob in done:
                continue
            pair = next((g for g in JOINT if knob in g and all(k in self.knobs for k in g)), (knob,))
            group = [k for k in pair if k not in done and tunable(k) is not None]
            done.update(pair)
            if not group:
                continue
            found = {k: tunable(k) for k in group}
            default = tuple(found[k][0] for k in group)
            times = {}
            for combo in itertools.product(*(found[k][1] for k in group)):
                for k, v in zip(group, combo):
                    ctx.tune_set(k, v)
                times[combo] = self._time(ctx, calls[group[0]])
            pick = choose(times, default, self.min_gain)
            for i, k in enumerate(group):
                ctx.tune_set(k, pick[i])
                chosen[k] = pick[i]
                # the knob's candidates with the others at their chosen values
                line = {c[i]: t for c, t in times.items() if all(c[j] == pick[j] for j in range(len(group)) if j != i)}
                report[k] = {"default": default[i], "chosen": pick[i],
                             "ms": {str(v): round(t, 4) for v, t in sorted(line.items())}}
            if len(group) > 1:
                report[group[0]]["grid"] = {",".join(map(str, c)): round(t, 4) for c, t in times.items()}
        key = shape_key(ctx, bool(surrogate))
        self.cache[key] = chosen
        self.reports[key] = report
        self._save()
        return chosen

    def tune_walkers(self, ctx, walkers, surrogate=False, force=False):
        """Apply the cached choices for ctx's shape, or tune with the walkers'
        Green's function (computed into temporaries: the walkers are not
        changed).  surrogate as in tune; force=True tunes even when a cached
        entry exists."""
        if not force:
            got = self.apply(ctx, bool(surrogate))
            if got is not None:
                return got
        dev = walkers.phia.device
        ga = torch.empty_like(walkers.Ghalfa)
        gb = torch.empty_like(walkers.Ghalfb)
        ovlp = torch.empty(walkers.phia.shape[0], dtype=torch.complex128, device=dev)
        ctx.greens_function(walkers.phia, walkers.phib, ga, gb, ovlp)
        return self.tune(ctx, ga, gb, surrogate, phia=walkers.phia, phib=walkers.phib)


def as_tuner(spec):
    """None / False: no tuning; True: the default cache file; a path: that
    cache file; an Autotuner: itself."""
    if spec is None or spec is False:
        return None
    if isinstance(spec, Autotuner):
        return spec
    if spec is True:
        return Autotuner(cache_path=DEFAULT_CACHE)
    return Autotuner(cache_path=str(spec))
