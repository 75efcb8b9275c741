"""Throughput of the MI355X phaseless-AFQMC step (walker-steps/s).

One step = one full ``propagate_walkers`` (GF, exp(-dt/2 H1), force bias,
fields, VHS, 6-term Taylor exp(VHS), exp(-dt/2 H1), overlap, weights) plus
the local energy of the walkers (e1b + Coulomb + exchange, from the step's own
Green's function) plus the QR re-orthogonalisation every 5 steps.  Inputs are
resident in HBM before the timed region; the timed region is K consecutive
steps bracketed by device synchronisation.

Prints one JSON line.  ``--gpus`` above 1 is refused: this build runs one GPU.
"""

import argparse
import json
import math
import os
import sys
import time

import numpy

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "walker-steps/s at N=400,X=2400: FP64 vs mixed+surrogate at <=0.1 mHa bias"
PEAK = {"f64_mfma_tflops": 78.6, "bf16_mfma_tflops": 2500.0, "fp8_mfma_tflops": 5000.0, "hbm_gbs": 8000.0}
REORTHO_EVERY = 5
SURROGATE_WARMUP = 100
BIAS_BUDGET_HA = 1e-4
# a control-variate mode qualifies for the headline only if the variance it adds to
# the per-step energy stays below this fraction of the step's own variance
VARIANCE_BUDGET_FRAC = 0.01
# modes run by default per workload: the headline compares the FP64 step with the
# mixed-precision modes (the control variate with its quadratic term in FP8 and
# in BF16: the FP8 form is faster but its 2% rounding of Q leaves a larger
# residual, so its paired bias estimate is noisier and may miss the budget by
# chance, the BF16 form then heads); the surrogate workload compares the control
# variate (with and without its MLP term, with its quadratic term in BF16
# instead of FP8, and without it) with the all-exact mixed-precision estimator
# (the ablations also run at n400 when named with --modes)
DEFAULT_MODES = {
    "n400": ("fp64", "mixed", "mixed+surrogate", "mixed+surrogate-bf16q"),
    "n300": ("mixed", "mixed+surrogate", "mixed+surrogate-bf16q", "mixed+surrogate-nomlp", "mixed+surrogate-linear"),
}


def quadratic_mode(mode):
    """The control variate's quadratic term for a surrogate mode: "none"
    (-linear), "bf16" (-bf16q) or "fp8"."""
    return "none" if mode.endswith("-linear") else ("bf16" if mode.endswith("-bf16q") else "fp8")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def step_flops(N, X, na, nb):
    """Algorithmic real flops of one walker-step (both spins, UHF), as ipie performs them.

    complex x complex multiply-add = 8 flops, real x complex = 4."""
    f = 0.0
    for n in (na, nb):
        f += 12 * N * n * n + 11 * n**3  # GF: O = phi^T psi*, LU/inverse, Ghalf
        f += 2 * 8 * N * N * n  # two one-body products
        f += 4 * X * n * N  # force bias
        f += 48 * N * N * n  # 6 Taylor terms
        f += 4 * N * n * n + 3 * n**3  # overlap
        f += 8 * n * N  # e1b
        f += 4 * X * n * N  # Coulomb contraction
        f += X * (4 * n * n * N + 8 * n * n)  # exchange
        f += 8 * N * n * n / REORTHO_EVERY  # QR
    f += 4 * N * (N + 1) / 2 * X  # VHS
    f += 8 * X  # Coulomb square
    return f


def exchange_flops_per_launch(N, X, n, W, form="t"):
    """Real flops of one exchange launch (one spin, W walkers) in the form the
    context uses: 't' forms every T_x = L_x Ghalf^T (ipie's contraction), 'quad'
    is Y = G Wq (complex G, real Wq of size (nN)^2) plus a dot per walker,
    'quad_sym' the same quadratic form over the (nN)(nN + 1) / 2 distinct
    entries of the symmetric Wq (the block-triangular split-BF16 form)."""
    if form == "quad":
        return float(W) * (4.0 * (n * N) ** 2 + 8.0 * n * N)
    if form == "quad_sym":
        return float(W) * (2.0 * n * N * (n * N + 1) + 8.0 * n * N)
    return float(X) * W * (4 * n * n * N + 8 * n * n)


def build_host(w, scale):
    from ipie_amd import Generic, synthetic
    from ipie_amd.hamiltonian import GenericRealChol
    from ipie_amd.trial import SingleDet

    t0 = time.time()
    h1e, chol, ecore = synthetic.make_hamiltonian(w.nbasis, w.nchol, w.decay, seed=w.seed)
    if scale != 1.0:
        chol *= scale
    psi = synthetic.make_trial(h1e, w.nalpha, w.nbeta)
    log(f"[bench] generated workload {w.name} in {time.time() - t0:.1f}s")
    t0 = time.time()
    system = Generic((w.nalpha, w.nbeta))
    ham = GenericRealChol(h1e, chol, ecore=ecore)
    trial = SingleDet(psi, (w.nalpha, w.nbeta), w.nbasis)
    trial.half_rotate(ham)
    trial.calculate_energy(system, ham)
    log(f"[bench] host set-up in {time.time() - t0:.1f}s, E_trial={trial.energy.real:.6f}")
    return dict(h1e=h1e, chol=chol, ecore=ecore, psi=psi, system=system, ham=ham, trial=trial)


def build_device(host, w, mode, tuner=None):
    import torch

    from ipie_amd.propagator import PhaselessGeneric
    from ipie_amd.walkers import UHFWalkers

    precision = "fp64" if mode == "fp64" else "mixed"
    surrogate = None
    if "surrogate" in mode:
        # "-nomlp": the physical control variate alone (train_iters = 0 turns the MLP term off);
        # "-linear": without its quadratic term Q(dG); "-bf16q": Q in BF16 instead of FP8
        surrogate = dict(fraction=0.125, warmup=SURROGATE_WARMUP, train_iters=0 if mode.endswith("-nomlp") else 4,
                         lr=1e-3, quadratic=quadratic_mode(mode))
    t0 = time.time()
    walkers = UHFWalkers(host["psi"], w.nalpha, w.nbeta, w.nbasis, w.nwalkers)
    walkers.build(host["trial"])
    prop = PhaselessGeneric(w.dt, seed=w.seed, precision=precision, surrogate=surrogate, autotune=tuner)
    prop.build(host["ham"], host["trial"], walkers)
    torch.cuda.synchronize()
    log(f"[bench] {mode}: device context + upload in {time.time() - t0:.1f}s"
        + (f", launch shapes {prop.tuning}" if prop.tuning is not None else ""))
    return dict(host, walkers=walkers, prop=prop)


def run_setup_timing(w, scale):
    """Set-up from the raw arrays until the propagator is built: on the host
    (packing, h1e_mod, half rotation, mean-field shift, expH1 in numpy / scipy,
    then upload) against the device set-up kernels (upload of chol, then
    afqmc_setup_*).  Prints one JSON line."""
    import torch

    from ipie_amd import synthetic
    from ipie_amd.hamiltonian import GenericRealChol
    from ipie_amd.propagator import PhaselessGeneric
    from ipie_amd.trial import SingleDet
    from ipie_amd.walkers import UHFWalkers

    h1e, chol, ecore = synthetic.make_hamiltonian(w.nbasis, w.nchol, w.decay, seed=w.seed)
    chol *= scale
    psi = synthetic.make_trial(h1e, w.nalpha, w.nbeta)
    times = {}
    for mode in ("host", "device", "host", "device"):  # second round timed (first one warms up)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ham = GenericRealChol(h1e, chol, ecore=ecore, host_setup=(mode == "host"))
        trial = SingleDet(psi, (w.nalpha, w.nbeta), w.nbasis)
        walkers = UHFWalkers(psi, w.nalpha, w.nbeta, w.nbasis, 8)
        prop = PhaselessGeneric(w.dt, precision="fp64")
        prop.build(ham, trial, walkers, device_setup=(mode == "device"))
        torch.cuda.synchronize()
        times[mode] = time.perf_counter() - t0
        log(f"[bench] set-up on the {mode}: {times[mode]:.2f}s")
        del ham, trial, walkers, prop
        torch.cuda.empty_cache()
    print(json.dumps({"metric": "set-up seconds from raw arrays to a built propagator", "workload": w.name,
                      "N": w.nbasis, "X": w.nchol, "host_s": times["host"], "device_s": times["device"],
                      "speedup": times["host"] / times["device"],
                      "host_threads": int(os.environ.get("OMP_NUM_THREADS", "0") or 0)}), flush=True)


def shape_key(ctx, surrogate):
    from ipie_amd.autotune import shape_key as key

    return key(ctx, surrogate)


def make_tuner(args):
    """The launch-shape autotuner (in memory: every run tunes afresh), or None."""
    if args.no_autotune:
        return None
    from ipie_amd.autotune import Autotuner

    return Autotuner(cache_path=None)


def run_multi(args, emit=True):
    """The uneven multi-system batch: 16 independent systems (N 64-256, X = 5 N,
    N / 10 occupied per spin, 64-512 walkers, mixed precision, local energy
    every step, QR every 5 steps), stepped (a) eagerly one system after
    another on one stream, (b) from one step graph per system, replayed one
    system after another, (c) from one multi-system graph with every system on
    its own stream.  Prints one JSON line; the value is (c)."""
    import torch

    from ipie_amd import synthetic
    from ipie_amd.batch import MultiSystemBatch

    ws = synthetic.multi_system_workloads()
    K = max(args.steps // REORTHO_EVERY, 1) * REORTHO_EVERY
    tuner = make_tuner(args)
    systems = []
    for w in ws:
        host = build_host(w, math.sqrt(12.0 / (w.nbasis * w.decay)))
        st = build_device(host, w, "mixed", tuner)
        st["energy"] = torch.empty(w.nwalkers, 3, dtype=torch.complex128, device=st["walkers"].device)
        st["eshift"] = float(host["trial"].energy.real)
        st["w"] = w
        systems.append(st)
    total_w = sum(w.nwalkers for w in ws)
    log(f"[bench] multi: {len(ws)} systems, N {min(w.nbasis for w in ws)}-{max(w.nbasis for w in ws)}, "
        f"{total_w} walkers")

    def eager_block(st):
        for _ in range(REORTHO_EVERY):
            st["prop"].propagate_walkers(st["walkers"], st["ham"], st["trial"], st["eshift"],
                                         energy_prev=st["energy"])
        st["walkers"].reortho()

    for st in systems:  # warm-up: workspace, then one block
        eager_block(st)
    torch.cuda.synchronize()
    # (a) eager, one system after another
    t0 = time.perf_counter()
    for _ in range(K // REORTHO_EVERY):
        for st in systems:
            eager_block(st)
    torch.cuda.synchronize()
    t_eager = time.perf_counter() - t0
    log(f"[bench] multi eager sequential: {K} steps in {t_eager:.3f}s")
    # (b) one graph per system, replayed one system after another
    for st in systems:
        st["prop"].capture_steps(st["walkers"], st["ham"], st["trial"], REORTHO_EVERY, reortho_every=REORTHO_EVERY,
                                 energy_prev=st["energy"])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(K // REORTHO_EVERY):
        for st in systems:
            st["prop"].replay_steps(st["walkers"], st["eshift"], 1)
    torch.cuda.synchronize()
    t_graphs = time.perf_counter() - t0
    log(f"[bench] multi per-system graphs: {K} steps in {t_graphs:.3f}s")
    # (c) one graph for all systems, each on its own stream
    batch = MultiSystemBatch([(st["prop"], st["walkers"], st["ham"], st["trial"]) for st in systems])
    batch.capture(REORTHO_EVERY, reortho_every=REORTHO_EVERY, energies=[st["energy"] for st in systems])
    torch.cuda.synchronize()
    shifts = [st["eshift"] for st in systems]
    t0 = time.perf_counter()
    batch.replay(shifts, replays=K // REORTHO_EVERY)
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    log(f"[bench] multi batched graph: {K} steps in {t_batch:.3f}s")
    finite = all(bool(torch.isfinite(st["walkers"].weight).all()) for st in systems)
    rate = lambda t: total_w * K / t  # noqa: E731
    out = {
        "metric": "aggregate walker-steps/s, 16 uneven systems in one step graph",
        "value": rate(t_batch),
        "unit": "walker-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": REORTHO_EVERY,
        "ms_per_step": 1e3 * t_batch / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (Green's function, weights, one-body, QR) + bf16x3 force bias, VHS, Taylor, exchange",
        "data": "synthetic",
        "config": {
            "workload": "multi16: 16 systems, N uniform [64, 256], X = 5N, N/10 occupied per spin, walkers uniform "
                        "[64, 512], D = 0.75 N, seed 7",
            "systems": [dict(N=w.nbasis, X=w.nchol, n=w.nalpha, W=w.nwalkers) for w in ws],
            "total_walkers": total_w,
            "parallelism": "single GPU, one stream per system inside one graph",
        },
        "sequential_eager_value": rate(t_eager),
        "sequential_graphs_value": rate(t_graphs),
        "speedup_vs_sequential_eager": t_eager / t_batch,
        "finite": finite,
        "launch_shapes": [st["prop"].tuning for st in systems] if tuner is not None else "heuristic",
    }
    if tuner is not None:
        out["autotune_buckets"] = len(tuner.reports)
    del systems, batch
    torch.cuda.empty_cache()
    if emit:
        print(json.dumps(out), flush=True)
    return out


def run_surrogate_workload(args, tuner):
    """The surrogate workload (N=300, X=1800, n=30, W=2048): the all-exact
    mixed-precision estimator, the control variate with its MLP term and the
    control variate alone, on one trajectory, over one block of at least 100
    steps after the 100 training steps.  Returns one record: rates, and for
    each control-variate mode its bias against the all-exact run, the variance
    it adds to the per-step energy, and the share of that variance the MLP
    removes."""
    import torch

    from ipie_amd import synthetic

    w = synthetic.WORKLOADS["n300"]
    host = build_host(w, math.sqrt(12.0 / (w.nbasis * w.decay)))
    K = max(args.steps, 100)
    modes = run_modes(args, w, host, list(DEFAULT_MODES["n300"]), K, SURROGATE_WARMUP + 5, tuner)
    compare_modes(modes, "mixed")
    keep = ("value", "ms_per_step", "timed", "energy_bias_Ha", "energy_bias_sem_Ha", "added_variance_per_step_Ha2",
            "reference_step_variance_Ha2", "within_0p1_mHa", "within_variance_budget", "surrogate", "roofline_kernel")
    rec = {
        "workload": f"{w.name}: N={w.nbasis} X={w.nchol} decay={w.decay} nalpha={w.nalpha} nbeta={w.nbeta} "
                    f"W={w.nwalkers} dt={w.dt}, normalised Cholesky scale",
        "steps": K,
        "warmup": SURROGATE_WARMUP + 5,
        "modes": {m: {k: r[k] for k in keep if k in r} for m, r in modes.items()},
    }
    a = modes.get("mixed+surrogate", {}).get("added_variance_per_step_Ha2")
    b = modes.get("mixed+surrogate-nomlp", {}).get("added_variance_per_step_Ha2")
    if a is not None and b:
        rec["mlp_variance_reduction"] = 1.0 - a / b
    if "mixed" in modes:
        rec["surrogate_speedup_vs_exact"] = modes.get("mixed+surrogate", {}).get("value", 0.0) / modes["mixed"]["value"]
    del host
    torch.cuda.empty_cache()
    return rec


def run_graph(st, w, steps, one_step, counter, energy, eshift):
    """`steps` steps replayed from one step graph: a block of REORTHO_EVERY
    propagate calls (each followed by the device sums of the weighted local
    energy of the incoming walkers) and a reortho, recorded once and launched
    steps / REORTHO_EVERY times.  Returns (seconds, per-step weighted energies)."""
    import torch

    walkers, prop = st["walkers"], st["prop"]
    ctx = walkers._engine.ctx
    dev = walkers.device
    while counter[0] % REORTHO_EVERY:  # blocks start right after a reortho
        one_step()
    nb = max(steps // REORTHO_EVERY, 1)
    acc = torch.zeros(REORTHO_EVERY, 3, dtype=torch.complex128, device=dev)
    series = torch.zeros(nb, REORTHO_EVERY, 3, dtype=torch.complex128, device=dev)
    wprev = torch.empty_like(walkers.weight)
    s = torch.cuda.Stream(device=dev)
    s.wait_stream(torch.cuda.current_stream(dev))
    with torch.cuda.stream(s):
        ctx.graph_begin()
        for j in range(REORTHO_EVERY):
            wprev.copy_(walkers.weight)
            ctx.propagate(walkers.phia, walkers.phib, walkers.Ghalfa, walkers.Ghalfb, walkers.weight, walkers.ovlp,
                          walkers.hybrid_energy, eshift, seed=prop.seed, counter=0, log_shift=walkers.log_shift,
                          energy_prev=energy)
            ctx.walker_sums(wprev, acc[j], values=energy)
        walkers.reortho()
        nrec = ctx.graph_end()
    assert nrec == REORTHO_EVERY, nrec
    torch.cuda.synchronize()
    log(f"[bench] step graph recorded ({nrec} steps + reortho); {nb} launches")
    t0 = time.perf_counter()
    with torch.cuda.stream(s):
        for r in range(nb):
            ctx.graph_launch(1, prop.counter + 1 + r * REORTHO_EVERY, eshift)
            series[r].copy_(acc)
            acc.zero_()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    prop.counter += nb * REORTHO_EVERY
    counter[0] += nb * REORTHO_EVERY
    ser = series.cpu().numpy().reshape(-1, 3)
    e_steps = ser[:, 2] / ser[:, 0].real
    log(f"[bench] graph: {nb * REORTHO_EVERY} steps in {elapsed:.3f}s; block E={numpy.mean(e_steps.real):.6f}")
    return elapsed, e_steps, nb * REORTHO_EVERY


def run_gpu(st, w, steps, warmup, graph=True):
    """Warm-up, then `steps` eager steps timed with per-phase HIP events, then
    (graph=True) the same number of steps replayed from a step graph.  Returns
    (eager seconds, per-phase profile, eager per-step weighted energies, finite
    flag, graph result (seconds, energies, steps) or None)."""
    import torch

    walkers, prop, ham, trial = st["walkers"], st["prop"], st["ham"], st["trial"]
    ctx = walkers._engine.ctx
    energy = torch.empty(w.nwalkers, 3, dtype=torch.complex128, device=walkers.device)
    eshift = float(trial.energy.real)
    counter = [0]

    def one_step():
        prop.propagate_walkers(walkers, ham, trial, eshift, energy_prev=energy)
        counter[0] += 1
        if counter[0] % REORTHO_EVERY == 0:
            walkers.reortho()

    for i in range(warmup):
        one_step()
        if (i + 1) % 25 == 0:
            log(f"[bench] warmup step {i + 1}/{warmup}")
    torch.cuda.synchronize()
    log(f"[bench] warmup done ({warmup} steps)")
    # weighted energy numerator / denominator of every timed step, kept on the device
    # (energy_prev is the energy of the walkers entering the step, so it pairs
    # with the weights before the step's update)
    numer = torch.zeros(steps, dtype=torch.complex128, device=walkers.device)
    denom = torch.zeros(steps, dtype=torch.float64, device=walkers.device)
    wprev = torch.empty_like(walkers.weight)
    ctx.profile(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        wprev.copy_(walkers.weight)
        one_step()
        numer[i] = torch.sum(wprev * energy[:, 0])
        denom[i] = torch.sum(wprev)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    prof = ctx.profile_read()
    ctx.profile(False)
    e_steps = (numer / denom).cpu().numpy()
    log(f"[bench] eager: {steps} steps in {elapsed:.3f}s; block E={numpy.mean(e_steps.real):.6f}")
    g = run_graph(st, w, steps, one_step, counter, energy, eshift) if graph else None
    wts = walkers.weight.cpu().numpy()
    finite = bool(numpy.all(numpy.isfinite(e_steps)) and numpy.all(numpy.isfinite(wts)))
    if g is not None:
        finite = finite and bool(numpy.all(numpy.isfinite(g[1])))
    return elapsed, prof, e_steps, finite, g


def cpu_baseline(st, w, sample_walkers, sample_steps):
    """Time the oracle (numpy restatement of ipie's CPU path) on a bounded sample,
    with every host thread the BLAS has and again with one thread."""
    try:
        from threadpoolctl import threadpool_limits
    except Exception:
        threadpool_limits = None
    rec = _cpu_baseline_once(st, w, sample_walkers, sample_steps)
    if threadpool_limits is not None:
        with threadpool_limits(limits=1):
            one = _cpu_baseline_once(st, w, sample_walkers, sample_steps)
        rec["single_core"] = {k: one[k] for k in ("value", "unit", "cores", "sample")}
    return rec


def _cpu_baseline_once(st, w, sample_walkers, sample_steps):
    from oracle import afqmc

    try:
        from threadpoolctl import threadpool_info

        info = threadpool_info()
        threads = max((d.get("num_threads", 1) for d in info), default=1)
        blas = ",".join(sorted({d.get("internal_api", "?") for d in info}))
    except Exception:
        threads, blas = os.cpu_count(), "unknown"
    t0 = time.time()
    orc = afqmc.build_state(st["h1e"], st["chol"], st["psi"], (w.nalpha, w.nbeta), sample_walkers, w.dt,
                            ecore=st["ecore"])
    log(f"[bench] oracle set-up {time.time() - t0:.1f}s")
    rs = numpy.random.RandomState(w.seed)
    orc.propagator.xi_source = lambda nw, nx: rs.normal(0.0, 1.0, nw * nx).reshape(nw, nx)
    eshift = float(orc.trial.energy.real)
    t0 = time.perf_counter()
    for _ in range(sample_steps):
        orc.propagator.propagate_walkers(orc.walkers, orc.hamiltonian, orc.trial, eshift)
        orc.trial.calc_greens_function(orc.walkers)
        afqmc.local_energy_batch(orc.system, orc.hamiltonian, orc.walkers, orc.trial)
    t_steps = time.perf_counter() - t0
    t0 = time.perf_counter()
    orc.walkers.reortho()
    t_qr = time.perf_counter() - t0
    per_step = t_steps / sample_steps + t_qr / REORTHO_EVERY
    rate = sample_walkers / per_step
    return {
        "value": rate,
        "unit": "walker-steps/s",
        "cores": int(threads),
        "kind": "port",
        "sample": f"{sample_walkers} walkers x {sample_steps} steps of {w.name} (propagate + GF + local energy, "
                  f"+1/{REORTHO_EVERY} reortho), numpy/{blas} {threads} threads, {per_step:.2f}s per step",
    }


def traffic_per_launch(mode, w):
    """HBM bytes per launch of the roofline kernel from the committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE per dispatch, each scaled by the same counter on a
    plain copy of known size), or None when it was not measured for this mode and
    workload."""
    path = os.path.join(ROOT, "profiles", "exchange_n400_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        line = line.strip()
        if line.startswith("{"):
            rec = json.loads(line)
            if rec.get("mode") == mode and rec.get("workload") == w.name and rec.get("nwalkers") == w.nwalkers:
                return rec["hbm_bytes_per_launch"]
    return None


def phase_rooflines(mode, phases, w):
    """Algorithmic rate of the step's other large phases (ms per step from the
    phase events of the eager pass; each phase is its kernels alone, launched
    back to back): flops as ipie performs them (step_flops) against the MFMA
    peak of the arithmetic the phase runs in, per algorithmic flop (split-BF16
    x3: a third of the BF16 peak; the Taylor series: three 3-product and three
    1-product terms, half of it)."""
    N, X, W = w.nbasis, w.nchol, w.nwalkers
    nsum = w.nalpha + w.nbeta
    f64 = mode == "fp64"
    bf3 = PEAK["f64_mfma_tflops"] if f64 else PEAK["bf16_mfma_tflops"] / 3.0
    rows = [
        ("vhs", "VHS, one launch per step (fp64: vhs_kernel; mixed: bf16_gemm_kernel<true, true, true> writing "
                "the split atoms)", 4.0 * X * N * (N + 1) / 2 * W, bf3),
        ("taylor", "Taylor series, six terms (fp64: batched_nn_kernel; mixed: taylor_bf16_kernel, three 3-product "
                   "and three 1-product terms)", 6 * 8.0 * N * N * nsum * W,
         PEAK["f64_mfma_tflops"] if f64 else PEAK["bf16_mfma_tflops"] / 2.0),
        ("one_body", "one body, two applications per step (mixed: bf16_gemm_kernel<true, true, true>, ADD epilogue, "
                     "and the transposed split of phi)", 2 * 8.0 * N * N * nsum * W, bf3),
        ("force_bias", "force bias (mixed: bf16_gemm_kernel<true, false, true>, split-K partials)",
         4.0 * X * N * nsum * W, bf3),
        ("gf", "Green's function: O, LU with the inverse, Ghalf (FP64 MFMA and one wave per LU)",
         sum(12.0 * N * n * n + 11.0 * n**3 for n in (w.nalpha, w.nbeta)) * W, PEAK["f64_mfma_tflops"]),
    ]
    out = []
    for key, name, fl, peak in rows:
        ms = phases.get(key)
        if not ms:
            continue
        ach = fl / (ms / 1e3) / 1e12
        out.append({"phase": key, "kernel": name, "ms_per_step": ms, "flops_per_step": fl, "achieved": ach,
                    "peak": peak, "unit": "TFLOP/s", "frac": ach / peak})
    return out


def compare_modes(modes, ref_mode):
    """Per-step weighted energies of every mode against ref_mode on the same
    trajectory (every timed step, eager and graph blocks): bias (mean
    difference) with its standard error, the variance the mode adds to the
    per-step estimate, and whether the bias is within 0.1 mHa."""
    if ref_mode not in modes:
        return
    ref = modes[ref_mode]["_e_steps"]
    ref_var = float(numpy.var(ref.real))
    for mode, rec in modes.items():
        if mode == ref_mode:
            continue
        d = (rec["_e_steps"] - ref).real
        rec["bias_reference_mode"] = ref_mode
        rec["bias_paired_steps"] = int(len(d))
        rec["energy_bias_Ha"] = float(numpy.mean(d))
        rec["energy_bias_sem_Ha"] = float(numpy.std(d) / numpy.sqrt(len(d)))
        rec["energy_maxdiff_per_step_Ha"] = float(numpy.abs(d).max())
        rec["added_variance_per_step_Ha2"] = float(numpy.var(d))
        rec["reference_step_variance_Ha2"] = ref_var
        rec["within_0p1_mHa"] = bool(abs(rec["energy_bias_Ha"]) <= BIAS_BUDGET_HA)
        rec["within_variance_budget"] = bool(rec["added_variance_per_step_Ha2"] <= VARIANCE_BUDGET_FRAC * ref_var)
        if ref_mode == "fp64":
            rec["energy_bias_vs_fp64_Ha"] = rec["energy_bias_Ha"]


def run_modes(args, w, host, mode_list, K, warmup, tuner):
    """Build and step every mode on one workload (same trajectory: the fields come
    from one counter-based stream and every mode runs the same steps before its
    timed block); per-mode records, each with its per-step energies (_e_steps)."""
    import torch

    modes = {}
    for mode in mode_list:
        st = build_device(host, w, mode, tuner)
        elapsed, prof, e_steps, finite, g = run_gpu(st, w, K, warmup, graph=not args.no_graph)
        log(f"[bench] {w.name} {mode}: done")
        ex_ms, ex_n = prof["exchange"]
        form = st["walkers"]._engine.ctx.exchange_form()
        if form == "quad" and mode != "fp64":
            form = "quad_sym"  # the split-BF16 planes hold the block-triangular Wq
        fl = 0.5 * (exchange_flops_per_launch(w.nbasis, w.nchol, w.nalpha, w.nwalkers, form)
                    + exchange_flops_per_launch(w.nbasis, w.nchol, w.nbeta, w.nwalkers, form))
        fl_t = 0.5 * (exchange_flops_per_launch(w.nbasis, w.nchol, w.nalpha, w.nwalkers, "t")
                      + exchange_flops_per_launch(w.nbasis, w.nchol, w.nbeta, w.nwalkers, "t"))
        ex_avg_s = ex_ms / max(ex_n, 1) / 1e3
        rec = dict(
            value=w.nwalkers * K / elapsed,
            ms_per_step=1e3 * elapsed / K,
            timed="eager",
            exchange_avg_launch_ms=ex_avg_s * 1e3,
            exchange_launches=ex_n,
            phases_ms_per_step={k: round(v[0] / K, 3) for k, v in prof.items()},
            block_energy=float(numpy.mean(e_steps.real)),
            finite=finite,
            _e_steps=e_steps,
            launch_shapes=st["prop"].tuning if tuner is not None else "heuristic",
        )
        if tuner is not None:
            rep = tuner.reports.get(shape_key(st["walkers"]._engine.ctx, "surrogate" in mode))
            if rep:
                rec["autotune_ms"] = rep
        if g is not None:
            g_s, g_e, g_k = g
            rec["eager_value"] = rec["value"]
            rec["eager_ms_per_step"] = rec["ms_per_step"]
            # the bias and variance comparisons pair both blocks (eager, then graph:
            # 2K consecutive steps of the one trajectory every mode follows)
            rec.update(value=w.nwalkers * g_k / g_s, ms_per_step=1e3 * g_s / g_k, timed="step graph",
                       graph_steps=g_k, block_energy=float(numpy.mean(g_e.real)),
                       _e_steps=numpy.concatenate([e_steps, g_e]))
        rec["exchange_form"] = form
        # the same launch in ipie's form (every T_x formed): the work it replaces
        rec["exchange_flops_reference_form_per_launch"] = fl_t
        if "surrogate" not in mode:
            rec["exchange_tflops"] = fl / ex_avg_s / 1e12
            rec["exchange_flops_per_launch"] = fl
            if form == "quad":
                kname = "Y = G Wq on FP64 MFMA (force_bias_kernel NT mainloop) + quad_reduce_kernel"
            elif form == "quad_sym":
                kname = ("Y = G W' over block-triangular W' split-BF16 x3 (bf16_gemm_kernel<true, false, true>, g.Y "
                         "in the epilogue) + quad_sum_kernel")
            else:
                kname = ("exchange_tiled_kernel (FP64" if mode == "fp64" else "exchange_bf16_kernel (split-BF16 x3")
                kname += ", every T_x formed)"
            rec["roofline_kernel"] = dict(
                kernel="exchange, one launch per spin: " + kname,
                avg_launch_ms=ex_avg_s * 1e3, flops_per_launch=fl, achieved=fl / ex_avg_s / 1e12,
                peak=PEAK["f64_mfma_tflops"] if mode == "fp64" else PEAK["bf16_mfma_tflops"] / 3.0,
                peak_note="FP64 MFMA 78.6 TFLOP/s" if mode == "fp64" else
                "dense BF16 MFMA 2.5 PFLOP/s / 3 products per algorithmic flop")
        elif mode.endswith("-linear"):
            # no control-variate GEMM: the dominant kernel is the VHS GEMM (one
            # launch per step), VHS[w] = i sqrt(dt) sum_x x[w, x] L_x on the
            # upper triangle: 4 X N(N+1)/2 real flops per walker, x3 BF16 products
            v_ms, v_n = prof["vhs"]
            v_avg_s = v_ms / max(v_n, 1) / 1e3
            vfl = 4.0 * w.nchol * w.nbasis * (w.nbasis + 1) / 2 * w.nwalkers
            rec["roofline_kernel"] = dict(
                kernel="VHS of every walker, one launch per step: split-BF16 x3 GEMM over the Cholesky planes in "
                       "16x16 atom order (bf16_gemm_kernel<true, true, true>, atom epilogue) + the split of the "
                       "shifted fields into planes",
                avg_launch_ms=v_avg_s * 1e3, flops_per_launch=vfl, achieved=vfl / v_avg_s / 1e12,
                peak=PEAK["bf16_mfma_tflops"] / 3.0,
                peak_note="dense BF16 MFMA 2.5 PFLOP/s / 3 products per algorithmic flop (useful upper-triangle "
                          "pairs only; the atom order computes 3.7% more at N=400)")
        if "surrogate" in mode:
            ss = st["prop"].surrogate_stats()
            rec["surrogate"] = {
                "fraction_exact": 0.125,
                "warmup_steps": SURROGATE_WARMUP,
                "last_subset": int(ss["last_k"]),
                "quadratic_term": quadratic_mode(mode),
                # mean |exact - g|^2 on the last subset (a constant offset in it cancels in the estimate)
                "mean_sq_residual_Ha2": ss["resid2"],
                # warm-up variance of exact - (E_T + 2<K,dG> [+ Q]), the MLP's target
                "control_variate_residual_var_Ha2": ss["target_var"],
                "mlp_gain": ss["gain"],
            }
        if "surrogate" in mode and not mode.endswith("-linear"):
            # the dominant kernel of this mode: the single-product exchange of the
            # control variate over all walkers (its own profiling phase)
            cv_ms, cv_n = prof["exchange_cv"]
            cv_avg_s = cv_ms / max(cv_n, 1) / 1e3
            fp8 = form == "quad_sym" and quadratic_mode(mode) == "fp8"
            if fp8:
                cv_kernel = ("Q = dG^T Wq dG over block-triangular W', single FP8 (E4M3) product, dG scaled per "
                             "walker and W' per spin by powers of two (bf16_gemm_kernel<false, false, true, Tile0, "
                             "true>, dG.Y in the epilogue) + quad_sum_kernel")
            elif form == "quad_sym":
                cv_kernel = ("Q = dG^T Wq dG over block-triangular W', single BF16 product (bf16_gemm_kernel<false, "
                             "false, true>, dG.Y in the epilogue) + quad_sum_kernel")
            else:
                cv_kernel = "exchange_bf16_kernel<false> (single BF16 product, every dT_x formed)"
            rec["roofline_kernel"] = dict(
                kernel="control-variate exchange of every walker, one launch per spin: " + cv_kernel,
                avg_launch_ms=cv_avg_s * 1e3, flops_per_launch=fl, achieved=fl / cv_avg_s / 1e12,
                peak=PEAK["fp8_mfma_tflops"] if fp8 else PEAK["bf16_mfma_tflops"],
                peak_note=("dense FP8 MFMA 5.0 PFLOP/s, one product per flop" if fp8
                           else "dense BF16 MFMA 2.5 PFLOP/s, one product per flop"))
        modes[mode] = rec
        del st
        torch.cuda.empty_cache()
    return modes


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="n400")
    ap.add_argument("--nwalkers", type=int, default=None)
    ap.add_argument("--cholesky-scale", default="normalised", choices=["normalised", "as-specified"])
    ap.add_argument("--modes", default="",
                    help="comma list of fp64, mixed, mixed+surrogate, mixed+surrogate-nomlp (default: per workload, "
                    "n400 fp64,mixed,mixed+surrogate; n300 mixed,mixed+surrogate,mixed+surrogate-nomlp); the "
                    "headline value is the fastest mode whose block energy is within 0.1 mHa of the FP64 mode "
                    "(or, without it, the all-exact mixed mode) on the same trajectory")
    ap.add_argument("--no-graph", action="store_true",
                    help="time eager steps only (default: the value is the step-graph replay, the per-phase "
                         "profile and the roofline kernel time come from the eager pass before it)")
    ap.add_argument("--setup-timing", action="store_true",
                    help="only time the set-up of the workload on the host and on the device, then exit")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-secondary", action="store_true",
                    help="with the n400 headline, skip the secondary records (the surrogate workload n300 at 2048 "
                    "walkers and the 16-system batch multi16)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the heuristic launch shapes (default: tune them per context at build time)")
    ap.add_argument("--cpu-walkers", type=int, default=4)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--calibration-copy-mb", type=int, default=0,
                    help="only copy a device buffer of this many MiB once and exit (the reference for the "
                         "FETCH_SIZE / WRITE_SIZE counters behind roofline.traffic)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs one GPU only (asked for --gpus {args.gpus}); sharding walkers over several "
              "GPUs is a design that is not built here.", file=sys.stderr)
        sys.exit(2)

    import dataclasses

    import torch

    from ipie_amd import synthetic

    if args.calibration_copy_mb > 0:
        n = args.calibration_copy_mb * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibration_copy_bytes": 8 * n}), flush=True)
        return

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.workload == "multi16":
        run_multi(args)
        return
    w = synthetic.WORKLOADS[args.workload]
    if args.nwalkers:
        w = dataclasses.replace(w, nwalkers=args.nwalkers)
    scale = 1.0 if args.cholesky_scale == "as-specified" else math.sqrt(12.0 / (w.nbasis * w.decay))
    if args.setup_timing:
        run_setup_timing(w, scale)
        return
    host = build_host(w, scale)
    K = args.steps
    mode_list = [m for m in args.modes.split(",") if m] if args.modes else list(DEFAULT_MODES.get(w.name, DEFAULT_MODES["n400"]))
    # every mode runs the same number of steps before its timed block, so all
    # modes time the same stretch of one trajectory (the fields come from the
    # same counter-based stream; the mixed modes' split-BF16 force bias and VHS
    # move the walkers by ~1e-6 relative per step, so the trajectories stay
    # close and the per-step energy difference measures the precision bias)
    warmup = args.warmup
    if any("surrogate" in m for m in mode_list):
        warmup = max(warmup, SURROGATE_WARMUP + 5)
    tuner = make_tuner(args)
    modes = run_modes(args, w, host, mode_list, K, warmup, tuner)
    compare_modes(modes, "fp64" if "fp64" in modes else "mixed")
    for rec in modes.values():
        rec.pop("_e_steps")
    # headline: the fastest mode whose block-energy bias vs FP64 is within 0.1 mHa
    # and whose added per-step variance is within VARIANCE_BUDGET_FRAC of the
    # step's own (without an FP64 run in the same invocation the bias is not
    # measured and the fastest mode is reported, flagged as such)
    ref_mode = "fp64" if "fp64" in modes else ("mixed" if "mixed" in modes else None)
    ok = ([m for m, r in modes.items()
           if m == ref_mode or (r.get("within_0p1_mHa") and r.get("within_variance_budget"))]
          if ref_mode else list(modes))
    head = max(ok, key=lambda m: modes[m]["value"])
    m = modes[head]
    rk = m["roofline_kernel"]
    step_fl = step_flops(w.nbasis, w.nchol, w.nalpha, w.nbeta)
    dtype = {"fp64": "f64",
             "mixed": "f64 (Green's function, weights, one-body, QR) + bf16x3 force bias, VHS, Taylor, exchange "
                      "(fp32 accumulate)",
             "mixed+surrogate": "f64 (Green's function, weights, one-body, QR) + bf16x3 force bias, VHS, Taylor (fp32 "
                                "accumulate); exchange estimated per walker as g + [sampled] 8 (exact - g) with "
                                "g = E_T + 2<K,dG> (f64) + Q(dG) (fp8 e4m3 x1, scaled) + gain * MLP (bf16 weights), "
                                "exact bf16x3 exchange on a random 1/8 of walkers",
             "mixed+surrogate-bf16q": "f64 (Green's function, weights, one-body, QR) + bf16x3 force bias, VHS, Taylor "
                                      "(fp32 accumulate); exchange estimated per walker as g + [sampled] 8 "
                                      "(exact - g) with g = E_T + 2<K,dG> (f64) + Q(dG) (bf16 x1) + gain * MLP (bf16 "
                                      "weights), exact bf16x3 exchange on a random 1/8 of walkers",
             "mixed+surrogate-nomlp": "f64 (Green's function, weights, one-body, QR) + bf16x3 force bias, VHS, Taylor "
                                      "(fp32 accumulate); exchange estimated per walker as g + [sampled] 8 "
                                      "(exact - g) with g = E_T + 2<K,dG> (f64) + Q(dG) (fp8 e4m3 x1, scaled), exact "
                                      "bf16x3 exchange on a random 1/8 of walkers",
             "mixed+surrogate-linear": "f64 (Green's function, weights, one-body, QR) + bf16x3 force bias, VHS, "
                                       "Taylor (fp32 accumulate); exchange estimated per walker as g + [sampled] 8 "
                                       "(exact - g) with g = E_T + 2<K,dG> (f64) + gain * MLP (bf16 weights), exact "
                                       "bf16x3 exchange on a random 1/8 of walkers"}[head]
    out = {
        "metric": METRIC,
        "value": m["value"],
        "unit": "walker-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": warmup,
        "warmup_note": (f"--warmup {args.warmup} raised to {warmup}: the control variate trains on "
                        f"{SURROGATE_WARMUP} all-exact steps before its timed block, and every mode runs the same "
                        "steps so all time one trajectory") if warmup != args.warmup else None,
        "ms_per_step": m["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic",
        "config": {
            "workload": f"{w.name}: N={w.nbasis} X={w.nchol} decay={w.decay} nalpha={w.nalpha} nbeta={w.nbeta} "
                        f"W={w.nwalkers} dt={w.dt}",
            "nbasis": w.nbasis,
            "nchol": w.nchol,
            "nalpha": w.nalpha,
            "nbeta": w.nbeta,
            "nwalkers": w.nwalkers,
            "cholesky_scale": args.cholesky_scale,
            "mode": head,
            "parallelism": "single GPU (walkers batched)",
        },
        "roofline": {
            "kernel": rk["kernel"],
            "bound": "mfma",
            "achieved": rk["achieved"],
            "peak": rk["peak"],
            "peak_note": rk["peak_note"],
            "unit": "TFLOP/s",
            "frac": rk["achieved"] / rk["peak"],
            "traffic": traffic_per_launch(head, w),
            "avg_launch_ms": rk["avg_launch_ms"],
            "flops_per_launch": rk["flops_per_launch"],
            "timing": "HIP events on the context's stream around each launch, over the eager pass of the same "
                      "number of steps that precedes the step-graph replay (events cannot bracket kernels "
                      "inside a replayed graph)" if m.get("timed") == "step graph" else
                      "HIP events on the context's stream around each launch, over the timed steps",
            # the step's other large phases, each against its own MFMA peak (the
            # VHS GEMM is the largest single launch of the step)
            "other_phases": phase_rooflines(head, m["phases_ms_per_step"], w),
        },
        "timed": m.get("timed"),
        "step_tflops_algorithmic": step_fl * m["value"] / 1e12,
        "modes": modes,
        "bias_measured": "fp64" in modes,
    }
    if "fp64" in modes:
        out["fp64_walker_steps_per_s"] = modes["fp64"]["value"]
        out["speedup_vs_fp64"] = m["value"] / modes["fp64"]["value"]
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(host, w, args.cpu_walkers, args.cpu_steps)
        log(f"[bench] cpu baseline {out['cpu_baseline']['value']:.3f} walker-steps/s")
    else:
        out["cpu_baseline"] = None
    if w.name == "n400" and not args.no_secondary:
        del host
        out["secondary"] = {"n300_surrogate": run_surrogate_workload(args, tuner),
                            "multi16": run_multi(args, emit=False)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
