"""Phaseless propagator (drop-in for ipie's PhaselessGeneric with GenericRealChol).

Mirrors the operator API of ipie/propagation/phaseless_base.py:242-388 and
phaseless_generic.py:28-111: ``build(hamiltonian, trial, walkers,
mpi_handler, verbose)``, ``propagate_walkers(walkers, hamiltonian, trial,
eshift)``, the one-/two-body halves, ``update_weight``, ``timer`` with the six
PropagatorTimer fields, ``nfb_trig``, ``dt``, ``sqrt_dt``, ``ebound``,
``fbbound``, ``mf_shift`` and ``expH1``.

``propagate_walkers`` is one C-ABI call (afqmc_propagate): GF -> exp(-dt/2 H1)
-> force bias -> fields -> VHS -> Taylor exp -> exp(-dt/2 H1) -> overlap ->
weights, all enqueued on the current stream with no host synchronisation.

``precision``: "fp64" (every contraction in FP64) or "mixed" (the exchange of
the local energy as split-BF16 products with FP32 accumulation; Green's
function, overlaps, weights, propagation and QR stay FP64).

``surrogate``: None, or True / a dict (fraction, warmup, train_iters, lr,
seed, quadratic) enabling the learned exchange control variate for the ``energy_prev``
output of ``propagate_walkers`` (an unbiased estimate per walker; exact on a
random subset of walkers, see afqmc_surrogate_configure).  ``local_energy_batch``
is always exact.

Auxiliary fields: ``xi_source`` (callable ``(nwalkers, nfields) -> [W, X]``
float64, numpy or torch) injects them exactly like ipie's
``xp.random.normal(0, 1, X * W).reshape(W, X)`` (phaseless_base.py:296);
without it the fields come from the device generator (Philox4x32-10, key
``seed``, one counter value per call).
"""

import numpy
import scipy.linalg
import torch

from ipie_amd import autotune as _autotune
from ipie_amd import engine as _engine


class PropagatorTimer:
    """The six fields of ipie/propagation/continuous_base.py:4-11, in seconds.

    ipie adds host wall time around each phase, with a device synchronisation
    after it (phaseless_base.py:274-338, phaseless_generic.py:39-58):
    tgf the Green's function, tgemm the two one-body halves and the Taylor
    series, tfbias the force bias and xbar, tvhs the VHS construction, tovlp the
    new overlap, tupdate the weight update.  Here the step is one fused C-ABI
    call, so the fields are read from the context's HIP-event time of the same
    phases (afqmc_profile_read; reading one synchronises the context's
    stream, as ipie's own timing does after every phase).  A field may be
    assigned (e.g. reset to 0); later reads add the device time since."""

    PHASES = {
        "tgf": ("gf",),
        "tgemm": ("one_body", "taylor"),
        "tfbias": ("force_bias", "fields"),
        "tvhs": ("vhs",),
        "tovlp": ("overlap",),
        "tupdate": ("weight",),
    }

    def __init__(self):
        self._ctx = None
        self._host = dict.fromkeys(self.PHASES, 0.0)
        self._zero = dict.fromkeys(self.PHASES, 0.0)

    def bind(self, ctx):
        """Time the phases of this device context (turns its profiling on)."""
        ctx.profile(True)
        self._ctx = ctx
        self._zero = dict.fromkeys(self.PHASES, 0.0)

    def _device(self, name):
        if self._ctx is None:
            return 0.0
        prof = self._ctx.profile_read()
        return 1e-3 * sum(prof[p][0] for p in self.PHASES[name])


def _timer_field(name):
    def get(self):
        return self._host[name] + self._device(name) - self._zero[name]

    def set_(self, value):
        self._host[name] = float(value)
        self._zero[name] = self._device(name)

    return property(get, set_)


for _name in PropagatorTimer.PHASES:
    setattr(PropagatorTimer, _name, _timer_field(_name))


def construct_mean_field_shift(hamiltonian, trial):
    """mf = i L^T Re vec(Ga + Gb) - L^T Im vec(Ga + Gb) (phaseless_base.py:153-172)."""
    g = (trial.G[0] + trial.G[1]).ravel()
    return 1j * hamiltonian.chol.T.dot(g.real) - hamiltonian.chol.T.dot(g.imag)


def construct_one_body_propagator(hamiltonian, mf_shift, dt):
    """expH1[s] = expm(-dt/2 (h1e_mod[s] - i sum_x L_x mf_x)) (phaseless_base.py:27-74)."""
    nb = hamiltonian.nbasis
    shift = 1j * hamiltonian.chol.dot(mf_shift).reshape(nb, nb)
    H1 = hamiltonian.h1e_mod - numpy.array([shift, shift])
    return numpy.array([scipy.linalg.expm(-0.5 * dt * H1[0]), scipy.linalg.expm(-0.5 * dt * H1[1])])


class PhaselessGeneric:
    def __init__(self, time_step, exp_nmax=6, verbose=False, seed=7, precision="fp64", surrogate=None,
                 autotune=None):
        self.dt = time_step
        self.autotune = autotune  # None / True / cache path / ipie_amd.autotune.Autotuner
        self.tuning = None
        self.precision = precision
        self.surrogate = surrogate
        self.verbose = verbose
        self.sqrt_dt = self.dt**0.5
        self.isqrt_dt = 1j * self.sqrt_dt
        self.exp_nmax = exp_nmax
        self.nhe_trig = 0
        self.ebound = (2.0 / self.dt) ** 0.5
        self.fbbound = 1.0
        self.timer = PropagatorTimer()
        self.mpi_handler = None
        self.vbias = None
        self.xi_source = None
        self.seed = int(seed)
        self.counter = 0
        self.last_xi = None
        self.keep_xi = False
        self._engine = None

    # ------------------------------------------------------------------ set-up
    def build(self, hamiltonian, trial=None, walkers=None, mpi_handler=None, verbose=False, device_setup=False):
        """Mean-field shift, one-body propagator and device context (phaseless_base.py:256-272).

        device_setup=True builds every static tensor (packed Cholesky vectors,
        h1e_mod, half rotation, mean-field shift, expH1) with the gfx950
        set-up kernels from hamiltonian.chol instead of on the host."""
        if walkers is None:
            raise ValueError("the MI355X propagator needs the walkers at build time (device context)")
        if device_setup:
            eng = self._setup_on_device(hamiltonian, trial, walkers)
        else:
            if not trial.half_rotated:
                trial.half_rotate(hamiltonian)
            self.mf_shift = construct_mean_field_shift(hamiltonian, trial)
            self.expH1 = construct_one_body_propagator(hamiltonian, self.mf_shift, self.dt)
            eng = _engine.Engine(trial, walkers, nchol=hamiltonian.nchol, exp_nmax=self.exp_nmax,
                                 precision=self.precision)
            eng.set_hamiltonian(hamiltonian, self.expH1, self.mf_shift, self.dt)
        opts = None
        if self.surrogate:
            opts = dict(fraction=0.125, warmup=100, train_iters=4, lr=1e-3, seed=self.seed)
            opts.update(self.surrogate if isinstance(self.surrogate, dict) else {})
        if opts is not None:
            quadratic = opts.pop("quadratic", "fp8")  # "fp8", "bf16" or "none" (True / False)
            eng.ctx.surrogate_configure(True, **opts)
            eng.ctx.surrogate_quadratic(quadratic)
        tuner = _autotune.as_tuner(self.autotune)
        if tuner is not None:  # times state-free calls only: no step runs, the model is not trained
            self.tuning = tuner.tune_walkers(eng.ctx, walkers, surrogate=opts is not None)
        walkers._engine = eng
        self._engine = eng
        self.timer.bind(eng.ctx)
        self.mpi_handler = mpi_handler

    def _setup_on_device(self, ham, trial, walkers):
        """Static tensors from the set-up kernels (generic.py:31-43, 64-72;
        half_rotate.py:11-116; phaseless_base.py:27-74, 153-172), for a real
        Hamiltonian and a real trial; the host keeps small copies of mf_shift,
        expH1 and h1e_mod."""
        psi = numpy.asarray(trial.psi)
        if numpy.iscomplexobj(psi) and numpy.abs(psi.imag).max(initial=0.0) > 0.0:
            raise NotImplementedError("device set-up needs a real trial")
        gcharge = numpy.asarray(trial.G[0] + trial.G[1]).ravel()
        if numpy.iscomplexobj(gcharge) and numpy.abs(gcharge.imag).max(initial=0.0) > 0.0:
            raise NotImplementedError("device set-up needs a real trial density")
        eng = _engine.Engine(trial, walkers, nchol=ham.nchol, exp_nmax=self.exp_nmax, precision=self.precision)
        ctx, dev = eng.ctx, walkers.device
        N, X = ham.nbasis, ham.nchol
        f64, c128 = torch.float64, torch.complex128
        chol = torch.from_numpy(numpy.ascontiguousarray(ham.chol, dtype=numpy.float64)).to(dev)
        h1e = torch.from_numpy(numpy.ascontiguousarray(ham.H1, dtype=numpy.float64)).to(dev)
        packed = torch.empty(N * (N + 1) // 2, X, dtype=f64, device=dev)
        ctx.setup_pack(chol, packed)
        h1e_mod = torch.empty(2, N, N, dtype=f64, device=dev)
        ctx.setup_h1e_mod(chol, h1e, h1e_mod)
        rot = {}
        for s, (orbs, n) in enumerate(((trial.psi0a, trial.nalpha), (trial.psi0b, trial.nbeta))):
            r = torch.zeros(X, n * N, dtype=f64, device=dev)
            rh = torch.zeros(n, N, dtype=c128, device=dev)
            if n > 0:
                p = torch.from_numpy(numpy.ascontiguousarray(numpy.real(orbs), dtype=numpy.float64)).to(dev)
                ctx.setup_half_rotate(s, p, chol, h1e[s].contiguous(), r, rh)
            rot[s] = (r, rh)
        mf = torch.empty(X, dtype=c128, device=dev)
        ctx.setup_mf_shift(chol, torch.from_numpy(gcharge.astype(numpy.complex128)).to(dev), mf)
        expH1 = torch.empty(2, N, N, dtype=f64, device=dev)
        ctx.setup_expH1(chol, h1e_mod, mf, self.dt, expH1)
        del chol
        key = str(dev)
        ham._dev[key] = {"chol_packed": packed}
        td = trial.device_tensors(dev)
        td.update(rchola=rot[0][0], rcholb=rot[1][0], rH1a=rot[0][1], rH1b=rot[1][1])
        trial.half_rotated = True
        trial._rchola = trial._rcholb = trial._rH1a = trial._rH1b = None  # on the device; fetched when read
        eng._set_static_partial(chol_packed=packed, rchola=rot[0][0], rcholb=rot[1][0], rH1a=rot[0][1],
                                rH1b=rot[1][1], expH1=expH1, mf_shift=mf, dt=float(self.dt), ecore=float(ham.ecore))
        self.mf_shift = mf.cpu().numpy()
        self.expH1 = expH1.cpu().numpy()
        if getattr(ham, "h1e_mod", None) is None:
            ham.h1e_mod = h1e_mod.cpu().numpy()
        return eng

    def cast_to_cupy(self, verbose=False):
        return None

    def surrogate_stats(self):
        return self._engine.ctx.surrogate_stats()

    @property
    def nfb_trig(self):
        return 0 if self._engine is None else self._engine.ctx.nfb()

    # ------------------------------------------------------------------ fields
    def _next_xi(self, walkers, hamiltonian):
        if self.xi_source is None:
            self.counter += 1
            return None
        xi = self.xi_source(walkers.nwalkers, hamiltonian.nfields)
        if not isinstance(xi, torch.Tensor):
            xi = torch.from_numpy(numpy.ascontiguousarray(xi, dtype=numpy.float64))
        xi = xi.to(device=walkers.device, dtype=torch.float64).contiguous()
        self.last_xi = xi
        return xi

    def _eng(self, walkers):
        eng = walkers._engine
        if eng is None or eng is not self._engine:
            raise RuntimeError("walkers are not bound to this propagator: call build(hamiltonian, trial, walkers)")
        return eng

    # ------------------------------------------------------------------ step
    def propagate_walkers(self, walkers, hamiltonian, trial, eshift, energy_prev=None):
        """One phaseless step (phaseless_base.py:312-338).

        ``energy_prev`` (optional [W, 3] complex128 device tensor) receives the
        local energy of the incoming walkers from the step's own Green's function.
        """
        eng = self._eng(walkers)
        xi = self._next_xi(walkers, hamiltonian)
        eng.ctx.propagate(
            walkers.phia, walkers.phib, walkers.Ghalfa, walkers.Ghalfb, walkers.weight, walkers.ovlp,
            walkers.hybrid_energy, float(numpy.real(eshift)), xi=xi, seed=self.seed, counter=self.counter,
            log_shift=walkers.log_shift, energy_prev=energy_prev,
        )

    # ------------------------------------------------------------ step graphs
    def capture_steps(self, walkers, hamiltonian, trial, nsteps, reortho_every=0, energy_prev=None):
        """Record `nsteps` propagate_walkers calls into one hipGraph (and
        walkers.reortho() after every `reortho_every`-th of them); nothing runs.
        The same calls must have run eagerly before (workspace, the control
        variate's warm-up).  Fields come from the device generator: a caller
        that injects them (xi_source) cannot capture.  Returns the number of
        recorded steps."""
        if self.xi_source is not None:
            raise RuntimeError("a step graph draws its fields on the device; unset xi_source to capture")
        eng = self._eng(walkers)
        dev = walkers.device
        if getattr(self, "_graph_stream", None) is None:
            self._graph_stream = torch.cuda.Stream(device=dev)
        s = self._graph_stream
        s.wait_stream(torch.cuda.current_stream(dev))
        with torch.cuda.stream(s):
            eng.ctx.graph_begin()
            try:
                for j in range(nsteps):
                    eng.ctx.propagate(
                        walkers.phia, walkers.phib, walkers.Ghalfa, walkers.Ghalfb, walkers.weight, walkers.ovlp,
                        walkers.hybrid_energy, 0.0, xi=None, seed=self.seed, counter=0,
                        log_shift=walkers.log_shift, energy_prev=energy_prev,
                    )
                    if reortho_every and (j + 1) % reortho_every == 0:
                        walkers.reortho()
            except BaseException:
                try:
                    eng.ctx.graph_end()
                except Exception:
                    pass
                raise
            steps = eng.ctx.graph_end()
        self._graph_steps = steps
        return steps

    def replay_steps(self, walkers, eshift, replays=1):
        """Replay the captured steps `replays` times with the given eshift; the
        fields continue the propagator's counter, as eager steps would."""
        eng = self._eng(walkers)
        dev = walkers.device
        s = self._graph_stream
        s.wait_stream(torch.cuda.current_stream(dev))
        with torch.cuda.stream(s):
            eng.ctx.graph_launch(replays, self.counter + 1, float(numpy.real(eshift)))
        torch.cuda.current_stream(dev).wait_stream(s)
        self.counter += replays * self._graph_steps

    def propagate_walkers_one_body(self, walkers):
        """phi_s <- expH1[s] phi_s (phaseless_base.py:274-280)."""
        self._eng(walkers).ctx.one_body(walkers.phia, walkers.phib)

    def propagate_walkers_two_body(self, walkers, hamiltonian, trial):
        """Force bias, fields and exp(VHS) from walkers.Ghalf (phaseless_base.py:282-310); returns (cmf, cfb)."""
        eng = self._eng(walkers)
        xi = self._next_xi(walkers, hamiltonian)
        nw = walkers.nwalkers
        cmf = torch.empty(nw, dtype=torch.complex128, device=walkers.device)
        cfb = torch.empty_like(cmf)
        self.vbias = torch.empty(nw, hamiltonian.nchol, dtype=torch.complex128, device=walkers.device)
        xi_out = None
        if self.keep_xi and xi is None:
            xi_out = torch.empty(nw, hamiltonian.nchol, dtype=torch.float64, device=walkers.device)
        eng.ctx.two_body(walkers.phia, walkers.phib, walkers.Ghalfa, walkers.Ghalfb, cmf, cfb, xi=xi, seed=self.seed,
                         counter=self.counter, xi_out=xi_out, vbias_out=self.vbias)
        if xi_out is not None:
            self.last_xi = xi_out
        return cmf, cfb

    def update_weight(self, walkers, ovlp, ovlp_new, cfb, cmf, eshift):
        """Hybrid energy and phaseless weight update (phaseless_base.py:340-380)."""
        self._eng(walkers).ctx.update_weight(ovlp, ovlp_new, cfb, cmf, float(numpy.real(eshift)), walkers.weight,
                                             walkers.hybrid_energy, walkers.ovlp)

    def propagate_walkers_stepwise(self, walkers, hamiltonian, trial, eshift):
        """The same step as separate entry points, in the reference's phase order."""
        ovlp = trial.calc_greens_function(walkers)
        self.propagate_walkers_one_body(walkers)
        cmf, cfb = self.propagate_walkers_two_body(walkers, hamiltonian, trial)
        self.propagate_walkers_one_body(walkers)
        ovlp_new = trial.calc_overlap(walkers)
        self.update_weight(walkers, ovlp, ovlp_new, cfb, cmf, eshift)


Phaseless = {"generic": PhaselessGeneric}
