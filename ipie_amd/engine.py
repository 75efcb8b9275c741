"""Binds a walker population to one libafqmc_hip context.

An engine is created lazily on the first device call that needs one: the
trial's Green's function only needs the trial orbitals, the propagator's
``build`` supplies the Hamiltonian tensors (mean-field shift, exp(-dt/2 H1),
packed and half-rotated Cholesky vectors).  The engine lives on the walkers
object (``walkers._engine``), so every drop-in entry point that receives the
walkers finds it.
"""

import numpy
import torch

from ipie_amd import _lib


class SingularOverlapError(numpy.linalg.LinAlgError):
    """A walker's overlap matrix with the trial was exactly singular."""


class Engine:
    def __init__(self, trial, walkers, nchol=0, exp_nmax=6, precision="fp64"):
        self.trial = trial
        self.precision = precision
        self.nchol = nchol
        self.exp_nmax = exp_nmax
        self.device = walkers.device
        self.nwalkers = walkers.nwalkers
        self.rhf = bool(walkers.rhf)
        self.ctx = _lib.Context(
            walkers.nbasis, nchol, walkers.nup, walkers.ndown, walkers.nwalkers, exp_nmax=exp_nmax, rhf=walkers.rhf,
            device=walkers.device, precision=precision,
        )
        self.static = {}
        self.check_singular = True
        self._set_static_partial()

    def _set_static_partial(self, **kw):
        t = self.trial.device_tensors(self.device)
        dev = self.device
        N, na, nb, X = self.ctx.dims.nbasis, self.ctx.dims.nalpha, self.ctx.dims.nbeta, self.ctx.dims.nchol
        f64, c128 = torch.float64, torch.complex128
        st = {
            "chol_packed": torch.zeros(N * (N + 1) // 2, X, dtype=f64, device=dev),
            "rchola": torch.zeros(X, na * N, dtype=f64, device=dev),
            "rcholb": torch.zeros(X, nb * N, dtype=f64, device=dev),
            "rH1a": torch.zeros(na, N, dtype=c128, device=dev),
            "rH1b": torch.zeros(nb, N, dtype=c128, device=dev),
            "expH1": torch.zeros(2, N, N, dtype=f64, device=dev),
            "mf_shift": torch.zeros(X, dtype=c128, device=dev),
            "psia": t["psia"],
            "psib": t["psib"],
            "dt": 1.0,
            "ecore": 0.0,
        }
        st.update(kw)
        self.static = st
        self.ctx.set_static(**st)

    def set_hamiltonian(self, hamiltonian, expH1, mf_shift, dt):
        """Upload the propagation tensors (called by the propagator's build)."""
        t = self.trial.device_tensors(self.device)
        h = hamiltonian.device_tensors(self.device)
        if numpy.iscomplexobj(expH1):
            if numpy.abs(numpy.imag(expH1)).max() > 0.0:
                raise NotImplementedError("complex one-body propagators (complex trial) are not on the MI355X path")
            expH1 = numpy.real(expH1)
        self._set_static_partial(
            chol_packed=h["chol_packed"],
            rchola=t["rchola"],
            rcholb=t["rcholb"],
            rH1a=t["rH1a"],
            rH1b=t["rH1b"],
            expH1=torch.from_numpy(numpy.ascontiguousarray(expH1, dtype=numpy.float64)).to(self.device),
            mf_shift=torch.from_numpy(numpy.ascontiguousarray(mf_shift, dtype=numpy.complex128)).to(self.device),
            dt=float(dt),
            ecore=float(hamiltonian.ecore),
        )

    def check_flags(self):
        """Raise like numpy.linalg would on a singular overlap (synchronises)."""
        if not self.check_singular:
            return 0
        f = self.ctx.flags(clear=True)
        if f & _lib.FLAG_SINGULAR:
            wf = self.ctx.walker_flags(self.nwalkers, clear=False).cpu()
            bad = torch.nonzero(wf & _lib.FLAG_SINGULAR).flatten().tolist()
            raise SingularOverlapError(f"Singular matrix: walkers {bad[:16]} have zero overlap with the trial")
        return f


def engine_for(trial, walkers, hamiltonian=None):
    """The engine attached to ``walkers`` (created for ``trial`` if absent)."""
    eng = getattr(walkers, "_engine", None)
    if eng is None or eng.trial is not trial:
        eng = Engine(trial, walkers, nchol=0 if hamiltonian is None else 0)
        walkers._engine = eng
    if hamiltonian is not None and eng.nchol == 0:
        raise RuntimeError("the propagator's build() must run before force-bias / energy calls")
    return eng
