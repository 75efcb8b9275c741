"""Shared builders for the parity tests: the same arrays feed the oracle and
the device path."""

import numpy

from oracle import afqmc


def rel_err(a, b):
    """max |a - b| / max |b| (both arrays)."""
    a = numpy.asarray(a)
    b = numpy.asarray(b)
    scale = max(numpy.abs(b).max(initial=0.0), 1e-300)
    return float(numpy.abs(a - b).max(initial=0.0) / scale)


def rel_err_elem(a, b, floor=1e-12):
    """max over entries of |a - b| / max(|b|, floor * max |b|): the relative
    error of every walker's value on its own scale (overlaps and weights span
    orders of magnitude), with an absolute floor so that entries the reference
    makes exactly zero (cosine-clipped weights) compare against a small
    multiple of the largest entry instead of zero."""
    a = numpy.asarray(a)
    b = numpy.asarray(b)
    big = max(numpy.abs(b).max(initial=0.0), 1e-300)
    den = numpy.maximum(numpy.abs(b), floor * big)
    return float((numpy.abs(a - b) / den).max(initial=0.0))


PER_WALKER = ("weight", "ovlp", "hybrid_energy")


def walker_err(attr, a, b):
    """Per-walker relative error for the per-walker scalars, max-norm relative
    error for arrays of matrices (phi, Ghalf)."""
    return rel_err_elem(a, b) if attr in PER_WALKER else rel_err(a, b)


def synthetic_arrays(N, X, decay, nalpha, nbeta, seed=7, scale=1.0):
    from ipie_amd import synthetic

    h1e, chol, ecore = synthetic.make_hamiltonian(N, X, decay, seed=seed)
    psi = synthetic.make_trial(h1e, nalpha, nbeta)
    return h1e, chol * scale, ecore, psi


def random_walkers(N, nalpha, nbeta, W, seed=3, amp=0.3, base=None):
    rs = numpy.random.RandomState(seed)
    n = nalpha + nbeta
    out = numpy.empty((W, N, n), dtype=numpy.complex128)
    for w in range(W):
        pert = rs.normal(size=(N, n)) + 1j * rs.normal(size=(N, n))
        b = base if base is not None else numpy.zeros((N, n))
        out[w] = b + amp * pert
    return out


class Pair:
    """Oracle objects and device objects built from one set of arrays."""

    def __init__(self, h1e, chol, psi, nelec, nwalkers, dt=0.005, ecore=0.0, init=None, rhf=False, exp_nmax=6,
                 precision="fp64", surrogate=None, autotune=None):
        import torch

        from ipie_amd import Generic as DevSystem
        from ipie_amd.hamiltonian import GenericRealChol
        from ipie_amd.propagator import PhaselessGeneric
        from ipie_amd.trial import SingleDet
        from ipie_amd.walkers import UHFWalkers

        self.torch = torch
        self.nelec = nelec
        self.N = h1e.shape[-1]
        self.W = nwalkers
        self.orc = afqmc.build_state(h1e, chol, psi, nelec, nwalkers, dt, ecore=ecore, exp_nmax=exp_nmax)
        self.system = DevSystem(nelec)
        self.ham = GenericRealChol(h1e, chol, ecore=ecore)
        self.trial = SingleDet(psi, nelec, self.N)
        self.trial.half_rotate(self.ham)
        self.trial.calculate_energy(self.system, self.ham)
        self.walkers = UHFWalkers(psi, nelec[0], nelec[1], self.N, nwalkers)
        self.walkers.rhf = rhf
        self.orc.walkers.rhf = rhf
        self.prop = PhaselessGeneric(dt, exp_nmax=exp_nmax, precision=precision, surrogate=surrogate,
                                     autotune=autotune)
        self.prop.build(self.ham, self.trial, self.walkers)
        if init is not None:
            self.set_phi(init[:, :, : nelec[0]], init[:, :, nelec[0]:])

    def set_phi(self, pa, pb):
        t = self.torch
        self.walkers.phia.copy_(t.from_numpy(numpy.ascontiguousarray(pa, dtype=numpy.complex128)))
        self.walkers.phib.copy_(t.from_numpy(numpy.ascontiguousarray(pb, dtype=numpy.complex128)))
        self.orc.walkers.phia = numpy.array(pa, dtype=numpy.complex128)
        self.orc.walkers.phib = numpy.array(pb, dtype=numpy.complex128)

    def sync_oracle_from_device(self):
        """Resync: the oracle continues from the device state."""
        w, o = self.walkers, self.orc.walkers
        o.phia = w.phia.cpu().numpy().copy()
        o.phib = w.phib.cpu().numpy().copy()
        o.weight = w.weight.cpu().numpy().copy()
        o.ovlp = w.ovlp.cpu().numpy().copy()
        o.hybrid_energy = w.hybrid_energy.cpu().numpy().copy()
