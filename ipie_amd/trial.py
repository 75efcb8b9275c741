"""Single-determinant trial (drop-in for ipie's SingleDet with a real Cholesky Hamiltonian).

Mirrors ipie/trial_wavefunction/single_det.py:30-211.  Set-up (G, Ghalf of the
trial, half rotation, trial energy) is host arithmetic done once, as in the
reference (single_det.py:31-86, half_rotate.py:11-116).  The per-step calls
``calc_greens_function``, ``calc_overlap`` and ``calc_force_bias`` run the
gfx950 kernels through the walkers' engine.
"""

import numpy
import torch

from ipie_amd import engine as _engine


def _gab(A, B):
    """(G, Ghalf) of bra A and ket B (ipie/estimators/utils.py:100-133)."""
    O = B.T.dot(A.conj())
    Ghalf = numpy.linalg.solve(O, B.T)
    G = A.conj().dot(Ghalf)
    return G, Ghalf


class SingleDet:
    def __init__(self, wavefunction, num_elec, num_basis, handler=None, verbose=False):
        assert isinstance(wavefunction, numpy.ndarray) and wavefunction.ndim == 2
        self.nalpha, self.nbeta = num_elec
        self.num_elec = num_elec
        self.nelec = num_elec
        self.nbasis = num_basis
        self.verbose = verbose
        self.handler = handler
        self.psi = wavefunction
        if numpy.sum(self.psi.imag.ravel() ** 2) <= 1e-8:
            self.psi = numpy.array(self.psi.real, dtype=numpy.float64)
        self.psi0a = numpy.ascontiguousarray(self.psi[:, : self.nalpha])
        self.psi0b = numpy.ascontiguousarray(self.psi[:, self.nalpha :])
        Ga, Gha = _gab(self.psi0a, self.psi0a)
        if self.nbeta > 0:
            Gb, Ghb = _gab(self.psi0b, self.psi0b)
        else:
            Gb = numpy.zeros_like(Ga)
            Ghb = numpy.zeros((0, num_basis), dtype=Gha.dtype)
        self.G = numpy.array([Ga, Gb])
        self.Ghalf = [Gha, Ghb]
        self.half_rotated = False
        self._num_dets = 1
        self._dev = {}

    @property
    def num_dets(self):
        return 1

    def build(self):
        pass

    def half_rotate(self, hamiltonian, comm=None):
        """rH1 and half-rotated Cholesky vectors (ipie/trial_wavefunction/half_rotate.py:11-116)."""
        M, X = hamiltonian.nbasis, hamiltonian.nchol
        chol3 = hamiltonian.chol.reshape(M, M, X)
        self._rH1a = self.psi0a.conj().T.dot(hamiltonian.H1[0])
        self._rH1b = self.psi0b.conj().T.dot(hamiltonian.H1[1])

        def rot(orbs):
            n = orbs.shape[1]
            if n == 0:
                return numpy.zeros((X, 0), dtype=numpy.result_type(orbs, chol3))
            r = orbs.conj().T.dot(chol3.reshape(M, M * X))  # [n, (m', x)]
            return numpy.ascontiguousarray(r.reshape(n, M, X).transpose(2, 0, 1).reshape(X, n * M))

        self._rchola = rot(self.psi0a)
        self._rcholb = rot(self.psi0b)
        self.half_rotated = True
        self._dev = {}

    def _host_rotated(self):
        """Host copies of rchol / rH1 after a device half rotation."""
        if self._rchola is None and self._dev:
            d = next(iter(self._dev.values()))
            self._rchola = d["rchola"].cpu().numpy()
            self._rcholb = d["rcholb"].cpu().numpy()
            self._rH1a = d["rH1a"].cpu().numpy()
            self._rH1b = d["rH1b"].cpu().numpy()

    def calculate_energy(self, system, hamiltonian):
        """Trial energy (ipie/trial_wavefunction/single_det.py:65-86)."""
        self._host_rotated()
        Ga, Gb = self.Ghalf
        self.e1b = numpy.sum(Ga * self._rH1a) + numpy.sum(Gb * self._rH1b) + hamiltonian.ecore
        Xa = self._rchola.dot(Ga.ravel()) if self.nalpha else 0.0
        Xb = self._rcholb.dot(Gb.ravel()) if self.nbeta else 0.0
        Xt = Xa + Xb
        self.ej = 0.5 * numpy.dot(Xt, Xt)
        ek = 0.0
        for r, G, n in ((self._rchola, Ga, self.nalpha), (self._rcholb, Gb, self.nbeta)):
            if n == 0:
                continue
            T = numpy.einsum("xim,jm->xij", r.reshape(-1, n, self.nbasis), G, optimize=True)
            ek = ek + 0.5 * numpy.einsum("xij,xji->", T, T)
        self.ek = ek
        self.e2b = self.ej - self.ek
        self.energy = self.e1b + self.e2b
        return self.energy

    # ------------------------------------------------------------------ device tensors
    def device_tensors(self, device):
        key = str(device)
        if key not in self._dev:
            c128 = torch.complex128
            d = {
                "psia": torch.from_numpy(numpy.ascontiguousarray(self.psi0a, dtype=numpy.complex128)).to(device),
                "psib": torch.from_numpy(numpy.ascontiguousarray(self.psi0b, dtype=numpy.complex128)).to(device),
            }
            if self.half_rotated:
                for nm in ("_rchola", "_rcholb"):
                    r = getattr(self, nm)
                    if numpy.iscomplexobj(r):
                        if numpy.abs(r.imag).max(initial=0.0) > 0.0:
                            raise NotImplementedError(
                                "complex half-rotated Cholesky vectors (complex trial) are not on the MI355X path"
                            )
                        r = r.real
                    d[nm[1:]] = torch.from_numpy(numpy.ascontiguousarray(r, dtype=numpy.float64)).to(device)
                d["rH1a"] = torch.from_numpy(numpy.ascontiguousarray(self._rH1a, dtype=numpy.complex128)).to(device)
                d["rH1b"] = torch.from_numpy(numpy.ascontiguousarray(self._rH1b, dtype=numpy.complex128)).to(device)
                assert d["rH1a"].dtype == c128
            self._dev[key] = d
        return self._dev[key]

    # ------------------------------------------------------------------ per-step calls
    def calc_greens_function(self, walkers, build_full=False):
        """Ghalf and overlap of every walker (greens_function_single_det.py:106-154)."""
        if build_full:
            raise NotImplementedError("the full N x N Green's function is not built on the MI355X path")
        eng = _engine.engine_for(self, walkers)
        ovlp = torch.empty(walkers.nwalkers, dtype=torch.complex128, device=walkers.device)
        eng.ctx.greens_function(walkers.phia, walkers.phib, walkers.Ghalfa, walkers.Ghalfb, ovlp, walkers.log_shift)
        eng.check_flags()
        return ovlp

    def calc_overlap(self, walkers):
        """Overlap only (ipie/propagation/overlap.py:29-59)."""
        eng = _engine.engine_for(self, walkers)
        ovlp = torch.empty(walkers.nwalkers, dtype=torch.complex128, device=walkers.device)
        eng.ctx.overlap(walkers.phia, walkers.phib, ovlp, walkers.log_shift)
        eng.check_flags()
        return ovlp

    def calc_force_bias(self, hamiltonian, walkers, mpi_handler=None):
        """vbias [W, X] (ipie/propagation/force_bias.py:79-121)."""
        eng = _engine.engine_for(self, walkers, hamiltonian=hamiltonian)
        vbias = torch.empty(walkers.nwalkers, hamiltonian.nchol, dtype=torch.complex128, device=walkers.device)
        eng.ctx.force_bias(walkers.Ghalfa, walkers.Ghalfb, vbias)
        return vbias
