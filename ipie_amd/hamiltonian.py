"""Real-Cholesky Hamiltonian container (drop-in for ipie's GenericRealChol).

Mirrors ipie/hamiltonians/generic.py:46-98: ``H1`` [2, N, N], ``chol``
[N*N, X], ``chol_packed`` [N(N+1)/2, X] in ``numpy.triu_indices`` order,
``h1e_mod`` = H1 - 1/2 sum_{k,x} L_{ik,x} L_{jk,x} (generic.py:31-43), and the
same attribute names.  The set-up arithmetic runs once on the host, as in the
reference; ``device_tensors`` uploads the packed Cholesky vectors once.
"""

import numpy
import torch


class GenericRealChol:
    def __init__(self, h1e, chol, ecore=0.0, verbose=False, host_setup=True):
        """host_setup=False leaves chol_packed and h1e_mod to the device set-up
        (PhaselessGeneric.build(..., device_setup=True))."""
        h1e = numpy.asarray(h1e)
        assert h1e.shape[0] == 2, "h1e must hold both spin components"
        chol = numpy.asarray(chol)
        assert chol.dtype == numpy.float64, "GenericRealChol needs float64 Cholesky vectors"
        self.H1 = h1e
        self.ecore = float(ecore)
        self.nbasis = h1e.shape[-1]
        self.chol = chol
        self.nchol = chol.shape[-1]
        self.nfields = self.nchol
        assert self.nbasis**2 == chol.shape[0]
        self.sym_idx = numpy.triu_indices(self.nbasis)
        self.sym_idx_i = self.sym_idx[0].copy()
        self.sym_idx_j = self.sym_idx[1].copy()
        self.chunked = False
        self._dev = {}
        self.chol_packed = None
        self.h1e_mod = None
        if host_setup:
            chol3 = chol.reshape(self.nbasis, self.nbasis, self.nchol)
            self.chol_packed = numpy.ascontiguousarray(chol3[self.sym_idx_i, self.sym_idx_j, :])
            # one-body term from re-ordering the two-body operator
            chol_view = chol.reshape(self.nbasis, self.nbasis * self.nchol)
            v0 = 0.5 * chol_view.dot(chol_view.T)
            self.h1e_mod = numpy.array([h1e[0] - v0, h1e[1] - v0])
        if verbose:
            print(f"# GenericRealChol: N={self.nbasis} X={self.nchol}")

    def device_tensors(self, device):
        """chol_packed as a device tensor (uploaded once per device)."""
        key = str(device)
        if key not in self._dev:
            if self.chol_packed is None:
                raise RuntimeError("chol_packed was left to the device set-up: build the propagator with "
                                   "device_setup=True")
            self._dev[key] = {
                "chol_packed": torch.from_numpy(self.chol_packed).to(device),
            }
        return self._dev[key]

    def hijkl(self, i, j, k, l):
        ik = i * self.nbasis + k
        jl = j * self.nbasis + l
        return numpy.dot(self.chol[ik], self.chol[jl])


def Generic(h1e, chol, ecore=0.0, verbose=False):
    """Factory mirroring ipie.hamiltonians.generic.Generic for real Cholesky vectors."""
    chol = numpy.asarray(chol)
    if chol.dtype != numpy.float64:
        raise NotImplementedError("only real (float64) Cholesky vectors run on the MI355X path")
    return GenericRealChol(h1e, chol, ecore=ecore, verbose=verbose)
