"""Synthetic restricted Hamiltonians used for the benchmark and parity workloads.

Generator (one seed, numpy legacy ``RandomState``; the same draws on every host):

* ``h1e``: symmetric, off-diagonal entries N(0, 0.1) (0.1 is the standard
  deviation; the matrix is (A + A^T)/sqrt(2) with A ~ N(0, 0.1)), diagonal
  drawn uniformly in [-5, -1].
* ``L^g`` for g = 0..X-1: (B + B^T)/sqrt(2) with B ~ N(0, 1), scaled by
  exp(-g / decay).  ipie layout ``chol[:, g] = vec(L^g)`` ([N*N, X]).
* trial: the ``nocc`` lowest eigenvectors of h1e (real), used for both spins.

The draws are, in order: A (N*N), the diagonal (N), then one B (N*N) per
Cholesky vector.
"""

from dataclasses import dataclass

import numpy


@dataclass(frozen=True)
class Workload:
    """A named synthetic AFQMC workload."""

    name: str
    nbasis: int
    nchol: int
    decay: float
    nalpha: int
    nbeta: int
    nwalkers: int
    dt: float = 0.005
    nsteps: int = 50
    seed: int = 7


# The five workloads: plain CPU-runnable, FP64 realistic, mixed-precision
# headline, learned-exchange, and (per system) the multi-system batch sizes.
WORKLOADS = {
    "n24": Workload("n24", 24, 72, 12.0, 5, 5, 10, nsteps=50),
    "n200": Workload("n200", 200, 1000, 150.0, 20, 20, 512, nsteps=200),
    "n400": Workload("n400", 400, 2400, 300.0, 40, 40, 1024, nsteps=100),
    "n300": Workload("n300", 300, 1800, 225.0, 30, 30, 2048, nsteps=100),
}


def multi_system_workloads(count=16, seed=7):
    """The uneven multi-system batch: `count` systems drawn from one seed with
    N uniform in [64, 256], X = 5 N, N / 10 occupied orbitals per spin
    (rounded), walkers uniform in [64, 512]; decay D = 0.75 N (as the other
    workloads' D / N); each system's generator seed is seed + 1 + its index."""
    rs = numpy.random.RandomState(seed)
    out = []
    for i in range(count):
        n = int(rs.randint(64, 257))
        w = int(rs.randint(64, 513))
        nocc = max(1, int(round(n / 10)))
        out.append(Workload(f"sys{i}", n, 5 * n, 0.75 * n, nocc, nocc, w, seed=seed + 1 + i))
    return out


def make_hamiltonian(nbasis, nchol, decay, seed=7, dtype=numpy.float64):
    """Return (h1e [2,N,N], chol [N*N, X], ecore) for the synthetic generator."""
    rs = numpy.random.RandomState(seed)
    a = rs.normal(0.0, 0.1, size=(nbasis, nbasis))
    h1 = (a + a.T) / numpy.sqrt(2.0)
    diag = rs.uniform(-5.0, -1.0, size=nbasis)
    h1[numpy.diag_indices(nbasis)] = diag
    chol = numpy.empty((nbasis * nbasis, nchol), dtype=dtype)
    inv_sqrt2 = 1.0 / numpy.sqrt(2.0)
    for g in range(nchol):
        b = rs.normal(0.0, 1.0, size=(nbasis, nbasis))
        lg = (b + b.T) * (inv_sqrt2 * numpy.exp(-g / decay))
        chol[:, g] = lg.ravel()
    h1e = numpy.array([h1, h1])
    return h1e, chol, 0.0


def make_trial(h1e, nalpha, nbeta):
    """RHF trial: the lowest eigenvectors of h1e[0] for both spins -> [N, na+nb]."""
    evals, evecs = numpy.linalg.eigh(h1e[0])
    psia = evecs[:, :nalpha]
    psib = evecs[:, :nbeta]
    return numpy.ascontiguousarray(numpy.hstack([psia, psib]))


def make_workload(w: Workload):
    """Return (h1e, chol, ecore, psi) for a workload."""
    h1e, chol, ecore = make_hamiltonian(w.nbasis, w.nchol, w.decay, seed=w.seed)
    psi = make_trial(h1e, w.nalpha, w.nbeta)
    return h1e, chol, ecore, psi
