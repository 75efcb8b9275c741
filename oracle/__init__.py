"""CPU oracle for the phaseless-AFQMC step (test infrastructure only).

This package is a plain numpy restatement of ipie's CPU path for
``PhaselessGeneric.propagate_walkers`` and ``local_energy_batch`` with a
``GenericRealChol`` Hamiltonian, a ``SingleDet`` trial and ``UHFWalkers``.
Every function cites the ipie file:line it restates (paths relative to the
ipie source tree, v0.7.1).

It is the *checker*: only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it.  The product package
``ipie_amd`` never imports it and never falls back to it.

Pinning status (how each function is checked independently of itself):

* fixtures + half-rotation + e1b/Coulomb/exchange: ipie's known answers
  (estimators/tests/test_generic.py:46-48, test_estimators.py:37-41) and a
  brute-force ERI contraction.
* Green's function / overlap: ``Ghalf psi* = I``, brute-force determinants.
* Taylor exponential: ``scipy.linalg.expm`` at small norm, exact for a
  nilpotent VHS.
* VHS build / packing: full (unpacked) Cholesky contraction; symmetry.
* force bias / mean-field shift: full-G contraction; ``xbar = 0`` at the trial.
* whole propagation (signs, sqrt(dt), mean-field terms, one-body split):
  ``-log <I>_xi / dt -> E_local`` as dt -> 0 with antithetic fields.
* weight update: closed form for a pure one-body propagator with
  eigenvector walkers (magnitude, the 1/2 average, eshift, cosine clip).
* reortho: Q^H Q = I, phi = Q R with positive real diag(R), GF invariance.

No function is "parity unpinned".
"""
