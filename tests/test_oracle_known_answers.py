"""The oracle reproduces the known answers hard-coded in ipie's own tests."""

import numpy
import pytest

from oracle import afqmc, fixtures


def test_local_energy_cholesky_known_answer():
    """ipie/estimators/tests/test_generic.py:32-69 (nmo=24, nelec=(4,2), seed 7)."""
    numpy.random.seed(7)
    nmo, nelec = 24, (4, 2)
    h1e, chol, enuc, eri = fixtures.generate_hamiltonian(nmo, nelec, cplx=False)
    system = afqmc.Generic(nelec)
    ham = afqmc.GenericRealChol(numpy.array([h1e, h1e]), chol.reshape((-1, nmo * nmo)).T.copy(), ecore=enuc)
    ci, wfn = fixtures.get_random_nomsd(system.nup, system.ndown, ham.nbasis, ndet=1, cplx=False)
    trial = afqmc.SingleDet(wfn[0], nelec, nmo)
    expected = (20.6826247016273, 23.0173528796140, -2.3347281779866)
    e_full = afqmc.local_energy_generic_cholesky(system, ham, trial.G)
    for got, ref in zip(e_full, expected):
        assert got == pytest.approx(ref, rel=1e-12)
    trial.half_rotate(ham)
    e_opt = afqmc.local_energy_cholesky_opt(trial, ham, trial.Ghalf)
    for got, ref in zip(e_opt, expected):
        assert got == pytest.approx(ref, rel=1e-12)


def test_energy_estimator_known_answer():
    """ipie/estimators/tests/test_estimators.py:28-41 (GF with a complex trial + UHF batch energy)."""
    system, ham, walkers, trial = fixtures.gen_random_test_instances(10, 8, 30, 10)
    trial.calc_greens_function(walkers)
    energy = afqmc.local_energy_batch(system, ham, walkers, trial)
    enumer = numpy.sum(walkers.weight * energy[:, 0].real)
    edenom = numpy.sum(walkers.weight)
    assert enumer == pytest.approx(-754.0373585215561, rel=1e-12)
    assert enumer / edenom == pytest.approx(-75.40373585215562, rel=1e-12)


def test_trial_energy_matches_full_g_form():
    """SingleDet.calculate_energy equals the full-G Cholesky energy at walker = trial."""
    numpy.random.seed(3)
    nmo, nelec = 9, (3, 2)
    h1e, chol, enuc, eri = fixtures.generate_hamiltonian(nmo, nelec, cplx=False)
    system = afqmc.Generic(nelec)
    ham = afqmc.GenericRealChol(numpy.array([h1e, h1e]), chol.reshape((-1, nmo * nmo)).T.copy(), ecore=enuc)
    trial, wfn, init = fixtures.build_random_single_det_trial(nelec, nmo)
    trial.half_rotate(ham)
    e = trial.calculate_energy(system, ham)
    e_full = afqmc.local_energy_generic_cholesky(system, ham, trial.G)
    assert e == pytest.approx(e_full[0], rel=1e-12)
