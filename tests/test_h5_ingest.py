"""HDF5 ingest (ipie_amd.h5read / ipie_amd.io) on ipie's h10 test file, and
the oracle's trial energy on that real molecule against ipie's recorded
block-0 estimate (walkers at the trial, ipie/qmc/tests/reference_data/
h10_cc-pvtz_batched/reference.json): a known answer from the reference's own
run, copied with the file by tests/golden/make_h10_fixture.py."""

import json
import os

import numpy
import pytest

from oracle import afqmc

HERE = os.path.dirname(os.path.abspath(__file__))
H5 = os.path.join(HERE, "golden", "h10_afqmc.h5")
REF = os.path.join(HERE, "golden", "h10_block0.json")


def test_reader_structure():
    from ipie_amd.h5read import H5File

    with H5File(H5) as f:
        assert f.keys() == ["Hamiltonian", "Wavefunction"]
        assert f.keys("Hamiltonian/DenseFactorized") == ["L"]
        assert f["Hamiltonian/DenseFactorized/L"].shape == (100, 27)
        assert list(f["Hamiltonian/dims"]) == [0, 0, 0, 10, 5, 5, 0, 27]
        assert "Wavefunction/NOMSD/PsiT_0" in f and "Wavefunction/PHMSD" not in f
        with pytest.raises(KeyError):
            f["Hamiltonian/nothing"]


def test_io_readers():
    from ipie_amd import io

    hcore, chol, ecore = io.read_integrals(H5)
    assert hcore.shape == (10, 10) and chol.shape == (100, 27)
    assert numpy.allclose(hcore, hcore.T, atol=1e-12)
    # L_x symmetric in its orbital pair
    L3 = chol.reshape(10, 10, 27)
    assert numpy.allclose(L3, L3.transpose(1, 0, 2), atol=1e-12)
    (coeffs, wfn), psi0, nelec = io.read_qmcpack_wfn_hdf(H5, get_nelec=True)
    assert nelec == (5, 5) and wfn.shape == (1, 10, 10) and coeffs.shape == (1,)
    # orthonormal orbitals, and the initial walker is the trial
    pa = wfn[0, :, :5]
    assert numpy.allclose(pa.conj().T @ pa, numpy.eye(5), atol=1e-10)
    assert numpy.abs(psi0 - wfn[0]).max() < 1e-12


def test_trial_energy_matches_reference_block0():
    from ipie_amd import io

    ref = json.load(open(REF))
    hcore, chol, ecore = io.read_integrals(H5)
    (coeffs, wfn), psi0, nelec = io.read_qmcpack_wfn_hdf(H5, get_nelec=True)
    ham = afqmc.GenericRealChol(numpy.array([hcore, hcore]), chol, ecore=ecore)
    trial = afqmc.SingleDet(wfn[0], nelec, 10)
    trial.half_rotate(ham)
    e = trial.calculate_energy(afqmc.Generic(nelec), ham)
    assert abs(e.real - ref["ETotal"]) < 1e-12
    assert abs(trial.e1b.real - ref["E1Body"]) < 1e-12
    assert abs(trial.e2b.real - ref["E2Body"]) < 1e-12
