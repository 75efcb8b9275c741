"""A real molecule through the device path: ipie's h10 test file (integrals
and trial) read by ipie_amd.io, static tensors built by the device set-up
kernels; the local energy of walkers at the trial reproduces ipie's recorded
block-0 estimate (reference.json), and propagation matches the oracle step by
step with the same fields (FP64, 1e-10)."""

import json
import os

import numpy
import pytest

from oracle import afqmc
from tests._helpers import rel_err

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
H5 = os.path.join(HERE, "golden", "h10_afqmc.h5")
REF = os.path.join(HERE, "golden", "h10_block0.json")


def test_h10_block0_energy_on_device():
    from ipie_amd import io
    from ipie_amd.estimators import local_energy_batch

    print("[case] h10 block-0 energy", flush=True)
    ref = json.load(open(REF))
    system, ham, trial, walkers, prop = io.load_generic(H5, nwalkers=8)
    trial.calc_greens_function(walkers)
    e = local_energy_batch(system, ham, walkers, trial).cpu().numpy()
    assert numpy.abs(e[:, 0].real - ref["ETotal"]).max() < 1e-10
    assert numpy.abs(e[:, 1].real - ref["E1Body"]).max() < 1e-10
    assert numpy.abs(e[:, 2].real - ref["E2Body"]).max() < 1e-10
    assert abs(trial.energy.real - ref["ETotal"]) < 1e-10


def test_h10_steps_match_oracle():
    import torch

    from ipie_amd import io

    print("[case] h10 propagation vs oracle", flush=True)
    W = 6
    system, ham, trial, walkers, prop = io.load_generic(H5, nwalkers=W)
    hcore, chol, ecore = io.read_integrals(H5)
    (coeffs, wfn), psi0, nelec = io.read_qmcpack_wfn_hdf(H5, get_nelec=True)
    orc = afqmc.build_state(numpy.array([hcore, hcore]), chol, wfn[0], nelec, W, prop.dt, ecore=ecore)
    rs = numpy.random.RandomState(3)
    trial.calc_greens_function(walkers)
    eshift = float(trial.energy.real)
    for step in range(4):
        xi = rs.normal(size=(W, chol.shape[1]))
        prop.xi_source = lambda nw, nx, xi=xi: xi
        orc.propagator.xi_source = lambda nw, nx, xi=xi: xi
        prop.propagate_walkers(walkers, ham, trial, eshift)
        orc.propagator.propagate_walkers(orc.walkers, orc.hamiltonian, orc.trial, eshift)
        for attr in ("phia", "phib", "weight", "ovlp"):
            err = rel_err(getattr(walkers, attr).cpu().numpy(), getattr(orc.walkers, attr))
            assert err < 1e-10, (step, attr, err)


def test_h10_energy_estimator_block0():
    """The EnergyEstimator drop-in (weighted sums on the device) gives ipie's
    block-0 row: ENumer / EDenom etc. after post_reduce_hook."""
    import torch

    from ipie_amd import io
    from ipie_amd.estimators import EnergyEstimator

    print("[case] h10 EnergyEstimator", flush=True)
    ref = json.load(open(REF))
    W = 25
    system, ham, trial, walkers, prop = io.load_generic(H5, nwalkers=W)
    rs = numpy.random.RandomState(1)
    w = rs.uniform(0.5, 1.5, size=W)
    walkers.weight.copy_(torch.from_numpy(w))
    est = EnergyEstimator(system, ham, trial)
    data = numpy.array(est.compute_estimator(system, walkers, ham, trial))
    assert abs(data[est.get_index("EDenom")] - w.sum()) < 1e-12 * w.sum()
    est.post_reduce_hook(data)
    assert abs(data[est.get_index("ETotal")].real - ref["ETotal"]) < 1e-10
    assert abs(data[est.get_index("E1Body")].real - ref["E1Body"]) < 1e-10
    assert abs(data[est.get_index("E2Body")].real - ref["E2Body"]) < 1e-10
