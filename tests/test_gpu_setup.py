"""Device set-up producers (afqmc_setup_*) against the oracle's set-up
functions on the same arrays: packing bit-exact; h1e_mod, half rotation,
mean-field shift and expH1 (scaled-and-squared Taylor against
scipy.linalg.expm) within 1e-12 of the largest entry; a propagator built with
device_setup=True steps like one built on the host (1e-10)."""

import numpy
import pytest

from oracle import afqmc
from tests._helpers import random_walkers, rel_err, synthetic_arrays

pytestmark = pytest.mark.gpu

TOL = 1e-12
CASES = {
    # name: (N, X, decay, nalpha, nbeta, W, scale, dt)
    "n24": (24, 72, 12.0, 5, 5, 6, 1.0, 0.005),
    "odd": (37, 53, 20.0, 6, 3, 5, 0.1, 0.005),
    "n48_big_dt": (48, 100, 30.0, 8, 6, 4, 0.3, 0.5),  # ||dt/2 H1|| > 1/2: scaling and squaring
}


def _build(name, device_setup):
    import torch  # noqa: F401

    from ipie_amd import Generic
    from ipie_amd.hamiltonian import GenericRealChol
    from ipie_amd.propagator import PhaselessGeneric
    from ipie_amd.trial import SingleDet
    from ipie_amd.walkers import UHFWalkers

    N, X, D, na, nb, W, scale, dt = CASES[name]
    print(f"[case] set-up {name} N={N} X={X} na={na} nb={nb} device_setup={device_setup}", flush=True)
    h1e, chol, ecore, psi = synthetic_arrays(N, X, D, na, nb, scale=scale)
    ham = GenericRealChol(h1e, chol, ecore=ecore, host_setup=not device_setup)
    trial = SingleDet(psi, (na, nb), N)
    walkers = UHFWalkers(psi, na, nb, N, W)
    prop = PhaselessGeneric(dt)
    prop.build(ham, trial, walkers, device_setup=device_setup)
    trial.calculate_energy(Generic((na, nb)), ham)
    return dict(h1e=h1e, chol=chol, ecore=ecore, psi=psi, ham=ham, trial=trial, walkers=walkers, prop=prop,
                dims=(N, X, na, nb, W, dt))


@pytest.mark.parametrize("name", list(CASES))
def test_setup_kernels_match_oracle(name):
    b = _build(name, device_setup=True)
    N, X, na, nb, W, dt = b["dims"]
    oham = afqmc.GenericRealChol(b["h1e"], b["chol"], ecore=b["ecore"])
    otrial = afqmc.SingleDet(b["psi"], (na, nb), N)
    otrial.half_rotate(oham)
    mf = afqmc.construct_mean_field_shift(oham, otrial)
    expH1 = afqmc.construct_one_body_propagator(oham, mf, dt)
    st = b["walkers"]._engine.static
    assert numpy.array_equal(st["chol_packed"].cpu().numpy(), oham.chol_packed)
    assert rel_err(b["ham"].h1e_mod, oham.h1e_mod) < TOL
    assert rel_err(st["rchola"].cpu().numpy(), otrial._rchola.real) < TOL
    assert rel_err(st["rcholb"].cpu().numpy(), otrial._rcholb.real) < TOL
    assert rel_err(st["rH1a"].cpu().numpy(), otrial._rH1a) < TOL
    assert rel_err(st["rH1b"].cpu().numpy(), otrial._rH1b) < TOL
    assert rel_err(b["prop"].mf_shift, mf) < TOL
    assert numpy.abs(numpy.imag(expH1)).max() < 1e-14
    assert rel_err(b["prop"].expH1, expH1.real) < TOL
    # the trial energy from the device half rotation
    otrial.calculate_energy(afqmc.Generic((na, nb)), oham)
    assert abs(b["trial"].energy - otrial.energy) < 1e-10 * abs(otrial.energy)


def test_device_setup_steps_like_host_setup():
    import torch

    hb, db = _build("odd", device_setup=False), _build("odd", device_setup=True)
    N, X, na, nb, W, dt = hb["dims"]
    init = random_walkers(N, na, nb, W, seed=4, amp=0.2, base=hb["psi"])
    for b in (hb, db):
        b["walkers"].phia.copy_(torch.from_numpy(numpy.ascontiguousarray(init[:, :, :na])))
        b["walkers"].phib.copy_(torch.from_numpy(numpy.ascontiguousarray(init[:, :, na:])))
        b["trial"].calc_greens_function(b["walkers"])
        for _ in range(3):
            b["prop"].propagate_walkers(b["walkers"], b["ham"], b["trial"], 0.0)
    for k in ("phia", "phib", "weight", "ovlp"):
        err = rel_err(getattr(db["walkers"], k).cpu().numpy(), getattr(hb["walkers"], k).cpu().numpy())
        assert err < 1e-10, (k, err)
