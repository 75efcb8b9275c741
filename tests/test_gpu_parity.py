"""Parity of the gfx950 path (through the C ABI) with the oracle on the same
inputs.  FP64 tolerance: 1e-10, checked step by step with the oracle
resynchronised to the device state.  Per-walker scalars (overlap, weight,
hybrid energy) are held to 1e-10 relative to each walker's own value; arrays of
matrices (phi, Ghalf) and energies to 1e-10 relative to their largest entry."""

import numpy
import pytest

from oracle import afqmc, fixtures
from tests._helpers import Pair, random_walkers, rel_err, rel_err_elem, synthetic_arrays, walker_err

TOL = 1e-10

pytestmark = pytest.mark.gpu


def _torch():
    import torch

    return torch


CASES = {
    # name: (N, X, decay, nalpha, nbeta, W, scale)
    "n24": (24, 72, 12.0, 5, 5, 10, 1.0),
    "n24_norm": (24, 72, 12.0, 5, 5, 10, (12.0 / (24 * 12.0)) ** 0.5),
    "odd": (37, 53, 20.0, 6, 3, 7, 0.1),
    "n10_k2": (10, 30, 8.0, 4, 2, 5, 0.3),
    "n64_n40": (64, 96, 40.0, 40, 40, 4, 0.05),
    "n200_n20": (200, 1000, 150.0, 20, 20, 6, (12.0 / (200 * 150.0)) ** 0.5),
    # fewer Cholesky vectors than basis functions: the exchange forms every T_x
    # instead of the quadratic form g^T Wq g
    "n48_x40": (48, 40, 20.0, 8, 6, 5, 0.1),
}


def make_pair(name, rhf=False, seed=3):
    N, X, D, na, nb, W, scale = CASES[name]
    print(f"[case] {name} N={N} X={X} na={na} nb={nb} W={W} rhf={rhf}", flush=True)
    h1e, chol, ecore, psi = synthetic_arrays(N, X, D, na, nb, scale=scale)
    init = random_walkers(N, na, nb, W, seed=seed, amp=0.2, base=psi)
    return Pair(h1e, chol, psi, (na, nb), W, init=init, rhf=rhf)


@pytest.mark.parametrize("name", list(CASES))
def test_greens_function(name):
    p = make_pair(name)
    ov_dev = p.trial.calc_greens_function(p.walkers).cpu().numpy()
    ov_orc = p.orc.trial.calc_greens_function(p.orc.walkers)
    assert rel_err_elem(ov_dev, ov_orc) < TOL
    assert rel_err(p.walkers.Ghalfa.cpu().numpy(), p.orc.walkers.Ghalfa) < TOL
    assert rel_err(p.walkers.Ghalfb.cpu().numpy(), p.orc.walkers.Ghalfb) < TOL
    ov2 = p.trial.calc_overlap(p.walkers).cpu().numpy()
    assert rel_err_elem(ov2, ov_orc) < TOL


@pytest.mark.parametrize("name", list(CASES))
def test_force_bias_and_local_energy(name):
    from ipie_amd.estimators import local_energy_batch

    p = make_pair(name)
    p.trial.calc_greens_function(p.walkers)
    p.orc.trial.calc_greens_function(p.orc.walkers)
    vb = p.trial.calc_force_bias(p.ham, p.walkers).cpu().numpy()
    vb_o = p.orc.trial.calc_force_bias(p.orc.hamiltonian, p.orc.walkers)
    assert rel_err(vb, vb_o) < TOL
    e = local_energy_batch(p.system, p.ham, p.walkers, p.trial).cpu().numpy()
    e_o = afqmc.local_energy_batch(p.orc.system, p.orc.hamiltonian, p.orc.walkers, p.orc.trial)
    for k in range(3):
        assert rel_err(e[:, k], e_o[:, k]) < TOL, k


def test_energy_estimator_known_answer_on_device():
    """ipie/estimators/tests/test_estimators.py:28-41 through the device GF + energy kernels."""
    import torch

    from ipie_amd import Generic
    from ipie_amd.engine import Engine
    from ipie_amd.estimators import local_energy_batch
    from ipie_amd.trial import SingleDet
    from ipie_amd.walkers import UHFWalkers

    system_o, ham_o, walkers_o, trial_o = fixtures.gen_random_test_instances(10, 8, 30, 10)
    # same draws as the fixture: the trial/initial walker is the complex random determinant
    numpy.random.seed(7)
    wfn = fixtures.get_random_nomsd(8, 8, 10, ndet=1)
    trial = SingleDet(wfn[1][0], (8, 8), 10)
    trial._rchola, trial._rcholb = trial_o._rchola, trial_o._rcholb
    trial._rH1a, trial._rH1b = trial_o._rH1a, trial_o._rH1b
    trial.half_rotated = True
    walkers = UHFWalkers(wfn[1][0], 8, 8, 10, 10)
    eng = Engine(trial, walkers, nchol=30)
    dev = walkers.device
    t = trial.device_tensors(dev)
    eng._set_static_partial(
        chol_packed=torch.from_numpy(ham_o.chol_packed).to(dev), rchola=t["rchola"], rcholb=t["rcholb"],
        rH1a=t["rH1a"], rH1b=t["rH1b"], dt=0.005,
    )
    walkers._engine = eng
    trial.calc_greens_function(walkers)
    energy = local_energy_batch(Generic((8, 8)), None, walkers, trial)
    enumer = torch.sum(walkers.weight * energy[:, 0].real).item()
    assert enumer == pytest.approx(-754.0373585215561, rel=1e-11)


@pytest.mark.parametrize("name", ["n24", "n24_norm", "odd", "n10_k2", "n64_n40", "n200_n20", "n48_x40"])
def test_propagation_steps_with_resync(name):
    torch = _torch()
    p = make_pair(name)
    N, X, D, na, nb, W, scale = CASES[name]
    rs = numpy.random.RandomState(11)
    nsteps = 4
    for step in range(nsteps):
        xi = rs.normal(0.0, 1.0, X * W).reshape(W, X)
        p.prop.xi_source = lambda w, x, xi=xi: xi
        p.orc.propagator.xi_source = lambda w, x, xi=xi: xi
        eshift = 0.0 if step < 2 else float(p.trial.energy.real)
        p.sync_oracle_from_device()
        p.prop.propagate_walkers(p.walkers, p.ham, p.trial, eshift)
        p.orc.propagator.propagate_walkers(p.orc.walkers, p.orc.hamiltonian, p.orc.trial, eshift)
        for attr in ("phia", "phib", "weight", "ovlp", "hybrid_energy", "Ghalfa", "Ghalfb"):
            d = getattr(p.walkers, attr).cpu().numpy()
            o = numpy.asarray(getattr(p.orc.walkers, attr))
            if not numpy.all(numpy.isfinite(o)):
                mask = numpy.isfinite(o)
                assert numpy.array_equal(mask, numpy.isfinite(d)), (attr, step)
                d, o = d[mask], o[mask]
            err = walker_err(attr, d, o)
            assert err < TOL, (attr, step, err)
    assert p.prop.nfb_trig == p.orc.propagator.nfb_trig


def test_stepwise_equals_fused():
    torch = _torch()
    a = make_pair("odd")
    b = make_pair("odd")
    N, X, D, na, nb, W, scale = CASES["odd"]
    xi = numpy.random.RandomState(2).normal(size=(W, X))
    for p in (a, b):
        p.prop.xi_source = lambda w, x: xi
    a.prop.propagate_walkers(a.walkers, a.ham, a.trial, 0.0)
    b.prop.propagate_walkers_stepwise(b.walkers, b.ham, b.trial, 0.0)
    for attr in ("phia", "phib", "weight", "ovlp", "hybrid_energy"):
        assert rel_err(getattr(a.walkers, attr).cpu().numpy(), getattr(b.walkers, attr).cpu().numpy()) < 1e-13


def test_fused_energy_matches_local_energy():
    torch = _torch()
    from ipie_amd.estimators import local_energy_batch

    p = make_pair("n24_norm")
    e_prev = torch.empty(p.W, 3, dtype=torch.complex128, device=p.walkers.device)
    p.trial.calc_greens_function(p.walkers)
    e_ref = local_energy_batch(p.system, p.ham, p.walkers, p.trial).cpu().numpy()
    p.prop.propagate_walkers(p.walkers, p.ham, p.trial, 0.0, energy_prev=e_prev)
    assert rel_err(e_prev.cpu().numpy(), e_ref) < 1e-13


@pytest.mark.parametrize("name", ["n24", "odd", "n64_n40"])
def test_reortho(name):
    p = make_pair(name)
    p.trial.calc_greens_function(p.walkers)
    ov = p.orc.trial.calc_greens_function(p.orc.walkers)
    p.walkers.ovlp.copy_(_torch().from_numpy(ov))
    p.orc.walkers.ovlp = ov.copy()
    detR = p.walkers.reortho().cpu().numpy()
    detR_o = numpy.array(p.orc.walkers.reortho())
    assert rel_err_elem(detR, detR_o) < TOL
    assert rel_err(p.walkers.phia.cpu().numpy(), p.orc.walkers.phia) < 1e-9
    assert rel_err(p.walkers.phib.cpu().numpy(), p.orc.walkers.phib) < 1e-9
    assert rel_err_elem(p.walkers.ovlp.cpu().numpy(), p.orc.walkers.ovlp) < TOL
    assert rel_err(p.walkers.log_detR.cpu().numpy(), numpy.array(p.orc.walkers.log_detR)) < TOL


def _ill_conditioned(N, n, W, kappa, rs):
    """W walkers [N, n] complex with singular values logspaced from 1 to 1/kappa."""
    out = numpy.empty((W, N, n), dtype=numpy.complex128)
    s = numpy.logspace(0.0, -numpy.log10(kappa), n)
    for w in range(W):
        u, _ = numpy.linalg.qr(rs.normal(size=(N, n)) + 1j * rs.normal(size=(N, n)))
        v, _ = numpy.linalg.qr(rs.normal(size=(n, n)) + 1j * rs.normal(size=(n, n)))
        out[w] = (u * s) @ v.conj().T
    return out


@pytest.mark.parametrize("name,kappa", [("n64_n40", 1e6), ("odd", 1e6), ("odd", 1e11)])
def test_reortho_ill_conditioned(name, kappa):
    """Walkers beyond the CholeskyQR2 bound (kappa 1e6: the unshifted Cholesky
    succeeds but R fails the bound; 1e11: the unshifted Cholesky breaks down)
    take the shifted three-pass path: Q orthonormal to 1e-12, X = Q R with R
    upper triangular, and Q, detR as numpy's QR to kappa-scaled tolerances."""
    p = make_pair(name)
    N, X, D, na, nb, W, scale = CASES[name]
    print(f"[case] ill-conditioned walkers kappa={kappa:g}", flush=True)
    rs = numpy.random.RandomState(17)
    xa, xb = _ill_conditioned(N, na, W, kappa, rs), _ill_conditioned(N, nb, W, kappa, rs)
    p.set_phi(xa, xb)
    p.walkers.ovlp.fill_(1.0)
    p.orc.walkers.ovlp = numpy.ones(W, dtype=numpy.complex128)
    detR = p.walkers.reortho().cpu().numpy()
    detR_o = numpy.array(p.orc.walkers.reortho())
    tol = max(1e-9, 1e3 * kappa * 1.1e-16)
    for x, q_dev, q_orc in ((xa, p.walkers.phia, p.orc.walkers.phia), (xb, p.walkers.phib, p.orc.walkers.phib)):
        q = q_dev.cpu().numpy()
        n = x.shape[2]
        for w in range(W):
            qq = q[w].conj().T @ q[w]
            assert numpy.abs(qq - numpy.eye(n)).max() < 1e-12, (kappa, w, "orthogonality")
            r = q[w].conj().T @ x[w]
            assert numpy.abs(numpy.tril(r, -1)).max() < 1e-12 * numpy.abs(r).max(), (kappa, w, "triangular")
            assert numpy.all(r.diagonal().real > 0.0), (kappa, w, "positive diagonal")
        assert rel_err(q, q_orc) < tol, (kappa, "Q vs numpy", rel_err(q, q_orc))
    assert rel_err(detR, detR_o) < tol, (kappa, "detR", rel_err(detR, detR_o))


def test_rhf_walkers():
    name = "n24_norm"
    p = make_pair(name, rhf=True)
    N, X, D, na, nb, W, scale = CASES[name]
    pa = p.walkers.phia.cpu().numpy()
    p.set_phi(pa, pa)
    xi = numpy.random.RandomState(4).normal(size=(W, X))
    p.prop.xi_source = lambda w, x: xi
    p.orc.propagator.xi_source = lambda w, x: xi
    p.prop.propagate_walkers(p.walkers, p.ham, p.trial, 0.0)
    p.orc.propagator.propagate_walkers(p.orc.walkers, p.orc.hamiltonian, p.orc.trial, 0.0)
    for attr in ("phia", "weight", "ovlp", "hybrid_energy"):
        assert walker_err(attr, getattr(p.walkers, attr).cpu().numpy(), getattr(p.orc.walkers, attr)) < TOL, attr


@pytest.mark.parametrize("name,nb", [("n24_norm", 5), ("odd", 3), ("n64_n40", 40), ("n10_k2", 0)])
def test_rhf_local_energy(name, nb):
    """walkers.rhf = True dispatches to the RHF energy (local_energy_batch.py:58-59 ->
    local_energy_sd.py:106-155 with the Coulomb of :44-73): every term from the
    alpha Green's function, doubled, whatever nbeta is (nbeta < nalpha included,
    and nbeta = 0, where the reference's force bias still doubles,
    force_bias.py:101-104).  The energy of afqmc_propagate's incoming walkers
    and its force bias take the same branch."""
    torch = _torch()
    from ipie_amd.estimators import local_energy_batch

    N, X, D, na, _, W, scale = CASES[name]
    print(f"[case] rhf energy {name} na={na} nb={nb}", flush=True)
    h1e, chol, ecore, psi = synthetic_arrays(N, X, D, na, nb, scale=scale)
    init = random_walkers(N, na, nb, W, seed=5, amp=0.2, base=psi)
    p = Pair(h1e, chol, psi, (na, nb), W, init=init, rhf=True)
    pa = init[:, :, :na]
    p.set_phi(pa, pa[:, :, :nb])
    p.trial.calc_greens_function(p.walkers)
    p.orc.trial.calc_greens_function(p.orc.walkers)
    assert rel_err(p.walkers.Ghalfa.cpu().numpy(), p.orc.walkers.Ghalfa) < TOL
    e = local_energy_batch(p.system, p.ham, p.walkers, p.trial).cpu().numpy()
    e_o = afqmc.local_energy_batch(p.orc.system, p.orc.hamiltonian, p.orc.walkers, p.orc.trial)
    e_rhf = afqmc.local_energy_single_det_rhf_batch(p.orc.system, p.orc.hamiltonian, p.orc.walkers, p.orc.trial)
    assert numpy.array_equal(e_o, e_rhf)
    for k in range(3):
        assert rel_err(e[:, k], e_o[:, k]) < TOL, k
    vb = p.trial.calc_force_bias(p.ham, p.walkers).cpu().numpy()
    vb_o = p.orc.trial.calc_force_bias(p.orc.hamiltonian, p.orc.walkers)
    assert rel_err(vb, vb_o) < TOL
    # the fused step's energy of the incoming walkers, and the step itself
    xi = numpy.random.RandomState(6).normal(size=(W, X))
    p.prop.xi_source = lambda w, x: xi
    p.orc.propagator.xi_source = lambda w, x: xi
    e_prev = torch.empty(W, 3, dtype=torch.complex128, device=p.walkers.device)
    p.prop.propagate_walkers(p.walkers, p.ham, p.trial, 0.0, energy_prev=e_prev)
    p.orc.propagator.propagate_walkers(p.orc.walkers, p.orc.hamiltonian, p.orc.trial, 0.0)
    assert rel_err(e_prev.cpu().numpy(), e_o) < TOL
    for attr in ("phia", "weight", "ovlp", "hybrid_energy"):
        err = walker_err(attr, getattr(p.walkers, attr).cpu().numpy(), getattr(p.orc.walkers, attr))
        assert err < TOL, (attr, err)


def test_device_normals():
    torch = _torch()
    from ipie_amd import _lib

    ctx = _lib.Context(8, 4, 2, 2, 4)
    out = torch.empty(1_000_001, dtype=torch.float64, device="cuda")
    ctx.draw_normals(out, seed=7, counter=1)
    x = out.cpu().numpy()
    assert abs(x.mean()) < 5e-3 and abs(x.var() - 1.0) < 5e-3
    out2 = torch.empty_like(out)
    ctx.draw_normals(out2, seed=7, counter=1)
    assert torch.equal(out, out2)
    ctx.draw_normals(out2, seed=7, counter=2)
    assert not torch.equal(out, out2)
