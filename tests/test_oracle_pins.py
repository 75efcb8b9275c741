"""Independent checks that pin every oracle function (closed forms, invariants,
brute force).  A planted mistake in the oracle must fail at least one of them;
``test_planted_mistakes_are_caught`` shows that for the propagation chain."""

import numpy
import pytest
import scipy.linalg

from oracle import afqmc, fixtures


def _random_case(nmo=8, nelec=(3, 2), nwalkers=3, seed=1, tol=1e-8):
    numpy.random.seed(seed)
    h1e, chol, enuc, eri = fixtures.generate_hamiltonian(nmo, nelec, cplx=False, tol=tol)
    chol = chol.reshape((-1, nmo * nmo)).T.copy()
    psi = numpy.linalg.qr(numpy.random.rand(nmo, sum(nelec)))[0]
    init = psi + 0.3 * (numpy.random.rand(nmo, sum(nelec)) + 1j * numpy.random.rand(nmo, sum(nelec)))
    st = afqmc.build_state(numpy.array([h1e, h1e]), chol, psi, nelec, nwalkers, 0.005, init=init, ecore=enuc)
    # give each walker a different state
    for iw in range(nwalkers):
        st.walkers.phia[iw] += 0.1 * iw * numpy.random.rand(nmo, nelec[0])
        st.walkers.phib[iw] += 0.1 * iw * numpy.random.rand(nmo, nelec[1])
    return st, eri


def test_local_energy_vs_brute_force_eri():
    """E1 = tr(h G), E2 = 1/2 sum (ij|kl) G_ij G_kl - 1/2 sum_s (ij|kl) G^s_il G^s_kj with (ij|kl) from L."""
    st, _ = _random_case()
    st.trial.calc_greens_function(st.walkers, build_full=True)
    energy = afqmc.local_energy_batch(st.system, st.hamiltonian, st.walkers, st.trial)
    N = st.hamiltonian.nbasis
    L = st.hamiltonian.chol.reshape(N, N, -1)
    eri = numpy.einsum("ijx,klx->ijkl", L, L)
    Ga, Gb = st.walkers.Ga, st.walkers.Gb
    G = Ga + Gb
    e1 = numpy.einsum("ij,wij->w", st.hamiltonian.H1[0], G) + st.hamiltonian.ecore
    ecoul = 0.5 * numpy.einsum("ijkl,wij,wkl->w", eri, G, G)
    exx = 0.5 * numpy.einsum("ijkl,wil,wkj->w", eri, Ga, Ga) + 0.5 * numpy.einsum("ijkl,wil,wkj->w", eri, Gb, Gb)
    numpy.testing.assert_allclose(energy[:, 1], e1, rtol=1e-11, atol=1e-11)
    numpy.testing.assert_allclose(energy[:, 2], ecoul - exx, rtol=1e-11, atol=1e-11)


def test_rhf_energy_equals_uhf_with_equal_spins():
    st, _ = _random_case(nelec=(3, 3))
    st.walkers.phib[:] = st.walkers.phia
    st.trial._rcholb = st.trial._rchola.copy()
    st.trial._rH1b = st.trial._rH1a.copy()
    st.trial.psi0b = st.trial.psi0a.copy()
    st.trial.calc_greens_function(st.walkers)
    e_uhf = afqmc.local_energy_batch(st.system, st.hamiltonian, st.walkers, st.trial)
    st.walkers.rhf = True
    e_rhf = afqmc.local_energy_batch(st.system, st.hamiltonian, st.walkers, st.trial)
    numpy.testing.assert_allclose(e_rhf, e_uhf, rtol=1e-12, atol=1e-12)


def test_greens_function_identities():
    st, _ = _random_case()
    ovlp = st.trial.calc_greens_function(st.walkers, build_full=True)
    for iw in range(st.walkers.nwalkers):
        numpy.testing.assert_allclose(st.walkers.Ghalfa[iw].dot(st.trial.psi0a.conj()), numpy.eye(3), atol=1e-12)
        numpy.testing.assert_allclose(st.walkers.Ghalfb[iw].dot(st.trial.psi0b.conj()), numpy.eye(2), atol=1e-12)
        ga = st.walkers.Ga[iw]
        numpy.testing.assert_allclose(ga.dot(ga), ga, atol=1e-11)  # idempotent
        da = numpy.linalg.det(st.walkers.phia[iw].T.dot(st.trial.psi0a.conj()))
        db = numpy.linalg.det(st.walkers.phib[iw].T.dot(st.trial.psi0b.conj()))
        assert ovlp[iw] == pytest.approx(da * db, rel=1e-12)
    numpy.testing.assert_allclose(st.trial.calc_overlap(st.walkers), ovlp, rtol=1e-12)


def test_taylor_exponential_vs_expm():
    """ipie/propagation/tests/test_operations.py:32-58: 6 terms match expm at small norm; exact if nilpotent."""
    rs = numpy.random.RandomState(0)
    N, n = 10, 4
    A = rs.rand(N, N) + 1j * rs.rand(N, N)
    VHS = 0.005 * (A + A.conj().T)
    phi = rs.rand(N, n) + 1j * rs.rand(N, n)
    out = afqmc.apply_exponential(phi.copy(), VHS, 6)
    numpy.testing.assert_allclose(out, scipy.linalg.expm(VHS).dot(phi), atol=1e-10)
    Nil = numpy.triu(rs.rand(N, N), k=1)  # nilpotent, Nil^N = 0
    out = afqmc.apply_exponential(phi.copy(), Nil, N)
    numpy.testing.assert_allclose(out, scipy.linalg.expm(Nil).dot(phi), rtol=1e-12, atol=1e-12)


def test_vhs_equals_unpacked_contraction():
    st, _ = _random_case()
    X = st.hamiltonian.nchol
    N = st.hamiltonian.nbasis
    rs = numpy.random.RandomState(2)
    xs = rs.normal(size=(X, 3)) + 1j * rs.normal(size=(X, 3))
    vhs = st.propagator.construct_VHS(st.hamiltonian, xs)
    L = st.hamiltonian.chol.reshape(N, N, X)
    ref = 1j * numpy.sqrt(st.propagator.dt) * numpy.einsum("ijx,xw->wij", L, xs)
    numpy.testing.assert_allclose(vhs, ref, rtol=1e-12, atol=1e-13)
    numpy.testing.assert_allclose(vhs, vhs.transpose(0, 2, 1), atol=0)


def test_force_bias_equals_full_g_form_and_vanishes_at_trial():
    """vbias = einsum('wi,ix->wx', Ga + Gb, chol) (propagation/tests/test_generic_complex.py:416)."""
    st, _ = _random_case()
    st.trial.calc_greens_function(st.walkers, build_full=True)
    vb = st.trial.calc_force_bias(st.hamiltonian, st.walkers)
    N = st.hamiltonian.nbasis
    G = (st.walkers.Ga + st.walkers.Gb).reshape(-1, N * N)
    numpy.testing.assert_allclose(vb, G.dot(st.hamiltonian.chol), rtol=1e-11, atol=1e-11)
    # at walker = trial: xbar = -sqrt(dt) (i vbias - mf_shift) = 0
    st.walkers.phia[:] = st.trial.psi0a
    st.walkers.phib[:] = st.trial.psi0b
    st.trial.calc_greens_function(st.walkers)
    vb = st.trial.calc_force_bias(st.hamiltonian, st.walkers)
    xbar = -st.propagator.sqrt_dt * (1j * vb - st.propagator.mf_shift)
    assert numpy.abs(xbar).max() < 1e-12


def _hybrid_estimate(dt, M=12000, plant=None, seed=11):
    numpy.random.seed(seed)
    nmo, nelec = 6, (3, 2)
    h1e, chol, enuc, eri = fixtures.generate_hamiltonian(nmo, nelec, cplx=False, tol=1e-8)
    chol = chol.reshape((-1, nmo * nmo)).T.copy()
    psi = numpy.linalg.qr(numpy.random.rand(nmo, 5))[0]
    init = psi + 0.3 * (numpy.random.rand(nmo, 5) + 1j * numpy.random.rand(nmo, 5))
    st = afqmc.build_state(numpy.array([h1e, h1e]), chol, psi, nelec, 2 * M, dt, init=init, ecore=enuc)
    st.trial.calc_greens_function(st.walkers)
    e_loc = afqmc.local_energy_batch(st.system, st.hamiltonian, st.walkers, st.trial)[0, 0]
    mf = st.propagator.mf_shift
    mf_core = enuc + 0.5 * numpy.dot(mf, mf)
    z = numpy.random.RandomState(5).normal(size=(M, st.hamiltonian.nchol))
    xi = numpy.empty((2 * M, z.shape[1]))
    xi[0::2], xi[1::2] = z, -z  # antithetic pairs cancel the O(sqrt(dt)) noise
    st.propagator.xi_source = lambda w, x: xi
    if plant is not None:
        plant(st)
    st.propagator.propagate_walkers(st.walkers, st.hamiltonian, st.trial, 0.0)
    est = -numpy.log(numpy.mean(numpy.exp(-dt * st.walkers.hybrid_energy))) / dt
    return est, e_loc - mf_core


@pytest.mark.slow
def test_hybrid_energy_converges_to_local_energy():
    """E_xi[exp(-dt E_hybrid)] = <psi|exp(-dt H)|phi>/<psi|phi> + O(dt^2): the whole step
    (one-body split, mean-field shift, force bias, VHS sign and scale, Taylor,
    cmf, cfb, overlap ratio) must give -log <I> / dt -> E_L - ecore - 1/2 sum mf^2."""
    est, target = _hybrid_estimate(1e-4)
    assert abs(est - target) < 0.01


@pytest.mark.slow
def test_planted_mistakes_are_caught():
    def flip_cmf(st):
        orig = st.propagator.propagate_walkers_two_body

        def two_body(w, h, t):
            cmf, cfb = orig(w, h, t)
            return -cmf, cfb

        st.propagator.propagate_walkers_two_body = two_body

    def drop_mean_field(st):
        st.propagator.mf_shift = 0.0 * st.propagator.mf_shift

    def wrong_vhs_sign(st):
        st.propagator.isqrt_dt = -st.propagator.isqrt_dt

    for plant in (flip_cmf, drop_mean_field, wrong_vhs_sign):
        est, target = _hybrid_estimate(1e-4, M=3000, plant=plant)
        assert abs(est - target) > 0.05, plant.__name__


def test_weight_update_closed_form():
    """No two-body term and eigenvector walkers: the overlap ratio is exp(-dt E0) exactly, so
    w_1 = exp(-dt (E0/2 - eshift)) cos-projected, w_2 = w_1 exp(-dt (E0 - eshift)); the hybrid
    energy cap and the cosine clip are checked with a complex diagonal shift."""
    N, nelec, dt = 6, (2, 2), 0.005
    rs = numpy.random.RandomState(4)
    A = rs.normal(size=(N, N))
    h = 0.5 * (A + A.T)
    evals, evecs = numpy.linalg.eigh(h)
    psi = numpy.hstack([evecs[:, :2], evecs[:, :2]])
    chol = numpy.zeros((N * N, 1))
    E0 = 2 * (evals[0] + evals[1])

    def run(c, eshift, nsteps):
        hc = h + 1j * c * numpy.eye(N)
        st = afqmc.build_state(numpy.array([hc, hc]), chol, psi, nelec, 1, dt)
        st.propagator.xi_source = lambda w, x: numpy.zeros((w, x))
        for _ in range(nsteps):
            st.propagator.propagate_walkers(st.walkers, st.hamiltonian, st.trial, eshift)
        return st.walkers.weight[0], st.walkers.hybrid_energy[0]

    w1, eh = run(0.0, E0 + 0.3, 1)
    assert eh == pytest.approx(E0, rel=1e-9)
    assert w1 == pytest.approx(numpy.exp(-dt * (0.5 * E0 - (E0 + 0.3))), rel=1e-9)
    w2, _ = run(0.0, E0 + 0.3, 2)
    assert w2 == pytest.approx(w1 * numpy.exp(-dt * (E0 - (E0 + 0.3))), rel=1e-9)
    # cap: E_h is clipped to eshift - sqrt(2/dt)
    es = E0 + 100.0
    w1, eh = run(0.0, es, 1)
    ecap = es - numpy.sqrt(2.0 / dt)
    assert eh.real == pytest.approx(ecap, rel=1e-12)
    assert w1 == pytest.approx(numpy.exp(-dt * (0.5 * ecap - es)), rel=1e-9)
    # phase: E_h gains i c n_el, the cosine factor is cos(dt c n_el) (clipped at 0)
    nel = 4
    c = numpy.arccos(0.5) / (dt * nel)
    w1, eh = run(c, E0 + 0.3, 1)
    assert eh.imag == pytest.approx(c * nel, rel=1e-9)
    mag = numpy.exp(-dt * (0.5 * E0 - (E0 + 0.3)))
    assert w1 == pytest.approx(mag * 0.5, rel=1e-7)
    c = 2.5 / (dt * nel)  # cos(2.5) < 0
    w1, _ = run(c, E0 + 0.3, 1)
    assert w1 == 0.0


def test_reortho_properties():
    st, _ = _random_case(nwalkers=2)
    phia0, phib0 = st.walkers.phia.copy(), st.walkers.phib.copy()
    ovlp0 = st.trial.calc_greens_function(st.walkers)
    st.walkers.ovlp = ovlp0.copy()
    ga0, gb0 = st.walkers.Ghalfa.copy(), st.walkers.Ghalfb.copy()
    detR = numpy.array(st.walkers.reortho())
    for iw in range(2):
        for phi0, phi in ((phia0[iw], st.walkers.phia[iw]), (phib0[iw], st.walkers.phib[iw])):
            n = phi.shape[1]
            numpy.testing.assert_allclose(phi.conj().T.dot(phi), numpy.eye(n), atol=1e-12)
            R = phi.conj().T.dot(phi0)
            numpy.testing.assert_allclose(numpy.tril(R, -1), 0.0, atol=1e-12)
            assert numpy.all(R.diagonal().real > 0) and numpy.abs(R.diagonal().imag).max() < 1e-12
        ra = numpy.abs(numpy.linalg.det(phia0[iw].conj().T.dot(phia0[iw]))) ** 0.5
        rb = numpy.abs(numpy.linalg.det(phib0[iw].conj().T.dot(phib0[iw]))) ** 0.5
        assert detR[iw] == pytest.approx(ra * rb, rel=1e-10)
    ovlp1 = st.trial.calc_greens_function(st.walkers)
    numpy.testing.assert_allclose(ovlp1, ovlp0 / detR, rtol=1e-10)
    numpy.testing.assert_allclose(st.walkers.ovlp, ovlp0 / detR, rtol=1e-12)
    numpy.testing.assert_allclose(st.walkers.Ghalfa, ga0, atol=1e-11)
    numpy.testing.assert_allclose(st.walkers.Ghalfb, gb0, atol=1e-11)
