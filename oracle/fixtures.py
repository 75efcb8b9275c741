"""ipie's own test-fixture generators, restated (test infrastructure only).

These reproduce the random draws of ipie's test helpers in the same order so
that the known answers hard-coded in ipie's tests can be re-derived here.
"""

import numpy

from oracle import afqmc


def modified_cholesky(M, tol=1e-6, cmax=20):
    """Pivoted (modified) Cholesky; restates ipie/utils/linalg.py:167-216."""
    assert len(M.shape) == 2
    delta = numpy.copy(M.diagonal())
    nchol_max = int(cmax * M.shape[0] ** 0.5)
    nu = numpy.argmax(numpy.abs(delta))
    delta_max = delta[nu]
    Mapprox = numpy.zeros(M.shape[0], dtype=M.dtype)
    chol_vecs = numpy.zeros((nchol_max, M.shape[0]), dtype=M.dtype)
    nchol = 0
    chol_vecs[0] = numpy.copy(M[:, nu]) / delta_max**0.5
    while abs(delta_max) > tol:
        Mapprox += chol_vecs[nchol] * chol_vecs[nchol].conj()
        delta = M.diagonal() - Mapprox
        nu = numpy.argmax(numpy.abs(delta))
        delta_max = numpy.abs(delta[nu])
        nchol += 1
        Munu0 = numpy.dot(chol_vecs[:nchol, nu].conj(), chol_vecs[:nchol, :])
        chol_vecs[nchol] = (M[:, nu] - Munu0) / (delta_max) ** 0.5
    return numpy.array(chol_vecs[:nchol])


def generate_hamiltonian(nmo, nelec, cplx=False, sym=8, tol=1e-3):
    """Random ERIs -> PSD -> Cholesky; restates ipie/utils/testing.py:48-89 (same draw order)."""
    h1e = numpy.random.random((nmo, nmo))
    if cplx:
        h1e = h1e + 1j * numpy.random.random((nmo, nmo))
    eri = numpy.random.normal(scale=0.01, size=(nmo, nmo, nmo, nmo))
    if cplx:
        eri = eri + 1j * numpy.random.normal(scale=0.01, size=(nmo, nmo, nmo, nmo))
    if sym >= 4:
        eri = eri + eri.transpose(2, 3, 0, 1)
        eri = eri + eri.transpose(3, 2, 1, 0).conj()
    if sym == 8:
        eri = eri + eri.transpose(1, 0, 2, 3)
    eri = eri.transpose((0, 1, 3, 2))
    eri = eri.reshape((nmo * nmo, nmo * nmo))
    eri = numpy.dot(eri, eri.conj().T)
    chol = modified_cholesky(eri, tol=tol, cmax=30)
    chol = chol.reshape((-1, nmo, nmo))
    enuc = numpy.random.rand()
    eri = eri.reshape((nmo, nmo, nmo, nmo))
    eri = eri.transpose((0, 1, 3, 2))
    return h1e, chol, enuc, eri


def get_random_nomsd(nup, ndown, nbasis, ndet=10, cplx=True, init=False):
    """Restates ipie/utils/testing.py:92-110 (same draw order)."""
    a = numpy.random.rand(ndet * nbasis * (nup + ndown))
    b = numpy.random.rand(ndet * nbasis * (nup + ndown))
    if cplx:
        wfn = (a + 1j * b).reshape((ndet, nbasis, nup + ndown))
        coeffs = numpy.random.rand(ndet) + 1j * numpy.random.rand(ndet)
    else:
        wfn = a.reshape((ndet, nbasis, nup + ndown))
        coeffs = numpy.random.rand(ndet)
    if init:
        a = numpy.random.rand(nbasis * (nup + ndown))
        b = numpy.random.rand(nbasis * (nup + ndown))
        if cplx:
            init_wfn = (a + 1j * b).reshape((nbasis, nup + ndown))
        else:
            init_wfn = a.reshape((nbasis, nup + ndown))
        return coeffs, wfn, init_wfn
    return coeffs, wfn


def shaped_normal(shape, cmplx=False):
    """Restates ipie/utils/testing.py:278-288."""
    size = int(numpy.prod(shape))
    if cmplx:
        arr_r = numpy.random.normal(size=size)
        arr_i = numpy.random.normal(size=size)
        arr = numpy.array(arr_r + 1j * arr_i, dtype=numpy.complex128)
    else:
        arr = numpy.array(numpy.random.normal(size=size), dtype=numpy.float64)
    return arr.reshape(shape)


def gen_random_test_instances(nmo, nocc, naux, nwalkers, seed=7):
    """Restates ipie/utils/testing.py:306-343 for ndets == 1."""
    numpy.random.seed(seed)
    wfn = get_random_nomsd(nocc, nocc, nmo, ndet=1)
    h1e = shaped_normal((nmo, nmo))
    system = afqmc.Generic((nocc, nocc))
    chol = shaped_normal((naux, nmo, nmo))
    ham = afqmc.GenericRealChol(numpy.array([h1e, h1e]), chol.reshape((naux, nmo * nmo)).T.copy(), ecore=0)
    trial = afqmc.SingleDet(wfn[1][0], (nocc, nocc), nmo)
    walkers = afqmc.UHFWalkers(wfn[1][0], nocc, nocc, nmo, nwalkers)
    walkers.build(trial)
    walkers.Ghalfa = shaped_normal((nwalkers, nocc, nmo), cmplx=True)
    walkers.Ghalfb = shaped_normal((nwalkers, nocc, nmo), cmplx=True)
    trial._rchola = shaped_normal((naux, nocc * nmo))
    trial._rcholb = shaped_normal((naux, nocc * nmo))
    trial._rH1a = shaped_normal((nocc, nmo))
    trial._rH1b = shaped_normal((nocc, nmo))
    trial.half_rotated = True
    return system, ham, walkers, trial


def build_random_single_det_trial(num_elec, num_basis, complex_trial=False, rhf_trial=False):
    """Restates ipie/utils/testing.py:388-401."""
    _, wfn, init = get_random_nomsd(num_elec[0], num_elec[1], num_basis, ndet=1, cplx=complex_trial, init=True)
    if rhf_trial:
        wfn[0, :, num_elec[0]:] = wfn[0, :, : num_elec[0]]
        init[:, num_elec[0]:] = init[:, : num_elec[0]]
    return afqmc.SingleDet(wfn[0], num_elec, num_basis), wfn[0], init
