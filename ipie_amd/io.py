"""Hamiltonian and trial-wavefunction files into the MI355X objects.

Mirrors ipie's readers (ipie/utils/io.py:46-57 ``read_hamiltonian``,
:150-167 ``read_single_det_wavefunction``, :273-337 ``read_qmcpack_wfn_hdf``
for NOMSD, :509-522 ``orbs_from_dset``; ipie/utils/hamiltonian_converter.py:
453-472 ``read_qmcpack_dense``; ipie/hamiltonians/generic.py:160-178
``read_integrals``) with the same dataset names and array conventions, read
by ``ipie_amd.h5read`` (h5py is not available here).  ``load_generic`` goes
from an integral file (and a trial file) to the device-ready objects.
"""

import numpy

from ipie_amd.h5read import H5File


def read_hamiltonian(filename):
    """ipie's own format: hcore [N, N], LXmn [X, N, N], e0 (io.py:46-57)."""
    with H5File(filename) as f:
        hcore = numpy.array(f["hcore"])
        LXmn = numpy.array(f["LXmn"])
        e0 = float(numpy.asarray(f["e0"]).ravel()[0])
    if hcore.ndim != 2:
        raise ValueError("Incorrect shape for hcore, expected 2-dimensional array")
    nmo = hcore.shape[0]
    naux = LXmn.size // (nmo * nmo)
    if LXmn.ndim != 3 or LXmn.shape[0] != naux:
        raise ValueError(f"Incorrect shape for LXmn: {LXmn.shape}")
    return hcore, LXmn, e0


def read_qmcpack_dense(filename):
    """QMCPACK dense factorisation: hcore, L [N*N, X], the core energy
    (hamiltonian_converter.py:453-472)."""
    with H5File(filename) as f:
        enuc = float(numpy.asarray(f["Hamiltonian/Energies"]).ravel()[0])
        hcore = numpy.array(f["Hamiltonian/hcore"])
        chol = numpy.array(f["Hamiltonian/DenseFactorized/L"])
    return hcore, chol, enuc


def read_integrals(filename):
    """(h1e, chol [N*N, X], ecore) from either format (generic.py:160-178;
    QMCPACK's sparse factorisation is not read)."""
    with H5File(filename) as f:
        sparse = "Hamiltonian/Factorized" in f
        dense = "Hamiltonian/DenseFactorized" in f
        own = "LXmn" in f
    if dense:
        return read_qmcpack_dense(filename)
    if own:
        hcore, LXmn, e0 = read_hamiltonian(filename)
        nbsf = hcore.shape[-1]
        return hcore, numpy.ascontiguousarray(LXmn.reshape(LXmn.shape[0], -1).T), e0
    if sparse:
        raise NotImplementedError("QMCPACK sparse Cholesky factorisations are not read")
    raise KeyError(f"{filename}: no Hamiltonian found")


def _from_qmcpack_complex(data, shape):
    """(re, im) pairs in a trailing axis of 2 -> complex (io.py:713-714)."""
    a = numpy.asarray(data)
    if numpy.iscomplexobj(a):
        return a.reshape(shape)
    return numpy.ascontiguousarray(a, dtype=numpy.float64).view(numpy.complex128).reshape(shape)


def _orbs_from_dset(f, group):
    """Orbitals stored as the CSR matrix of A^H (io.py:509-522); returns A."""
    dims = numpy.asarray(f[group + "/dims"])
    rows, cols, nnz = int(dims[0]), int(dims[1]), int(dims[2])
    data = _from_qmcpack_complex(f[group + "/data_"], (nnz,))
    idx = numpy.asarray(f[group + "/jdata_"])
    begin = numpy.asarray(f[group + "/pointers_begin_"])
    end = numpy.asarray(f[group + "/pointers_end_"])
    ah = numpy.zeros((rows, cols), dtype=numpy.complex128)
    for r in range(rows):
        for e in range(int(begin[r]), int(end[r])):
            ah[r, int(idx[e])] += data[e]
    return ah.conj().T.copy()


def read_qmcpack_wfn_hdf(filename, nelec=None, get_nelec=False):
    """NOMSD wavefunctions: ((ci_coeffs, wfn [nci, N, na+nb]), psi0 [N, na+nb])
    (io.py:273-337; particle-hole expansions are not read)."""
    with H5File(filename) as f:
        if "Wavefunction/NOMSD" not in f:
            if "Wavefunction/PHMSD" in f:
                raise NotImplementedError("particle-hole multi-determinant wavefunctions are not on this path")
            raise KeyError(f"{filename}: no wavefunction found")
        g = "Wavefunction/NOMSD"
        dims = numpy.asarray(f[g + "/dims"])
        nmo, na, nb, walker_type, nci = (int(v) for v in dims[:5])
        if nelec is not None and (na, nb) != tuple(nelec):
            raise ValueError(f"Number of electrons does not match wavefunction: {(na, nb)} vs {tuple(nelec)}")
        uhf = walker_type == 2
        coeffs = _from_qmcpack_complex(f[g + "/ci_coeffs"], (nci,))
        psi0a = _from_qmcpack_complex(f[g + "/Psi0_alpha"], (nmo, na))
        psi0 = numpy.zeros((nmo, na + nb), dtype=numpy.complex128)
        psi0[:, :na] = psi0a
        psi0[:, na:] = _from_qmcpack_complex(f[g + "/Psi0_beta"], (nmo, nb)) if uhf else psi0a[:, :nb]
        wfn = numpy.zeros((nci, nmo, na + nb), dtype=numpy.complex128)
        for idet in range(nci):
            ix = 2 * idet if uhf else idet
            pa = _orbs_from_dset(f, f"{g}/PsiT_{ix:d}")
            wfn[idet, :, :na] = pa
            wfn[idet, :, na:] = _orbs_from_dset(f, f"{g}/PsiT_{ix + 1:d}") if uhf else pa[:, :nb]
    if get_nelec:
        return (coeffs, wfn), psi0, (na, nb)
    return (coeffs, wfn), psi0


def read_single_det_wavefunction(filename):
    """ipie's own single-determinant format: ([psia, psib], [phi0a, phi0b])
    (io.py:150-167)."""
    with H5File(filename) as f:
        psia = numpy.array(f["psi_T_alpha"])
        phi0a = numpy.array(f["phi0_alpha"])
        if "psi_T_beta" in f:
            return [psia, numpy.array(f["psi_T_beta"])], [phi0a, numpy.array(f["phi0_beta"])]
    return [psia, psia], [phi0a, phi0a]


def load_generic(integrals, wavefunction=None, nwalkers=1, dt=0.005, precision="fp64", device_setup=True,
                 device=None, **prop_kw):
    """From files to built device objects: (system, hamiltonian, trial,
    walkers, propagator).  A single-determinant trial is required (NOMSD with
    one determinant, or ipie's single-determinant format); walkers start at
    its psi0.  The static tensors are built on the device unless
    device_setup=False."""
    from ipie_amd import Generic
    from ipie_amd.hamiltonian import GenericRealChol
    from ipie_amd.propagator import PhaselessGeneric
    from ipie_amd.trial import SingleDet
    from ipie_amd.walkers import UHFWalkers

    hcore, chol, ecore = read_integrals(integrals)
    nbasis = hcore.shape[-1]
    wfile = wavefunction or integrals
    with H5File(wfile) as f:
        qmcpack = "Wavefunction/NOMSD" in f or "Wavefunction/PHMSD" in f
    if qmcpack:
        (coeffs, wfn), psi0, nelec = read_qmcpack_wfn_hdf(wfile, get_nelec=True)
        if wfn.shape[0] != 1:
            raise NotImplementedError("multi-determinant trials are not on the single-determinant path")
        psi = wfn[0]
    else:
        (psia, psib), (phi0a, phi0b) = read_single_det_wavefunction(wfile)
        nelec = (psia.shape[-1], psib.shape[-1])
        psi = numpy.hstack([psia, psib])
        psi0 = numpy.hstack([phi0a, phi0b])
    system = Generic(nelec)
    h1e = numpy.array([hcore, hcore]) if hcore.ndim == 2 else numpy.asarray(hcore)
    ham = GenericRealChol(h1e, numpy.ascontiguousarray(chol, dtype=numpy.float64), ecore=ecore,
                          host_setup=not device_setup)
    trial = SingleDet(psi, nelec, nbasis)
    walkers = UHFWalkers(psi0, nelec[0], nelec[1], nbasis, nwalkers, device=device)
    prop = PhaselessGeneric(dt, precision=precision, **prop_kw)
    prop.build(ham, trial, walkers, device_setup=device_setup)
    trial.calculate_energy(system, ham)
    return system, ham, trial, walkers, prop
