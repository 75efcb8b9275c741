"""numpy restatement of ipie's CPU phaseless-AFQMC step (test oracle).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.

The classes mirror the reference objects the hot path reads and writes
(attribute names kept identical) and each function cites the ipie source it
restates.  ``xi`` (the auxiliary fields) can be injected so that a GPU step and
this oracle consume exactly the same random numbers.
"""

import numpy
import scipy.linalg


# --------------------------------------------------------------------------
# Hamiltonian: ipie/hamiltonians/generic.py
# --------------------------------------------------------------------------
def construct_h1e_mod(chol, h1e):
    """h1e_mod = h1e - 1/2 sum_{k,x} L_{ik,x} L_{jk,x}^*.

    Restates ipie/hamiltonians/generic.py:31-43.
    """
    nbasis = h1e.shape[-1]
    nchol = chol.shape[-1]
    chol_view = chol.reshape((nbasis, nbasis * nchol))
    v0 = 0.5 * numpy.dot(chol_view, chol_view.T.conj())
    h1e_mod = numpy.zeros(h1e.shape, dtype=h1e.dtype)
    h1e_mod[0] = h1e[0] - v0
    h1e_mod[1] = h1e[1] - v0
    return h1e_mod


def pack_cholesky(idx_i, idx_j, chol3):
    """chol_packed[p, x] = L[idx_i[p], idx_j[p], x] (upper triangle, row major).

    Restates ipie/utils/pack_numba.py:36-45 (``numpy.triu_indices`` order,
    ipie/hamiltonians/generic.py:64-72).
    """
    return numpy.ascontiguousarray(chol3[idx_i, idx_j, :])


def unpack_VHS_batch(idx_i, idx_j, VHS_packed, nbasis):
    """Complex-symmetric unpack of [W, P] into [W, N, N].

    Restates ipie/utils/pack_numba.py:22-33.
    """
    nwalkers = VHS_packed.shape[0]
    VHS = numpy.zeros((nwalkers, nbasis, nbasis), dtype=VHS_packed.dtype)
    VHS[:, idx_i, idx_j] = VHS_packed
    VHS[:, idx_j, idx_i] = VHS_packed
    return VHS


class GenericRealChol:
    """Real-Cholesky Hamiltonian; restates ipie/hamiltonians/generic.py:46-98.

    ``h1e`` is [2, N, N], ``chol`` is [N*N, X] (float64), ``ecore`` a float.
    """

    def __init__(self, h1e, chol, ecore=0.0):
        assert h1e.shape[0] == 2
        assert chol.dtype == numpy.dtype("float64")
        self.H1 = numpy.array(h1e)
        self.ecore = ecore
        self.nbasis = h1e.shape[-1]
        self.chol = chol
        self.nchol = chol.shape[-1]
        self.nfields = self.nchol
        assert self.nbasis**2 == chol.shape[0]
        self.sym_idx = numpy.triu_indices(self.nbasis)
        self.sym_idx_i = self.sym_idx[0].copy()
        self.sym_idx_j = self.sym_idx[1].copy()
        chol3 = chol.reshape((self.nbasis, self.nbasis, self.nchol))
        self.chol_packed = pack_cholesky(self.sym_idx_i, self.sym_idx_j, chol3)
        self.chunked = False
        self.h1e_mod = construct_h1e_mod(self.chol, self.H1)


class Generic:
    """System carrying the electron numbers; restates ipie/systems/generic.py:23-62."""

    def __init__(self, nelec):
        self.nup, self.ndown = nelec
        self.nelec = nelec
        self.ne = self.nup + self.ndown


# --------------------------------------------------------------------------
# Trial: ipie/trial_wavefunction/single_det.py, half_rotate.py
# --------------------------------------------------------------------------
def gab_mod(A, B):
    """(G, Ghalf) for bra A and ket B; restates ipie/estimators/utils.py:100-133."""
    O = numpy.dot(B.T, A.conj())
    GHalf = numpy.dot(scipy.linalg.inv(O), B.T)
    G = numpy.dot(A.conj(), GHalf)
    return G, GHalf


def gab_spin(A, B, na, nb):
    """Restates ipie/estimators/utils.py:136-142."""
    GA, GAH = gab_mod(A[:, :na], B[:, :na])
    if nb > 0:
        GB, GBH = gab_mod(A[:, na:], B[:, na:])
    else:
        GB = numpy.zeros(GA.shape, dtype=GA.dtype)
        GBH = numpy.zeros((0, GAH.shape[1]), dtype=GAH.dtype)
    return numpy.array([GA, GB]), [GAH, GBH]


def half_rotate_generic(ham, orbsa, orbsb):
    """rH1 and half-rotated Cholesky; restates ipie/trial_wavefunction/half_rotate.py:11-116.

    rH1s[i, q] = sum_p conj(psi[p, i]) H1[p, q];
    rchols[x, (i, n)] = sum_m conj(psi[m, i]) L[m, n, x]   (half_rotate.py:91-106).
    """
    M = ham.nbasis
    nchol = ham.nchol
    chol = ham.chol.reshape((M, M, nchol))
    na = orbsa.shape[-1]
    nb = orbsb.shape[-1]
    rH1a = numpy.einsum("pi,pq->iq", orbsa.conj(), ham.H1[0], optimize=True)
    rH1b = numpy.einsum("pi,pq->iq", orbsb.conj(), ham.H1[1], optimize=True)
    rchola = numpy.einsum("mi,mnx->xin", orbsa.conj(), chol, optimize=True).reshape(nchol, na * M)
    rcholb = numpy.einsum("mi,mnx->xin", orbsb.conj(), chol, optimize=True).reshape(nchol, nb * M)
    return (rH1a, rH1b), (numpy.ascontiguousarray(rchola), numpy.ascontiguousarray(rcholb))


class SingleDet:
    """Single-determinant trial; restates ipie/trial_wavefunction/single_det.py:30-211."""

    def __init__(self, wavefunction, num_elec, num_basis):
        assert len(wavefunction.shape) == 2
        self.psi = wavefunction
        self.nalpha, self.nbeta = num_elec
        self.nbasis = num_basis
        imag_norm = numpy.sum(self.psi.imag.ravel() * self.psi.imag.ravel())
        if imag_norm <= 1e-8:
            self.psi = numpy.array(self.psi.real, dtype=numpy.float64)
        self.psi0a = numpy.ascontiguousarray(self.psi[:, : self.nalpha])
        self.psi0b = numpy.ascontiguousarray(self.psi[:, self.nalpha :])
        self.G, self.Ghalf = gab_spin(self.psi, self.psi, self.nalpha, self.nbeta)
        self.half_rotated = False

    def half_rotate(self, hamiltonian):
        """Restates single_det.py:88-112."""
        (self._rH1a, self._rH1b), (self._rchola, self._rcholb) = half_rotate_generic(
            hamiltonian, self.psi0a, self.psi0b
        )
        self.half_rotated = True

    def calculate_energy(self, system, hamiltonian):
        """Trial energy; restates single_det.py:65-86."""
        self.e1b = (
            numpy.sum(self.Ghalf[0] * self._rH1a)
            + numpy.sum(self.Ghalf[1] * self._rH1b)
            + hamiltonian.ecore
        )
        self.ej, self.ek = half_rotated_cholesky_jk_uhf(self, self.Ghalf)
        self.e2b = self.ej - self.ek
        self.energy = self.e1b + self.e2b
        return self.energy

    def calc_greens_function(self, walkers, build_full=False):
        return greens_function_single_det(walkers, self, build_full=build_full)

    def calc_overlap(self, walkers):
        return calc_overlap_single_det_uhf(walkers, self)

    def calc_force_bias(self, hamiltonian, walkers, mpi_handler=None):
        return construct_force_bias_batch_single_det(hamiltonian, walkers, self._rchola, self._rcholb)


# --------------------------------------------------------------------------
# Walkers: ipie/walkers/uhf_walkers.py, base_walkers.py
# --------------------------------------------------------------------------
class UHFWalkers:
    """Restates the state of ipie/walkers/uhf_walkers.py:27-99 and base_walkers.py:86-143."""

    def __init__(self, initial_walker, nup, ndown, nbasis, nwalkers):
        self.nup = nup
        self.ndown = ndown
        self.nbasis = nbasis
        self.nwalkers = nwalkers
        self.weight = numpy.ones(nwalkers)
        self.unscaled_weight = self.weight.copy()
        self.phase = numpy.ones(nwalkers, dtype=numpy.complex128)
        self.ovlp = numpy.ones(nwalkers)
        self.sgn_ovlp = numpy.ones(nwalkers)
        self.log_ovlp = numpy.zeros(nwalkers)
        self.hybrid_energy = numpy.zeros(nwalkers)
        self.detR = [1.0] * nwalkers
        self.detR_shift = numpy.zeros(nwalkers)
        self.log_detR = [0.0] * nwalkers
        self.log_shift = numpy.zeros(nwalkers)
        self.phia = numpy.array(
            [initial_walker[:, :nup].copy() for _ in range(nwalkers)], dtype=numpy.complex128
        )
        self.phib = numpy.array(
            [initial_walker[:, nup:].copy() for _ in range(nwalkers)], dtype=numpy.complex128
        )
        self.Ga = numpy.zeros((nwalkers, nbasis, nbasis), dtype=numpy.complex128)
        self.Gb = numpy.zeros((nwalkers, nbasis, nbasis), dtype=numpy.complex128)
        self.Ghalfa = numpy.zeros((nwalkers, nup, nbasis), dtype=numpy.complex128)
        self.Ghalfb = numpy.zeros((nwalkers, ndown, nbasis), dtype=numpy.complex128)
        self.rhf = False

    def build(self, trial):
        """Restates uhf_walkers.py:101-102."""
        self.ovlp = trial.calc_greens_function(self)

    def reortho(self):
        """QR with positive R diagonal; restates uhf_walkers.py:108-143.

        LAPACK (scipy economic QR) returns a real diagonal R; multiplying Q by
        sign(diag R) makes diag(R) positive, so phi <- the unique Q whose R has
        a positive real diagonal.
        """
        detR = []
        for iw in range(self.nwalkers):
            q, r = scipy.linalg.qr(self.phia[iw], mode="economic")
            rdiag = numpy.diag(r)
            signs = numpy.sign(rdiag)
            self.phia[iw] = numpy.dot(q, numpy.diag(signs))
            log_det = numpy.sum(numpy.log(numpy.abs(rdiag)))
            if self.ndown > 0:
                q, r = scipy.linalg.qr(self.phib[iw], mode="economic")
                rdiag = numpy.diag(r)
                signs = numpy.sign(rdiag)
                self.phib[iw] = numpy.dot(q, numpy.diag(signs))
                log_det += numpy.sum(numpy.log(numpy.abs(rdiag)))
            detR.append(numpy.exp(log_det - self.detR_shift[iw]))
            self.log_detR[iw] += numpy.log(detR[iw])
            self.detR[iw] = detR[iw]
            self.ovlp[iw] = self.ovlp[iw] / detR[iw]
        return detR

    def orthogonalise(self):
        """Restates base_walkers.py:166-179 (phaseless branch)."""
        return self.reortho()


# --------------------------------------------------------------------------
# Green's function and overlap
# --------------------------------------------------------------------------
def greens_function_single_det(walker_batch, trial, build_full=False):
    """Per-walker GF; restates ipie/estimators/greens_function_single_det.py:62-103."""
    ndown = walker_batch.ndown
    det = []
    for iw in range(walker_batch.nwalkers):
        ovlp = numpy.dot(walker_batch.phia[iw].T, trial.psi0a.conj())
        ovlp_inv = numpy.linalg.inv(ovlp)
        walker_batch.Ghalfa[iw] = numpy.dot(ovlp_inv, walker_batch.phia[iw].T)
        if not trial.half_rotated or build_full:
            walker_batch.Ga[iw] = numpy.dot(trial.psi0a.conj(), walker_batch.Ghalfa[iw])
        sign_a, log_ovlp_a = numpy.linalg.slogdet(ovlp)
        if ndown > 0 and not walker_batch.rhf:
            ovlp = numpy.dot(walker_batch.phib[iw].T, trial.psi0b.conj())
            sign_b, log_ovlp_b = numpy.linalg.slogdet(ovlp)
            walker_batch.Ghalfb[iw] = numpy.dot(numpy.linalg.inv(ovlp), walker_batch.phib[iw].T)
            if not trial.half_rotated or build_full:
                walker_batch.Gb[iw] = numpy.dot(trial.psi0b.conj(), walker_batch.Ghalfb[iw])
            det.append(sign_a * sign_b * numpy.exp(log_ovlp_a + log_ovlp_b - walker_batch.log_shift[iw]))
        elif ndown > 0 and walker_batch.rhf:
            det.append(sign_a * sign_a * numpy.exp(log_ovlp_a + log_ovlp_a - walker_batch.log_shift[iw]))
        else:
            det.append(sign_a * numpy.exp(log_ovlp_a - walker_batch.log_shift[iw]))
    return numpy.array(det, dtype=numpy.complex128)


def calc_overlap_single_det_uhf(walkers, trial):
    """Restates ipie/propagation/overlap.py:29-59."""
    ovlp_a = numpy.einsum("wmi,mj->wij", walkers.phia, trial.psi0a.conj(), optimize=True)
    sign_a, log_ovlp_a = numpy.linalg.slogdet(ovlp_a)
    if walkers.ndown > 0 and not walkers.rhf:
        ovlp_b = numpy.einsum("wmi,mj->wij", walkers.phib, trial.psi0b.conj(), optimize=True)
        sign_b, log_ovlp_b = numpy.linalg.slogdet(ovlp_b)
        return sign_a * sign_b * numpy.exp(log_ovlp_a + log_ovlp_b - walkers.log_shift)
    elif walkers.ndown > 0 and walkers.rhf:
        return sign_a * sign_a * numpy.exp(log_ovlp_a + log_ovlp_a - walkers.log_shift)
    return sign_a * numpy.exp(log_ovlp_a - walkers.log_shift)


# --------------------------------------------------------------------------
# Propagation: ipie/propagation/*.py
# --------------------------------------------------------------------------
def construct_force_bias_batch_single_det(hamiltonian, walkers, rchola, rcholb):
    """vbias[w, x]; restates ipie/propagation/force_bias.py:79-121.

    Real and imaginary parts are separate real products exactly as in the
    reference (so a complex rchol keeps only the real part of each product).
    """
    nw = walkers.nwalkers
    nb = hamiltonian.nbasis
    if walkers.rhf:
        Ga = walkers.Ghalfa.reshape(nw, walkers.nup * nb)
        real = 2.0 * rchola.dot(Ga.T.real)
        imag = 2.0 * rchola.dot(Ga.T.imag)
    else:
        Ga = walkers.Ghalfa.reshape(nw, walkers.nup * nb)
        Gb = walkers.Ghalfb.reshape(nw, walkers.ndown * nb)
        real = rchola.dot(Ga.T.real) + rcholb.dot(Gb.T.real)
        imag = rchola.dot(Ga.T.imag) + rcholb.dot(Gb.T.imag)
    vbias = numpy.empty((nw, hamiltonian.nchol), dtype=numpy.complex128)
    vbias.real = numpy.real(real).T
    vbias.imag = numpy.real(imag).T
    return vbias


def construct_mean_field_shift(hamiltonian, trial):
    """mf_shift = i L^T Re(vec G) - L^T Im(vec G); restates ipie/propagation/phaseless_base.py:153-172."""
    Gcharge = (trial.G[0] + trial.G[1]).ravel()
    tmp_real = numpy.dot(hamiltonian.chol.T, Gcharge.real)
    tmp_imag = numpy.dot(hamiltonian.chol.T, Gcharge.imag)
    return 1.0j * tmp_real - tmp_imag


def construct_one_body_propagator(hamiltonian, mf_shift, dt):
    """expH1[s] = expm(-dt/2 (h1e_mod[s] - i sum_x L_x mf_x)); restates phaseless_base.py:27-74."""
    nb = hamiltonian.nbasis
    shift = 1j * numpy.einsum("mx,x->m", hamiltonian.chol, mf_shift).reshape(nb, nb)
    H1 = hamiltonian.h1e_mod - numpy.array([shift, shift])
    return numpy.array([scipy.linalg.expm(-0.5 * dt * H1[0]), scipy.linalg.expm(-0.5 * dt * H1[1])])


def propagate_one_body(phi, bt2):
    """phi[w] <- bt2 . phi[w]; restates ipie/propagation/operations.py:26-55 (CPU loop)."""
    for iw in range(phi.shape[0]):
        phi[iw] = numpy.dot(bt2, phi[iw])
    return phi


def apply_exponential(phi, VHS, exp_nmax):
    """Taylor series sum_{n<=exp_nmax} VHS^n phi / n!; restates operations.py:58-82."""
    Temp = phi.copy()
    for n in range(1, exp_nmax + 1):
        Temp = VHS.dot(Temp) / n
        phi += Temp
    return phi


class PropagatorTimer:
    """Restates ipie/propagation/continuous_base.py:4-11."""

    def __init__(self):
        self.tfbias = 0.0
        self.tovlp = 0.0
        self.tupdate = 0.0
        self.tgf = 0.0
        self.tvhs = 0.0
        self.tgemm = 0.0


class PhaselessGeneric:
    """Restates ipie/propagation/phaseless_base.py:242-381 and phaseless_generic.py:28-97.

    ``xi_source`` (callable ``(nwalkers, nfields) -> [W, X] float64``) replaces
    ``numpy.random.normal`` at phaseless_base.py:296; by default it draws from
    the global numpy RandomState in the reference's order.
    """

    def __init__(self, time_step, exp_nmax=6):
        self.dt = time_step
        self.sqrt_dt = self.dt**0.5
        self.isqrt_dt = 1j * self.sqrt_dt
        self.exp_nmax = exp_nmax
        self.nfb_trig = 0
        self.nhe_trig = 0
        self.ebound = (2.0 / self.dt) ** 0.5
        self.fbbound = 1.0
        self.timer = PropagatorTimer()
        self.xi_source = None
        self.last_xi = None

    def build(self, hamiltonian, trial=None, walkers=None, mpi_handler=None, verbose=False):
        """Restates phaseless_base.py:256-272."""
        self.mf_shift = construct_mean_field_shift(hamiltonian, trial)
        self.expH1 = construct_one_body_propagator(hamiltonian, self.mf_shift, self.dt)
        self.vbias = None

    def propagate_walkers_one_body(self, walkers):
        """Restates phaseless_base.py:274-280."""
        walkers.phia = propagate_one_body(walkers.phia, self.expH1[0])
        if walkers.ndown > 0 and not walkers.rhf:
            walkers.phib = propagate_one_body(walkers.phib, self.expH1[1])

    def draw_xi(self, nwalkers, nfields):
        if self.xi_source is not None:
            xi = numpy.asarray(self.xi_source(nwalkers, nfields), dtype=numpy.float64)
        else:
            xi = numpy.random.normal(0.0, 1.0, nfields * nwalkers).reshape(nwalkers, nfields)
        self.last_xi = xi
        return xi

    def propagate_walkers_two_body(self, walkers, hamiltonian, trial):
        """Restates phaseless_base.py:282-310."""
        self.vbias = trial.calc_force_bias(hamiltonian, walkers)
        xbar = -self.sqrt_dt * (1j * self.vbias - self.mf_shift)
        xbar = self.apply_bound_force_bias(xbar, self.fbbound)
        xi = self.draw_xi(walkers.nwalkers, hamiltonian.nfields)
        xshifted = xi - xbar
        cmf = -self.sqrt_dt * numpy.einsum("wx,x->w", xshifted, self.mf_shift)
        cfb = numpy.einsum("wx,wx->w", xi, xbar) - 0.5 * numpy.einsum("wx,wx->w", xbar, xbar)
        xshifted = xshifted.T.copy()
        self.apply_VHS(walkers, hamiltonian, xshifted)
        return cmf, cfb

    def construct_VHS(self, hamiltonian, xshifted):
        """Restates phaseless_generic.py:67-97 (CPU branch)."""
        nwalkers = xshifted.shape[-1]
        VHS_packed = hamiltonian.chol_packed.dot(xshifted.real).astype(xshifted.dtype)
        VHS_packed += 1.0j * hamiltonian.chol_packed.dot(xshifted.imag)
        VHS_packed = self.isqrt_dt * VHS_packed.T.reshape(nwalkers, hamiltonian.chol_packed.shape[0])
        return unpack_VHS_batch(hamiltonian.sym_idx_i, hamiltonian.sym_idx_j, VHS_packed, hamiltonian.nbasis)

    def apply_VHS(self, walkers, hamiltonian, xshifted):
        """Restates phaseless_generic.py:35-58 (CPU branch)."""
        assert walkers.nwalkers == xshifted.shape[-1]
        VHS = self.construct_VHS(hamiltonian, xshifted)
        for iw in range(walkers.nwalkers):
            walkers.phia[iw] = apply_exponential(walkers.phia[iw], VHS[iw], self.exp_nmax)
            if walkers.ndown > 0 and not walkers.rhf:
                walkers.phib[iw] = apply_exponential(walkers.phib[iw], VHS[iw], self.exp_nmax)

    def propagate_walkers(self, walkers, hamiltonian, trial, eshift):
        """One phaseless step; restates phaseless_base.py:312-338."""
        ovlp = trial.calc_greens_function(walkers)
        self.propagate_walkers_one_body(walkers)
        cmf, cfb = self.propagate_walkers_two_body(walkers, hamiltonian, trial)
        self.propagate_walkers_one_body(walkers)
        ovlp_new = trial.calc_overlap(walkers)
        self.update_weight(walkers, ovlp, ovlp_new, cfb, cmf, eshift)

    def update_weight(self, walkers, ovlp, ovlp_new, cfb, cmf, eshift):
        """Restates phaseless_base.py:340-358."""
        ovlp_ratio = ovlp_new / ovlp
        hybrid_energy = -(numpy.log(ovlp_ratio) + cfb + cmf) / self.dt
        hybrid_energy = self.apply_bound_hybrid(hybrid_energy, eshift)
        importance_function = numpy.exp(-self.dt * (0.5 * (hybrid_energy + walkers.hybrid_energy) - eshift))
        magn = numpy.abs(importance_function)
        walkers.hybrid_energy = hybrid_energy
        dtheta = (-self.dt * hybrid_energy - cfb).imag
        cosine_fac = numpy.cos(dtheta)
        numpy.clip(cosine_fac, a_min=0.0, a_max=None, out=cosine_fac)
        walkers.weight = walkers.weight * magn * cosine_fac
        walkers.ovlp = ovlp_new

    def apply_bound_force_bias(self, xbar, max_bound=1.0):
        """Restates phaseless_base.py:360-368."""
        absxbar = numpy.abs(xbar)
        idx_to_rescale = absxbar > max_bound
        nonzeros = absxbar > 1e-13
        xbar_rescaled = xbar.copy()
        xbar_rescaled[nonzeros] = xbar_rescaled[nonzeros] / absxbar[nonzeros]
        xbar = numpy.where(idx_to_rescale, xbar_rescaled, xbar)
        self.nfb_trig += int(numpy.sum(idx_to_rescale))
        return xbar

    def apply_bound_hybrid(self, ehyb, eshift):
        """Restates phaseless_base.py:370-380."""
        if abs(eshift) < 1e-10:
            return ehyb
        ehyb = numpy.array(ehyb, dtype=numpy.complex128)
        emax = eshift.real + self.ebound
        emin = eshift.real - self.ebound
        ehyb.real = numpy.clip(ehyb.real, emin, emax)
        return ehyb


# --------------------------------------------------------------------------
# Local energy: ipie/estimators/local_energy_sd.py, generic.py
# --------------------------------------------------------------------------
def ecoul_kernel_batch_real_rchol_uhf(rchola, rcholb, Ghalfa_batch, Ghalfb_batch):
    """Ecoul[w] = 1/2 sum_x (Xa + Xb)^2; restates local_energy_sd.py:158-201."""
    nw = Ghalfa_batch.shape[0]
    Ga = Ghalfa_batch.reshape(nw, -1)
    Gb = Ghalfb_batch.reshape(nw, -1)
    Xa = Ga.real.dot(rchola.T) + 1j * Ga.imag.dot(rchola.T)
    Xb = Gb.real.dot(rcholb.T) + 1j * Gb.imag.dot(rcholb.T)
    X = Xa + Xb
    return 0.5 * numpy.einsum("wx,wx->w", X, X)


def ecoul_kernel_batch_real_rchol_rhf(rchola, Ghalfa_batch):
    """Ecoul[w] = 2 sum_x Xa^2; restates local_energy_sd.py:44-73."""
    nw = Ghalfa_batch.shape[0]
    Ga = Ghalfa_batch.reshape(nw, -1)
    X = Ga.real.dot(rchola.T) + 1j * Ga.imag.dot(rchola.T)
    return 2.0 * numpy.einsum("wx,wx->w", X, X)


def exx_kernel_batch_real_rchol(rchol, Ghalf_batch):
    """Exx[w] = 1/2 sum_x tr(T_x T_x), T_x = L_x Ghalf^T; restates local_energy_sd.py:263-303."""
    nwalkers, nocc, nbasis = Ghalf_batch.shape
    nchol = rchol.shape[0]
    exx = numpy.zeros(nwalkers, dtype=numpy.complex128)
    L = rchol.reshape(nchol, nocc, nbasis)
    for iw in range(nwalkers):
        GT = Ghalf_batch[iw].T
        T = numpy.einsum("xim,mj->xij", L, GT, optimize=True)
        exx[iw] = numpy.einsum("xij,xji->", T, T)
    return 0.5 * exx


def local_energy_single_det_uhf_batch(system, hamiltonian, walkers, trial):
    """[W, 3] = (E, E1, E2); restates local_energy_sd.py:358-403."""
    nwalkers = walkers.Ghalfa.shape[0]
    Ga = walkers.Ghalfa.reshape((nwalkers, -1))
    Gb = walkers.Ghalfb.reshape((nwalkers, -1))
    e1b = Ga.dot(trial._rH1a.ravel()) + Gb.dot(trial._rH1b.ravel()) + hamiltonian.ecore
    ecoul = ecoul_kernel_batch_real_rchol_uhf(trial._rchola, trial._rcholb, walkers.Ghalfa, walkers.Ghalfb)
    exx = exx_kernel_batch_real_rchol(trial._rchola, walkers.Ghalfa) + exx_kernel_batch_real_rchol(
        trial._rcholb, walkers.Ghalfb
    )
    e2b = ecoul - exx
    energy = numpy.zeros((nwalkers, 3), dtype=numpy.complex128)
    energy[:, 0] = e1b + e2b
    energy[:, 1] = e1b
    energy[:, 2] = e2b
    return energy


def local_energy_single_det_rhf_batch(system, hamiltonian, walkers, trial):
    """RHF form; restates local_energy_sd.py:106-155."""
    nwalkers = walkers.Ghalfa.shape[0]
    Ga = walkers.Ghalfa.reshape(nwalkers, -1)
    e1b = 2.0 * Ga.dot(trial._rH1a.ravel()) + hamiltonian.ecore
    ecoul = ecoul_kernel_batch_real_rchol_rhf(trial._rchola, walkers.Ghalfa)
    exx = 2.0 * exx_kernel_batch_real_rchol(trial._rchola, walkers.Ghalfa)
    e2b = ecoul - exx
    energy = numpy.zeros((nwalkers, 3), dtype=numpy.complex128)
    energy[:, 0] = e1b + e2b
    energy[:, 1] = e1b
    energy[:, 2] = e2b
    return energy


def local_energy_batch(system, hamiltonian, walkers, trial):
    """CPU dispatch; restates ipie/estimators/local_energy_batch.py:36-75."""
    if walkers.rhf:
        return local_energy_single_det_rhf_batch(system, hamiltonian, walkers, trial)
    return local_energy_single_det_uhf_batch(system, hamiltonian, walkers, trial)


def half_rotated_cholesky_jk_uhf(trial, Ghalf):
    """Single-walker (J, K); restates ipie/estimators/generic.py:198-262, 377-405."""
    Ghalfa, Ghalfb = Ghalf
    Xa = trial._rchola.dot(Ghalfa.real.ravel()) + 1j * trial._rchola.dot(Ghalfa.imag.ravel())
    Xb = trial._rcholb.dot(Ghalfb.real.ravel()) + 1j * trial._rcholb.dot(Ghalfb.imag.ravel())
    X = Xa + Xb
    ecoul = 0.5 * numpy.dot(X, X)
    exx = 0.0j
    for rchol, G in ((trial._rchola, Ghalfa), (trial._rcholb, Ghalfb)):
        nocc, nbasis = G.shape
        if nocc == 0:
            continue
        L = rchol.reshape(-1, nocc, nbasis)
        T = numpy.einsum("xim,jm->xij", L, G, optimize=True)
        exx += 0.5 * numpy.einsum("xij,xji->", T, T)
    return ecoul, exx


def local_energy_cholesky_opt(trial, hamiltonian, Ghalf):
    """Restates ipie/estimators/generic.py:143-168."""
    e1b = numpy.sum(trial._rH1a * Ghalf[0]) + numpy.sum(trial._rH1b * Ghalf[1])
    eJ, eK = half_rotated_cholesky_jk_uhf(trial, Ghalf)
    e2b = eJ - eK
    return (e1b + e2b + hamiltonian.ecore, e1b + hamiltonian.ecore, e2b)


def local_energy_generic_cholesky(system, hamiltonian, G):
    """Full-G local energy; restates ipie/estimators/generic.py:28-95 (real chol)."""
    e1b = numpy.sum(hamiltonian.H1[0] * G[0]) + numpy.sum(hamiltonian.H1[1] * G[1])
    nbasis = hamiltonian.nbasis
    Ga, Gb = G[0], G[1]
    Xa = hamiltonian.chol.T.dot(Ga.real.ravel()) + 1j * hamiltonian.chol.T.dot(Ga.imag.ravel())
    Xb = hamiltonian.chol.T.dot(Gb.real.ravel()) + 1j * hamiltonian.chol.T.dot(Gb.imag.ravel())
    ecoul = numpy.dot(Xa, Xa) + numpy.dot(Xb, Xb) + 2 * numpy.dot(Xa, Xb)
    exx = 0.0j
    for x in range(hamiltonian.nchol):
        Lmn = hamiltonian.chol[:, x].reshape((nbasis, nbasis))
        T = Ga.T.dot(Lmn)
        exx += numpy.trace(T.dot(T))
        T = Gb.T.dot(Lmn)
        exx += numpy.trace(T.dot(T))
    e2b = 0.5 * (ecoul - exx)
    return (e1b + e2b + hamiltonian.ecore, e1b + hamiltonian.ecore, e2b)


# --------------------------------------------------------------------------
# Helpers used by tests and the CPU baseline.
# --------------------------------------------------------------------------
class OracleState:
    """Bundle of (system, hamiltonian, trial, walkers, propagator) built the way
    ipie's AFQMC.build does (ipie/qmc/afqmc.py:297-392)."""

    def __init__(self, system, hamiltonian, trial, walkers, propagator):
        self.system = system
        self.hamiltonian = hamiltonian
        self.trial = trial
        self.walkers = walkers
        self.propagator = propagator


def build_state(h1e, chol, psi, nelec, nwalkers, dt, init=None, ecore=0.0, exp_nmax=6):
    """Build the oracle objects from raw arrays (h1e [2,N,N], chol [N^2,X], psi [N, na+nb])."""
    nbasis = h1e.shape[-1]
    system = Generic(nelec)
    ham = GenericRealChol(h1e, chol, ecore=ecore)
    trial = SingleDet(psi, nelec, nbasis)
    trial.half_rotate(ham)
    trial.calculate_energy(system, ham)
    walkers = UHFWalkers(psi if init is None else init, nelec[0], nelec[1], nbasis, nwalkers)
    walkers.build(trial)
    prop = PhaselessGeneric(dt, exp_nmax=exp_nmax)
    prop.build(ham, trial, walkers)
    return OracleState(system, ham, trial, walkers, prop)
