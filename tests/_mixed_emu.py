"""CPU emulation of the mixed-precision step's arithmetic (test infrastructure).

The device's mixed mode (AFQMC_PREC_MIXED) splits every FP64 operand x of the
large contractions as hi = bf16(fp32(x)), lo = bf16(fp32(x - hi)) (round to
nearest even; split_bf16_pair in ipie_amd/csrc/afqmc_common.hpp) and forms a
real product A B as A_hi B_hi + A_hi B_lo + A_lo B_hi with FP32 accumulation.
This module repeats that rounding with numpy (the FP32 sums in sgemm order,
not the MFMA order, which changes only the last FP32 bits) on top of the
oracle's FP64 step, so that the mixed tests' tolerances come from the
arithmetic itself rather than from the device's own output:

  one body     P = phi + (expH1 - I) phi, the product split, the sum FP64
  force bias   vbias = L~ Ghalf per spin, split, FP32
  VHS          i sqrt(dt) sum_x xshifted_x L_x, split, kept in FP32
  Taylor       term k = VHS term_{k-1} / k: terms 1-3 split (three products),
               4-6 hi * hi only; terms 1-5 kept as their split planes (hi + lo),
               Q = P + their sum + term 6 in FP64
  exchange     g^T Wq g per spin (quadratic form, Wq = sum_x L_x (x) L_x) or
               sum_x tr(T_x T_x) (T-form, X < N), split, FP32 accumulation,
               the final dot / trace in FP64
  Coulomb      1/2 sum_x vbias^2 from the FP32 force bias

Green's function, overlap, fields and weights are the oracle's FP64 code.
"""

import numpy

from oracle import afqmc


# EXACT = True replaces every rounding by FP64 arithmetic (lo = 0 is then not
# used): the emulated step must then equal the oracle's, which checks that the
# emulation drops no term and transposes nothing.
EXACT = False


def bf16(x):
    """float64 -> float32 -> bfloat16 (round to nearest even), as float32 values."""
    f = numpy.ascontiguousarray(x, dtype=numpy.float32)
    u = f.view(numpy.uint32).astype(numpy.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(numpy.uint32).view(numpy.float32)


def split(x):
    """(hi, lo) with x ~ hi + lo to about 2^-17 relative."""
    x = numpy.asarray(x, dtype=numpy.float64)
    if EXACT:
        return x, numpy.zeros_like(x)
    hi = bf16(x)
    lo = bf16(x - hi.astype(numpy.float64))
    return hi, lo


def mm(a, b, products=3):
    """Real a @ b in split BF16 with FP32 accumulation (float64 result).
    a, b: float64 arrays, or (hi, lo) pairs already split."""
    ah, al = a if isinstance(a, tuple) else split(a)
    bh, bl = b if isinstance(b, tuple) else split(b)
    acc = ah @ bh
    if EXACT:
        return acc
    if products == 3:
        acc = acc + ah @ bl
        acc = acc + al @ bh
    return acc.astype(numpy.float64)


def cmm_real(a, b, products=3):
    """complex a @ real b."""
    bs = split(b)
    return mm(a.real, bs, products) + 1j * mm(a.imag, bs, products)


def real_cmm(a, b, products=3):
    """real a @ complex b."""
    a_s = split(a)
    return mm(a_s, b.real, products) + 1j * mm(a_s, b.imag, products)


def cmm(a, b, products=3):
    """complex a @ complex b: four real products, combined in FP32."""
    ar, ai, br, bi = split(a.real), split(a.imag), split(b.real), split(b.imag)
    f32 = numpy.float64 if EXACT else numpy.float32
    re = mm(ar, br, products).astype(f32) - mm(ai, bi, products).astype(f32)
    im = mm(ar, bi, products).astype(f32) + mm(ai, br, products).astype(f32)
    return re.astype(numpy.float64) + 1j * im.astype(numpy.float64)


def to_c64(z):
    if EXACT:
        return numpy.asarray(z, dtype=numpy.complex128)
    return numpy.asarray(z, dtype=numpy.complex64).astype(numpy.complex128)


# ---------------------------------------------------------------- contractions
def one_body(phi, expH1):
    """phi + (expH1 - I) phi per walker, product in split BF16, sum in FP64."""
    d = numpy.asarray(expH1) - numpy.eye(expH1.shape[0])
    if numpy.iscomplexobj(d):
        assert numpy.abs(d.imag).max(initial=0.0) == 0.0, "the device takes a real one-body propagator"
        d = d.real
    out = numpy.empty_like(phi)
    for w in range(phi.shape[0]):
        out[w] = phi[w] + real_cmm(d, phi[w])
    return out


def force_bias(rchola, rcholb, Ga, Gb, rhf=False):
    """vbias [W, X]: per spin L~ G in split BF16 (FP32), spins summed in FP64."""
    nw = Ga.shape[0]
    vb = cmm_real(Ga.reshape(nw, -1), rchola.T)
    if rhf:
        return 2.0 * vb
    if Gb.shape[1] > 0:
        vb = vb + cmm_real(Gb.reshape(nw, -1), rcholb.T)
    return vb


def vhs(chol_packed, xshifted_wx, sqrt_dt, idx_i, idx_j, nbasis):
    """VHS [W, N, N] complex: i sqrt(dt) sum_x xshifted[w, x] L_x, split BF16, FP32 values."""
    packed = cmm_real(xshifted_wx, chol_packed.T)  # [W, P]
    packed = to_c64(1j * sqrt_dt * packed)
    return afqmc.unpack_VHS_batch(idx_i, idx_j, packed, nbasis)


def planes_value(x):
    """The value a term's split planes hold: hi + lo."""
    if EXACT:
        return numpy.asarray(x, dtype=numpy.float64)
    hi, lo = split(x)
    return hi.astype(numpy.float64) + lo.astype(numpy.float64)


def taylor(phi, VHS, exp_nmax=6, split_terms=3):
    """P + sum_k term_k per walker (see the module docstring for the precision of each term)."""
    out = numpy.empty_like(phi)
    for w in range(phi.shape[0]):
        t = phi[w]
        terms = []
        for k in range(1, exp_nmax + 1):
            t = cmm(VHS[w], t, products=3 if k <= split_terms else 1) / k
            if k < exp_nmax:
                terms.append(planes_value(t.real) + 1j * planes_value(t.imag))
        acc = numpy.zeros(phi[w].shape, dtype=numpy.complex128)
        for term in reversed(terms):
            acc = acc + term
        out[w] = phi[w] + (t + acc)
    return out


def quad_matrix(rchol, nocc, nbasis):
    """Wq[(i, m'), (j, m)] = sum_x L_x[i][m] L_x[j][m'], so that g^T Wq g = sum_x tr(T_x T_x)."""
    L = rchol.reshape(-1, nocc, nbasis)
    X = L.shape[0]
    A = L.reshape(X, nocc * nbasis)  # [x, (i, m)]
    W = A.T @ A  # [(i, m), (j, m')]
    return numpy.ascontiguousarray(
        W.reshape(nocc, nbasis, nocc, nbasis).transpose(0, 3, 2, 1).reshape(nocc * nbasis, nocc * nbasis))


def exchange_quad(Wq, G):
    """sum_x tr(T_x T_x) per walker as g^T Wq g, the product g Wq in split BF16 (FP32)."""
    nw = G.shape[0]
    g = G.reshape(nw, -1)
    Y = cmm_real(g, Wq)
    return numpy.einsum("wk,wk->w", g, Y)


def exchange_t(rchol, G):
    """sum_x tr(T_x T_x) per walker with T_x = L_x G^T in split BF16 (FP32), trace in FP64."""
    nw, nocc, nbasis = G.shape
    L = rchol.reshape(-1, nbasis)  # [(x, i), m]
    X = L.shape[0] // nocc
    ls = split(L)
    out = numpy.empty(nw, dtype=numpy.complex128)
    for w in range(nw):
        gt = G[w].T  # [m, j]
        T = (mm(ls, gt.real) + 1j * mm(ls, gt.imag)).reshape(X, nocc, nocc)
        out[w] = numpy.einsum("xij,xji->", T, T)
    return out


# ---------------------------------------------------------------- step and energy
class MixedEmulator:
    """The mixed step on an oracle state (afqmc.build_state), with the oracle's
    FP64 Green's function, fields and weight update around emulated contractions.

    exchange: "quad" (g^T Wq g, the device's form when X > N) or "t" (every T_x)."""

    def __init__(self, orc, exchange="quad"):
        self.orc = orc
        self.exchange = exchange
        t = orc.trial
        self.Wq = None
        if exchange == "quad":
            self.Wq = [quad_matrix(t._rchola, t.nalpha, orc.hamiltonian.nbasis)]
            if t.nbeta > 0:
                self.Wq.append(quad_matrix(t._rcholb, t.nbeta, orc.hamiltonian.nbasis))

    def local_energy(self):
        """[W, 3] of the oracle walkers' current Ghalf (UHF form)."""
        o = self.orc
        w, t, ham = o.walkers, o.trial, o.hamiltonian
        nw = w.Ghalfa.shape[0]
        e1 = (w.Ghalfa.reshape(nw, -1).dot(t._rH1a.ravel()) + w.Ghalfb.reshape(nw, -1).dot(t._rH1b.ravel())
              + ham.ecore)
        vb = force_bias(t._rchola, t._rcholb, w.Ghalfa, w.Ghalfb)
        ecoul = 0.5 * numpy.einsum("wx,wx->w", vb, vb)
        ex = 0.0
        for s, (r, G) in enumerate(((t._rchola, w.Ghalfa), (t._rcholb, w.Ghalfb))):
            if G.shape[1] == 0:
                continue
            ex = ex + (exchange_quad(self.Wq[s], G) if self.exchange == "quad" else exchange_t(r, G))
        e2 = ecoul - 0.5 * ex
        e = numpy.empty((nw, 3), dtype=numpy.complex128)
        e[:, 0], e[:, 1], e[:, 2] = e1 + e2, e1, e2
        return e

    def propagate(self, xi, eshift):
        """One step with fields xi [W, X] (phaseless_base.py:312-338 with the
        contractions emulated); mutates the oracle walkers."""
        o = self.orc
        w, t, ham, p = o.walkers, o.trial, o.hamiltonian, o.propagator
        ovlp = t.calc_greens_function(w)
        w.phia = one_body(w.phia, p.expH1[0])
        w.phib = one_body(w.phib, p.expH1[1])
        vbias = force_bias(t._rchola, t._rcholb, w.Ghalfa, w.Ghalfb)
        xbar = -p.sqrt_dt * (1j * vbias - p.mf_shift)
        xbar = p.apply_bound_force_bias(xbar, p.fbbound)
        xshifted = xi - xbar
        cmf = -p.sqrt_dt * numpy.einsum("wx,x->w", xshifted, p.mf_shift)
        cfb = numpy.einsum("wx,wx->w", xi, xbar) - 0.5 * numpy.einsum("wx,wx->w", xbar, xbar)
        V = vhs(ham.chol_packed, xshifted, p.sqrt_dt, ham.sym_idx_i, ham.sym_idx_j, ham.nbasis)
        w.phia = taylor(w.phia, V, p.exp_nmax)
        w.phib = taylor(w.phib, V, p.exp_nmax)
        w.phia = one_body(w.phia, p.expH1[0])
        w.phib = one_body(w.phib, p.expH1[1])
        ovlp_new = t.calc_overlap(w)
        p.update_weight(w, ovlp, ovlp_new, cfb, cmf, eshift)
