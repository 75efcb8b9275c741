"""Full-size parity cases checked through a subset of walkers.

The device steps the whole population at the workload's real shape (launch
grids over all W walkers, the full quadratic-form exchange matrices); the
oracle steps a fixed subset of those walkers, fed the same rows of the
auxiliary fields.  Walkers are independent within a step
(ipie/propagation/phaseless_base.py:312-338: every product, the fields and the
weight update act walker by walker), so the subset's rows must agree with the
oracle run on those rows alone.  The oracle's static tensors (packed and
half-rotated Cholesky vectors, h1e_mod) are handed to the device objects so
the set-up arithmetic runs once.
"""

import functools

import numpy

from oracle import afqmc
from tests._helpers import random_walkers

# (N, X, decay D, n per spin, W): the headline workload and the FP64 parity workload
SHAPES = {
    "n400": (400, 2400, 300.0, 40, 1024),
    "n300": (300, 1800, 225.0, 30, 2048),
    "n200": (200, 1000, 150.0, 20, 512),
}


def normalised_scale(N, D):
    """Cholesky scale of the benchmark's normalised generator (sqrt(12 / (N D)))."""
    return (12.0 / (N * D)) ** 0.5


@functools.lru_cache(maxsize=2)
def arrays(name, scale_kind="normalised"):
    """(h1e, chol, ecore, psi) of a workload; cached across tests of one process."""
    from ipie_amd import synthetic

    N, X, D, n, W = SHAPES[name]
    h1e, chol, ecore = synthetic.make_hamiltonian(N, X, D, seed=7)
    if scale_kind == "normalised":
        chol *= normalised_scale(N, D)
    psi = synthetic.make_trial(h1e, n, n)
    return h1e, chol, ecore, psi


def spread_subset(W, k=8):
    """k walker indices covering the first and last walkers, both sides of the
    128-walker row blocks of the GEMM grids, and the middle of the population."""
    cand = [0, 1, 127, 128, W // 2 - 1, W // 2 + 5, W - 2, W - 1]
    out = sorted(set(i for i in cand if 0 <= i < W))
    return out[:k]


def device_state(name, precision="mixed", surrogate=None, amp=0.01, seed=3):
    """Device objects only (no oracle) for a full-size workload, the static
    tensors built by the device set-up kernels: (system, ham, trial, walkers, prop)."""
    import torch

    from ipie_amd import Generic as DevSystem
    from ipie_amd.hamiltonian import GenericRealChol
    from ipie_amd.propagator import PhaselessGeneric
    from ipie_amd.trial import SingleDet
    from ipie_amd.walkers import UHFWalkers

    N, X, D, n, W = SHAPES[name]
    print(f"[case] {name} device only N={N} X={X} n={n} W={W} precision={precision} surrogate={surrogate}",
          flush=True)
    h1e, chol, ecore, psi = arrays(name)
    ham = GenericRealChol(h1e, chol, ecore=ecore, host_setup=False)
    trial = SingleDet(psi, (n, n), N)
    walkers = UHFWalkers(psi, n, n, N, W)
    prop = PhaselessGeneric(0.005, precision=precision, surrogate=surrogate)
    prop.build(ham, trial, walkers, device_setup=True)
    trial.calculate_energy(DevSystem((n, n)), ham)
    init = random_walkers(N, n, n, W, seed=seed, amp=amp, base=psi)
    walkers.phia.copy_(torch.from_numpy(numpy.ascontiguousarray(init[:, :, :n])))
    walkers.phib.copy_(torch.from_numpy(numpy.ascontiguousarray(init[:, :, n:])))
    walkers.ovlp.copy_(trial.calc_greens_function(walkers))
    return DevSystem((n, n)), ham, trial, walkers, prop


class SubsetPair:
    """Device objects over all W walkers, oracle objects over a subset."""

    def __init__(self, name, precision="fp64", scale_kind="normalised", amp=0.01, seed=3, surrogate=None):
        import torch

        from ipie_amd import Generic as DevSystem
        from ipie_amd.hamiltonian import GenericRealChol
        from ipie_amd.propagator import PhaselessGeneric
        from ipie_amd.trial import SingleDet
        from ipie_amd.walkers import UHFWalkers

        N, X, D, n, W = SHAPES[name]
        print(f"[case] {name} N={N} X={X} n={n} W={W} precision={precision} scale={scale_kind}", flush=True)
        self.torch = torch
        self.N, self.X, self.W, self.n = N, X, W, n
        h1e, chol, ecore, psi = arrays(name, scale_kind)
        self.subset = spread_subset(W)
        k = len(self.subset)
        self.orc = afqmc.build_state(h1e, chol, psi, (n, n), k, 0.005, ecore=ecore)
        o = self.orc
        self.system = DevSystem((n, n))
        self.ham = GenericRealChol(h1e, chol, ecore=ecore, host_setup=False)
        self.ham.chol_packed = o.hamiltonian.chol_packed
        self.ham.h1e_mod = o.hamiltonian.h1e_mod
        self.trial = SingleDet(psi, (n, n), N)
        self.trial._rchola, self.trial._rcholb = o.trial._rchola, o.trial._rcholb
        self.trial._rH1a, self.trial._rH1b = o.trial._rH1a, o.trial._rH1b
        self.trial.half_rotated = True
        self.trial.energy = o.trial.energy
        self.walkers = UHFWalkers(psi, n, n, N, W)
        self.prop = PhaselessGeneric(0.005, precision=precision, surrogate=surrogate)
        self.prop.build(self.ham, self.trial, self.walkers)
        init = random_walkers(N, n, n, W, seed=seed, amp=amp, base=psi)
        self.walkers.phia.copy_(torch.from_numpy(numpy.ascontiguousarray(init[:, :, :n])))
        self.walkers.phib.copy_(torch.from_numpy(numpy.ascontiguousarray(init[:, :, n:])))
        ovlp = self.trial.calc_greens_function(self.walkers)
        self.walkers.ovlp.copy_(ovlp)

    def rows(self, attr):
        """The subset's rows of a device walker array, on the host."""
        t = getattr(self.walkers, attr)
        idx = self.torch.tensor(self.subset, device=t.device)
        return t.index_select(0, idx).cpu().numpy()

    def sync_oracle(self):
        """The oracle continues from the device state of its walkers."""
        o = self.orc.walkers
        for attr in ("phia", "phib", "weight", "ovlp", "hybrid_energy"):
            setattr(o, attr, self.rows(attr).copy())

    def step(self, xi_full, eshift, energy=True):
        """One device step over all walkers and one oracle step over the subset
        (after resynchronising it); returns (device energy rows of the incoming
        walkers, oracle energy of the same walkers) when `energy`."""
        torch = self.torch
        self.sync_oracle()
        self.prop.xi_source = lambda w, x: xi_full
        self.orc.propagator.xi_source = lambda w, x: xi_full[self.subset]
        e_dev = None
        if energy:
            e_dev = torch.empty(self.W, 3, dtype=torch.complex128, device=self.walkers.device)
        self.prop.propagate_walkers(self.walkers, self.ham, self.trial, eshift, energy_prev=e_dev)
        e_orc = None
        if energy:
            self.orc.trial.calc_greens_function(self.orc.walkers)
            e_orc = afqmc.local_energy_batch(self.orc.system, self.orc.hamiltonian, self.orc.walkers, self.orc.trial)
        self.orc.propagator.propagate_walkers(self.orc.walkers, self.orc.hamiltonian, self.orc.trial, eshift)
        if energy:
            idx = torch.tensor(self.subset, device=e_dev.device)
            e_dev = e_dev.index_select(0, idx).cpu().numpy()
        return e_dev, e_orc
