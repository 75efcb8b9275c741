"""UHF walker population held in device memory (drop-in for ipie's UHFWalkers).

Mirrors ipie/walkers/uhf_walkers.py:27-161 and base_walkers.py:86-224: the
same attribute names (phia, phib, Ghalfa, Ghalfb, weight, unscaled_weight,
phase, ovlp, sgn_ovlp, log_ovlp, hybrid_energy, detR, detR_shift, log_detR,
log_shift, rhf, buff_names) as torch tensors on the GPU.  ``reortho`` runs
the gfx950 QR kernel and follows the CPU convention (positive real diag R).
"""

import numpy
import torch

from ipie_amd import engine as _engine


class UHFWalkers:
    def __init__(self, initial_walker, nup, ndown, nbasis, nwalkers, mpi_handler=None, verbose=False,
                 device=None):
        assert initial_walker.ndim == 2
        self.nup = nup
        self.ndown = ndown
        self.nbasis = nbasis
        self.nwalkers = nwalkers
        self.mpi_handler = mpi_handler
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
        dev, f64, c128 = self.device, torch.float64, torch.complex128
        init = numpy.asarray(initial_walker, dtype=numpy.complex128)
        pa = torch.from_numpy(numpy.ascontiguousarray(init[:, :nup])).to(dev)
        pb = torch.from_numpy(numpy.ascontiguousarray(init[:, nup:])).to(dev)
        self.phia = pa.unsqueeze(0).repeat(nwalkers, 1, 1).contiguous()
        self.phib = pb.unsqueeze(0).repeat(nwalkers, 1, 1).contiguous()
        self.Ghalfa = torch.zeros(nwalkers, nup, nbasis, dtype=c128, device=dev)
        self.Ghalfb = torch.zeros(nwalkers, ndown, nbasis, dtype=c128, device=dev)
        self.weight = torch.ones(nwalkers, dtype=f64, device=dev)
        self.unscaled_weight = self.weight.clone()
        self.phase = torch.ones(nwalkers, dtype=c128, device=dev)
        self.ovlp = torch.ones(nwalkers, dtype=c128, device=dev)
        self.sgn_ovlp = torch.ones(nwalkers, dtype=f64, device=dev)
        self.log_ovlp = torch.zeros(nwalkers, dtype=f64, device=dev)
        self.eloc = torch.zeros(nwalkers, dtype=f64, device=dev)
        # ipie starts with float zeros and turns complex after the first update;
        # here it is complex from the start (same values)
        self.hybrid_energy = torch.zeros(nwalkers, dtype=c128, device=dev)
        self.detR = torch.ones(nwalkers, dtype=f64, device=dev)
        self.detR_shift = torch.zeros(nwalkers, dtype=f64, device=dev)
        self.log_detR = torch.zeros(nwalkers, dtype=f64, device=dev)
        self.log_shift = torch.zeros(nwalkers, dtype=f64, device=dev)
        self.log_detR_shift = torch.zeros(nwalkers, dtype=f64, device=dev)
        self.buff_names = ["weight", "unscaled_weight", "phase", "hybrid_energy", "ovlp", "sgn_ovlp", "log_ovlp",
                           "phia", "phib"]
        self.rhf = False
        self._engine = None

    def status_flags(self, clear=True):
        """Per-walker status, int32 [W] on the device: the OR of
        ipie_amd._lib.FLAG_SINGULAR (zero overlap with the trial), FLAG_NONFINITE
        (non-finite weight) and FLAG_CHOLESKY (rank-deficient in reortho) each
        walker met since the last call (afqmc_walker_flags; no synchronisation)."""
        if self._engine is None:
            return torch.zeros(self.nwalkers, dtype=torch.int32, device=self.device)
        return self._engine.ctx.walker_flags(self.nwalkers, clear=clear)

    def build(self, trial):
        """Restates uhf_walkers.py:101-102."""
        self.ovlp = trial.calc_greens_function(self)

    def reortho(self):
        """QR re-orthogonalisation of every walker (uhf_walkers.py:108-143); returns detR."""
        eng = self._engine
        if eng is None or eng.nchol == 0:
            raise RuntimeError("reortho needs the engine created by the propagator's build()")
        eng.ctx.reortho(self.phia, self.phib, self.ovlp, self.detR, self.log_detR, self.detR_shift)
        return self.detR

    def reortho_batched(self):
        return self.reortho()

    def orthogonalise(self, free_projection=False):
        """Restates base_walkers.py:166-179 (phaseless branch)."""
        if free_projection:
            raise NotImplementedError("free projection is not part of the phaseless step")
        return self.reortho()

    def cast_to_cupy(self, verbose=False):
        """The population already lives on the device."""
        return None


class WalkerAccumulator:
    """Per-block walker sums (ipie/walkers/base_walkers.py:32-83) kept on the
    device: ``update`` adds (sum weight, sum unscaled_weight, sum weight *
    hybrid_energy) with one kernel and no host synchronisation (the reference
    reads three sums back every step); ``buffer`` copies them to the host when
    it is read (print_block, once per block)."""

    def __init__(self, names, nsteps):
        self.names = names
        self.size = len(names)
        self._data_index = {k: i for i, k in enumerate(self.names)}
        self.nsteps_per_block = nsteps
        self._eshift = 0.0
        self._dev = None
        self._host = numpy.zeros((self.size,), dtype=numpy.complex128)

    def update(self, walkers):
        eng = walkers._engine
        if eng is None:
            raise RuntimeError("WalkerAccumulator.update needs walkers built against a trial (walkers.build)")
        if self._dev is None:
            self._dev = torch.zeros(3, dtype=torch.complex128, device=walkers.weight.device)
            self._dev.copy_(torch.from_numpy(self._host[:3].copy()))
        eng.ctx.walker_sums(walkers.weight, self._dev, walkers.unscaled_weight, walkers.hybrid_energy)

    @property
    def buffer(self):
        if self._dev is not None:
            self._host[:3] = self._dev.cpu().numpy()
        return self._host

    def zero(self):
        self._host.fill(0.0j)
        if self._dev is not None:
            self._dev.zero_()

    def get_index(self, name):
        index = self._data_index.get(name, None)
        if index is None:
            raise RuntimeError(f"Unknown walker property {name}")
        return index

    @property
    def eshift(self):
        return self._eshift.real

    @eshift.setter
    def eshift(self, value):
        self._eshift = value

    def post_reduce_hook(self, vals, block):
        """Block averages from the reduced sums (base_walkers.py:70-80)."""
        assert len(vals) == len(self.names)
        factor = 1 if block == 0 else self.nsteps_per_block
        nume = self.get_index("HybridEnergy")
        deno = self.get_index("Weight")
        vals[nume] = vals[nume] / vals[deno]
        vals[deno] = vals[deno] / factor
        ix = self.get_index("WeightFactor")
        vals[ix] = vals[ix] / factor

    def to_text(self, vals):
        """Fixed-width text of the real parts (ipie/utils/io.py:197-198 format)."""
        return " ".join(f"{f: .16e}" for f in numpy.asarray(vals).real)


def clip_weights(walkers, wbound):
    """The driver's clip of the weights to [-wbound, wbound] after each step
    (ipie/qmc/afqmc.py:584-588), in place on the device."""
    eng = walkers._engine
    if eng is None:
        raise RuntimeError("clip_weights needs walkers built against a trial (walkers.build)")
    eng.ctx.clip_weights(walkers.weight, wbound)
