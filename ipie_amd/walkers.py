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
