"""Local energy (drop-in for ipie's local_energy_batch) and the energy estimator.

``local_energy_batch(system, hamiltonian, walkers, trial)`` mirrors
ipie/estimators/local_energy_batch.py:36-75: it reads the walkers'
precomputed ``Ghalfa``/``Ghalfb`` and returns [W, 3] complex128 =
(E1 + E2, E1, E2), with E1 = Ghalf . rH1 + ecore, E2 = Coulomb - exchange
(local_energy_sd.py:106-155 for RHF walkers, 358-403 otherwise).  The
Coulomb term comes out of the force-bias contraction and the exchange out of
the fused MFMA exchange kernel.
"""

import torch

from ipie_amd import engine as _engine


def local_energy_batch(system, hamiltonian, walkers, trial):
    eng = _engine.engine_for(trial, walkers)
    if eng.nchol == 0:
        raise RuntimeError("the propagator's build() must run before local energies (device context)")
    energy = torch.empty(walkers.nwalkers, 3, dtype=torch.complex128, device=walkers.device)
    eng.ctx.local_energy(walkers.Ghalfa, walkers.Ghalfb, energy)
    return energy


def local_energy(system, hamiltonian, walkers, trial):
    """ipie/estimators/energy.py:54-61 dispatch for (Generic, GenericRealChol, UHFWalkers, SingleDet)."""
    return local_energy_batch(system, hamiltonian, walkers, trial)


class EnergyEstimator:
    """Weighted energy sums of one block (ipie/estimators/energy.py:133-181)."""

    def __init__(self, system=None, ham=None, trial=None, filename=None):
        self.names = ["ENumer", "EDenom", "ETotal", "E1Body", "E2Body"]
        self._data = {k: 0.0j for k in self.names}
        self.print_to_stdout = True
        self.ascii_filename = filename

    @property
    def data(self):
        return [self._data[k] for k in self.names]

    def __getitem__(self, key):
        return self._data[key]

    def get_index(self, name):
        return self.names.index(name)

    def compute_estimator(self, system=None, walkers=None, hamiltonian=None, trial=None):
        """Weighted sums on the device (afqmc_walker_sums, fixed order), one
        read-back for the block."""
        trial.calc_greens_function(walkers)
        energy = local_energy(system, hamiltonian, walkers, trial)
        ctx = _engine.engine_for(trial, walkers).ctx
        acc = torch.zeros(3, 3, dtype=torch.complex128, device=walkers.device)
        for k in range(3):
            ctx.walker_sums(walkers.weight, acc[k], values=energy, column=k)
        s = acc.cpu().numpy()
        self._data["ENumer"] = complex(s[0, 2].real)
        self._data["EDenom"] = complex(s[0, 0].real)
        self._data["E1Body"] = complex(s[1, 2].real)
        self._data["E2Body"] = complex(s[2, 2].real)
        return self.data

    def post_reduce_hook(self, data):
        ip, nn, dd = self.get_index("ETotal"), self.get_index("ENumer"), self.get_index("EDenom")
        data[ip] = data[nn] / data[dd]
        data[self.get_index("E1Body")] = data[self.get_index("E1Body")] / data[dd]
        data[self.get_index("E2Body")] = data[self.get_index("E2Body")] / data[dd]
