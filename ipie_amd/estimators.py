"""Local energy (drop-in for ipie's local_energy_batch) and the energy estimator.

``local_energy_batch(system, hamiltonian, walkers, trial)`` mirrors
ipie/estimators/local_energy_batch.py:36-75: it reads the walkers'
precomputed ``Ghalfa``/``Ghalfb`` and returns [W, 3] complex128 =
(E1 + E2, E1, E2), with E1 = Ghalf . rH1 + ecore, E2 = Coulomb - exchange
(local_energy_sd.py:106-155 for RHF walkers, 358-403 otherwise).  The
Coulomb term comes out of the force-bias contraction and the exchange out of
the fused MFMA exchange kernel.
"""

import os

import numpy
import torch

from ipie_amd import engine as _engine
from ipie_amd import h5write


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
        self.scalar_estimator = True
        self.shape = (len(self.names),)
        self.size = len(self.names)
        self._ascii_filename = None
        self.ascii_filename = filename

    @property
    def ascii_filename(self):
        return self._ascii_filename

    @ascii_filename.setter
    def ascii_filename(self, name):
        """Text output file, started with the header (estimator_base.py:52-63)."""
        self._ascii_filename = name
        if name is not None:
            with open(name, "w") as fh:
                fh.write(self.header_to_text + "\n")

    @property
    def data(self):
        return [self._data[k] for k in self.names]

    @property
    def header_to_text(self):
        return " ".join(f"{s:>23}" for s in self.names)

    def data_to_text(self, vals):
        """Fixed-width text of the real parts (estimator_base.py:109-117)."""
        if self._ascii_filename is None and not self.print_to_stdout:
            return ""
        return " ".join(f"{f: .16e}" for f in numpy.asarray(vals).real)

    def to_ascii_file(self, vals):
        if self._ascii_filename is not None:
            with open(self._ascii_filename, "a") as fh:
                fh.write(vals + "\n")

    def zero(self):
        for k in self.names:
            self._data[k] = 0.0j

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


class EstimatorHandler:
    """Block estimates and their output (ipie/estimators/handler.py:41-248)
    for one rank and the energy observable.

    The file has ipie's layout (handler.py:148-181, estimators/utils.py:24-97):
    ``metadata`` (the JSON string), and under ``block_size_1``: ``data/<chunk>``
    complex128 [buffer_size, walker props + estimates] rows, ``max_block/<chunk>``,
    ``num_walker_props``, ``walker_prop_header``, and per observable ``shape``,
    ``size``, ``scalar``, ``names``, ``offset``.  It is written with
    ``ipie_amd.h5write`` (h5py is absent) and rewritten whole at every block.
    """

    def __init__(self, comm, system, hamiltonian, trial, walker_state=None, verbose=False, filename=None,
                 block_size=1, basename="estimates", overwrite=True, observables=("energy",), index=0):
        if comm is not None and getattr(comm, "size", 1) != 1:
            raise NotImplementedError("the estimator handler covers one rank (comm.size == 1)")
        self.basename = basename
        self.filename = filename
        self.index = 0
        if self.filename is None:
            self.filename = f"{self.basename}.{self.index}.h5"
            while os.path.isfile(self.filename) and not overwrite:
                self.index = int(self.filename.split(".")[1]) + 1
                self.filename = f"{self.basename}.{self.index}.h5"
        self.buffer_size = 1000
        if walker_state is not None:
            self.num_walker_props = walker_state.size
            self.walker_header = list(walker_state.names)
        else:
            self.num_walker_props = 0
            self.walker_header = []
        self._estimators = {}
        self._offsets = {}
        self.json_string = "{}"
        self._rows = []
        for obs in observables:
            if obs != "energy":
                raise RuntimeError(f"unknown observable: {obs}")
            self[obs] = EnergyEstimator(system=system, ham=hamiltonian, trial=trial)
        if verbose:
            print(f"# Writing estimator data to {self.filename}")

    def __setitem__(self, name, estimator):
        if name not in self._estimators:
            self._offsets[name] = sum(e.size for e in self._estimators.values())
        self._estimators[name] = estimator

    def __getitem__(self, key):
        return self._estimators[key]

    def items(self):
        return self._estimators.items()

    def get_offset(self, name):
        return self._offsets[name]

    @property
    def size(self):
        return sum(e.size for e in self._estimators.values())

    def initialize(self, comm=None, print_header=True):
        n = self.size + self.num_walker_props
        self.local_estimates = numpy.zeros(n, dtype=numpy.complex128)
        self.global_estimates = numpy.zeros(n, dtype=numpy.complex128)
        header = f"{'Block':>17s}  " + " ".join(f"{s:>23}" for s in self.walker_header) + " "
        for _, e in self.items():
            if e.print_to_stdout:
                header += e.header_to_text
        self._write()
        if print_header:
            print(header)
        return header

    def compute_estimators(self, system=None, hamiltonian=None, trial=None, walker_batch=None):
        offset = self.num_walker_props
        for k, e in self.items():
            e.compute_estimator(system=system, walkers=walker_batch, hamiltonian=hamiltonian, trial=trial)
            start = offset + self.get_offset(k)
            self.local_estimates[start:start + e.size] += numpy.asarray(e.data, dtype=numpy.complex128)

    def print_block(self, comm, block, walker_factors, div_factor=None, print_line=True):
        """Walker sums + estimates of the block: hooks, text line, file row
        (handler.py:207-242; the reduction over ranks is the identity here)."""
        self.local_estimates[: walker_factors.size] = walker_factors.buffer
        self.global_estimates[:] = self.local_estimates
        offset = walker_factors.size
        walker_factors.post_reduce_hook(self.global_estimates[:offset], block)
        out = " " + walker_factors.to_text(self.global_estimates[:offset]) + " "
        for k, e in self.items():
            start = offset + self.get_offset(k)
            est = self.global_estimates[start:start + e.size]
            e.post_reduce_hook(est)
            text = e.data_to_text(est)
            e.to_ascii_file(text)
            if e.print_to_stdout:
                out += text
        walker_factors.eshift = self.global_estimates[walker_factors.get_index("HybridEnergy")]
        self._rows.append(self.global_estimates.copy())
        self._write()
        line = f"{block:>17d} " + out
        if print_line:
            print(line)
        self.zero()
        return line

    def _write(self):
        n = self.size + self.num_walker_props
        data, max_block = {}, {}
        for c in range(0, max(len(self._rows), 1), self.buffer_size):
            chunk = numpy.zeros((self.buffer_size, n), dtype=numpy.complex128)
            rows = self._rows[c:c + self.buffer_size]
            if rows:
                chunk[: len(rows)] = numpy.array(rows)
            key = f"{c // self.buffer_size:09d}"
            data[key] = chunk
            max_block[key] = numpy.int64(max(len(rows) - 1, 0))
        blk = {
            "num_walker_props": numpy.int64(self.num_walker_props),
            "walker_prop_header": numpy.array([h.encode() for h in self.walker_header] or [b""]),
            "shape": {k: numpy.array(e.shape, dtype=numpy.int64) for k, e in self.items()},
            "size": {k: numpy.int64(e.size) for k, e in self.items()},
            "scalar": {k: numpy.int64(int(e.scalar_estimator)) for k, e in self.items()},
            "names": {k: " ".join(e.names) for k, e in self.items()},
            "offset": {k: numpy.int64(self.num_walker_props + self.get_offset(k)) for k, e in self.items()},
        }
        if self._rows:
            blk["data"] = data
            blk["max_block"] = max_block
        h5write.write_file(self.filename, {"metadata": self.json_string, "block_size_1": blk})

    def dump_metadata(self):
        self._write()

    def zero(self):
        self.local_estimates[:] = 0.0
        self.global_estimates[:] = 0.0
        for _, e in self.items():
            e.zero()
