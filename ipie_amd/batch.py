"""Several independent systems stepped together from one hipGraph.

Each system is its own (propagator, walkers, hamiltonian, trial) with its own
device context, of any size.  ``MultiSystemBatch.capture`` records a block of
steps of every system into one graph, each system on its own stream (forked
from and joined to the first), so the kernels of small systems overlap on the
chip; ``replay`` launches the block, continuing each propagator's field
counter exactly as its eager steps would.
"""

import ctypes

import numpy
import torch

from ipie_amd import _lib


class MultiSystemBatch:
    def __init__(self, systems):
        """systems: list of (propagator, walkers, hamiltonian, trial), each
        built (propagator.build) on the same device."""
        if not systems:
            raise ValueError("a batch needs at least one system")
        self.systems = list(systems)
        self.ctx = [p._eng(w).ctx for p, w, _, _ in self.systems]
        self.device = self.ctx[0].device
        if any(c.device != self.device for c in self.ctx):
            raise ValueError("the systems of a batch share one device")
        self.lib = _lib.load()
        arr = (ctypes.c_void_p * len(self.ctx))(*[c.h.value for c in self.ctx])
        h = ctypes.c_void_p()
        st = self.lib.afqmc_batch_create(ctypes.byref(h), arr, len(self.ctx))
        self._check(st, "afqmc_batch_create")
        self.h = h
        self.streams = [torch.cuda.Stream(device=self.device) for _ in self.ctx]
        self.steps = None

    def _check(self, st, name):
        if st != _lib.AFQMC_OK:
            msg = self.lib.afqmc_last_error(self.ctx[0].h)
            raise _lib.AfqmcError(f"{name}: {_lib.STATUS_NAMES.get(st, st)}: {msg.decode() if msg else ''}")

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            try:
                self.lib.afqmc_batch_destroy(h)
            except Exception:
                pass
            self.h = None

    def capture(self, nsteps, reortho_every=0, energies=None):
        """Record `nsteps` propagate_walkers calls of every system (and a reortho
        after every `reortho_every`-th); energies: optional list of [W_i, 3]
        complex128 tensors receiving each system's local energy per step.
        Every system must have run the same calls eagerly before."""
        cur = torch.cuda.current_stream(self.device)
        for s in self.streams:
            s.wait_stream(cur)
        for c, s in zip(self.ctx, self.streams):
            c.bind_stream(s)
        self.steps = None  # a failed capture leaves nothing to replay
        self._check(self.lib.afqmc_batch_begin(self.h), "afqmc_batch_begin")
        try:
            for j in range(nsteps):
                for i, (prop, walkers, ham, trial) in enumerate(self.systems):
                    if prop.xi_source is not None:
                        raise RuntimeError("a step graph draws its fields on the device; unset xi_source")
                    with torch.cuda.stream(self.streams[i]):
                        self.ctx[i].propagate(
                            walkers.phia, walkers.phib, walkers.Ghalfa, walkers.Ghalfb, walkers.weight,
                            walkers.ovlp, walkers.hybrid_energy, 0.0, seed=prop.seed, counter=0,
                            log_shift=walkers.log_shift, energy_prev=None if energies is None else energies[i])
                        if reortho_every and (j + 1) % reortho_every == 0:
                            walkers.reortho()
        except BaseException:
            # close the capture (joins every forked stream) before the error propagates,
            # so no stream is left recording into a graph that will never be launched
            self.lib.afqmc_batch_end(self.h)
            raise
        self._check(self.lib.afqmc_batch_end(self.h), "afqmc_batch_end")
        steps = (_lib._i32 * len(self.ctx))()
        self._check(self.lib.afqmc_batch_steps(self.h, steps), "afqmc_batch_steps")
        self.steps = [int(v) for v in steps]
        return self.steps

    def replay(self, eshifts, replays=1):
        """Launch the recorded block `replays` times; eshifts: one per system."""
        if self.steps is None:
            raise RuntimeError("capture() first")
        n = len(self.ctx)
        counters = (_lib._u64 * n)(*[p.counter + 1 for p, _, _, _ in self.systems])
        es = (_lib._f64 * n)(*[float(numpy.real(e)) for e in eshifts])
        s0 = self.streams[0]
        s0.wait_stream(torch.cuda.current_stream(self.device))
        self.ctx[0].bind_stream(s0)
        self._check(self.lib.afqmc_batch_launch(self.h, int(replays), counters, es), "afqmc_batch_launch")
        torch.cuda.current_stream(self.device).wait_stream(s0)
        for (prop, _, _, _), k in zip(self.systems, self.steps):
            prop.counter += replays * k
