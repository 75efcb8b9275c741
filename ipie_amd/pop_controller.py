"""Population control of the device walkers (drop-in for ipie's PopController).

Mirrors ipie/walkers/pop_controller.py:34-126 for one rank: the same
constructor, ``pop_control(walkers, comm)``, ``total_weight`` (read by the
driver's weight clipping, ipie/qmc/afqmc.py:584-588) and timer attributes.
The decisions run in one gfx950 workgroup (``afqmc_pop_plan``) and the walker
buffers (``walkers.buff_names``) move on the device (``afqmc_pop_gather``);
the host reads back the total weight and the pair count only.

Random numbers come from ``numpy.random`` exactly as in the reference: one
``rand()`` per branched pair (pair_branch), one ``random()`` per comb, one
``rand()`` per reconfiguration, so a seeded run consumes the global stream the
same way and takes the same decisions.
"""

import time

import numpy
import torch

from ipie_amd import _lib

METHODS = ("pair_branch", "comb", "stochastic_reconfiguration")


class PopControllerTimer:
    """Timer attributes of ipie/walkers/pop_controller.py:10-31.  On one rank
    there is no communication: everything counts as non-communication."""

    def __init__(self):
        self.start_time_const = 0.0
        self.communication_time = 0.0
        self.non_communication_time = 0.0
        self.recv_time = 0.0
        self.send_time = 0.0

    def start_time(self):
        self.start_time_const = time.time()

    def add_non_communication(self):
        self.non_communication_time += time.time() - self.start_time_const

    def add_communication(self):
        self.communication_time += time.time() - self.start_time_const


class PopController:
    """ipie/walkers/pop_controller.py:34-76; ``mpi_handler`` may be None or
    describe one rank."""

    def __init__(self, num_walkers_local, num_steps, mpi_handler=None, pop_control_method="pair_branch",
                 min_weight=0.1, max_weight=4, reconfiguration_freq=50, verbose=False):
        if pop_control_method not in METHODS:
            raise ValueError(f"unknown population control method {pop_control_method!r}")
        size = 1 if mpi_handler is None else mpi_handler.size
        if size != 1:
            raise NotImplementedError("device population control covers one rank (comm.size == 1)")
        self.verbose = verbose
        self.num_walkers_local = num_walkers_local
        self.num_steps = num_steps
        self.mpi_handler = mpi_handler
        self.method = pop_control_method
        self.min_weight = min_weight
        self.max_weight = max_weight
        self.reconfiguration_counter = 0
        self.reconfiguration_freq = reconfiguration_freq
        self.size = size
        self.ntot_walkers = num_walkers_local * size
        self.target_weight = self.ntot_walkers
        self.total_weight = self.ntot_walkers
        self.timer = PopControllerTimer()
        self.last_npairs = 0  # walkers branched by the last pair_branch call
        self._bufs = {}
        if verbose:
            print(f"# Using {self.method} population control algorithm.")
            print(f"# target weight is {self.target_weight}")

    def _scratch(self, nw, device):
        key = (nw, device)
        b = self._bufs.get(key)
        if b is None:
            b = {
                "src": torch.empty(nw, dtype=torch.int32, device=device),
                "new_weight": torch.empty(nw, dtype=torch.float64, device=device),
                "total": torch.empty(1, dtype=torch.float64, device=device),
                "npairs": torch.empty(1, dtype=torch.int32, device=device),
            }
            self._bufs = {key: b}
        return b

    def pop_control(self, walkers, comm=None):
        """Rescale the weights to the target and apply the method
        (ipie/walkers/pop_controller.py:78-126).  Raises ValueError when the
        total weight is below 1e-8, before any walker is touched."""
        if comm is not None and getattr(comm, "size", 1) != 1:
            raise NotImplementedError("device population control covers one rank (comm.size == 1)")
        self.timer.start_time()
        if self.ntot_walkers == 1:
            return
        eng = getattr(walkers, "_engine", None)
        if eng is None:
            raise RuntimeError("pop_control needs walkers built against a trial (walkers.build)")
        ctx = eng.ctx
        nw = walkers.nwalkers
        dev = walkers.weight.device
        b = self._scratch(nw, dev)

        method = self.method
        if method == "stochastic_reconfiguration":
            self.reconfiguration_counter += 1
            if self.reconfiguration_counter % self.reconfiguration_freq == 0:
                self.reconfiguration_counter = 0
            else:
                method = "rescale"

        state = None
        uniforms = None
        if method == "pair_branch":
            # at most nw // 2 pairs; the stream is rewound afterwards and
            # advanced by the pairs actually branched
            state = numpy.random.get_state()
            uniforms = numpy.random.rand(max(nw // 2, 1))
        elif method == "comb":
            uniforms = numpy.array([numpy.random.random()])
        elif method == "stochastic_reconfiguration":
            uniforms = numpy.array([numpy.random.rand()])
        u_dev = None if uniforms is None else torch.from_numpy(uniforms).to(dev)

        ctx.pop_plan(method, walkers.weight, self.target_weight, self.min_weight, self.max_weight, u_dev, b["src"],
                     b["new_weight"], b["total"], b["npairs"])
        total = float(b["total"].item())
        self.last_npairs = 0
        if state is not None:
            numpy.random.set_state(state)
            npairs = int(b["npairs"].item())
            self.last_npairs = npairs
            if npairs:
                numpy.random.rand(npairs)
        if total < 1e-8:
            raise ValueError(f"total weight is {total:13.8e}")
        self.total_weight = total

        walkers.unscaled_weight.copy_(walkers.weight)
        if method != "rescale":
            in_place = method != "stochastic_reconfiguration"
            for name in walkers.buff_names:
                if name == "weight":
                    continue
                t = getattr(walkers, name)
                if in_place:
                    ctx.pop_gather(b["src"], t, t)
                else:
                    tmp = torch.empty_like(t)
                    ctx.pop_gather(b["src"], t, tmp)
                    t.copy_(tmp)
        walkers.weight.copy_(b["new_weight"])
        self.timer.add_non_communication()
