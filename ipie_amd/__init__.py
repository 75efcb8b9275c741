"""MI355X-native phaseless-AFQMC step: drop-ins for ipie's
``PhaselessGeneric.propagate_walkers`` and ``local_energy_batch``.

Host code is Python; the step itself runs in hand-written gfx950 HIP kernels
behind the C ABI of ``libafqmc_hip.so`` (include/afqmc_hip.h).
"""

__version__ = "0.1.0"


class Generic:
    """System carrying the electron numbers (ipie/systems/generic.py:23-62)."""

    def __init__(self, nelec, verbose=False):
        self.nup, self.ndown = nelec
        self.nelec = nelec
        self.ne = self.nup + self.ndown


def load_library():
    """Load libafqmc_hip.so (raises if it was not built)."""
    from ipie_amd import _lib

    return _lib.load()
