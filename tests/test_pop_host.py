"""Host-side checks of the PopController drop-in (no GPU needed)."""

import pytest


class _Handler:
    def __init__(self, size):
        self.size = size


def test_constructor_mirrors_reference():
    from ipie_amd.pop_controller import PopController

    ctl = PopController(32, 10)
    assert ctl.method == "pair_branch"
    assert (ctl.min_weight, ctl.max_weight, ctl.reconfiguration_freq) == (0.1, 4, 50)
    assert ctl.target_weight == ctl.total_weight == ctl.ntot_walkers == 32
    for attr in ("communication_time", "non_communication_time", "send_time", "recv_time"):
        assert getattr(ctl.timer, attr) == 0.0


def test_constructor_rejects_what_it_does_not_cover():
    from ipie_amd.pop_controller import PopController

    with pytest.raises(ValueError):
        PopController(8, 10, pop_control_method="branch_and_bound")
    with pytest.raises(NotImplementedError):
        PopController(8, 10, mpi_handler=_Handler(2))
    ctl = PopController(8, 10, mpi_handler=_Handler(1))
    with pytest.raises(NotImplementedError):
        ctl.pop_control(object(), _Handler(4))


def test_walker_accumulator_host_logic():
    import numpy

    from ipie_amd.walkers import WalkerAccumulator

    acc = WalkerAccumulator(["Weight", "WeightFactor", "HybridEnergy"], 10)
    assert acc.get_index("HybridEnergy") == 2
    with pytest.raises(RuntimeError):
        acc.get_index("Nope")
    vals = numpy.array([20.0, 30.0, -50.0 + 1j])
    acc.post_reduce_hook(vals, 3)
    assert vals[2] == (-50.0 + 1j) / 20.0 and vals[0] == 2.0 and vals[1] == 3.0
    vals = numpy.array([20.0, 30.0, -50.0])
    acc.post_reduce_hook(vals, 0)  # block 0 is not divided by the block length
    assert vals[0] == 20.0 and vals[2] == -2.5
    assert acc.to_text([1.5, -2.0]) == " 1.5000000000000000e+00 -2.0000000000000000e+00"
    acc.eshift = -1.25 + 0j
    assert acc.eshift == -1.25
