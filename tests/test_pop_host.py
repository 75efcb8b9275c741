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
