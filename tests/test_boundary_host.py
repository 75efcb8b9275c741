"""Host-side parts of the drop-in boundary (no GPU): the binding refuses
strided views (the C ABI takes dense arrays only), the PropagatorTimer fields
follow the device phase times, and a multi-system capture that fails midway
is closed before the error propagates."""

import ctypes

import pytest
import torch

from ipie_amd import _lib
from ipie_amd.propagator import PropagatorTimer


def test_non_contiguous_view_is_refused():
    t = torch.zeros(4, 6, 3, dtype=torch.complex128)
    cpu = torch.device("cpu")
    _lib._check_tensor("phia", t, torch.complex128, (4, 6, 3), cpu)
    with pytest.raises(ValueError, match="contiguous"):
        _lib._check_tensor("phia", t.transpose(1, 2), torch.complex128, (4, 3, 6), cpu)
    with pytest.raises(ValueError, match="contiguous"):
        _lib._check_tensor("weight", torch.zeros(8, dtype=torch.float64)[::2], torch.float64, (4,), cpu)
    with pytest.raises(TypeError):
        _lib._check_tensor("weight", torch.zeros(4, dtype=torch.float32), torch.float64, (4,), cpu)


class _FakeCtx:
    """Accumulated phase times as afqmc_profile_read reports them (ms, launches)."""

    def __init__(self):
        self.ms = dict.fromkeys(_lib.PHASES, 0.0)
        self.enabled = False

    def profile(self, enable=True):
        self.enabled = bool(enable)
        self.ms = dict.fromkeys(_lib.PHASES, 0.0)

    def profile_read(self):
        return {k: (v, 1) for k, v in self.ms.items()}


def test_timer_fields_follow_device_phases():
    timer = PropagatorTimer()
    assert timer.tgf == 0.0 and timer.tgemm == 0.0  # unbound: zeros, as ipie's fresh timer
    ctx = _FakeCtx()
    timer.bind(ctx)
    assert ctx.enabled
    ctx.ms.update(gf=2.0, one_body=1.0, taylor=3.0, force_bias=0.5, fields=0.25, vhs=4.0, overlap=0.75, weight=0.1,
                  exchange=9.0, energy=1.0)
    assert timer.tgf == pytest.approx(2e-3)
    assert timer.tgemm == pytest.approx(4e-3)  # both one-body halves and the Taylor series
    assert timer.tfbias == pytest.approx(0.75e-3)
    assert timer.tvhs == pytest.approx(4e-3)
    assert timer.tovlp == pytest.approx(0.75e-3)
    assert timer.tupdate == pytest.approx(0.1e-3)
    # the local energy is not a propagation phase: no field carries it
    total = sum(getattr(timer, f) for f in PropagatorTimer.PHASES)
    assert total == pytest.approx(11.6e-3)
    timer.tgf = 0.0  # reset, as a driver may do
    assert timer.tgf == 0.0
    ctx.ms["gf"] += 1.5
    assert timer.tgf == pytest.approx(1.5e-3)
    timer.tgemm += 1.0  # a host-side addition is kept beside the device time
    ctx.ms["taylor"] += 2.0
    assert timer.tgemm == pytest.approx(4e-3 + 1.0 + 2e-3)


class _StubLib:
    def __init__(self, end_status=0):
        self.calls = []
        self.end_status = end_status

    def afqmc_batch_begin(self, h):
        self.calls.append("begin")
        return 0

    def afqmc_batch_end(self, h):
        self.calls.append("end")
        return self.end_status

    def afqmc_batch_steps(self, h, steps):
        self.calls.append("steps")
        return 0

    def afqmc_last_error(self, h):
        return b"multi-system capture failed"


class _StubCtx:
    def __init__(self, fail):
        self.fail = fail
        self.h = ctypes.c_void_p(1)

    def bind_stream(self, s):
        pass

    def propagate(self, *a, **k):
        if self.fail:
            raise _lib.AfqmcError("afqmc_propagate: AFQMC_ERR_HIP: launch failed")


class _StubStream:
    def wait_stream(self, other):
        pass


class _Walkers:
    phia = phib = Ghalfa = Ghalfb = weight = ovlp = hybrid_energy = log_shift = None


class _Prop:
    xi_source = None
    seed = 7
    counter = 0


def _stub_batch(monkeypatch, fail_index, end_status=0):
    from ipie_amd import batch as batch_mod

    monkeypatch.setattr(torch.cuda, "current_stream", lambda device=None: _StubStream())

    class _NoStream:
        def __init__(self, s):
            pass

        def __enter__(self):
            return self

        def __exit__(self, *a):
            return False

    monkeypatch.setattr(torch.cuda, "stream", _NoStream)
    b = object.__new__(batch_mod.MultiSystemBatch)
    b.systems = [(_Prop(), _Walkers(), None, None) for _ in range(3)]
    b.ctx = [_StubCtx(i == fail_index) for i in range(3)]
    b.device = torch.device("cpu")
    b.lib = _StubLib(end_status)
    b.h = ctypes.c_void_p(0)
    b.streams = [_StubStream() for _ in range(3)]
    b.steps = [1, 1, 1]
    return b


def test_failed_multi_system_capture_is_closed(monkeypatch):
    b = _stub_batch(monkeypatch, fail_index=1)
    with pytest.raises(_lib.AfqmcError, match="launch failed"):
        b.capture(2)
    assert b.lib.calls == ["begin", "end"]  # the capture is ended exactly once, then the error propagates
    assert b.steps is None
    with pytest.raises(RuntimeError, match="capture"):
        b.replay([0.0, 0.0, 0.0])


def test_capture_end_error_is_raised(monkeypatch):
    b = _stub_batch(monkeypatch, fail_index=-1, end_status=4)
    with pytest.raises(_lib.AfqmcError, match="afqmc_batch_end: AFQMC_ERR_HIP: multi-system capture failed"):
        b.capture(1)
    assert b.steps is None
