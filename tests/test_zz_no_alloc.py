"""afqmc_greens_function_noalloc (static scratch, no context, no allocation)
against afqmc_greens_function through a context, called through the C ABI.

Two calls in a row share the GPU's static scratch array; the second has fewer
walkers and a different spin split, so a stale entry from the first call would
show up in the second's results.  The oversize case is refused on the host,
before anything is launched."""

import ctypes

import numpy
import pytest

from tests._helpers import Pair, random_walkers, rel_err, rel_err_elem, synthetic_arrays

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _call(lib, nw, N, na, nb, rhf, psia, psib, phia, phib, ga, gb, ovlp, flags=None, log_shift=None):
    import torch

    from ipie_amd._lib import ptr

    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    return lib.afqmc_greens_function_noalloc(nw, N, na, nb, int(rhf), ptr(psia), ptr(psib), ptr(phia), ptr(phib),
                                             ptr(ga), ptr(gb), ptr(ovlp), ptr(log_shift), ptr(flags), stream)


def _case(N, X, na, nb, W, seed):
    h1e, chol, _, psi = synthetic_arrays(N, X, 12.0, na, nb, scale=(12.0 / (N * 12.0)) ** 0.5)
    init = random_walkers(N, na, nb, W, seed=seed, amp=0.3, base=psi)
    return Pair(h1e, chol, psi, (na, nb), W, init=init)


def _check(p, nw):
    """Context entry point vs the no-allocation one on the first nw walkers."""
    import torch

    from ipie_amd import _lib
    from ipie_amd import engine as _engine

    lib = _lib.load()
    dev = p.walkers.device
    N, (na, nb) = p.N, p.nelec
    w = p.walkers
    phia, phib = w.phia[:nw].contiguous(), w.phib[:nw].contiguous()
    log_shift = torch.linspace(-0.5, 0.5, nw, dtype=torch.float64, device=dev)
    # reference: the context's Green's function over the same walkers
    ga_ref = torch.empty(nw, na, N, dtype=torch.complex128, device=dev)
    gb_ref = torch.empty(nw, nb, N, dtype=torch.complex128, device=dev)
    ov_ref = torch.empty(nw, dtype=torch.complex128, device=dev)
    eng = _engine.engine_for(p.trial, p.walkers)
    eng.ctx.greens_function(phia, phib, ga_ref, gb_ref, ov_ref, log_shift)
    ov_only_ref = torch.empty(nw, dtype=torch.complex128, device=dev)
    eng.ctx.overlap(phia, phib, ov_only_ref, log_shift)
    # no-allocation entry point: outputs start as NaN so unwritten entries show
    t = p.trial.device_tensors(dev)
    ga = torch.full((nw, na, N), float("nan"), dtype=torch.complex128, device=dev)
    gb = torch.full((nw, nb, N), float("nan"), dtype=torch.complex128, device=dev)
    ov = torch.full((nw,), float("nan"), dtype=torch.complex128, device=dev)
    ov_only = torch.full((nw,), float("nan"), dtype=torch.complex128, device=dev)
    flags = torch.zeros(nw, dtype=torch.int32, device=dev)
    assert lib.afqmc_noalloc_scratch_bytes(nw, na, nb) <= 134217728
    st = _call(lib, nw, N, na, nb, False, t["psia"], t["psib"], phia, phib, ga, gb, ov, flags, log_shift)
    assert st == 0, lib.afqmc_noalloc_last_error()
    st = _call(lib, nw, N, na, nb, False, t["psia"], t["psib"], phia, phib, None, None, ov_only, None, log_shift)
    assert st == 0, lib.afqmc_noalloc_last_error()
    torch.cuda.synchronize()
    print(f"[noalloc] N={N} na={na} nb={nb} nw={nw}", flush=True)
    assert rel_err_elem(ov.cpu().numpy(), ov_ref.cpu().numpy()) < TOL
    assert rel_err_elem(ov_only.cpu().numpy(), ov_only_ref.cpu().numpy()) < TOL
    assert rel_err(ga.cpu().numpy(), ga_ref.cpu().numpy()) < TOL
    assert rel_err(gb.cpu().numpy(), gb_ref.cpu().numpy()) < TOL
    assert int(flags.abs().sum().item()) == 0


def test_noalloc_gf_two_calls_second_smaller():
    p1 = _case(56, 112, 9, 7, 40, seed=5)
    _check(p1, 40)
    p2 = _case(40, 80, 5, 4, 24, seed=6)
    _check(p2, 11)


def test_noalloc_gf_refuses_oversize_on_host():
    import torch

    from ipie_amd import _lib

    lib = _lib.load()
    dev = torch.device("cuda", torch.cuda.current_device())
    x = torch.zeros(4, dtype=torch.complex128, device=dev)
    nw = 1 << 20  # 2 * nw * 2 * 64^2 * 16 bytes >> 128 MiB
    assert lib.afqmc_noalloc_scratch_bytes(nw, 64, 64) > 134217728
    st = _call(lib, nw, 128, 64, 64, False, x, x, x, x, x, x, x)
    assert st == 2  # AFQMC_ERR_SHAPE
    assert b"exceeds the static" in lib.afqmc_noalloc_last_error()
    numpy.testing.assert_array_equal(x.cpu().numpy(), numpy.zeros(4, dtype=numpy.complex128))
