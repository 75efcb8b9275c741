"""ctypes binding of libafqmc_hip.so (the C ABI declared in include/afqmc_hip.h).

Loading fails loudly when the library is missing: there is no CPU fallback.
Tensors are passed as raw device pointers; every wrapper checks dtype, device,
contiguity and shape before the call.
"""

import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libafqmc_hip.so")

AFQMC_OK = 0
STATUS_NAMES = {
    1: "AFQMC_ERR_INVALID_ARG",
    2: "AFQMC_ERR_SHAPE",
    3: "AFQMC_ERR_UNSUPPORTED",
    4: "AFQMC_ERR_HIP",
    5: "AFQMC_ERR_ALLOC",
}
FLAG_SINGULAR = 1
FLAG_NONFINITE = 2
FLAG_CHOLESKY = 4
PREC_FP64 = 0
PREC_MIXED = 1
PRECISIONS = {"fp64": PREC_FP64, "mixed": PREC_MIXED}

# Every symbol the header declares, with (restype, argtypes).
_dp = ctypes.c_void_p
_i32 = ctypes.c_int32
_u64 = ctypes.c_uint64
_i64 = ctypes.c_int64
_f64 = ctypes.c_double


class Dims(ctypes.Structure):
    _fields_ = [
        ("nbasis", _i32),
        ("nchol", _i32),
        ("nalpha", _i32),
        ("nbeta", _i32),
        ("nwalkers", _i32),
        ("exp_nmax", _i32),
        ("rhf", _i32),
        ("precision", _i32),
    ]


class Static(ctypes.Structure):
    _fields_ = [
        ("chol_packed", _dp),
        ("rchola", _dp),
        ("rcholb", _dp),
        ("rH1a", _dp),
        ("rH1b", _dp),
        ("expH1", _dp),
        ("mf_shift", _dp),
        ("psia", _dp),
        ("psib", _dp),
        ("dt", _f64),
        ("ecore", _f64),
    ]


SIGNATURES = {
    "afqmc_create": (ctypes.c_int, [ctypes.POINTER(_dp), ctypes.c_int, ctypes.POINTER(Dims)]),
    "afqmc_destroy": (None, [_dp]),
    "afqmc_set_stream": (ctypes.c_int, [_dp, _dp]),
    "afqmc_set_static": (ctypes.c_int, [_dp, ctypes.POINTER(Static)]),
    "afqmc_last_error": (ctypes.c_char_p, [_dp]),
    "afqmc_sync": (ctypes.c_int, [_dp]),
    "afqmc_get_flags": (ctypes.c_int, [_dp, ctypes.POINTER(_i32), ctypes.c_int]),
    "afqmc_get_nfb": (ctypes.c_int, [_dp, ctypes.POINTER(_u64)]),
    "afqmc_walker_flags": (ctypes.c_int, [_dp, _i32, _dp, ctypes.c_int]),
    "afqmc_greens_function": (ctypes.c_int, [_dp, _i32, _dp, _dp, _dp, _dp, _dp, _dp]),
    "afqmc_overlap": (ctypes.c_int, [_dp, _i32, _dp, _dp, _dp, _dp]),
    "afqmc_force_bias": (ctypes.c_int, [_dp, _i32, _dp, _dp, _dp]),
    "afqmc_one_body": (ctypes.c_int, [_dp, _i32, _dp, _dp]),
    "afqmc_two_body": (ctypes.c_int, [_dp, _i32, _dp, _dp, _dp, _dp, _dp, _u64, _u64, _dp, _dp, _dp, _dp]),
    "afqmc_update_weight": (ctypes.c_int, [_dp, _i32, _dp, _dp, _dp, _dp, _f64, _dp, _dp, _dp]),
    "afqmc_propagate": (
        ctypes.c_int,
        [_dp, _i32, _dp, _dp, _dp, _dp, _dp, _u64, _u64, _f64, _dp, _dp, _dp, _dp, _dp],
    ),
    "afqmc_local_energy": (ctypes.c_int, [_dp, _i32, _dp, _dp, _dp]),
    "afqmc_reortho": (ctypes.c_int, [_dp, _i32, _dp, _dp, _dp, _dp, _dp, _dp]),
    "afqmc_draw_normals": (ctypes.c_int, [_dp, _dp, _i64, _u64, _u64]),
    "afqmc_profile": (ctypes.c_int, [_dp, ctypes.c_int]),
    "afqmc_surrogate_configure": (ctypes.c_int, [_dp, ctypes.c_int, _f64, ctypes.c_int, ctypes.c_int, _f64, _u64]),
    "afqmc_surrogate_stats": (ctypes.c_int, [_dp, ctypes.POINTER(_f64), ctypes.c_int]),
    "afqmc_profile_read": (ctypes.c_int, [_dp, ctypes.POINTER(_f64), ctypes.POINTER(_i64)]),
    "afqmc_tune_get": (ctypes.c_int, [_dp, _i32, ctypes.POINTER(_i32), ctypes.POINTER(_i32), ctypes.POINTER(_i32)]),
    "afqmc_tune_set": (ctypes.c_int, [_dp, _i32, _i32]),
    "afqmc_exchange_form": (ctypes.c_int, [_dp]),
    "afqmc_pop_plan": (ctypes.c_int, [_dp, _i32, _i32, _dp, _f64, _f64, _f64, _dp, _dp, _dp, _dp, _dp]),
    "afqmc_pop_gather": (ctypes.c_int, [_dp, _i32, _dp, _dp, _dp, _i64]),
    "afqmc_setup_pack": (ctypes.c_int, [_dp, _dp, _dp]),
    "afqmc_setup_h1e_mod": (ctypes.c_int, [_dp, _dp, _dp, _dp]),
    "afqmc_setup_half_rotate": (ctypes.c_int, [_dp, _i32, _dp, _i32, _dp, _dp, _dp, _dp]),
    "afqmc_setup_mf_shift": (ctypes.c_int, [_dp, _dp, _dp, _dp]),
    "afqmc_setup_expH1": (ctypes.c_int, [_dp, _dp, _dp, _dp, _f64, _dp]),
    "afqmc_walker_sums": (ctypes.c_int, [_dp, _i32, _dp, _dp, _dp, _i64, _dp]),
    "afqmc_clip_weights": (ctypes.c_int, [_dp, _i32, _dp, _f64]),
    "afqmc_graph_begin": (ctypes.c_int, [_dp]),
    "afqmc_graph_end": (ctypes.c_int, [_dp, ctypes.POINTER(_i32)]),
    "afqmc_graph_launch": (ctypes.c_int, [_dp, _i32, _u64, _f64]),
    "afqmc_batch_create": (ctypes.c_int, [ctypes.POINTER(_dp), ctypes.POINTER(_dp), _i32]),
    "afqmc_batch_destroy": (None, [_dp]),
    "afqmc_batch_begin": (ctypes.c_int, [_dp]),
    "afqmc_batch_end": (ctypes.c_int, [_dp]),
    "afqmc_batch_steps": (ctypes.c_int, [_dp, ctypes.POINTER(_i32)]),
    "afqmc_batch_launch": (ctypes.c_int, [_dp, _i32, ctypes.POINTER(_u64), ctypes.POINTER(_f64)]),
    "afqmc_control_variate": (ctypes.c_int, [_dp, _i32, _dp, _dp, _dp]),
    "afqmc_surrogate_quadratic": (ctypes.c_int, [_dp, ctypes.c_int]),
    "afqmc_noalloc_scratch_bytes": (ctypes.c_size_t, [_i32, _i32, _i32]),
    "afqmc_noalloc_last_error": (ctypes.c_char_p, []),
    "afqmc_greens_function_noalloc": (
        ctypes.c_int,
        [_i32, _i32, _i32, _i32, _i32, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp],
    ),
}

POP_METHODS = {"pair_branch": 0, "comb": 1, "stochastic_reconfiguration": 2, "rescale": 3}

# launch-shape knobs of afqmc_tune_get / afqmc_tune_set
TUNE_KNOBS = {"fb_split": 1, "quad_wgs": 2, "exx_split": 3, "tile_fb": 4, "tile_vhs": 5, "tile_one_body": 6,
              "quad_rgroup": 7, "taylor": 8}

PHASES = ["gf", "one_body", "force_bias", "fields", "vhs", "taylor", "overlap", "weight", "exchange", "energy",
          "reortho", "exchange_cv", "split"]

_lib = None


def load():
    """Load libafqmc_hip.so once; raise if it is absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m ipie_amd.build` (hipcc, gfx950). "
                "There is no CPU fallback."
            )
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


class AfqmcError(RuntimeError):
    pass


def ptr(t):
    """Device pointer of a tensor (None -> NULL)."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def _check_tensor(name, t, dtype, shape, device):
    if t is None:
        raise ValueError(f"{name} is required")
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch tensor, got {type(t)}")
    if t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype}, got {t.dtype}")
    if t.device != device:
        raise ValueError(f"{name} must live on {device}, got {t.device}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} has shape {tuple(t.shape)}, expected {tuple(shape)}")
    return t


class Context:
    """Owns one afqmc_ctx (dims, workspace, borrowed static tensors) on one device."""

    def __init__(self, nbasis, nchol, nalpha, nbeta, nwalkers, exp_nmax=6, rhf=False, device=None, precision="fp64"):
        self.lib = load()
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(PRECISIONS)}")
        self.precision = precision
        self.dims = Dims(nbasis, nchol, nalpha, nbeta, nwalkers, exp_nmax, int(bool(rhf)), PRECISIONS[precision])
        h = _dp()
        with torch.cuda.device(self.device):
            st = self.lib.afqmc_create(ctypes.byref(h), self.device.index or 0, ctypes.byref(self.dims))
        if st != AFQMC_OK:
            raise AfqmcError(f"afqmc_create failed: {STATUS_NAMES.get(st, st)} (dims={self.describe()})")
        self.h = h
        self._static_refs = None

    def describe(self):
        d = self.dims
        return f"N={d.nbasis} X={d.nchol} na={d.nalpha} nb={d.nbeta} W={d.nwalkers} rhf={d.rhf}"

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            try:
                self.lib.afqmc_destroy(h)
            except Exception:
                pass
            self.h = None

    # ------------------------------------------------------------------ plumbing
    def _call(self, name, *args):
        stream = torch.cuda.current_stream(self.device).cuda_stream
        self.lib.afqmc_set_stream(self.h, ctypes.c_void_p(stream))
        st = getattr(self.lib, name)(self.h, *args)
        if st != AFQMC_OK:
            msg = self.lib.afqmc_last_error(self.h)
            raise AfqmcError(f"{name}: {STATUS_NAMES.get(st, st)}: {msg.decode() if msg else ''}")

    def set_static(self, *, chol_packed, rchola, rcholb, rH1a, rH1b, expH1, mf_shift, psia, psib, dt, ecore):
        d = self.dims
        dev = self.device
        f64, c128 = torch.float64, torch.complex128
        N, X, na, nb = d.nbasis, d.nchol, d.nalpha, d.nbeta
        _check_tensor("chol_packed", chol_packed, f64, (N * (N + 1) // 2, X), dev)
        _check_tensor("rchola", rchola, f64, (X, na * N), dev)
        _check_tensor("rcholb", rcholb, f64, (X, nb * N), dev)
        _check_tensor("rH1a", rH1a, c128, (na, N), dev)
        _check_tensor("rH1b", rH1b, c128, (nb, N), dev)
        _check_tensor("expH1", expH1, f64, (2, N, N), dev)
        _check_tensor("mf_shift", mf_shift, c128, (X,), dev)
        _check_tensor("psia", psia, c128, (N, na), dev)
        _check_tensor("psib", psib, c128, (N, nb), dev)
        s = Static(
            ptr(chol_packed), ptr(rchola), ptr(rcholb), ptr(rH1a), ptr(rH1b), ptr(expH1), ptr(mf_shift),
            ptr(psia), ptr(psib), float(dt), float(ecore),
        )
        # on the current stream: the context reads the trial orbitals (real or complex) after their upload
        self._call("afqmc_set_static", ctypes.byref(s))
        # the context borrows these: keep them alive
        self._static_refs = (chol_packed, rchola, rcholb, rH1a, rH1b, expH1, mf_shift, psia, psib)

    def _walker_shapes(self, nw):
        d = self.dims
        return (nw, d.nbasis, d.nalpha), (nw, d.nbasis, d.nbeta), (nw, d.nalpha, d.nbasis), (nw, d.nbeta, d.nbasis)

    def _chk_phi(self, phia, phib):
        nw = phia.shape[0]
        sa, sb, _, _ = self._walker_shapes(nw)
        _check_tensor("phia", phia, torch.complex128, sa, self.device)
        _check_tensor("phib", phib, torch.complex128, sb, self.device)
        return nw

    def _chk_ghalf(self, ga, gb, nw):
        _, _, ga_s, gb_s = self._walker_shapes(nw)
        _check_tensor("Ghalfa", ga, torch.complex128, ga_s, self.device)
        _check_tensor("Ghalfb", gb, torch.complex128, gb_s, self.device)

    def _chk_vec(self, name, t, nw, dtype):
        return _check_tensor(name, t, dtype, (nw,), self.device)

    # ------------------------------------------------------------------ entry points
    def greens_function(self, phia, phib, ghalfa, ghalfb, ovlp, log_shift=None):
        nw = self._chk_phi(phia, phib)
        self._chk_ghalf(ghalfa, ghalfb, nw)
        self._chk_vec("ovlp", ovlp, nw, torch.complex128)
        if log_shift is not None:
            self._chk_vec("log_shift", log_shift, nw, torch.float64)
        self._call("afqmc_greens_function", nw, ptr(phia), ptr(phib), ptr(ghalfa), ptr(ghalfb), ptr(ovlp),
                   ptr(log_shift))

    def overlap(self, phia, phib, ovlp, log_shift=None):
        nw = self._chk_phi(phia, phib)
        self._chk_vec("ovlp", ovlp, nw, torch.complex128)
        if log_shift is not None:
            self._chk_vec("log_shift", log_shift, nw, torch.float64)
        self._call("afqmc_overlap", nw, ptr(phia), ptr(phib), ptr(ovlp), ptr(log_shift))

    def force_bias(self, ghalfa, ghalfb, vbias):
        nw = ghalfa.shape[0]
        self._chk_ghalf(ghalfa, ghalfb, nw)
        _check_tensor("vbias", vbias, torch.complex128, (nw, self.dims.nchol), self.device)
        self._call("afqmc_force_bias", nw, ptr(ghalfa), ptr(ghalfb), ptr(vbias))

    def one_body(self, phia, phib):
        nw = self._chk_phi(phia, phib)
        self._call("afqmc_one_body", nw, ptr(phia), ptr(phib))

    def two_body(self, phia, phib, ghalfa, ghalfb, cmf, cfb, xi=None, seed=0, counter=0, xi_out=None,
                 vbias_out=None):
        nw = self._chk_phi(phia, phib)
        self._chk_ghalf(ghalfa, ghalfb, nw)
        self._chk_vec("cmf", cmf, nw, torch.complex128)
        self._chk_vec("cfb", cfb, nw, torch.complex128)
        X = self.dims.nchol
        if xi is not None:
            _check_tensor("xi", xi, torch.float64, (nw, X), self.device)
        if xi_out is not None:
            _check_tensor("xi_out", xi_out, torch.float64, (nw, X), self.device)
        if vbias_out is not None:
            _check_tensor("vbias_out", vbias_out, torch.complex128, (nw, X), self.device)
        self._call("afqmc_two_body", nw, ptr(phia), ptr(phib), ptr(ghalfa), ptr(ghalfb), ptr(xi), seed, counter,
                   ptr(cmf), ptr(cfb), ptr(xi_out), ptr(vbias_out))

    def update_weight(self, ovlp, ovlp_new, cfb, cmf, eshift, weight, hybrid_energy, walker_ovlp):
        nw = ovlp.shape[0]
        for nm, t in (("ovlp", ovlp), ("ovlp_new", ovlp_new), ("cfb", cfb), ("cmf", cmf),
                      ("hybrid_energy", hybrid_energy), ("walker_ovlp", walker_ovlp)):
            self._chk_vec(nm, t, nw, torch.complex128)
        self._chk_vec("weight", weight, nw, torch.float64)
        self._call("afqmc_update_weight", nw, ptr(ovlp), ptr(ovlp_new), ptr(cfb), ptr(cmf), float(eshift),
                   ptr(weight), ptr(hybrid_energy), ptr(walker_ovlp))

    def propagate(self, phia, phib, ghalfa, ghalfb, weight, ovlp, hybrid_energy, eshift, xi=None, seed=0,
                  counter=0, log_shift=None, energy_prev=None):
        nw = self._chk_phi(phia, phib)
        self._chk_ghalf(ghalfa, ghalfb, nw)
        self._chk_vec("weight", weight, nw, torch.float64)
        self._chk_vec("ovlp", ovlp, nw, torch.complex128)
        self._chk_vec("hybrid_energy", hybrid_energy, nw, torch.complex128)
        if xi is not None:
            _check_tensor("xi", xi, torch.float64, (nw, self.dims.nchol), self.device)
        if log_shift is not None:
            self._chk_vec("log_shift", log_shift, nw, torch.float64)
        if energy_prev is not None:
            _check_tensor("energy_prev", energy_prev, torch.complex128, (nw, 3), self.device)
        self._call("afqmc_propagate", nw, ptr(phia), ptr(phib), ptr(ghalfa), ptr(ghalfb), ptr(xi), seed, counter,
                   float(eshift), ptr(weight), ptr(ovlp), ptr(hybrid_energy), ptr(log_shift), ptr(energy_prev))

    def local_energy(self, ghalfa, ghalfb, energy):
        nw = ghalfa.shape[0]
        self._chk_ghalf(ghalfa, ghalfb, nw)
        _check_tensor("energy", energy, torch.complex128, (nw, 3), self.device)
        self._call("afqmc_local_energy", nw, ptr(ghalfa), ptr(ghalfb), ptr(energy))

    def control_variate(self, ghalfa, ghalfb, cv):
        """cv [W] complex: the physical control variate of each walker's exchange
        sum (afqmc_control_variate); needs the control variate configured."""
        nw = ghalfa.shape[0]
        self._chk_ghalf(ghalfa, ghalfb, nw)
        self._chk_vec("cv", cv, nw, torch.complex128)
        self._call("afqmc_control_variate", nw, ptr(ghalfa), ptr(ghalfb), ptr(cv))

    def reortho(self, phia, phib, ovlp, detR, log_detR, detR_shift=None):
        nw = self._chk_phi(phia, phib)
        self._chk_vec("ovlp", ovlp, nw, torch.complex128)
        self._chk_vec("detR", detR, nw, torch.float64)
        self._chk_vec("log_detR", log_detR, nw, torch.float64)
        if detR_shift is not None:
            self._chk_vec("detR_shift", detR_shift, nw, torch.float64)
        self._call("afqmc_reortho", nw, ptr(phia), ptr(phib), ptr(ovlp), ptr(detR), ptr(log_detR),
                   ptr(detR_shift))

    # ------------------------------------------------------------ set-up producers
    def setup_pack(self, chol, chol_packed):
        d, dev = self.dims, self.device
        N, X = d.nbasis, d.nchol
        _check_tensor("chol", chol, torch.float64, (N * N, X), dev)
        _check_tensor("chol_packed", chol_packed, torch.float64, (N * (N + 1) // 2, X), dev)
        self._call("afqmc_setup_pack", ptr(chol), ptr(chol_packed))

    def setup_h1e_mod(self, chol, h1e, h1e_mod):
        d, dev = self.dims, self.device
        N, X = d.nbasis, d.nchol
        _check_tensor("chol", chol, torch.float64, (N * N, X), dev)
        _check_tensor("h1e", h1e, torch.float64, (2, N, N), dev)
        _check_tensor("h1e_mod", h1e_mod, torch.float64, (2, N, N), dev)
        self._call("afqmc_setup_h1e_mod", ptr(chol), ptr(h1e), ptr(h1e_mod))

    def setup_half_rotate(self, spin, psi, chol, h1, rchol, rh1):
        d, dev = self.dims, self.device
        N, X = d.nbasis, d.nchol
        n = d.nalpha if spin == 0 else d.nbeta
        _check_tensor("psi", psi, torch.float64, (N, n), dev)
        _check_tensor("chol", chol, torch.float64, (N * N, X), dev)
        _check_tensor("h1", h1, torch.float64, (N, N), dev)
        _check_tensor("rchol", rchol, torch.float64, (X, n * N), dev)
        _check_tensor("rh1", rh1, torch.complex128, (n, N), dev)
        self._call("afqmc_setup_half_rotate", int(spin), ptr(psi), n, ptr(chol), ptr(h1), ptr(rchol), ptr(rh1))

    def setup_mf_shift(self, chol, gcharge, mf):
        d, dev = self.dims, self.device
        N, X = d.nbasis, d.nchol
        _check_tensor("chol", chol, torch.float64, (N * N, X), dev)
        _check_tensor("gcharge", gcharge, torch.complex128, (N * N,), dev)
        _check_tensor("mf", mf, torch.complex128, (X,), dev)
        self._call("afqmc_setup_mf_shift", ptr(chol), ptr(gcharge), ptr(mf))

    def setup_expH1(self, chol, h1e_mod, mf, dt, expH1):
        d, dev = self.dims, self.device
        N, X = d.nbasis, d.nchol
        _check_tensor("chol", chol, torch.float64, (N * N, X), dev)
        _check_tensor("h1e_mod", h1e_mod, torch.float64, (2, N, N), dev)
        _check_tensor("mf", mf, torch.complex128, (X,), dev)
        _check_tensor("expH1", expH1, torch.float64, (2, N, N), dev)
        self._call("afqmc_setup_expH1", ptr(chol), ptr(h1e_mod), ptr(mf), float(dt), ptr(expH1))

    def walker_sums(self, weight, acc, unscaled_weight=None, values=None, column=0):
        """acc (complex128 [3]) += (sum weight, sum unscaled_weight, sum weight * values)
        (afqmc_walker_sums).  values: complex128 [W], or [W, k] read at `column`."""
        nw = weight.shape[0]
        self._chk_vec("weight", weight, nw, torch.float64)
        _check_tensor("acc", acc, torch.complex128, (3,), self.device)
        if unscaled_weight is not None:
            self._chk_vec("unscaled_weight", unscaled_weight, nw, torch.float64)
        stride = 1
        if values is not None:
            if values.dtype != torch.complex128 or values.device != self.device or values.shape[0] != nw:
                raise ValueError("values must be complex128 on the context's device, walkers first")
            if values.dim() == 2:
                if not values.is_contiguous():
                    raise ValueError("values must be contiguous")
                stride = values.shape[1]
                if not 0 <= column < stride:
                    raise ValueError(f"column {column} outside [0, {stride})")
            elif values.dim() != 1 or not values.is_contiguous() or column != 0:
                raise ValueError("values must be [W] or [W, k] contiguous")
        vptr = None if values is None else ctypes.c_void_p(values.data_ptr() + 16 * column)
        self._call("afqmc_walker_sums", nw, ptr(weight), ptr(unscaled_weight), vptr, stride, ptr(acc))

    def clip_weights(self, weight, bound):
        nw = weight.shape[0]
        self._chk_vec("weight", weight, nw, torch.float64)
        self._call("afqmc_clip_weights", nw, ptr(weight), float(bound))

    def bind_stream(self, stream):
        """Set the context's stream (a torch.cuda.Stream) without making a call."""
        self.lib.afqmc_set_stream(self.h, ctypes.c_void_p(stream.cuda_stream))

    def graph_begin(self):
        """Start recording this context's calls into a step graph (afqmc_graph_begin);
        torch's current stream must not be the null stream."""
        self._call("afqmc_graph_begin")

    def graph_end(self):
        """Close the recording and instantiate it; returns the recorded step count."""
        n = _i32(0)
        self._call("afqmc_graph_end", ctypes.byref(n))
        return int(n.value)

    def graph_launch(self, replays, counter, eshift):
        """Replay the step graph `replays` times; the first replayed step draws the
        field stream `counter`."""
        self._call("afqmc_graph_launch", int(replays), int(counter), float(eshift))

    def draw_normals(self, out, seed, counter):
        _check_tensor("out", out, torch.float64, None, self.device)
        self._call("afqmc_draw_normals", ptr(out), out.numel(), seed, counter)

    def flags(self, clear=True):
        f = _i32(0)
        st = self.lib.afqmc_get_flags(self.h, ctypes.byref(f), int(clear))
        if st != AFQMC_OK:
            raise AfqmcError(f"afqmc_get_flags: {STATUS_NAMES.get(st, st)}")
        return int(f.value)

    def walker_flags(self, nwalkers, out=None, clear=True):
        """Per-walker flags (int32 [nwalkers], OR of FLAG_* since the last take):
        a singular overlap matrix, a non-finite weight, a rank-deficient walker
        in reortho.  Enqueued on the current stream; no synchronisation."""
        if out is None:
            out = torch.empty(nwalkers, dtype=torch.int32, device=self.device)
        _check_tensor("out", out, torch.int32, (nwalkers,), self.device)
        self._call("afqmc_walker_flags", int(nwalkers), ptr(out), int(bool(clear)))
        return out

    def nfb(self):
        v = _u64(0)
        st = self.lib.afqmc_get_nfb(self.h, ctypes.byref(v))
        if st != AFQMC_OK:
            raise AfqmcError(f"afqmc_get_nfb: {STATUS_NAMES.get(st, st)}")
        return int(v.value)

    def surrogate_configure(self, enable=True, fraction=0.125, warmup=100, train_iters=4, lr=1e-3, seed=11):
        self._call("afqmc_surrogate_configure", int(bool(enable)), float(fraction), int(warmup), int(train_iters),
                   float(lr), int(seed))

    def surrogate_quadratic(self, mode="fp8"):
        """The control variate's quadratic term Q(G - G_T) (afqmc_surrogate_quadratic):
        "fp8" (or True: the default), "bf16", or "none" (or False)."""
        modes = {"none": 0, "bf16": 1, "fp8": 2, False: 0, True: 2}
        if mode not in modes:
            raise ValueError(f"quadratic mode {mode!r}: one of 'fp8', 'bf16', 'none'")
        self._call("afqmc_surrogate_quadratic", modes[mode])

    def exchange_form(self):
        """'quad' (g^T Wq g, Wq built once) or 't' (every T_x = L_x Ghalf^T, ipie's contraction)."""
        return "quad" if self.lib.afqmc_exchange_form(self.h) == 1 else "t"

    def surrogate_stats(self):
        out = (_f64 * 7)()
        st = self.lib.afqmc_surrogate_stats(self.h, out, 7)
        if st != AFQMC_OK:
            raise AfqmcError(f"afqmc_surrogate_stats: {self.lib.afqmc_last_error(self.h)}")
        keys = ["steps", "warmup", "last_k", "train_loss", "resid2", "target_var", "gain"]
        return dict(zip(keys, [float(v) for v in out]))

    def tune_get(self, knob):
        """(value, lo, hi) of a launch-shape knob; hi == 0 when it does not apply."""
        v, lo, hi = _i32(), _i32(), _i32()
        st = self.lib.afqmc_tune_get(self.h, TUNE_KNOBS[knob], ctypes.byref(v), ctypes.byref(lo), ctypes.byref(hi))
        if st != AFQMC_OK:
            raise AfqmcError(f"afqmc_tune_get: {STATUS_NAMES.get(st, st)}: {self.lib.afqmc_last_error(self.h)}")
        return v.value, lo.value, hi.value

    def tune_set(self, knob, value):
        with torch.cuda.device(self.device):
            st = self.lib.afqmc_tune_set(self.h, TUNE_KNOBS[knob], int(value))
        if st != AFQMC_OK:
            msg = self.lib.afqmc_last_error(self.h)
            raise AfqmcError(f"afqmc_tune_set: {STATUS_NAMES.get(st, st)}: {msg.decode() if msg else ''}")

    def profile(self, enable=True):
        st = self.lib.afqmc_profile(self.h, int(bool(enable)))
        if st != AFQMC_OK:
            raise AfqmcError(f"afqmc_profile: {STATUS_NAMES.get(st, st)}")

    def profile_read(self):
        """{phase: (milliseconds, launches)} accumulated since profile(True) (synchronises)."""
        n = len(PHASES)
        ms = (_f64 * n)()
        cnt = (_i64 * n)()
        st = self.lib.afqmc_profile_read(self.h, ms, cnt)
        if st != AFQMC_OK:
            raise AfqmcError(f"afqmc_profile_read: {self.lib.afqmc_last_error(self.h)}")
        return {PHASES[i]: (float(ms[i]), int(cnt[i])) for i in range(n)}

    def pop_plan(self, method, weight, target_weight, min_weight, max_weight, uniforms, src, new_weight,
                 total_weight, npairs):
        """Population-control decisions for the walkers (afqmc_pop_plan): rescaling
        to target_weight, then src (int32 [W]: slot i takes walker src[i]) and
        new_weight (float64 [W]); total_weight (float64 [1]) and npairs (int32
        [1]) on the device."""
        nw = weight.shape[0]
        dev = self.device
        _check_tensor("weight", weight, torch.float64, (nw,), dev)
        _check_tensor("src", src, torch.int32, (nw,), dev)
        _check_tensor("new_weight", new_weight, torch.float64, (nw,), dev)
        _check_tensor("total_weight", total_weight, torch.float64, (1,), dev)
        _check_tensor("npairs", npairs, torch.int32, (1,), dev)
        if uniforms is not None:
            _check_tensor("uniforms", uniforms, torch.float64, (uniforms.shape[0],), dev)
        self._call("afqmc_pop_plan", POP_METHODS[method], nw, ptr(weight), float(target_weight), float(min_weight),
                   float(max_weight), ptr(uniforms), ptr(src), ptr(new_weight), ptr(total_weight), ptr(npairs))

    def pop_gather(self, src, inp, out):
        """out[i] = inp[src[i]] along the walker (first) axis (afqmc_pop_gather)."""
        nw = src.shape[0]
        if inp.shape != out.shape or inp.dtype != out.dtype or inp.shape[0] != nw:
            raise ValueError("pop_gather: input and output must have the same shape, walkers first")
        for name, t in (("src", src), ("input", inp), ("output", out)):
            if not t.is_contiguous() or t.device != self.device:
                raise ValueError(f"{name} must be contiguous on {self.device}")
        self._call("afqmc_pop_gather", nw, ptr(src), ptr(inp), ptr(out), inp.element_size() * inp[0].numel())

    def sync(self):
        st = self.lib.afqmc_sync(self.h)
        if st != AFQMC_OK:
            raise AfqmcError(f"afqmc_sync: {self.lib.afqmc_last_error(self.h)}")
