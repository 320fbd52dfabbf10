"""Loader and ctypes declarations of libowl_amd.so (include/owl_amd.h).

The library is loaded from this package directory only.  There is no
fallback: if the shared library is missing, every entry point raises.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libowl_amd.so")

OWL_OK = 0
OWL_ERR_ARG = -1
OWL_ERR_CFL = -2
OWL_ERR_DEVICE = -3
OWL_ERR_NOMEM = -4
OWL_ERR_STATE = -5
OWL_ERR_NONFINITE = -6

OWL_EXPLOSION, OWL_FORCE = 0, 1
OWL_RICKER, OWL_GAUSSIAN, OWL_GAUSSIAN_DERIV, OWL_RICKER_DERIV = 0, 1, 2, 3
OWL_GRAD_OWL, OWL_GRAD_EXACT = 0, 1
OWL_CKPT_FULL, OWL_CKPT_UNIFORM, OWL_CKPT_BINOMIAL = 0, 1, 2
OWL_STORE_F32, OWL_STORE_F16, OWL_STORE_BF16 = 0, 1, 2
OWL_ANISO_ISO, OWL_ANISO_THOMSEN, OWL_ANISO_AT = 0, 1, 2
OWL_MISFIT_WAVEFORM, OWL_MISFIT_CORR, OWL_MISFIT_ENVELOPE, OWL_MISFIT_PHASE = 0, 1, 2, 3
OWL_MISFIT_ADAPTIVE, OWL_MISFIT_LOCAL_ADAPTIVE, OWL_MISFIT_ADAPTIVE_SPACETIME = 4, 5, 6
MISFITS = {"waveform": OWL_MISFIT_WAVEFORM, "corr": OWL_MISFIT_CORR, "envelope": OWL_MISFIT_ENVELOPE,
           "phase": OWL_MISFIT_PHASE, "adaptive": OWL_MISFIT_ADAPTIVE, "local-adaptive": OWL_MISFIT_LOCAL_ADAPTIVE,
           "adaptive-spacetime": OWL_MISFIT_ADAPTIVE_SPACETIME}
PENALTIES = {"linear": 0, "power": 1, "gaussian": 2, "exp": 3}

MECHANISMS = {"explosion": OWL_EXPLOSION, "force": OWL_FORCE}
WAVELETS = {"ricker": OWL_RICKER, "gaussian_deriv_deriv": OWL_RICKER, "gaussian_deriv2": OWL_RICKER,
            "gaussian": OWL_GAUSSIAN, "gaussian_deriv": OWL_GAUSSIAN_DERIV,
            "gaussian_deriv1": OWL_GAUSSIAN_DERIV, "ricker_deriv": OWL_RICKER_DERIV,
            "ricker_deriv1": OWL_RICKER_DERIV, "gaussian_deriv3": OWL_RICKER_DERIV,
            "gaussian_deriv_deriv_deriv": OWL_RICKER_DERIV}


class OwlError(RuntimeError):
    """A failed library call; .code is the owl_status."""

    def __init__(self, code, msg):
        super().__init__(f"owl_amd error {code}: {msg}")
        self.code = code


class CflError(OwlError):
    pass


class owl_misfit_params(C.Structure):
    _fields_ = [("dt", C.c_float), ("tlag_max", C.c_float), ("penalty", C.c_int), ("penalty_power", C.c_float),
                ("deconv_eps", C.c_float), ("lawi_sigma", C.c_float), ("adaptive_hw", C.c_int)]


class owl_source(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float), ("f0", C.c_float),
                ("amp", C.c_float), ("t0", C.c_float), ("polar", C.c_float),
                ("azimuth", C.c_float), ("mechanism", C.c_int), ("wavelet", C.c_int)]


class owl_receiver(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float), ("weight", C.c_float)]


class owl_shot(C.Structure):
    _fields_ = [("id", C.c_int), ("ns", C.c_int), ("sources", C.POINTER(owl_source)),
                ("nr", C.c_int), ("receivers", C.POINTER(owl_receiver)),
                ("nxbeg", C.c_int), ("nxend", C.c_int), ("nybeg", C.c_int), ("nyend", C.c_int),
                ("nzbeg", C.c_int), ("nzend", C.c_int)]


class owl_elastic_surface(C.Structure):
    _fields_ = [("dz_refine", C.c_float), ("dz_max", C.c_float)]


class owl_acoustic_config(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("dx", C.c_float), ("dy", C.c_float), ("dz", C.c_float),
                ("ox", C.c_float), ("oy", C.c_float), ("oz", C.c_float),
                ("dt", C.c_float), ("tmax", C.c_float), ("data_dt", C.c_float),
                ("data_tmax", C.c_float), ("pml", C.c_int), ("fdhalf", C.c_int),
                ("free_surface", C.c_int), ("f0factor", C.c_float),
                ("cc_step_interval", C.c_int), ("trace_discard_threshold", C.c_float)]

    def __init__(self, *args, trace_discard_threshold=1.0e-6, **kw):
        super().__init__(*args, **kw)
        if len(args) < len(self._fields_):
            self.trace_discard_threshold = trace_discard_threshold


_P = C.c_void_p
_SIGNATURES = {
    "owl_version": (C.c_char_p, []),
    "owl_last_error": (C.c_char_p, []),
    "owl_dt_stable": (C.c_float, [C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float]),
    "owl_locate_point": (C.c_int, [C.c_int] + [C.c_float] * 9 + [_P, _P]),
    "owl_source_time_function": (C.c_int, [C.c_int, C.c_float, C.c_float, C.c_float, _P, C.c_int, _P]),
    "owl_set_adaptive_range": (C.c_int, [_P, C.c_int, _P, C.c_int, _P, _P, _P]),
    "owl_resample_traces": (C.c_int, [C.c_int, C.c_int, C.c_float, _P, C.c_int, C.c_float, C.c_float, _P]),
    "owl_ginterp": (C.c_int, [C.c_int, _P, _P, C.c_int, _P, _P, C.c_int]),
    "owl_cfspml_axis": (C.c_int, [C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float,
                                  _P, _P, _P, _P]),
    "owl_acoustic2d_create": (C.c_int, [_P, _P, _P, C.c_int, _P, _P]),
    "owl_acoustic2d_destroy": (None, [_P]),
    "owl_acoustic2d_nt": (C.c_int, [_P]),
    "owl_acoustic2d_set_model": (C.c_int, [_P, _P, _P]),
    "owl_acoustic2d_forward": (C.c_int, [_P, _P, C.c_int]),
    "owl_acoustic2d_forward_static": (C.c_int, [_P, _P, _P, C.c_int, _P, _P]),
    "owl_acoustic2d_misfit": (C.c_int, [_P, _P, _P, _P]),
    "owl_acoustic2d_set_energy_precond": (C.c_int, [_P, C.c_int]),
    "owl_tti2d_cij": (C.c_int, [C.c_int, C.c_longlong, _P, _P, _P, _P, _P, _P, _P]),
    "owl_adjoint_source": (C.c_int, [C.c_int, C.c_int, C.c_int, _P, _P, _P, C.c_float, _P, _P]),
    "owl_adjoint_source_ex": (C.c_int, [C.c_int, _P, C.c_int, C.c_int, _P, _P, _P, C.c_float, _P, _P]),
    "owl_tti2d_param_gradient": (C.c_int, [C.c_int, C.c_longlong, _P, _P, _P, _P, _P, _P, _P, _P]),
    "owl_acoustic2d_gradient": (C.c_int, [_P, _P, C.c_int, C.c_int, C.c_int, C.c_int, _P, _P, _P, _P]),
    "owl_acoustic2d_adjoint": (C.c_int, [_P, _P, _P, _P]),
    "owl_acoustic2d_shot_dims": (C.c_int, [_P, C.c_int, _P, _P, _P]),
    "owl_acoustic2d_last_stats": (C.c_int, [_P, _P, _P]),
    "owl_acoustic2d_set_observed": (C.c_int, [_P, _P]),
    "owl_acoustic2d_observe": (C.c_int, [_P]),
    "owl_acoustic2d_set_graphs": (C.c_int, [_P, C.c_int]),
    "owl_acoustic2d_set_step_kernel": (C.c_int, [_P, C.c_int, C.c_int]),
    "owl_acoustic2d_set_adjoint_rows": (C.c_int, [_P, C.c_int]),
    "owl_acoustic2d_set_autotune": (C.c_int, [_P, C.c_int]),
    "owl_acoustic2d_fused_rows": (C.c_int, [_P, _P, _P]),
    "owl_acoustic2d_set_profiling": (C.c_int, [_P, C.c_int]),
    "owl_acoustic2d_kernel_stats": (C.c_int, [_P, C.c_int, _P, _P, _P]),
    "owl_acoustic2d_forward_store": (C.c_int, [_P, _P, C.c_int, C.c_int, C.c_int]),
    "owl_acoustic2d_backward": (C.c_int, [_P, _P, _P, _P]),
    "owl_acoustic2d_store_id": (C.c_int, [_P, _P]),
    "owl_acoustic3d_create": (C.c_int, [_P, _P, _P, C.c_int, _P, _P]),
    "owl_acoustic3d_destroy": (None, [_P]),
    "owl_acoustic3d_nt": (C.c_int, [_P]),
    "owl_acoustic3d_forward": (C.c_int, [_P, _P, C.c_int]),
    "owl_acoustic3d_adjoint": (C.c_int, [_P, _P, _P, _P]),
    "owl_acoustic3d_gradient": (C.c_int, [_P, _P, _P, _P, _P, _P]),
    "owl_acoustic3d_set_launch": (C.c_int, [_P, C.c_int, C.c_int]),
    "owl_acoustic3d_set_graphs": (C.c_int, [_P, C.c_int]),
    "owl_acoustic3d_autotune": (C.c_int, [_P, C.c_int, _P, _P]),
    "owl_acoustic3d_set_profiling": (C.c_int, [_P, C.c_int]),
    "owl_acoustic3d_kernel_stats": (C.c_int, [_P, C.c_int, _P, _P, _P]),
    "owl_acoustic3d_last_stats": (C.c_int, [_P, _P, _P]),
    "owl_elastic2d_create": (C.c_int, [_P, _P, _P, _P, C.c_int, _P, _P]),
    "owl_elastic2d_create_surface": (C.c_int, [_P, _P, _P, _P, C.c_int, _P, _P, _P]),
    "owl_elastic2d_rows": (C.c_int, [_P, C.c_int]),
    "owl_elastic2d_destroy": (None, [_P]),
    "owl_elastic2d_nt": (C.c_int, [_P]),
    "owl_elastic2d_forward": (C.c_int, [_P, _P, C.c_int]),
    "owl_elastic2d_adjoint": (C.c_int, [_P, _P, _P, _P, _P]),
    "owl_elastic2d_gradient": (C.c_int, [_P, _P, _P, _P, _P, _P, _P]),
    "owl_elastic2d_set_launch": (C.c_int, [_P, C.c_int, C.c_int]),
    "owl_elastic2d_set_record_storage": (C.c_int, [_P, C.c_int]),
    "owl_elastic2d_autotune": (C.c_int, [_P, C.c_int, _P, _P]),
    "owl_elastic2d_set_graphs": (C.c_int, [_P, C.c_int]),
    "owl_elastic2d_set_profiling": (C.c_int, [_P, C.c_int]),
    "owl_elastic2d_kernel_stats": (C.c_int, [_P, C.c_int, _P, _P, _P]),
    "owl_elastic2d_last_stats": (C.c_int, [_P, _P, _P]),
    "owl_fsg2d_create": (C.c_int, [_P, _P, _P, C.c_int, C.c_int, _P, _P]),
    "owl_fsg2d_destroy": (None, [_P]),
    "owl_fsg2d_nt": (C.c_int, [_P]),
    "owl_fsg2d_damping_ratios": (C.c_int, [_P, C.c_int, _P, _P, _P, _P]),
    "owl_fsg2d_damp_ratio": (C.c_int, [_P, C.c_float, C.c_int, _P]),
    "owl_fsg2d_set_restore_all": (C.c_int, [_P, C.c_int]),
    "owl_fsg2d_forward": (C.c_int, [_P, _P, C.c_int]),
    "owl_fsg2d_adjoint": (C.c_int, [_P, _P, _P]),
    "owl_fsg2d_gradient": (C.c_int, [_P, _P, _P, _P, _P]),
    "owl_fsg2d_set_launch": (C.c_int, [_P, C.c_int, C.c_int]),
    "owl_fsg2d_set_model": (C.c_int, [_P, _P, _P]),
    "owl_fsg2d_set_record_storage": (C.c_int, [_P, C.c_int]),
    "owl_fsg2d_autotune": (C.c_int, [_P, C.c_int, _P, _P]),
    "owl_fsg2d_set_graphs": (C.c_int, [_P, C.c_int]),
    "owl_fsg2d_set_profiling": (C.c_int, [_P, C.c_int]),
    "owl_fsg2d_kernel_stats": (C.c_int, [_P, C.c_int, _P, _P, _P]),
    "owl_fsg2d_last_stats": (C.c_int, [_P, _P, _P]),
    "owl_acoustic2d_set_model_device": (C.c_int, [_P, _P, _P, C.c_int]),
    "owl_acoustic2d_forward_store_device": (C.c_int, [_P, _P, C.c_int, C.c_int, C.c_int]),
    "owl_acoustic2d_backward_device": (C.c_int, [_P, _P, _P, _P, C.c_int]),
    "owl_fsg2d_create_surface": (C.c_int, [_P, _P, _P, C.c_int, C.c_int, _P, _P, _P]),
    "owl_fsg2d_surface_mesh": (C.c_int, [_P, _P, _P, C.c_float, _P, C.c_int, _P, _P] + [_P] * 9),
    "owl_fsg2d_rows": (C.c_int, [_P, C.c_int]),
    "owl_elastic3d_create": (C.c_int, [_P, _P, _P, _P, C.c_int, _P, _P]),
    "owl_elastic3d_destroy": (None, [_P]),
    "owl_elastic3d_nt": (C.c_int, [_P]),
    "owl_elastic3d_forward": (C.c_int, [_P, _P, C.c_int]),
    "owl_elastic3d_adjoint": (C.c_int, [_P, _P, _P, _P, _P]),
    "owl_elastic3d_gradient": (C.c_int, [_P, _P, _P, _P, _P, _P, _P]),
    "owl_elastic3d_set_launch": (C.c_int, [_P, C.c_int, C.c_int]),
    "owl_elastic3d_set_record_storage": (C.c_int, [_P, C.c_int]),
    "owl_elastic3d_autotune": (C.c_int, [_P, C.c_int, _P, _P]),
    "owl_elastic3d_set_graphs": (C.c_int, [_P, C.c_int]),
    "owl_elastic3d_set_profiling": (C.c_int, [_P, C.c_int]),
    "owl_elastic3d_kernel_stats": (C.c_int, [_P, C.c_int, _P, _P, _P]),
    "owl_elastic3d_last_stats": (C.c_int, [_P, _P, _P]),
}

_lib = None


def lib():
    """The loaded library (raises OSError if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is missing: build it with `make -C owl_amd/csrc` "
                          "(or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return sorted(_SIGNATURES)


def check(status):
    """Raise OwlError for a non-zero owl_status."""
    if status != OWL_OK:
        msg = lib().owl_last_error().decode(errors="replace")
        if status == OWL_ERR_CFL:
            raise CflError(status, msg)
        raise OwlError(status, msg)
