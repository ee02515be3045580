"""ctypes declarations of the C ABI in include/lettuce_hip.h.

Loads lettuce_amd/liblettuce_hip.so and nothing else: there is no CPU fallback.
If the library is missing or fails to load, importing raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblettuce_hip.so")

c_ctx = ctypes.c_void_p
c_field = ctypes.c_void_p
c_flow = ctypes.c_void_p
D = ctypes.POINTER(ctypes.c_double)
I = ctypes.c_int

# name -> (argtypes); every entry point returns int status
SIGNATURES = {
    "lt_version": [ctypes.POINTER(I), ctypes.POINTER(I)],
    "lt_ctx_create": [I, ctypes.POINTER(c_ctx)],
    "lt_ctx_destroy": [c_ctx],
    "lt_ctx_synchronize": [c_ctx],
    "lt_ctx_stream": [c_ctx, ctypes.POINTER(ctypes.c_void_p)],
    "lt_ctx_profile": [c_ctx, I],
    "lt_ctx_profile_query": [c_ctx, ctypes.c_char_p, ctypes.POINTER(I), D],
    "lt_field_create": [c_ctx, I, I, I, I, ctypes.POINTER(c_field)],
    "lt_field_destroy": [c_field],
    "lt_field_shape": [c_field, ctypes.POINTER(I)],
    "lt_field_upload": [c_field, D],
    "lt_field_download": [c_field, D],
    "lt_field_copy": [c_field, c_field],
    "lt_field_swap": [c_field, c_field],
    "lt_field_set_identity": [c_field],
    "lt_field_set_haar": [c_field, ctypes.c_uint64],
    "lt_field_reunitarize": [c_field],
    "lt_field_check": [c_field, D, ctypes.POINTER(ctypes.c_longlong)],
    "lt_field_set_rng": [c_field, ctypes.c_uint64, ctypes.c_uint64],
    "lt_field_get_rng": [c_field, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)],
    "lt_field_set_tile": [c_field, I, I, I, I],
    "lt_field_get_tile": [c_field, ctypes.POINTER(I)],
    "lt_staple_field": [c_field, c_field],
    "lt_or_sweep": [c_field, I, I],
    "lt_heatbath_sweep": [c_field, ctypes.c_double, I, I],
    "lt_update_sweeps": [c_field, ctypes.c_double, I, I, I, I, D],
    "lt_plaquette_sum": [c_field, D],
    "lt_wilson_action": [c_field, ctypes.c_double, D],
    "lt_topcharge_clover": [c_field, D, D],
    "lt_topcharge_clover_density": [c_field, D],
    "lt_energy": [c_field, D, D, D],
    "lt_observables": [c_field, D],
    "lt_observables_noalloc": [c_field, D],
    "lt_stout": [c_field, c_field, ctypes.c_double, I],
    "lt_flow_create": [c_field, c_field, I, ctypes.c_double, ctypes.POINTER(c_flow)],
    "lt_flow_destroy": [c_flow],
    "lt_flow_run": [c_flow, I],
    "lt_flow_resume": [c_flow, I],
    "lt_matrix_field_create": [c_ctx, I, I, I, I, ctypes.POINTER(c_field)],
    "lt_hmc_random_momentum": [c_field, ctypes.c_uint64, ctypes.c_uint64],
    "lt_hmc_update_momenta": [c_field, c_field, ctypes.c_double, ctypes.c_double],
    "lt_hmc_update_fields": [c_field, c_field, ctypes.c_double],
    "lt_hmc_hamiltonian": [c_field, c_field, ctypes.c_double, D],
    "lt_hmc_integrate": [c_field, c_field, ctypes.c_double, I, ctypes.c_double, I],
    "lt_hmc_trajectory": [c_field, c_field, c_field, ctypes.c_double, I, ctypes.c_double, I, I, D,
                          ctypes.POINTER(I)],
    "lt_cv_create": [c_ctx, I, I, I, I, I, ctypes.c_double, ctypes.POINTER(ctypes.c_void_p)],
    "lt_cv_create_checkpointed": [c_ctx, I, I, I, I, I, ctypes.c_double, I, ctypes.POINTER(ctypes.c_void_p)],
    "lt_cv_bytes": [ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t)],
    "lt_ctx_mem_info": [c_ctx, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)],
    "lt_ctx_read_rate": [c_ctx, ctypes.c_size_t, ctypes.POINTER(ctypes.c_double)],
    "lt_cv_destroy": [ctypes.c_void_p],
    "lt_cv_forward": [ctypes.c_void_p, c_field, D],
    "lt_cv_backward": [ctypes.c_void_p, ctypes.c_double, c_field],
    "lt_bias_create": [ctypes.c_double, ctypes.c_double, I, ctypes.c_double, ctypes.c_double, I, ctypes.c_double,
                       ctypes.c_double, ctypes.POINTER(ctypes.c_void_p)],
    "lt_bias_destroy": [ctypes.c_void_p],
    "lt_bias_deposit": [ctypes.c_void_p, ctypes.c_double, I],
    "lt_bias_potential": [ctypes.c_void_p, ctypes.c_double, D],
    "lt_bias_derivative": [ctypes.c_void_p, ctypes.c_double, D],
    "lt_bias_grid": [ctypes.c_void_p, D, I, I],
    "lt_bias_cv": [ctypes.c_void_p, D, I],
    "lt_metad_hmc_trajectory": [c_field, c_field, c_field, c_field, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                I, ctypes.c_double, I, I, I, D, ctypes.POINTER(I), D],
    "lt_cv_path_samples": [ctypes.c_void_p, I, D, I, ctypes.POINTER(I)],
    "lt_clover_derivative": [c_field, c_field, ctypes.c_double],
    "lt_hmc_reversibility_test": [c_field, c_field, ctypes.c_double, I, ctypes.c_double, I, D, D],
    "lt_metad_reversibility_test": [c_field, c_field, c_field, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, I,
                                    ctypes.c_double, I, D, D],
    "lt_integrator_schedule": [I, ctypes.c_double, I, ctypes.c_char_p, D, ctypes.POINTER(I), I, ctypes.POINTER(I)],
    "lt_cv_log_path": [ctypes.c_void_p, I, ctypes.c_char_p],
    "lt_batch_create": [c_ctx, I, ctypes.POINTER(I), I, ctypes.c_double, ctypes.POINTER(ctypes.c_void_p)],
    "lt_batch_destroy": [ctypes.c_void_p],
    "lt_batch_size": [ctypes.c_void_p, ctypes.POINTER(I)],
    "lt_batch_field": [ctypes.c_void_p, I, ctypes.POINTER(c_field)],
    "lt_batch_set_rng": [ctypes.c_void_p, I, ctypes.c_uint64, ctypes.c_uint64],
    "lt_batch_get_rng": [ctypes.c_void_p, I, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)],
    "lt_batch_set_ladder": [ctypes.c_void_p, D, ctypes.POINTER(I), ctypes.c_uint64, ctypes.c_uint64],
    "lt_batch_set_bias": [ctypes.c_void_p, ctypes.c_void_p],
    "lt_batch_get_bias": [ctypes.c_void_p, ctypes.c_void_p],
    "lt_batch_use_graph": [ctypes.c_void_p, I],
    "lt_batch_metad_hmc": [ctypes.c_void_p, ctypes.c_double, I, ctypes.c_double, I, I, I, I, I],
    "lt_batch_results": [ctypes.c_void_p, D, ctypes.POINTER(I), D, ctypes.POINTER(I)],
    "lt_batch_samples": [ctypes.c_void_p, D, I],
    "lt_batch_swap_stats": [ctypes.c_void_p, ctypes.POINTER(ctypes.c_longlong), I],
    "lt_autotune_field": [c_field, I, ctypes.c_char_p, ctypes.POINTER(I), D],
    "lt_parity_update": [c_field, c_field],
    "lt_staple_field_action": [c_field, ctypes.c_double, ctypes.c_double, c_field],
    "lt_or_sweep_action": [c_field, ctypes.c_double, ctypes.c_double, I, I],
    "lt_heatbath_sweep_action": [c_field, ctypes.c_double, ctypes.c_double, ctypes.c_double, I, I],
    "lt_metad_tempering_swap": [c_field, c_field, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64,
                                ctypes.POINTER(I), D],
    "lt_action_rectangular": [c_field, ctypes.c_double, ctypes.c_double, D],
    "lt_topcharge_plaquette": [c_field, D],
    "lt_wilson_loop": [c_field, c_field, I, I, I, D],
    "lt_polyakov_loop": [c_field, D, D],
    "lt_field_save_bmw": [c_field, ctypes.c_char_p, I, I, ctypes.c_char_p, I],
    "lt_field_load_bmw": [c_field, ctypes.c_char_p, I, ctypes.POINTER(I), D],
    "lt_field_bmw_checksum": [c_field, I, ctypes.POINTER(ctypes.c_ulonglong)],
    "lt_bias_save_parameters": [ctypes.c_void_p, ctypes.c_char_p, I],
    "lt_bias_save_potential": [ctypes.c_void_p, ctypes.c_char_p, I],
    "lt_bias_load_potential": [ctypes.c_void_p, ctypes.c_char_p],
    "lt_bias_info": [ctypes.c_void_p, D, ctypes.POINTER(I), ctypes.POINTER(ctypes.c_ulonglong)],
    "lt_batch_autotune": [ctypes.c_void_p, ctypes.c_char_p],
}


class LettuceError(RuntimeError):
    def __init__(self, fn, code, msg):
        super().__init__(f"{fn} failed ({code}): {msg}")
        self.code = code


def load(path=LIB_PATH):
    if not os.path.exists(path):
        raise ImportError(
            f"{path} is missing: build it with `python -m lettuce_amd.build` (hipcc, gfx950); "
            "there is no CPU fallback")
    lib = ctypes.CDLL(path)
    lib.lt_last_error.restype = ctypes.c_char_p
    lib.lt_last_error.argtypes = []
    for name, args in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = ctypes.c_int
        f.argtypes = args
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = load()
    return _lib


def call(name, *args):
    rc = getattr(lib(), name)(*args)
    if rc != 0:
        raise LettuceError(name, rc, lib().lt_last_error().decode(errors="replace"))
    return rc


def exported_symbols():
    return ["lt_last_error"] + list(SIGNATURES)
