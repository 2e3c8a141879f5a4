"""ctypes binding of libmopachip.so (the C ABI declared in include/mopac_hip.h).

The library is built in-tree by mopac_gpu_amd.build; nothing here falls back
to a CPU implementation: if the shared object is missing, loading fails.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmopachip.so")
_lib = None

c_int = ctypes.c_int
c_ll = ctypes.c_longlong
c_double = ctypes.c_double
c_void_p = ctypes.c_void_p
c_char = ctypes.c_char


class ScfOpts(ctypes.Structure):
    _fields_ = [
        ("nclose", c_int),
        ("pulay", c_int),
        ("itrmax", c_int),
        ("scfcrt", c_double),
        ("last", c_int),
        ("precision", c_int),
        ("pseudo_diag", c_int),
        ("adaptive_eigen", c_int),
        ("enuclr", c_double),
        ("atheat", c_double),
        ("pdiag", c_void_p),
        ("p_init", c_void_p),
        ("block_start", c_int),
        ("fragment", c_void_p),
        ("start_from_guess", c_int),
    ]


class MozymeSystemC(ctypes.Structure):
    _fields_ = [("numat", c_int), ("norbs", c_int), ("mpack", c_ll), ("nfirst", c_void_p), ("nlast", c_void_p),
                ("nijbo", c_void_p), ("xyz", c_void_p), ("am", c_void_p), ("dd", c_void_p), ("ad", c_void_p),
                ("h", c_void_p), ("w", c_void_p), ("nw", c_ll), ("feather", c_int), ("trunc1", c_double),
                ("trunc2", c_double)]


class GuessModelC(ctypes.Structure):
    _fields_ = [("x_mean", c_void_p), ("x_std", c_void_p), ("y_mean", c_void_p), ("y_std", c_void_p)]


class ScfResult(ctypes.Structure):
    _fields_ = [
        ("escf", c_double),
        ("ee", c_double),
        ("niter", c_int),
        ("converged", c_int),
        ("n_full_diag", c_int),
        ("n_pseudo_diag", c_int),
        ("seconds", c_double),
        ("jacobi_sweeps", c_int),
        ("n_fragment_diag", c_int),
        ("n_split_iter", c_int),
    ]


class LmoSetC(ctypes.Structure):
    _fields_ = [("numat", c_int), ("norbs", c_int), ("nocc", c_int), ("nvir", c_int), ("mpack", c_ll),
                ("nfirst", c_void_p), ("nlast", c_void_p), ("nijbo", c_void_p),
                ("ncf", c_void_p), ("nncf", c_void_p), ("ncocc", c_void_p), ("icocc", c_void_p),
                ("icocc_dim", c_int), ("cocc", c_void_p), ("cocc_dim", c_int),
                ("nce", c_void_p), ("nnce", c_void_p), ("ncvir", c_void_p), ("icvir", c_void_p),
                ("icvir_dim", c_int), ("cvir", c_void_p), ("cvir_dim", c_int)]


class MozymeOptsC(ctypes.Structure):
    _fields_ = [("scfcrt", c_double), ("itrmax", c_int), ("thresh", c_double), ("damp", c_double),
                ("fmo_dim", c_int), ("enuclr", c_double), ("atheat", c_double), ("print_level", c_int),
                ("heats", c_void_p), ("heats_cap", c_int), ("order_free", c_int), ("fp32_fmo", c_int)]


class MozymeResultC(ctypes.Structure):
    _fields_ = [("escf", c_double), ("ee", c_double), ("niter", c_int), ("converged", c_int),
                ("ovmax", c_double), ("seconds", c_double)]


class Diagg1State(ctypes.Structure):
    _fields_ = [("reset", c_int), ("oldens", c_int), ("fref", c_double), ("safety", c_double),
                ("oldlim", c_double), ("tiny", c_double), ("sumt", c_double), ("ovmax", c_double),
                ("nfmo", c_void_p), ("fp32_accumulate", c_int)]


class Diagg2State(ctypes.Structure):
    _fields_ = [("reset", c_int), ("tiny", c_double), ("sumb", c_double), ("damp", c_double),
                ("shift", c_double), ("thresh", c_double), ("nrejct", c_int * 2), ("order_free", c_int)]


# name -> (restype, argtypes)
SIGNATURES = {
    "mhip_init": (c_int, [c_int, ctypes.POINTER(c_void_p)]),
    "mhip_finalize": (c_int, [c_void_p]),
    "mhip_last_error": (ctypes.c_char_p, []),
    "mhip_device_info": (c_int, [c_void_p, ctypes.c_char_p, c_int, ctypes.POINTER(c_double), ctypes.POINTER(c_int)]),
    "mhip_upload_system": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_int, c_ll, c_void_p, c_void_p]),
    "mhip_density": (c_int, [c_void_p, c_void_p, c_int, c_double, c_void_p]),
    "mhip_fock2": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int]),
    "mhip_helect": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, ctypes.POINTER(c_double)]),
    "mhip_pulay_sequence": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    "mhip_helect_noalloc": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "mhip_helect_noalloc_max_rows": (c_int, []),
    "mhip_eigen": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int]),
    "mhip_pseudo_diag": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_int, c_int, ctypes.POINTER(c_ll)]),
    "mhip_scf": (c_int, [c_void_p, ctypes.POINTER(ScfOpts), ctypes.POINTER(ScfResult), c_void_p, c_void_p, c_void_p, c_void_p]),
    "mhip_call_gemm": (None, [c_char, c_char, c_int, c_int, c_int, c_double, c_void_p, c_int, c_void_p, c_int, c_double, c_void_p, c_int]),
    "mhip_gemm_split_bf16": (None, [c_char, c_char, c_int, c_int, c_int, c_double, c_void_p, c_int, c_void_p, c_int, c_double, c_void_p, c_int]),
    "mhip_call_syrk": (None, [c_char, c_char, c_int, c_int, c_double, c_void_p, c_int, c_double, c_void_p, c_int]),
    "mhip_dsyevd": (None, [c_int, c_void_p, c_int, c_void_p, ctypes.POINTER(c_int)]),
    "mhip_call_rot": (None, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_double, c_double]),
    "mhip_getGPUInfo": (None, [c_void_p] * 9),
    "mhip_setDevice": (None, [c_int, c_void_p]),
    "mhip_set_mozyme_gpu_pair": (None, [c_int, c_int]),
    "mhip_selected_device": (c_int, []),
    "mhip_destroy_resources": (None, []),
    "mhip_set_tuning": (c_int, [c_void_p, c_int, c_int, c_double]),
    "mhip_set_jacobi_schedule": (c_int, [c_int]),
    "mhip_mozyme_stats": (c_int, [c_int, ctypes.POINTER(c_double), ctypes.POINTER(c_ll), ctypes.POINTER(c_ll),
                                  ctypes.POINTER(c_ll), ctypes.POINTER(c_double), ctypes.POINTER(c_ll)]),
    "mhip_mozyme_diagg2_one_workgroup": (c_int, [c_int]),
    "mhip_mozyme_tidy": (c_int, [c_void_p, c_void_p, c_int, c_double]),
    "mhip_mozyme_check": (c_int, [c_void_p, c_void_p, c_int, ctypes.POINTER(c_double)]),
    "mhip_mozyme_eimp": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_ll, c_void_p, c_void_p]),
    "mhip_mozyme_diagg2_wide_rounds": (c_int, [c_int]),
    "mhip_mozyme_diagg2_list_cap": (c_int, [c_int]),
    "mhip_mozyme_diagg2_coef_cap": (c_int, [c_int]),
    "mhip_last_sweeps": (c_int, [c_void_p]),
    "mhip_mozyme_density": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
    "mhip_mozyme_diagg1": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p,
                                   c_int, c_void_p, c_void_p, c_void_p]),
    "mhip_mozyme_diagg2": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int,
                                   c_void_p]),
    "mhip_mozyme_helecz": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_ll, c_void_p, c_void_p, c_void_p,
                                   ctypes.POINTER(c_double)]),
    "mhip_mozyme_cnvgz": (c_int, [c_void_p, c_int, c_ll, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                                  c_void_p, c_int, ctypes.POINTER(c_double)]),
    "mhip_batch_create": (c_int, [c_int, c_void_p, ctypes.POINTER(c_void_p)]),
    "mhip_batch_set_tiles": (c_int, [c_void_p, c_void_p]),
    "mhip_batch_set_eig_lds": (c_int, [c_void_p, c_int]),
    "mhip_batch_eig_profile": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    "mhip_batch_run": (c_int, [c_void_p, c_int, c_int, ctypes.POINTER(c_double), ctypes.POINTER(c_int)]),
    "mhip_batch_results": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "mhip_batch_density": (c_int, [c_void_p, c_int, c_void_p]),
    "mhip_batch_destroy": (None, [c_void_p]),
    "mhip_kernel_stats": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "mhip_kernel_class_name": (ctypes.c_char_p, [c_int]),
    "mhip_kernel_timing": (c_int, [c_void_p, c_int, ctypes.POINTER(c_double), ctypes.POINTER(c_ll),
                                   ctypes.POINTER(c_double), ctypes.POINTER(c_ll),
                                   ctypes.POINTER(c_double), ctypes.POINTER(c_ll),
                                   ctypes.POINTER(c_ll), ctypes.POINTER(c_ll), ctypes.POINTER(c_ll)]),
    "mhip_mlp_create": (c_int, [c_int, c_void_p, ctypes.POINTER(c_void_p)]),
    "mhip_mlp_destroy": (None, [c_void_p]),
    "mhip_mlp_set_params": (c_int, [c_void_p, c_void_p]),
    "mhip_mlp_get_params": (c_int, [c_void_p, c_void_p]),
    "mhip_mlp_forward": (c_int, [c_void_p, c_int, c_void_p, c_void_p]),
    "mhip_mlp_gradients": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, ctypes.POINTER(ctypes.c_float)]),
    "mhip_mlp_train_step": (c_int, [c_void_p, c_int, c_void_p, c_void_p, ctypes.c_float, ctypes.POINTER(ctypes.c_float)]),
    "mhip_guess_water": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "mhip_guess_density": (c_int, [c_void_p, c_void_p]),
    "mhip_mozyme_setup": (c_int, [c_void_p, c_void_p]),
    "mhip_pm7_pair": (c_int, [c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                               c_void_p]),
    "mhip_pm7_corrections": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "mhip_den_write": (c_int, [ctypes.c_char_p, c_int, c_int, c_void_p, c_int]),
    "mhip_den_read": (c_int, [ctypes.c_char_p, c_int, c_int, c_void_p, c_int]),
    "mhip_pm7_element": (c_int, [c_int, c_void_p]),
    "mhip_pm7_hamiltonian_mozyme": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_ll, c_double, c_double, c_void_p,
                                             ctypes.POINTER(c_double)]),
    "mhip_pm7_atoms": (c_int, [c_int, c_void_p, c_int, c_void_p, c_void_p, ctypes.POINTER(c_double), c_void_p]),
    "mhip_pm7_integrals_mozyme": (c_int, [c_int, c_void_p, c_void_p, c_void_p, ctypes.POINTER(c_ll), c_void_p]),
    "mhip_mozyme_fock2z": (c_int, [c_void_p, c_void_p, c_void_p]),
    "mhip_mozyme_scf": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "mhip_pm7_sizes": (c_int, [c_int, c_void_p, ctypes.POINTER(c_int), ctypes.POINTER(c_ll)]),
    "mhip_pm7_hamiltonian": (c_int, [c_int, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, ctypes.POINTER(c_double), ctypes.POINTER(c_double), c_void_p]),
    "mhip_ev_to_kcal": (c_double, []),
}


def load():
    """Load libmopachip.so and attach signatures; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m mopac_gpu_amd.build` "
            "(there is no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols():
    """Names declared in include/mopac_hip.h that the ABI must export."""
    hdr = os.path.join(os.path.dirname(_HERE), "include", "mopac_hip.h")
    import re
    names = []
    with open(hdr) as fh:
        for line in fh:
            m = re.match(r"^\s*(?:const\s+)?(?:int|void|double|char)\s*\*?\s*(mhip_\w+)\s*\(", line)
            if m:
                names.append(m.group(1))
    return names


def check(rc):
    if rc != 0:
        msg = load().mhip_last_error()
        raise RuntimeError((msg or b"unknown error").decode())


def ptr(a):
    return a.ctypes.data_as(c_void_p) if a is not None else None
