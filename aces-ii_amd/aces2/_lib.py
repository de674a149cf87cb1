"""ctypes binding of the C-ABI library ``libaces2_amd.so`` (include/aces2_amd.h).

The library is the product path; there is no fallback.  Loading fails loudly
when it has not been built (``__graft_entry__.build()`` or ``make`` in
``aces-ii_amd/csrc``).
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libaces2_amd.so')

_lib = None

c_int = ctypes.c_int
c_long = ctypes.c_long
c_double = ctypes.c_double
c_vp = ctypes.c_void_p
c_dp = ctypes.POINTER(ctypes.c_double)
c_char_p = ctypes.c_char_p

# name -> (restype, argtypes)
SIGNATURES = {
    'aces_last_error': (c_char_p, []),
    'aces_gemm': (c_int, [c_int, c_int, c_int, c_int, c_vp, c_vp, c_long, c_int, c_vp, c_vp, c_long, c_int,
                          c_vp, c_long, c_double, c_double, c_int, c_long, c_long, c_long, c_vp]),
    'aces_gemm_grouped': (c_int, [c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_int, c_vp, c_vp, c_vp,
                                   c_int, c_vp, c_vp, c_double, c_double, c_vp, c_long, c_vp]),
    'aces_gemm_grouped_noalloc': (c_int, [c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_int, c_vp, c_vp,
                                           c_vp, c_int, c_vp, c_vp, c_double, c_double, c_long, c_vp]),
    'aces_convert_f32': (c_int, [c_int, c_int, c_vp, c_long, c_vp, c_long, c_vp]),
    'aces_split_bf16': (c_int, [c_int, c_int, c_vp, c_long, c_vp, c_vp, c_long, c_vp]),
    'aces_ccsd_create': (c_int, [c_int, c_int, c_vp, ctypes.POINTER(c_vp)]),
    'aces_ccsd_destroy': (c_int, [c_vp]),
    'aces_ccsd_setup': (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    'aces_ccsd_set_amplitudes': (c_int, [c_vp, c_vp, c_vp]),
    'aces_ccsd_get_amplitudes': (c_int, [c_vp, c_vp, c_vp]),
    'aces_ccsd_start_mp1': (c_int, [c_vp]),
    'aces_ccsd_increments': (c_int, [c_vp, c_int, c_vp, c_vp]),
    'aces_ccsd_cycle': (c_int, [c_vp, c_int, c_int, c_double, c_dp]),
    'aces_ccsd_set_diis_order': (c_int, [c_vp, c_int]),
    'aces_ccsd_reset_diis': (c_int, [c_vp]),
    'aces_ccsd_reset_anchor': (c_int, [c_vp]),
    'aces_ccsd_set_lambda_precision': (c_int, [c_vp, c_int]),
    'aces_ccsd_energy': (c_int, [c_vp, c_vp, c_vp, c_dp]),
    'aces_ccsd_device_bytes': (c_long, [c_vp]),
    'aces_ccsd_set_profiling': (c_int, [c_vp, c_int]),
    'aces_ccsd_profile': (c_int, [c_vp, c_dp, c_int]),
    'aces_ccsd_set_fock': (c_int, [c_vp, c_vp]),
    'aces_ccsd_triples': (c_int, [c_vp, c_int, c_dp, c_vp]),
    'aces_ccsd_eom': (c_int, [c_vp, c_vp, c_int, c_int, c_double, c_int, c_dp, c_dp, c_vp, c_vp]),
    'aces_ccsd_adjoint': (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    'aces_ccsd_lambda_start': (c_int, [c_vp]),
    'aces_ccsd_lambda_cycle': (c_int, [c_vp, c_int, c_double, c_dp]),
    'aces_ccsd_set_lambda_diis_order': (c_int, [c_vp, c_int]),
    'aces_ccsd_get_lambda': (c_int, [c_vp, c_vp, c_vp, c_int]),
    'aces_ccsd_set_lambda': (c_int, [c_vp, c_vp, c_vp, c_int]),
    'aces_ccsd_density': (c_int, [c_vp, c_vp]),
    'aces_sym_create': (c_int, [c_int, c_vp, c_vp, c_vp, ctypes.POINTER(c_vp)]),
    'aces_sym_destroy': (c_int, [c_vp]),
    'aces_sym_setup': (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    'aces_sym_set_amplitudes': (c_int, [c_vp, c_vp, c_vp]),
    'aces_sym_get_amplitudes': (c_int, [c_vp, c_vp, c_vp]),
    'aces_sym_start_mp1': (c_int, [c_vp]),
    'aces_sym_increments': (c_int, [c_vp, c_int, c_vp, c_vp]),
    'aces_sym_cycle': (c_int, [c_vp, c_int, c_int, c_double, c_dp]),
    'aces_sym_set_diis_order': (c_int, [c_vp, c_int]),
    'aces_sym_reset_diis': (c_int, [c_vp]),
    'aces_sym_reset_anchor': (c_int, [c_vp]),
    'aces_sym_energy': (c_int, [c_vp, c_dp]),
    'aces_sym_adjoint': (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    'aces_sym_lambda_start': (c_int, [c_vp]),
    'aces_sym_set_lambda_precision': (c_int, [c_vp, c_int]),
    'aces_sym_lambda_cycle': (c_int, [c_vp, c_int, c_double, c_dp]),
    'aces_sym_set_lambda_diis_order': (c_int, [c_vp, c_int]),
    'aces_sym_get_lambda': (c_int, [c_vp, c_vp, c_vp, c_int]),
    'aces_sym_density': (c_int, [c_vp, c_vp]),
    'aces_sym_set_profiling': (c_int, [c_vp, c_int]),
    'aces_sym_profile': (c_int, [c_vp, c_dp, c_int]),
    'aces_sym_device_bytes': (c_long, [c_vp]),
    'aces_sym_packed_t2': (c_long, [c_vp]),
    'aces_init': (c_int, [c_vp]),
    'aces_fin': (c_int, []),
    'aces_set_symmetry': (c_int, [c_int, c_vp, c_vp, c_vp, c_vp]),
    'aces_irpdpd': (c_long, [c_int, c_int]),
    'aces_isymoff': (c_long, [c_int, c_int, c_int]),
    'aces_inipck': (c_int, [c_int, c_int, c_int, c_int]),
    'aces_updmoi': (c_int, [c_long, c_long, c_int, c_int]),
    'aces_list_shape': (c_int, [c_int, c_int, ctypes.POINTER(c_long), ctypes.POINTER(c_long)]),
    'aces_list_isytyp': (c_int, [c_int, ctypes.POINTER(c_int), ctypes.POINTER(c_int)]),
    'aces_getlst': (c_int, [c_vp, c_long, c_long, c_int, c_int, c_int]),
    'aces_putlst': (c_int, [c_vp, c_long, c_long, c_int, c_int, c_int]),
    'aces_getall': (c_int, [c_vp, c_long, c_int, c_int, c_int]),
    'aces_putall': (c_int, [c_vp, c_long, c_int, c_int, c_int]),
    'aces_sumsym2': (c_int, [c_vp, c_int, c_int, c_int]),
    'aces_list_device_ptr': (c_vp, [c_int, c_int]),
    'aces_list_c1': (c_int, [c_int, c_int, c_int, c_int, c_vp, c_int]),
    'aces_intprc_df': (c_int, [c_vp, c_int, c_vp]),
    'aces_small_eig': (c_int, [c_int, c_vp, c_vp, c_vp]),
    'aces_small_eigvec': (c_int, [c_int, c_vp, c_double, c_vp]),
    'aces_jobarc_read': (c_int, [c_char_p]),
    'aces_jobarc_write': (c_int, [c_char_p]),
    'aces_lists_read': (c_int, [c_char_p]),
    'aces_lists_write': (c_int, [c_char_p]),
    'aces_putrec': (c_int, [c_char_p, c_long, c_vp]),
    'aces_getrec': (c_int, [c_int, c_char_p, ctypes.POINTER(c_long), c_vp]),
    'aces_vcc_mp1': (c_int, []),
    'aces_vcc_set_diis_order': (c_int, [c_int]),
    'aces_vcc_set_precision': (c_int, [c_int]),
    'aces_genint': (c_int, []),
    'aces_dre3en': (c_int, []),
    'aces_e4s': (c_int, []),
    'aces_newt2': (c_int, []),
    'aces_cmpeng': (c_int, [c_int, c_int, c_dp, c_dp, c_dp]),
    'aces_dodiis0': (c_int, [c_int, c_double, ctypes.POINTER(c_int), c_dp]),
    'aces_drmove': (c_int, []),
    'aces_drmove_lists': (c_int, [c_int, c_int]),
}


class AcesError(RuntimeError):
    pass


def _init_torch_runtime():
    """PyTorch-ROCm ships its own HIP runtime; when both it and the system one
    that the library links serve one process, the GPU must be opened through
    torch's first (the other order leaves torch without devices), so torch's
    runtime is initialised before the library's first call."""
    try:
        import torch
    except ImportError:
        return
    if torch.cuda.is_available():
        torch.cuda.init()


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise AcesError('libaces2_amd.so is not built (%s); run __graft_entry__.build()' % LIB_PATH)
        _init_torch_runtime()
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status, what=''):
    if status != 0:
        msg = lib().aces_last_error()
        raise AcesError('%s failed (status %d): %s' % (what, status, msg.decode() if msg else ''))


def call(name, *args):
    check(getattr(lib(), name)(*args), name)
