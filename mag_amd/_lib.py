"""ctypes binding of libmag_amd.so (the C ABI declared in include/mag_amd.h).

The library is built in-tree (``make -C mag_amd``); importing this module
loads it and fails loudly when it is missing. There is no CPU fallback.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmag_amd.so")

# status codes
OK = 0
(ERR_ARG, ERR_GEOMETRY, ERR_OOM, ERR_DEVICE, ERR_NAN, ERR_UNSUPPORTED, ERR_BUDGET, ERR_STATE, ERR_IO,
 ERR_CHECKSUM) = range(-1, -11, -1)
STATUS_NAMES = {0: "MAG_OK", -1: "MAG_ERR_ARG", -2: "MAG_ERR_GEOMETRY", -3: "MAG_ERR_OOM", -4: "MAG_ERR_DEVICE",
                -5: "MAG_ERR_NAN", -6: "MAG_ERR_UNSUPPORTED", -7: "MAG_ERR_BUDGET", -8: "MAG_ERR_STATE",
                -9: "MAG_ERR_IO", -10: "MAG_ERR_CHECKSUM"}
# field kinds / precisions / actions
LINK18, LINK12, ALG8, REAL1 = 0, 1, 2, 3
F32, F64 = 0, 1
WILSON, IWASAKI = 0, 1


class Action(C.Structure):
    _fields_ = [("type", C.c_int), ("beta", C.c_double), ("c1", C.c_double), ("mag", C.c_int),
                ("n_stout", C.c_int), ("rho", C.c_double)]


class HmcParams(C.Structure):
    _fields_ = [("trajectory_length", C.c_int), ("dt", C.c_double), ("integrator", C.c_int),
                ("num_forced_accept", C.c_int), ("accept_mode", C.c_int), ("demon_emax", C.c_double),
                ("seed", C.c_uint64), ("reunitarize", C.c_int), ("precision", C.c_int), ("recon12", C.c_int)]


class TrajResult(C.Structure):
    _fields_ = [("h_old", C.c_double), ("h_new", C.c_double), ("dh", C.c_double), ("die_roll", C.c_double),
                ("accept", C.c_int), ("avg_plaq", C.c_double), ("s_gauge", C.c_double),
                ("demon_e", C.c_double * 2), ("reunit_dev", C.c_double), ("status", C.c_int)]


class NerscHeader(C.Structure):
    _fields_ = [("dims", C.c_int * 4), ("skip_third", C.c_int), ("link_trace", C.c_double), ("plaquette", C.c_double),
                ("beta", C.c_double), ("sequence_number", C.c_int), ("ensemble_label", C.c_char * 128),
                ("checksum", C.c_uint32), ("crc32", C.c_uint32), ("checksum_swapped", C.c_int)]


class ServeRequest(C.Structure):
    _fields_ = [("dims", C.c_int * 4), ("act", Action), ("hmc", HmcParams), ("ntraj", C.c_int),
                ("links_seed", C.c_uint64)]


class MultiParams(C.Structure):
    _fields_ = [("rho", C.c_double), ("xi0", C.c_double), ("md_alpha", C.c_double)]


class MultiResult(C.Structure):
    _fields_ = [("h_old", C.c_double), ("h_new", C.c_double), ("dh", C.c_double), ("die_roll", C.c_double),
                ("accept", C.c_int), ("avg_plaq_fine", C.c_double), ("avg_plaq_coarse", C.c_double),
                ("parts_old", C.c_double * 3), ("parts_new", C.c_double * 3), ("reunit_dev", C.c_double * 2),
                ("status", C.c_int)]


_P = C.c_void_p
_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)
_L = C.POINTER(C.c_long)

SIGNATURES = {
    "mag_ctx_create": [C.c_int, C.POINTER(_P)],
    "mag_ctx_destroy": [_P],
    "mag_last_error": [_P],
    "mag_ctx_synchronize": [_P],
    "mag_ctx_memory": [_P, C.POINTER(C.c_size_t)],
    "mag_ctx_set_timing": [_P, C.c_int],
    "mag_ctx_set_timing_family": [_P, C.c_char_p],
    "mag_ctx_kernel_time": [_P, C.c_char_p, _D, _L],
    "mag_ctx_reset_timing": [_P],
    "mag_field_create": [_P, _I, C.c_int, C.c_int, C.POINTER(_P)],
    "mag_field_destroy": [_P],
    "mag_field_upload_aos": [_P, _D],
    "mag_field_download_aos": [_P, _D],
    "mag_field_copy": [_P, _P],
    "mag_field_info": [_P, _I, _I, _I, C.POINTER(C.c_size_t)],
    "mag_unit_links": [_P],
    "mag_haar_links": [_P, C.c_uint64],
    "mag_init_momentum": [_P, C.c_uint64, C.c_uint64, C.c_double],
    "mag_reunitarize": [_P, _D],
    "mag_expiQ_batch": [_P, _D, _D, C.c_long, C.c_int],
    "mag_compute_Lambda_batch": [_P, _D, _D, _D, _D, C.c_long, C.c_int],
    "mag_gauge_force": [_P, _P, C.POINTER(Action)],
    "mag_evolve_gauge": [_P, _P, C.c_double, C.c_int],
    "mag_evolve_momentum": [_P, _P, C.c_double, C.c_double],
    "mag_stout_smear": [_P, _P, C.c_double, C.c_int],
    "mag_fg_integrate": [_P, _P, C.POINTER(Action), C.c_int, C.c_double, C.c_int, _D],
    "mag_leapfrog_integrate": [_P, _P, C.POINTER(Action), C.c_int, C.c_double],
    "mag_total_plaq": [_P, _D],
    "mag_total_rectangular": [_P, _D],
    "mag_avg_plaquette": [_P, _D],
    "mag_gauge_action": [_P, C.POINTER(Action), _D],
    "mag_kinetic_energy": [_P, C.c_int, _D],
    "mag_hamiltonian": [_P, _P, C.POINTER(Action), _D, _D],
    "mag_avg_wilson_loop": [_P, _I, C.c_int, _D],
    "mag_total_plaq_async": [_P, _P],
    "mag_total_rectangular_async": [_P, _P],
    "mag_hamiltonian_async": [_P, _P, C.POINTER(Action), _P],
    "mag_avg_wilson_loop_async": [_P, _I, C.c_int, _P],
    "mag_hamiltonian_noalloc": [_P, _P, C.POINTER(Action), _P],
    "mag_DexpiQ_batch": [_P, _D, _D, _D, C.c_long, C.c_int],
    "mag_chair_staple_dagger": [_P, _P],
    "mag_twist_staple_dagger": [_P, _P],
    "mag_multi_get_Q": [_P, _P, C.POINTER(MultiParams)],
    "mag_chain_create": [_P, _I, C.POINTER(Action), C.POINTER(HmcParams), C.POINTER(_P)],
    "mag_chain_destroy": [_P],
    "mag_chain_set_links_aos": [_P, _D],
    "mag_chain_get_links_aos": [_P, _D],
    "mag_chain_haar_links": [_P, C.c_uint64],
    "mag_chain_links": [_P, C.POINTER(_P)],
    "mag_chain_reset_demon": [_P],
    "mag_chain_run": [_P, C.c_int, C.POINTER(TrajResult)],
    "mag_chain_reversibility": [_P, _D, _D],
    "mag_chain_demon_state": [_P, _D],
    "mag_chain_traj_index": [_P, _L],
    "mag_demon_microcanonical": [_P, C.c_int, C.c_double, C.c_uint64, _D, _L],
    "mag_demon_checkerboard": [_P, _P, _P, _P, _P, C.c_int, C.c_double, C.c_uint64, _L],
    "mag_local_metropolis": [_P, C.c_double, C.c_int, C.c_int, C.c_int, C.c_double, C.c_uint64, C.c_int, _L, _D,
                             _D],
    "mag_coarsen_links": [_P, C.c_int, _P],
    "mag_update_constrain": [_P, _P, C.c_int],
    "mag_constrained_plaquette": [_P, C.c_int, _D],
    "mag_lattice_expansion": [_P, C.c_int, _P],
    "mag_xi_energy": [_P, _P, _P, C.POINTER(MultiParams), _D],
    "mag_multi_gb": [_P, _P, C.POINTER(MultiParams)],
    "mag_multi_fforce": [_P, _P, _P, _P, C.POINTER(Action), C.POINTER(MultiParams)],
    "mag_multi_cforce": [_P, _P, _P, C.POINTER(MultiParams)],
    "mag_multi_heatbath": [_P, _P, C.POINTER(MultiParams), C.c_uint64, C.c_uint64],
    "mag_multi_zeta_T": [_P, _P, _P, C.c_double, C.c_int, C.POINTER(MultiParams)],
    "mag_multi_integrate": [_P, _P, _P, _P, C.POINTER(Action), C.POINTER(MultiParams), C.c_int, C.c_int,
                            C.c_double, _D],
    "mag_multi_hamiltonian": [_P, _P, _P, _P, C.POINTER(Action), C.POINTER(MultiParams), _D, _D],
    "mag_multi_create": [_P, _I, C.POINTER(Action), C.POINTER(HmcParams), C.POINTER(MultiParams), C.POINTER(_P)],
    "mag_multi_destroy": [_P],
    "mag_multi_fields": [_P, C.POINTER(_P), C.POINTER(_P)],
    "mag_multi_run": [_P, C.c_int, C.POINTER(MultiResult)],
    "mag_nersc_write": [C.c_char_p, _D, C.POINTER(NerscHeader)],
    "mag_nersc_read": [C.c_char_p, _D, C.POINTER(NerscHeader)],
    "mag_export_config_nersc": [_P, C.c_char_p, C.c_double, C.c_int, C.c_char_p, C.c_int],
    "mag_import_config_nersc": [_P, C.c_char_p, C.POINTER(NerscHeader)],
    "mag_server_footprint": [C.POINTER(ServeRequest), C.POINTER(C.c_size_t)],
    "mag_server_create": [_P, C.c_size_t, C.c_int, C.c_int, C.POINTER(_P)],
    "mag_server_destroy": [_P],
    "mag_server_submit": [_P, C.POINTER(ServeRequest), _I],
    "mag_server_cancel": [_P, C.c_int],
    "mag_server_step": [_P, _I, _I],
    "mag_server_status": [_P, C.c_int, _I, _I],
    "mag_server_results": [_P, C.c_int, C.POINTER(TrajResult), _D, C.c_int],
    "mag_server_info": [_P, C.POINTER(C.c_size_t), _L, _D],
    "mag_server_host_times": [_P, _D, _L],
}


def exported_names():
    """Entry points declared in include/mag_amd.h, in declaration order."""
    return list(SIGNATURES)


_lib = None


def lib():
    """Load libmag_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `make -C mag_amd` "
                              "or __graft_entry__.build(); there is no CPU fallback")
        l = C.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            if not hasattr(l, name):
                continue
            f = getattr(l, name)
            f.argtypes = args
            f.restype = C.c_char_p if name == "mag_last_error" else C.c_int
        _lib = l
    return _lib


class MagError(RuntimeError):
    def __init__(self, code, msg=""):
        self.code = code
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {msg}")
