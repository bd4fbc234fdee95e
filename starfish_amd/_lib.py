"""ctypes binding of libstarfish_hip.so (the C-ABI declared in include/starfish_hip.h).

The library is loaded from this package directory.  There is no fallback: if the shared
library is missing the import of this module raises, so no caller can silently run a CPU path.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libstarfish_hip.so")

SF_OK = 0
SF_WARN_NOT_CONVERGED = 1
SF_WARN_MOMENT_RANGE = 2
SF_NOALLOC_MAX_ADD = 65536   # particles per sf_particles_add_noalloc call

DOMAIN = {"XY": 0, "RZ": 1, "ZR": 2}
FACE = {"OPEN": -1, "DIRICHLET": 0, "NEUMANN": 1, "PERIODIC": 2, "SYMMETRY": 3, "MESH": 4, "SINK": 5,
        "CIRCUIT": 6}
NODE = {"UNKNOWN": -2, "OPEN": -1, "DIRICHLET": 0, "INSULATOR": 1, "INTERNAL": 2}
BND = {"DIRICHLET": 0, "OPEN": 1, "SINK": 2, "VIRTUAL": 3, "INSULATOR": 4}

# species field selectors of sf_species_field
FIELD = {"nd": 0, "u": 1, "v": 2, "w": 3, "count-sum": 4, "u-sum": 5, "v-sum": 6, "w-sum": 7, "uu-sum": 8,
         "vv-sum": 9, "ww-sum": 10, "mpc-sum": 11, "t": 12, "t1": 13, "t2": 14, "t3": 15, "nd-ave": 16,
         "u-ave": 17, "v-ave": 18, "w-ave": 19, "p": 20, "mpc": 21}


class MeshDesc(C.Structure):
    _fields_ = [("domain", C.c_int), ("ni", C.c_int), ("nj", C.c_int), ("x0", C.c_double * 2),
                ("dh", C.c_double * 2)]


class SpeciesDesc(C.Structure):
    _fields_ = [("charge", C.c_double), ("molwt", C.c_double), ("spwt", C.c_double), ("nd_back", C.c_double),
                ("vmax", C.c_double), ("capacity", C.c_int64), ("track_ids", C.c_int), ("init_T", C.c_double)]


class PoissonDesc(C.Structure):
    _fields_ = [("linear", C.c_int), ("n0", C.c_double), ("Te0", C.c_double), ("phi0", C.c_double),
                ("eps_r", C.c_double), ("max_it", C.c_int), ("rel_tol", C.c_double), ("nl_max_it", C.c_int),
                ("nl_tol", C.c_double)]


class SolveInfo(C.Structure):
    _fields_ = [("newton_iterations", C.c_int), ("pcg_iterations", C.c_int), ("rel_residual", C.c_double),
                ("ref_norm", C.c_double), ("ms", C.c_double)]


class MccDesc(C.Structure):
    _fields_ = [("source_species", C.c_int), ("model", C.c_int), ("sigma", C.c_double), ("target_molwt", C.c_double),
                ("frequency", C.c_int), ("stop_it", C.c_int), ("max_target_temp", C.c_double),
                ("seed", C.c_uint64), ("tag", C.c_uint32)]


class DsmcDesc(C.Structure):
    _fields_ = [("species_a", C.c_int), ("species_b", C.c_int), ("sigma_type", C.c_int), ("sigma_c0", C.c_double),
                ("diam_a", C.c_double), ("diam_b", C.c_double), ("ref_temp_a", C.c_double), ("ref_temp_b", C.c_double),
                ("visc_temp_index_a", C.c_double), ("visc_temp_index_b", C.c_double), ("vss_alpha_a", C.c_double),
                ("vss_alpha_b", C.c_double), ("sig_cr_max0", C.c_double), ("frequency", C.c_int),
                ("seed", C.c_uint64), ("tag", C.c_uint32)]


P = C.c_void_p
DP = C.POINTER(C.c_double)
IP = C.POINTER(C.c_int)
FP = C.POINTER(C.c_float)
I64 = C.c_int64

# name -> (restype, argtypes); every symbol declared in include/starfish_hip.h
SIGNATURES = {
    "sf_version": (C.c_char_p, []),
    "sf_ctx_create": (C.c_int, [C.c_int, C.POINTER(P)]),
    "sf_ctx_destroy": (C.c_int, [P]),
    "sf_last_error": (C.c_char_p, [P]),
    "sf_ctx_set_dt": (C.c_int, [P, C.c_double]),
    "sf_ctx_set_rng": (C.c_int, [P, I64, C.c_int]),
    "sf_ctx_get_rng_state": (C.c_int, [P, C.POINTER(I64)]),
    "sf_ctx_synchronize": (C.c_int, [P]),
    "sf_ctx_set_it": (C.c_int, [P, C.c_int]),
    "sf_geometry_build": (C.c_int, [P, C.POINTER(MeshDesc), IP, DP, C.c_int, IP, DP, IP, IP, DP, IP,
                                    C.POINTER(P)]),
    "sf_geometry_sizes": (C.c_int, [P, IP, IP]),
    "sf_geometry_export": (C.c_int, [P, IP, DP, DP, DP, IP, IP, IP, IP]),
    "sf_geometry_segment_boundaries": (C.c_int, [P, IP]),
    "sf_mesh_segment_boundaries": (C.c_int, [P, C.c_int, IP, C.c_int]),
    "sf_geometry_free": (C.c_int, [P]),
    "sf_mesh_setup": (C.c_int, [P, C.POINTER(MeshDesc), IP, DP, IP, DP, DP, C.c_int, DP, IP, IP, IP, IP]),
    "sf_species_create": (C.c_int, [P, C.POINTER(SpeciesDesc), IP]),
    "sf_particles_add": (C.c_int, [P, C.c_int, I64, DP, DP, DP, DP, DP, IP, C.c_int]),
    "sf_particles_add_noalloc": (C.c_int, [P, C.c_int, I64, DP, DP, DP, DP, DP, IP, C.c_int]),
    "sf_particles_count": (C.c_int, [P, C.c_int, C.POINTER(I64)]),
    "sf_particles_download": (C.c_int, [P, C.c_int, DP, DP, DP, DP, DP, DP, DP, IP]),
    "sf_species_restart_size": (C.c_int, [P, C.c_int, C.POINTER(C.c_int64)]),
    "sf_species_restart_write": (C.c_int, [P, C.c_int, C.c_void_p, C.c_int64]),
    "sf_species_restart_read": (C.c_int, [P, C.c_int, C.c_void_p, C.c_int64, C.POINTER(C.c_int64)]),
    "sf_particles_preload": (C.c_int, [P, C.c_int, I64, DP, C.c_double, C.c_uint64]),
    "sf_push_deposit": (C.c_int, [P, C.c_int]),
    "sf_species_set_sort_interval": (C.c_int, [P, C.c_int, C.c_int]),
    "sf_species_field": (C.c_int, [P, C.c_int, C.c_int, DP]),
    "sf_species_clear_samples": (C.c_int, [P, C.c_int]),
    "sf_species_stats": (C.c_int, [P, C.c_int, DP]),
    "sf_species_boundary_hits": (C.c_int, [P, C.c_int, C.c_int, C.POINTER(C.c_int64), C.c_int]),
    "sf_poisson_setup": (C.c_int, [P, C.POINTER(PoissonDesc)]),
    "sf_rho_update": (C.c_int, [P]),
    "sf_rho_set": (C.c_int, [P, DP]),
    "sf_rho_get": (C.c_int, [P, DP]),
    "sf_poisson_solve": (C.c_int, [P, DP, C.POINTER(SolveInfo)]),
    "sf_phi_set": (C.c_int, [P, DP]),
    "sf_phi_get": (C.c_int, [P, DP]),
    "sf_boltzmann_density": (C.c_int, [P, C.c_double, C.c_double, C.c_double, DP]),
    "sf_efield_update": (C.c_int, [P]),
    "sf_efield_set": (C.c_int, [P, DP, DP]),
    "sf_efield_get": (C.c_int, [P, DP, DP]),
    "sf_bfield_set": (C.c_int, [P, DP, DP]),
    "sf_mcc_create": (C.c_int, [P, C.POINTER(MccDesc), C.POINTER(C.c_void_p)]),
    "sf_mcc_destroy": (C.c_int, [P]),
    "sf_mcc_set_target": (C.c_int, [P, DP, DP, DP, DP, DP]),
    "sf_mcc_perform": (C.c_int, [P, C.c_int, C.c_double, C.POINTER(C.c_int64)]),
    "sf_mcc_field": (C.c_int, [P, C.c_int, DP]),
    "sf_mcc_clear_samples": (C.c_int, [P, C.c_int]),
    "sf_dsmc_create": (C.c_int, [P, C.POINTER(DsmcDesc), C.POINTER(C.c_void_p)]),
    "sf_dsmc_destroy": (C.c_int, [P]),
    "sf_dsmc_perform": (C.c_int, [P, C.c_int, C.POINTER(C.c_int64)]),
    "sf_dsmc_field": (C.c_int, [P, C.c_int, DP]),
    "sf_dsmc_clear_samples": (C.c_int, [P]),
    "sf_operator_apply": (C.c_int, [P, DP, DP]),
    "sf_source_maxwellian": (C.c_int, [P, C.c_int, C.c_int, DP, C.c_int, C.c_double, C.c_double, C.c_double,
                                       IP]),
    "sf_sources_sample": (C.c_int, [P, C.POINTER(I64)]),
    "sf_sources_sample_device": (C.c_int, [P, C.c_uint64, C.POINTER(C.c_int64)]),
    "sf_source_added": (C.c_int, [P, C.c_int, C.POINTER(C.c_int64)]),
    "sf_source_options": (C.c_int, [P, C.c_int, C.c_double, C.c_int, C.c_int]),
    "sf_last_timings": (C.c_int, [P, DP]),
    "sf_last_push_timings": (C.c_int, [P, DP]),
    "sf_last_push_queue": (C.c_int, [P, DP]),
    "sf_set_timing": (C.c_int, [P, C.c_int]),
    "sf_cnn_create": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_uint64, C.POINTER(P)]),
    "sf_cnn_destroy": (C.c_int, [P]),
    "sf_cnn_dims": (C.c_int, [P, IP, IP, C.POINTER(I64)]),
    "sf_cnn_snapshot": (C.c_int, [P, P]),
    "sf_cnn_add_sample": (C.c_int, [P, FP, FP, C.POINTER(C.c_uint8)]),
    "sf_cnn_get_sample": (C.c_int, [P, C.c_int, FP, FP, C.POINTER(C.c_uint8)]),
    "sf_cnn_num_samples": (C.c_int, [P]),
    "sf_cnn_train": (C.c_int, [P, C.c_int, C.c_int, C.c_double, DP]),
    "sf_cnn_loss_and_grads": (C.c_int, [P, C.c_int, DP, FP, FP, C.POINTER(I64)]),
    "sf_cnn_params": (C.c_int, [P, C.c_int, FP, C.POINTER(I64)]),
    "sf_cnn_predict": (C.c_int, [P, P]),
    "sf_poisson_solve_learned": (C.c_int, [P, P, C.c_int, C.POINTER(SolveInfo), IP]),
    "sf_cnn_activation": (C.c_int, [P, C.c_int, FP, C.POINTER(I64)]),
    "sf_cnn_mfma_probe": (C.c_int, [P, FP, FP, FP]),
}

_lib = None


def load():
    """Loads libstarfish_hip.so and binds every entry point; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libstarfish_hip.so is not built (expected at %s); run `python -m starfish_amd.build`"
                          % LIB_PATH)
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class StarfishError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("starfish_hip error %d: %s" % (code, msg))
        self.code = code


def check(ctx, rc, allow_warning=True):
    """Maps a C-ABI status to an exception (errors) or a returned status (warnings)."""
    if rc < 0:
        msg = load().sf_last_error(ctx).decode() if ctx else "no context"
        raise StarfishError(rc, msg)
    if rc > 0 and not allow_warning:
        raise StarfishError(rc, load().sf_last_error(ctx).decode() if ctx else "")
    return rc
