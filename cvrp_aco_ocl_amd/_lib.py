"""ctypes binding of libcvrpaco.so (the C ABI declared in include/cvrp_aco.h).

The shared library is built in-tree by ``cvrp_aco_ocl_amd.build``.  Loading
fails loudly when it is missing: there is no CPU fallback for any entry
point of the solver.
"""
from __future__ import annotations

import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
# CVRP_ACO_LIB selects another build of the same library (e.g. an
# instrumented one); the default is the in-tree build.
LIB_PATH = os.environ.get("CVRP_ACO_LIB") or os.path.join(PKG, "libcvrpaco.so")

ACO_OK = 0
ERRORS = {-1: "ACO_ERR_ARG", -2: "ACO_ERR_HIP", -3: "ACO_ERR_STATE", -4: "ACO_ERR_INFEASIBLE",
          -5: "ACO_ERR_NODEV", -6: "ACO_ERR_IO"}

SEL_IROULETTE, SEL_KNN = 0, 1
RNG_PARK_MILLER, RNG_PHILOX = 0, 1
LS_2OPT, LS_SWAP, LS_RELOCATE, LS_2OPT_RANDOM = 1, 2, 4, 8
TAU0_FIRST_COLONY, TAU0_NN_TOUR = 0, 1
DIST_FLOAT, DIST_NINT, DIST_CEIL, DIST_GEO, DIST_ATT = 0, 1, 2, 3, 4


class AcoError(RuntimeError):
    def __init__(self, code: int, where: str, msg: str):
        super().__init__(f"{where}: {ERRORS.get(code, code)}: {msg}")
        self.code = code


class AcoProblem(C.Structure):
    _fields_ = [
        ("n", C.c_int32), ("dist", C.c_void_p), ("demand", C.c_void_p), ("capacity", C.c_int32),
        ("max_dist", C.c_float), ("service_time", C.c_float), ("nn_list", C.c_void_p),
        ("nn_width", C.c_int32), ("prior", C.c_void_p),
    ]


class AcoParams(C.Structure):
    _fields_ = [
        ("alpha", C.c_float), ("beta", C.c_float), ("rho", C.c_float), ("n_ants", C.c_int32),
        ("ras_ranks", C.c_int32), ("k", C.c_int32), ("nn_ls", C.c_int32),
        ("selection", C.c_int32), ("rng", C.c_int32), ("ls_flags", C.c_int32),
        ("tau0_rule", C.c_int32), ("seed", C.c_uint64), ("bsf_capacity", C.c_int32),
        ("use_graph", C.c_int32),
    ]


class AcoDecomposeParams(C.Structure):
    _fields_ = [
        ("subs", C.c_int32), ("sub_iterations", C.c_int32), ("rotation", C.c_int32),
        ("pheromone_back", C.c_int32), ("sub_params", AcoParams),
    ]


class AcoPriorTrainParams(C.Structure):
    _fields_ = [
        ("epochs", C.c_int32), ("batch", C.c_int32), ("lr", C.c_float), ("lr_final", C.c_float),
        ("pos_weight", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float),
        ("shuffle_seed", C.c_uint64),
    ]


PRIOR_FEATURES = 12
PRIOR_HIDDEN = 64
PRIOR_PARAMS = 64 * 12 + 64 + 64 * 64 + 64 + 64 + 1


class AcoLaunchShape(C.Structure):
    _fields_ = [("ls_waves", C.c_int32), ("ls_ring", C.c_int32), ("ls_prefetch", C.c_int32),
                ("construct_apb", C.c_int32), ("ls_row16", C.c_int32), ("fused_iteration", C.c_int32)]


class AcoTuneResult(C.Structure):
    _fields_ = [("shape", AcoLaunchShape), ("stage", C.c_int32), ("ms", C.c_float)]


class AcoBsfRecord(C.Structure):
    _fields_ = [("iter", C.c_int32), ("length", C.c_float), ("time", C.c_float)]


class AcoCounters(C.Structure):
    _fields_ = [("ls_evals", C.c_int64), ("ls_moves", C.c_int64), ("fallbacks", C.c_int64),
                ("iterations", C.c_int64), ("bsf_overflow", C.c_int32)]


_P = C.c_void_p
_SIGNATURES = {
    "aco_default_params": (None, [C.POINTER(AcoParams)]),
    "aco_create": (C.c_int, [C.POINTER(AcoProblem), C.POINTER(AcoParams), C.POINTER(C.c_void_p)]),
    "aco_destroy": (C.c_int, [_P]),
    "aco_create_batch": (C.c_int, [C.POINTER(AcoProblem), C.c_int32, C.POINTER(AcoParams),
                                   C.POINTER(C.c_void_p)]),
    "aco_batch_size": (C.c_int32, [_P]),
    "aco_get_best_at": (C.c_int, [_P, C.c_int32, _P, C.c_int32, C.POINTER(C.c_int32),
                                  C.POINTER(C.c_float)]),
    "aco_get_ants_at": (C.c_int, [_P, C.c_int32, _P, C.c_int32, _P, _P]),
    "aco_set_ants_at": (C.c_int, [_P, C.c_int32, _P, C.c_int32, _P, _P]),
    "aco_get_bsf_at": (C.c_int, [_P, C.c_int32, _P, C.c_int32, C.POINTER(C.c_int32)]),
    "aco_get_pheromone_at": (C.c_int, [_P, C.c_int32, _P, _P]),
    "aco_set_pheromone_at": (C.c_int, [_P, C.c_int32, _P, _P]),
    "aco_init": (C.c_int, [_P]),
    "aco_iteration": (C.c_int, [_P]),
    "aco_run": (C.c_int, [_P, C.c_int32]),
    "aco_construct": (C.c_int, [_P]),
    "aco_local_search": (C.c_int, [_P]),
    "aco_update_statistics": (C.c_int, [_P]),
    "aco_pheromone_update": (C.c_int, [_P]),
    "aco_pheromone_init": (C.c_int, [_P, C.c_float]),
    "aco_synchronize": (C.c_int, [_P]),
    "aco_get_iteration": (C.c_int, [_P, C.POINTER(C.c_int32)]),
    "aco_get_best": (C.c_int, [_P, _P, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_float)]),
    "aco_get_ants": (C.c_int, [_P, _P, C.c_int32, _P, _P]),
    "aco_set_ants": (C.c_int, [_P, _P, C.c_int32, _P, _P]),
    "aco_get_bsf": (C.c_int, [_P, _P, C.c_int32, C.POINTER(C.c_int32)]),
    "aco_get_pheromone": (C.c_int, [_P, _P, _P]),
    "aco_set_pheromone": (C.c_int, [_P, _P, _P]),
    "aco_get_rng_state": (C.c_int, [_P, _P]),
    "aco_set_rng_state": (C.c_int, [_P, _P]),
    "aco_get_ls_seed": (C.c_int, [_P, C.POINTER(C.c_int32)]),
    "aco_set_ls_seed": (C.c_int, [_P, C.c_int32]),
    "aco_get_counters": (C.c_int, [_P, C.POINTER(AcoCounters)]),
    "aco_profile_enable": (C.c_int, [_P, C.c_int32]),
    "aco_profile_get": (C.c_int, [_P, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "aco_get_launch_shape": (C.c_int, [_P, C.POINTER(AcoLaunchShape)]),
    "aco_set_launch_shape": (C.c_int, [_P, C.POINTER(AcoLaunchShape)]),
    "aco_autotune": (C.c_int, [_P, C.c_int32, _P, C.c_int32, C.POINTER(C.c_int32)]),
    "aco_last_error": (C.c_char_p, []),
    "aco_compute_distances": (C.c_int, [C.c_int32, _P, _P, C.c_int32, _P]),
    "aco_compute_nn_lists": (C.c_int, [C.c_int32, _P, C.c_int32, C.c_int32, _P]),
    "aco_compute_distances_batch": (C.c_int, [C.c_int32, _P, _P, _P, C.c_int32, _P]),
    "aco_get_setup_at": (C.c_int, [_P, C.c_int32, _P, _P, _P, _P, _P, _P, _P, _P]),
    "aco_set_candidates_at": (C.c_int, [_P, C.c_int32, _P, C.c_int32]),
    "aco_compute_nn_lists_batch": (C.c_int, [C.c_int32, _P, _P, _P, C.c_int32, _P]),
    "aco_compute_nn_lists_batch_noalloc": (C.c_int, [C.c_int32, _P, _P, _P, C.c_int32, _P, _P]),
    "aco_read_vrplib": (C.c_int, [C.c_char_p, C.c_int32, C.POINTER(C.c_int32), _P, _P, _P,
                                  C.POINTER(C.c_int32), C.POINTER(C.c_float),
                                  C.POINTER(C.c_float), C.POINTER(C.c_float),
                                  C.POINTER(C.c_int32)]),
    "aco_check_solution": (C.c_int, [C.c_int32, _P, _P, C.c_int32, C.c_float, C.c_float, _P,
                                     C.c_int32]),
    "aco_nn_tour": (C.c_int, [C.c_int32, _P, _P, C.c_int32, C.c_float, C.c_float, _P, C.c_int32,
                              C.POINTER(C.c_int32), C.POINTER(C.c_float)]),
    # learned prior (cvrp_prior.h)
    "aco_prior_default_train_params": (None, [C.POINTER(AcoPriorTrainParams)]),
    "aco_prior_create": (C.c_int, [C.c_uint64, C.POINTER(C.c_void_p)]),
    "aco_prior_destroy": (C.c_int, [_P]),
    "aco_prior_get_params": (C.c_int, [_P, _P, C.c_int32]),
    "aco_prior_set_params": (C.c_int, [_P, _P, C.c_int32]),
    "aco_prior_features": (C.c_int, [C.c_int32, _P, _P, _P, C.c_int32, _P, _P, C.c_int32,
                                     C.c_int32, _P]),
    "aco_prior_forward": (C.c_int, [_P, _P, C.c_int64, _P]),
    "aco_prior_gradient": (C.c_int, [_P, _P, _P, C.c_int64, C.c_float, _P,
                                     C.POINTER(C.c_float)]),
    "aco_prior_adam_step": (C.c_int, [_P, _P, C.c_float, C.c_float, C.c_float, C.c_float]),
    "aco_prior_train": (C.c_int, [_P, _P, _P, C.c_int64, C.POINTER(AcoPriorTrainParams), _P]),
    "aco_prior_build": (C.c_int, [_P, C.c_int32, _P, _P, _P, C.c_int32, _P, _P, C.c_int32,
                                  C.c_int32, C.c_float, C.c_float, _P]),
    "aco_prior_apply": (C.c_int, [_P, _P, _P, _P, C.c_int32, C.c_float, C.c_float]),
    # sub-problem decomposition (include/cvrp_decompose.h)
    "aco_route_centers": (C.c_int, [C.c_int32, _P, _P, _P, _P, C.c_int32, C.c_int32, _P, _P, _P, _P, _P,
                                    C.POINTER(C.c_int32)]),
    "aco_partition_routes": (C.c_int, [C.c_int32, _P, _P, _P, _P, C.c_int32, C.c_int32, C.c_int32, _P, _P,
                                       _P, _P, C.POINTER(C.c_int32)]),
    "aco_sub_problem": (C.c_int, [C.c_int32, _P, _P, C.c_int32, _P, _P, _P]),
    "aco_merge_sub_tours": (C.c_int, [C.c_int32, _P, C.c_int32, _P, _P, _P, _P, _P, _P, _P, C.c_int32,
                                      C.POINTER(C.c_int32), C.POINTER(C.c_float), _P]),
    "aco_set_ant_at": (C.c_int, [_P, C.c_int32, C.c_int32, _P, C.c_int32]),
    "aco_pheromone_from_subs": (C.c_int, [_P, C.c_int32, _P, _P, _P, _P]),
    "aco_decompose_round": (C.c_int, [_P, C.c_int32, C.POINTER(AcoProblem), _P, _P,
                                      C.POINTER(AcoDecomposeParams), _P, C.c_int32, C.POINTER(C.c_int32),
                                      C.POINTER(C.c_float), C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
}

EXPORTED = tuple(_SIGNATURES)


def load():
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"libcvrpaco.so not found at {LIB_PATH}: build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'` (no CPU fallback exists)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = load()


def check(rc: int, where: str) -> None:
    if rc != ACO_OK:
        msg = lib.aco_last_error()
        raise AcoError(rc, where, msg.decode() if msg else "")
