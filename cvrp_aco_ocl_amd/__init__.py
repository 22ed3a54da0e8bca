"""MI355X-native ant colony optimizer for the CVRP.

The per-iteration loop of i-sunny/cvrp_aco_ocl (construct_solutions ->
local_search -> update_statistics -> pheromone_update,
cvrp_ocl/gpu/g_aco.cpp:122-134) as hand-written HIP kernels for gfx950
behind the C ABI of ``include/cvrp_aco.h`` (libcvrpaco.so).  This package
is the Python side of that boundary: ``GpuACO`` mirrors ``g_ACO``,
``Instance`` the reference's ``Problem``.
"""
from ._lib import (  # noqa: F401
    DIST_ATT, DIST_CEIL, DIST_FLOAT, DIST_GEO, DIST_NINT, LS_2OPT, LS_2OPT_RANDOM, LS_RELOCATE, LS_SWAP,
    RNG_PARK_MILLER, RNG_PHILOX, SEL_IROULETTE, SEL_KNN, TAU0_FIRST_COLONY, TAU0_NN_TOUR,
    AcoError, LIB_PATH,
)
from .solver import GpuACO, Instance  # noqa: F401
from . import instances  # noqa: F401
from . import decompose  # noqa: F401
