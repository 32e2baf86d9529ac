"""ctypes binding of libatmo_hip.so (the C ABI declared in include/atmo_hip.h).

The library must have been built (``python trixiatmo.jl_amd/build_lib.py`` or
``__graft_entry__.build()``); there is no CPU fallback: every compute call
goes to the HIP kernels and fails loudly when they are unavailable.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libatmo_hip.so")

c_int32_p = ctypes.POINTER(ctypes.c_int32)
c_double_p = ctypes.POINTER(ctypes.c_double)
c_int64_p = ctypes.POINTER(ctypes.c_int64)

ATMO_OK = 0
STATUS_NAMES = {
    -1: "ATMO_ERR_INVALID",
    -2: "ATMO_ERR_DEVICE",
    -3: "ATMO_ERR_NONFINITE",
    -4: "ATMO_ERR_NEWTON",
    -5: "ATMO_ERR_MESH",
}


class AtmoConfig(ctypes.Structure):
    _fields_ = [
        ("equations", ctypes.c_int32),
        ("polydeg", ctypes.c_int32),
        ("n_elements", ctypes.c_int32),
        ("volume_flux", ctypes.c_int32),
        ("volume_noncons", ctypes.c_int32),
        ("surface_flux", ctypes.c_int32),
        ("surface_noncons", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("c_p", ctypes.c_double),
        ("c_v", ctypes.c_double),
        ("gravity", ctypes.c_double),
        ("lmars_c", ctypes.c_double),
        ("source_terms", ctypes.c_int32),
        ("rotation_rate", ctypes.c_double),
        ("background_p0", ctypes.c_double),
        ("background_T0", ctypes.c_double),
    ]


class AtmoGeometry3D(ctypes.Structure):
    _fields_ = [
        ("contravariant_vectors", c_double_p),
        ("inverse_jacobian", c_double_p),
        ("n_interfaces", ctypes.c_int32),
        ("interface_neighbor_ids", c_int32_p),
        ("interface_node_indices", c_int32_p),
        ("n_boundaries", ctypes.c_int32),
        ("boundary_element", c_int32_p),
        ("boundary_node_indices", c_int32_p),
        ("boundary_condition", c_int32_p),
        ("node_coordinates", c_double_p),
        ("period", c_double_p),
        ("boundary_condition_by_kind", ctypes.c_int32 * 3),
    ]


class AtmoSweConfig(ctypes.Structure):
    _fields_ = [
        ("polydeg", ctypes.c_int32),
        ("n_elements", ctypes.c_int32),
        ("volume_flux", ctypes.c_int32),
        ("surface_flux", ctypes.c_int32),
        ("surface_noncons", ctypes.c_int32),
        ("source_terms", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("gravity", ctypes.c_double),
        ("rotation_rate", ctypes.c_double),
    ]


class AtmoGeometry2D(ctypes.Structure):
    _fields_ = [
        ("aux_node_vars", c_double_p),
        ("node_coordinates", c_double_p),
        ("n_interfaces", ctypes.c_int32),
        ("interface_neighbor_ids", c_int32_p),
        ("interface_node_indices", c_int32_p),
        ("n_members", ctypes.c_int32),
        ("member_first_element", c_int32_p),
        ("christoffel_collocation", ctypes.c_int32),
        ("bottom_topography", c_double_p),
    ]


class AtmoRainyConfig(ctypes.Structure):
    _fields_ = [
        ("polydeg", ctypes.c_int32),
        ("cells_x", ctypes.c_int32),
        ("cells_y", ctypes.c_int32),
        ("x_min", ctypes.c_double),
        ("y_min", ctypes.c_double),
        ("x_max", ctypes.c_double),
        ("y_max", ctypes.c_double),
        ("periodic_y", ctypes.c_int32),
        ("volume_integral", ctypes.c_int32),
        ("surface_flux", ctypes.c_int32),
        ("source_terms", ctypes.c_int32),
        ("newton_tolerance", ctypes.c_double),
        ("newton_max_iterations", ctypes.c_int32),
    ]


class AtmoNewtonStats(ctypes.Structure):
    _fields_ = [("nodes", ctypes.c_int64), ("updates", ctypes.c_int64),
                ("max_updates", ctypes.c_int64), ("failures", ctypes.c_int64),
                ("fallbacks", ctypes.c_int64)]


c_float_p = ctypes.POINTER(ctypes.c_float)

# (name, restype, argtypes) of every exported symbol of include/atmo_hip.h
SIGNATURES = [
    ("atmo_lgl_basis", ctypes.c_int, [ctypes.c_int, c_double_p, c_double_p, c_double_p, c_double_p]),
    ("atmo_mesh3d_cartesian", ctypes.c_int,
     [ctypes.c_int, c_int32_p, c_double_p, c_double_p, c_int32_p, ctypes.POINTER(ctypes.c_void_p)]),
    ("atmo_mesh3d_cubed_sphere_shell", ctypes.c_int,
     [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
      ctypes.POINTER(ctypes.c_void_p)]),
    ("atmo_mesh3d_sizes", ctypes.c_int, [ctypes.c_void_p, c_int32_p, c_int32_p, c_int32_p]),
    ("atmo_mesh3d_arrays", ctypes.c_int,
     [ctypes.c_void_p] + [ctypes.POINTER(c_double_p)] * 3 + [ctypes.POINTER(c_int32_p)] * 5),
    ("atmo_mesh3d_destroy", None, [ctypes.c_void_p]),
    ("atmo_mesh2d_cubed_sphere", ctypes.c_int,
     [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p)]),
    ("atmo_mesh2d_sizes", ctypes.c_int, [ctypes.c_void_p, c_int32_p, c_int32_p]),
    ("atmo_mesh2d_arrays", ctypes.c_int,
     [ctypes.c_void_p] + [ctypes.POINTER(c_double_p)] * 2 + [ctypes.POINTER(c_int32_p)] * 2),
    ("atmo_mesh2d_destroy", None, [ctypes.c_void_p]),
    ("atmo_swe_create", ctypes.c_int,
     [ctypes.POINTER(AtmoSweConfig), ctypes.POINTER(AtmoGeometry2D), ctypes.POINTER(ctypes.c_void_p)]),
    ("atmo_max_dt_members", ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, c_double_p]),
    ("atmo_step_members", ctypes.c_int, [ctypes.c_void_p, c_double_p, ctypes.c_int32]),
    ("atmo_autotune", ctypes.c_int, [ctypes.c_void_p, c_int32_p, c_double_p]),
    ("atmo_set_kernel_variant", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    ("atmo_rainy_create", ctypes.c_int,
     [ctypes.POINTER(AtmoRainyConfig), ctypes.POINTER(ctypes.c_void_p)]),
    ("atmo_set_time", ctypes.c_int, [ctypes.c_void_p, ctypes.c_double]),
    ("atmo_saturation_newton", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.POINTER(AtmoNewtonStats)]),
    ("atmo_enable_newton_stats", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    ("atmo_newton_stats_read", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(AtmoNewtonStats)]),
    ("atmo_rainy_features", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("atmo_rainy_attach_mlp", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ("atmo_rainy_seed_stage_limiter", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    ("atmo_mlp_create", ctypes.c_int, [ctypes.c_uint64, ctypes.POINTER(ctypes.c_void_p)]),
    ("atmo_mlp_destroy", None, [ctypes.c_void_p]),
    ("atmo_mlp_set_normalization", ctypes.c_int, [ctypes.c_void_p, c_float_p]),
    ("atmo_mlp_set_params", ctypes.c_int, [ctypes.c_void_p, c_float_p]),
    ("atmo_mlp_get_params", ctypes.c_int, [ctypes.c_void_p, c_float_p]),
    ("atmo_mlp_loss_grad", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, c_float_p, c_float_p]),
    ("atmo_mlp_train", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
      ctypes.c_int32, ctypes.c_int32, ctypes.c_float, c_float_p]),
    ("atmo_mlp_forward", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]),
    ("atmo_create", ctypes.c_int,
     [ctypes.POINTER(AtmoConfig), ctypes.POINTER(AtmoGeometry3D), ctypes.POINTER(ctypes.c_void_p)]),
    ("atmo_destroy", None, [ctypes.c_void_p]),
    ("atmo_last_error", ctypes.c_char_p, [ctypes.c_void_p]),
    ("atmo_state_length", ctypes.c_int64, [ctypes.c_void_p]),
    ("atmo_rhs", ctypes.c_int, [ctypes.c_void_p, c_double_p, c_double_p, ctypes.c_double]),
    ("atmo_rhs_device", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double]),
    ("atmo_set_state", ctypes.c_int, [ctypes.c_void_p, c_double_p]),
    ("atmo_get_state", ctypes.c_int, [ctypes.c_void_p, c_double_p]),
    ("atmo_step", ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_int32]),
    ("atmo_max_dt", ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, c_double_p]),
    ("atmo_mesh3d_connectivity_device", ctypes.c_int,
     [ctypes.c_int, ctypes.c_int32, ctypes.c_void_p, c_double_p, ctypes.c_void_p, ctypes.c_void_p,
      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, c_int32_p, c_int32_p]),
    ("atmo_mesh2d_connectivity_device", ctypes.c_int,
     [ctypes.c_int, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
      c_int32_p]),
    ("atmo_covariant_aux_device", ctypes.c_int, [ctypes.c_int, ctypes.c_int32, ctypes.c_void_p,
                                                 ctypes.c_int32, ctypes.c_void_p,
                                                 ctypes.c_void_p]),
    ("atmo_max_dt3d_noalloc", ctypes.c_int, [ctypes.POINTER(AtmoConfig), ctypes.c_void_p,
                                             ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                             ctypes.c_void_p, ctypes.c_void_p]),
    ("atmo_solution_variables", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int32, c_double_p, c_int32_p]),
    ("atmo_metric_terms3d", ctypes.c_int,
     [ctypes.c_int, ctypes.c_int32, c_double_p, c_double_p, c_double_p]),
    ("atmo_metric_terms3d_device", ctypes.c_int,
     [ctypes.c_int, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("atmo_solve", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_int32,
      ctypes.POINTER(ctypes.c_int64), c_double_p]),
    ("atmo_integrals", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, c_double_p]),
    ("atmo_synchronize", ctypes.c_int, [ctypes.c_void_p]),
    ("atmo_enable_timing", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    ("atmo_kernel_times", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, c_int64_p, c_double_p]),
]

_lib = None


class AtmoError(RuntimeError):
    """A libatmo_hip call returned a negative status."""

    def __init__(self, status, message):
        self.status = status
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")


def lib():
    """Load libatmo_hip.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with trixiatmo.jl_amd/build_lib.py "
                "(there is no CPU fallback)")
        # torch, when installed, brings its own copy of the HIP runtime; loading
        # it first lets this library bind to the same runtime, so device
        # tensors handed across (training samples, iteration counts) are valid
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status, handle=None):
    if status != ATMO_OK:
        msg = lib().atmo_last_error(handle)
        raise AtmoError(status, msg.decode() if msg else "")
    return status


def dptr(a):
    return a.ctypes.data_as(c_double_p)


def iptr(a):
    return a.ctypes.data_as(c_int32_p)
