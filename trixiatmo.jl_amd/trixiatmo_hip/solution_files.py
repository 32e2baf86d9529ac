"""SaveSolutionCallback / SaveRestartCallback files in Trixi's HDF5 layout.

`save_solution_file` writes what Trixi's and TrixiAtmo's `save_solution_file`
write (src/callbacks_step/save_solution_2d_manifold_in_3d_cartesian.jl:44-120
for the covariant form on P4estMesh{2}; Trixi's P4estMesh{3} and
StructuredMesh writers [ext] store the same attributes and datasets): root
attributes ndims, equations, polydeg, n_vars, n_elements, mesh_type,
mesh_file, time, dt, timestep and one dataset `variables_<v>` per output
variable with its `name`.  The output variables (`solution_variables`) are
computed on the GPU from the resident state.  `save_restart_file` /
`load_restart_file` write and read the conservative state with the same
layout (Trixi's SaveRestartCallback files), so a run can be resumed.

The files are written by libatmo_h5.so over the HDF5 C library; where that
library is not built, these functions raise ImportError.
"""
import ctypes
import os

import numpy as np

from ._lib import check, lib

H5_PATH = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "libatmo_h5.so")


class H5Attributes(ctypes.Structure):
    _fields_ = [("ndims", ctypes.c_int64), ("polydeg", ctypes.c_int64),
                ("n_vars", ctypes.c_int64), ("n_elements", ctypes.c_int64),
                ("timestep", ctypes.c_int64), ("time", ctypes.c_double), ("dt", ctypes.c_double),
                ("equations", ctypes.c_char * 128), ("mesh_type", ctypes.c_char * 64),
                ("mesh_file", ctypes.c_char * 256)]


_h5 = None


def h5lib():
    global _h5
    if _h5 is None:
        if not os.path.exists(H5_PATH):
            raise ImportError(f"{H5_PATH} is missing (built by build_lib.py where HDF5 exists)")
        L = ctypes.CDLL(H5_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        L.atmo_h5_write_solution.restype = ctypes.c_int
        L.atmo_h5_write_solution.argtypes = [ctypes.c_char_p, ctypes.POINTER(H5Attributes),
                                             ctypes.c_int64, dp, ctypes.POINTER(ctypes.c_char_p)]
        L.atmo_h5_read_attributes.restype = ctypes.c_int
        L.atmo_h5_read_attributes.argtypes = [ctypes.c_char_p, ctypes.POINTER(H5Attributes)]
        L.atmo_h5_read_variable.restype = ctypes.c_int
        L.atmo_h5_read_variable.argtypes = [ctypes.c_char_p, ctypes.c_int32, ctypes.c_int64, dp,
                                            ctypes.c_char_p, ctypes.c_int32]
        L.atmo_h5_last_error.restype = ctypes.c_char_p
        L.atmo_h5_last_error.argtypes = []
        _h5 = L
    return _h5


def _h5check(status):
    if status != 0:
        raise OSError(h5lib().atmo_h5_last_error().decode())


# Trixi.varnames of each equation system's output choices
VARNAMES = {
    ("CompressibleEulerEnergyEquationsWithGravity3D", "cons2cons"):
        ("rho", "rho_v1", "rho_v2", "rho_v3", "rho_e", "phi"),
    ("CompressibleEulerEnergyEquationsWithGravity3D", "cons2prim"):
        ("rho", "v1", "v2", "v3", "p", "phi"),
    ("covariant", "cons2cons"): ("h", "h_vcon1", "h_vcon2"),
    ("covariant", "cons2prim_and_vorticity"): ("H", "v1", "v2", "v3", "b", "vorticity"),
    ("CompressibleRainyEulerEquations2D", "cons2cons"):
        ("rho_dry", "rho_moist", "rho_rain", "rho_v1", "rho_v2", "energy_density",
         "rho_vapour", "rho_cloud", "temperature"),
    ("CompressibleRainyEulerEquations2D", "cons2prim"):
        ("rho_dry", "rho_moist", "rho_rain", "v1", "v2", "energy_density", "rho_vapour",
         "rho_cloud", "temperature"),
}
# the kind argument of atmo_solution_variables
KIND = {"cons2cons": 0, "cons2prim": 1, "cons2prim_and_vorticity": 1}


def _family(semi):
    name = type(semi.equations).__name__
    return "covariant" if getattr(semi, "is_covariant", False) else name


def solution_variables(semi, solution_variables="cons2cons"):
    """the output variables of the resident state, computed on the GPU:
    [n_vars, n_nodes] in Trixi's node order, and their names"""
    key = (_family(semi), solution_variables)
    if key not in VARNAMES:
        raise ValueError(f"{solution_variables} is not an output of {key[0]}")
    names = VARNAMES[key]
    n_nodes = semi.ndofs
    out = np.empty(len(names) * n_nodes)
    nv = ctypes.c_int32()
    check(lib().atmo_solution_variables(semi.handle, KIND[solution_variables],
                                        out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                        ctypes.byref(nv)), semi.handle)
    assert nv.value == len(names)
    return out.reshape(len(names), n_nodes), names


def _attributes(semi, n_vars, time, dt, timestep, mesh_file):
    mesh = semi.mesh
    ndims = 3 if type(mesh).__name__ == "P4estMesh3D" else 2
    mesh_type = "StructuredMesh" if type(mesh).__name__ == "StructuredMesh2D" else "P4estMesh"
    n_elements = semi.shape[-1]
    return H5Attributes(ndims=ndims, polydeg=mesh.polydeg, n_vars=n_vars, n_elements=n_elements,
                        timestep=int(timestep), time=float(time), dt=float(dt),
                        equations=type(semi.equations).__name__.encode(),
                        mesh_type=mesh_type.encode(), mesh_file=mesh_file.encode())


def write_file(filename, attributes, data, names):
    data = np.ascontiguousarray(data, dtype=np.float64)
    arr = (ctypes.c_char_p * len(names))(*[n.encode() for n in names])
    _h5check(h5lib().atmo_h5_write_solution(
        filename.encode(), ctypes.byref(attributes), data.shape[1] if data.ndim == 2 else 0,
        data.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), arr))
    return filename


def save_solution_file(semi, time, dt, timestep, solution_variables="cons2cons",
                       output_directory="out", system="", mesh_file=""):
    """Trixi.save_solution_file of the resident state: `solution_XXXXXXXXX.h5`
    (or `solution_<system>_XXXXXXXXX.h5`) in output_directory; returns the
    file name."""
    data, names = globals()["solution_variables"](semi, solution_variables)
    os.makedirs(output_directory, exist_ok=True)
    stem = f"solution_{system}_{timestep:09d}.h5" if system else f"solution_{timestep:09d}.h5"
    filename = os.path.join(output_directory, stem)
    return write_file(filename, _attributes(semi, len(names), time, dt, timestep, mesh_file),
                      data, names)


def save_restart_file(semi, time, dt, timestep, output_directory="out", mesh_file=""):
    """Trixi.save_restart_file of the resident state (conservative
    variables): `restart_XXXXXXXXX.h5`; returns the file name."""
    data, names = solution_variables(semi, "cons2cons")
    os.makedirs(output_directory, exist_ok=True)
    filename = os.path.join(output_directory, f"restart_{timestep:09d}.h5")
    return write_file(filename, _attributes(semi, len(names), time, dt, timestep, mesh_file),
                      data, names)


def read_file(filename):
    """(attributes as a dict, data [n_vars, n_values], names) of a solution
    or restart file"""
    a = H5Attributes()
    _h5check(h5lib().atmo_h5_read_attributes(filename.encode(), ctypes.byref(a)))
    attrs = {f: getattr(a, f) for f, _ in H5Attributes._fields_}
    for f in ("equations", "mesh_type", "mesh_file"):
        attrs[f] = attrs[f].decode()
    return attrs


def load_solution_file(filename, n_values):
    """attributes, data [n_vars, n_values] and variable names of a file"""
    attrs = read_file(filename)
    data = np.empty((attrs["n_vars"], n_values))
    names = []
    buf = ctypes.create_string_buffer(256)
    for v in range(attrs["n_vars"]):
        _h5check(h5lib().atmo_h5_read_variable(
            filename.encode(), v + 1, n_values,
            data[v].ctypes.data_as(ctypes.POINTER(ctypes.c_double)), buf, 256))
        names.append(buf.value.decode())
    return attrs, data, names


def load_restart_file(semi, filename):
    """Trixi.load_restart_file into the resident state and its time; checks
    that the file matches the semidiscretization (equations, polydeg,
    n_elements, variables); returns the file's attributes."""
    attrs = read_file(filename)
    n_nodes = semi.ndofs
    want = VARNAMES[(_family(semi), "cons2cons")]
    if (attrs["equations"] != type(semi.equations).__name__ or
            attrs["polydeg"] != semi.mesh.polydeg or attrs["n_elements"] != semi.shape[-1] or
            attrs["n_vars"] != len(want)):
        raise ValueError(f"{filename} does not match this semidiscretization: {attrs}")
    _, data, names = load_solution_file(filename, n_nodes)
    if tuple(names) != want:
        raise ValueError(f"{filename} holds {names}, not the conservative variables {want}")
    # variable-major file data -> the node-major resident state
    semi.set_state(np.ascontiguousarray(data.T).ravel())
    semi.set_time(attrs["time"])
    return attrs
