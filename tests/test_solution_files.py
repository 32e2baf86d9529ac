"""Trixi-compatible HDF5 solution and restart files (libatmo_h5, CPU).

The layout is the one of save_solution_file
(src/callbacks_step/save_solution_2d_manifold_in_3d_cartesian.jl:74-97): the
root attributes ndims, equations, polydeg, n_vars, n_elements, mesh_type,
mesh_file, time, dt, timestep and datasets variables_1..n_vars with a `name`
attribute, with the HDF5 types HDF5.jl writes for Int64 (64-bit little-endian
integers), Float64 (IEEE FP64) and String (scalar fixed-length UTF-8 of the
string's byte length).  Checked by a write/read round trip and by h5dump,
the HDF5 distribution's own reader, on the written file."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from trixiatmo_hip import solution_files as SF

H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump"
                                    if os.path.exists("/opt/conda/bin/h5dump") else None)


def _write(tmp_path, n_vars=3, n_values=1000, seed=0):
    rng = np.random.default_rng(seed)
    data = rng.standard_normal((n_vars, n_values))
    names = ["h", "h_vcon1", "h_vcon2"][:n_vars]
    attrs = SF.H5Attributes(ndims=2, polydeg=3, n_vars=n_vars, n_elements=n_values // 16,
                            timestep=42, time=86400.5, dt=123.25,
                            equations=b"SplitCovariantShallowWaterEquations2D",
                            mesh_type=b"P4estMesh", mesh_file=b"mesh.h5")
    fn = str(tmp_path / "solution_000000042.h5")
    SF.write_file(fn, attrs, data, names)
    return fn, data, names


def test_round_trip(tmp_path):
    fn, data, names = _write(tmp_path)
    attrs, back, back_names = SF.load_solution_file(fn, data.shape[1])
    np.testing.assert_array_equal(back, data)
    assert back_names == names
    assert attrs["ndims"] == 2 and attrs["polydeg"] == 3 and attrs["n_vars"] == 3
    assert attrs["n_elements"] == data.shape[1] // 16 and attrs["timestep"] == 42
    assert attrs["time"] == 86400.5 and attrs["dt"] == 123.25
    assert attrs["equations"] == "SplitCovariantShallowWaterEquations2D"
    assert attrs["mesh_type"] == "P4estMesh" and attrs["mesh_file"] == "mesh.h5"


def test_wrong_size_and_missing_files_are_errors(tmp_path):
    fn, data, _ = _write(tmp_path)
    with pytest.raises(OSError):
        SF.load_solution_file(fn, data.shape[1] + 1)
    with pytest.raises(OSError):
        SF.read_file(str(tmp_path / "absent.h5"))


@pytest.mark.skipif(H5DUMP is None, reason="h5dump is not installed")
def test_layout_read_by_h5dump(tmp_path):
    fn, data, names = _write(tmp_path, n_vars=2, n_values=64)
    header = subprocess.run([H5DUMP, "-H", fn], capture_output=True, text=True, check=True).stdout
    flat = " ".join(header.split())
    for a in ("ndims", "polydeg", "n_vars", "n_elements", "timestep"):
        assert f'ATTRIBUTE "{a}" {{ DATATYPE H5T_STD_I64LE DATASPACE SCALAR' in flat, a
    for a in ("time", "dt"):
        assert f'ATTRIBUTE "{a}" {{ DATATYPE H5T_IEEE_F64LE DATASPACE SCALAR' in flat, a
    eq = "SplitCovariantShallowWaterEquations2D"
    assert (f'ATTRIBUTE "equations" {{ DATATYPE H5T_STRING {{ STRSIZE {len(eq)}; '
            'STRPAD H5T_STR_NULLTERM; CSET H5T_CSET_UTF8; CTYPE H5T_C_S1; } '
            'DATASPACE SCALAR') in flat
    for v in (1, 2):
        assert (f'DATASET "variables_{v}" {{ DATATYPE H5T_IEEE_F64LE '
                'DATASPACE SIMPLE { ( 64 ) / ( 64 ) }') in flat
    values = subprocess.run([H5DUMP, "-d", "/variables_2", "-A", "0", "-m", "%.17g", fn],
                            capture_output=True, text=True, check=True).stdout
    body = values.split("DATA {", 1)[1].split("}", 1)[0]
    nums = [float(t.split(":")[-1]) for t in body.replace(",", " ").split() if ":" not in t or
            t.split(":")[-1]]
    np.testing.assert_array_equal(np.array(nums[-64:]), data[1])
    name = subprocess.run([H5DUMP, "-a", "/variables_2/name", fn], capture_output=True, text=True,
                          check=True).stdout
    assert f'"{names[1]}"' in name
