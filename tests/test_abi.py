"""The C-ABI library loads on a machine without a GPU, exports every function
declared in include/atmo_hip.h, and rejects bad configurations before it
touches a device."""
import ctypes
import os
import re

import numpy as np
import pytest

import trixiatmo_hip as T
from trixiatmo_hip import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(atmo_[a-z0-9_]+)\s*\(", text))
    return names


def test_every_declared_symbol_is_exported():
    L = _lib.lib()
    names = declared_functions()
    assert len(names) >= 20
    for name in sorted(names):
        assert hasattr(L, name), name
    assert names == {s[0] for s in _lib.SIGNATURES}


def test_invalid_configurations_are_rejected():
    mesh, eq, solver, bcs, u0 = T.workloads.periodic_cube(trees=2, polydeg=2)
    bad = [dict(polydeg=9), dict(volume_flux=7), dict(surface_flux=9), dict(precision=5),
           dict(c_v=2000.0), dict(n_elements=0), dict(source_terms=4)]
    for kw in bad:
        cfg = _lib.AtmoConfig(equations=1, polydeg=2, n_elements=mesh.n_elements, c_p=1004.0,
                              c_v=717.0, gravity=9.81)
        for k, v in kw.items():
            setattr(cfg, k, v)
        geo = _lib.AtmoGeometry3D()
        h = ctypes.c_void_p()
        rc = _lib.lib().atmo_create(ctypes.byref(cfg), ctypes.byref(geo), ctypes.byref(h))
        assert rc == -1, kw
        assert h.value is None


def test_malformed_connectivity_rejected():
    mesh, eq, solver, bcs, u0 = T.workloads.periodic_cube(trees=2, polydeg=2)
    cfg = _lib.AtmoConfig(equations=1, polydeg=2, n_elements=mesh.n_elements, c_p=1004.0,
                          c_v=717.0, gravity=9.81)
    cv = np.ascontiguousarray(mesh.contravariant_vectors.ravel(order="F"))
    ij = np.ascontiguousarray(mesh.inverse_jacobian.ravel(order="F"))
    ids = np.ascontiguousarray(mesh.interface_neighbor_ids.ravel(order="F"))
    idx = np.ascontiguousarray(mesh.interface_node_indices.ravel(order="F")).copy()
    for mutate in ("drop", "code", "elem"):
        i2, x2 = ids.copy(), idx.copy()
        n_if = mesh.n_interfaces
        if mutate == "drop":
            n_if -= 1
        elif mutate == "code":
            x2[0] = 9
        else:
            i2[0] = 1000
        geo = _lib.AtmoGeometry3D(contravariant_vectors=_lib.dptr(cv), inverse_jacobian=_lib.dptr(ij),
                                  n_interfaces=n_if, interface_neighbor_ids=_lib.iptr(i2),
                                  interface_node_indices=_lib.iptr(x2), n_boundaries=0)
        h = ctypes.c_void_p()
        assert _lib.lib().atmo_create(ctypes.byref(cfg), ctypes.byref(geo), ctypes.byref(h)) == -1


def test_mesh_geometry_closed_forms():
    """affine brick: exact metric terms; shell: volume of the spherical shell"""
    m = T.P4estMesh((2, 3, 4), polydeg=3, coordinates_min=(0, 0, 0), coordinates_max=(1, 3, 2),
                    periodicity=(True, False, True))
    hx, hy, hz = 0.5 / 2, 1.0 / 2, 0.5 / 2  # half element widths
    J = hx * hy * hz
    np.testing.assert_allclose(1 / m.inverse_jacobian, J, rtol=1e-13)
    cv = m.contravariant_vectors
    np.testing.assert_allclose(cv[0, 0], hy * hz, rtol=1e-12)
    np.testing.assert_allclose(cv[1, 1], hx * hz, rtol=1e-12)
    np.testing.assert_allclose(cv[2, 2], hx * hy, rtol=1e-12)
    for a, b in [(0, 1), (0, 2), (1, 0), (1, 2), (2, 0), (2, 1)]:
        assert np.abs(cv[a, b]).max() < 1e-13
    assert m.n_boundaries == 2 * 2 * 4  # the non-periodic y faces
    s = T.P4estMeshCubedSphereTopography(4, 3, 6.371e6, 30000.0, polydeg=5)
    _, w, _, _ = T.lgl_basis(5)
    W = w[:, None, None] * w[None, :, None] * w[None, None, :]
    vol = (W[..., None] / s.inverse_jacobian).sum()
    exact = 4 / 3 * np.pi * ((6.371e6 + 30000) ** 3 - 6.371e6 ** 3)
    assert abs(vol / exact - 1) < 1e-9
    r = np.sqrt((s.node_coordinates ** 2).sum(0))
    assert r.min() > 6.371e6 * (1 - 1e-15) and r.max() < (6.371e6 + 30000) * (1 + 1e-15)
    assert s.n_elements == 6 * 16 * 3
    assert s.n_boundaries == 2 * 6 * 16
    assert s.n_interfaces == (6 * 16 * 3 * 6 - s.n_boundaries) // 2
