"""The C-ABI libraries load on a machine without a GPU and export every
function their headers declare (include/atmo_hip.h: libatmo_hip.so;
include/atmo_h5.h: libatmo_h5.so), and libatmo_hip rejects bad
configurations before it touches a device."""
import ctypes
import os
import re

import numpy as np
import pytest

import trixiatmo_hip as T
from trixiatmo_hip import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header="atmo_hip.h"):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(atmo_[a-z0-9_]+)\s*\(", text))


def test_every_declared_symbol_is_exported():
    assert sorted(f for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h")) == \
        ["atmo_h5.h", "atmo_hip.h"]
    L = _lib.lib()
    names = declared_functions()
    assert len(names) >= 20
    for name in sorted(names):
        assert hasattr(L, name), name
    assert names == {s[0] for s in _lib.SIGNATURES}


def test_h5_library_exports_its_header():
    from trixiatmo_hip import solution_files
    L = solution_files.h5lib()
    names = declared_functions("atmo_h5.h")
    assert names == {"atmo_h5_write_solution", "atmo_h5_read_attributes",
                     "atmo_h5_read_variable", "atmo_h5_last_error"}
    for name in sorted(names):
        assert hasattr(L, name), name


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


def _swe_geometry(mesh, aux, ids=None, idx=None, n_if=None, first=None):
    keep = [np.ascontiguousarray(aux.ravel(order="F")),
            np.ascontiguousarray(mesh.node_coordinates.ravel(order="F")),
            np.ascontiguousarray((mesh.interface_neighbor_ids if ids is None else ids)
                                 .ravel(order="F"), dtype=np.int32),
            np.ascontiguousarray((mesh.interface_node_indices if idx is None else idx)
                                 .ravel(order="F"), dtype=np.int32)]
    geo = _lib.AtmoGeometry2D(aux_node_vars=_lib.dptr(keep[0]), node_coordinates=_lib.dptr(keep[1]),
                              n_interfaces=mesh.n_interfaces if n_if is None else n_if,
                              interface_neighbor_ids=_lib.iptr(keep[2]),
                              interface_node_indices=_lib.iptr(keep[3]), n_members=0,
                              member_first_element=None)
    if first is not None:
        keep.append(np.asarray(first, dtype=np.int32))
        geo.n_members = len(first) - 1
        geo.member_first_element = _lib.iptr(keep[-1])
    return geo, keep


def test_covariant_invalid_configurations_and_connectivity_rejected():
    mesh = T.P4estMeshCubedSphere2D(2, 1.0, polydeg=3)
    aux = mesh.data(False).aux_node_vars
    good = dict(polydeg=3, n_elements=mesh.n_elements, volume_flux=0, surface_flux=0,
                surface_noncons=0, source_terms=1, precision=0, gravity=9.81,
                rotation_rate=7.292e-5)
    bad = [dict(polydeg=8), dict(volume_flux=2), dict(surface_flux=3), dict(surface_noncons=2),
           dict(precision=1), dict(gravity=0.0), dict(n_elements=0), dict(source_terms=2),
           dict(volume_flux=3), dict(surface_flux=4), dict(source_terms=3),
           # the standard form's weak form, flux and sources do not mix with the split form's
           dict(volume_flux=2, surface_flux=3, source_terms=1),
           dict(volume_flux=2, surface_flux=1, source_terms=2)]
    geo, keep = _swe_geometry(mesh, aux)
    for kw in bad:
        cfg = _lib.AtmoSweConfig(**{**good, **kw})
        h = ctypes.c_void_p()
        assert _lib.lib().atmo_swe_create(ctypes.byref(cfg), ctypes.byref(geo), ctypes.byref(h)) == -1, kw
        assert h.value is None
    cfg = _lib.AtmoSweConfig(**good)
    ids = mesh.interface_neighbor_ids.copy()
    idx = mesh.interface_node_indices.copy()
    i_bad = ids.copy()
    i_bad[0, 0] = mesh.n_elements
    x_bad = idx.copy()
    x_bad[0, 0, 0] = 7
    cases = [dict(n_if=mesh.n_interfaces - 1), dict(ids=i_bad), dict(idx=x_bad),
             # members must cover the elements, and no interface may join two members
             dict(first=[0, mesh.n_elements - 1]), dict(first=[0, 12, mesh.n_elements])]
    for kw in cases:
        geo, keep = _swe_geometry(mesh, aux, **kw)
        h = ctypes.c_void_p()
        assert _lib.lib().atmo_swe_create(ctypes.byref(cfg), ctypes.byref(geo), ctypes.byref(h)) == -1, kw
        msg = _lib.lib().atmo_last_error(None)
        assert msg, kw


def test_covariant_mesh_connectivity_closed_forms():
    """6 k^2 elements, 12 k^2 interfaces, each element face used once; the
    node coordinates of the two sides of every interface coincide."""
    for k in (1, 2, 3):
        mesh = T.P4estMeshCubedSphere2D(k, 2.0, polydeg=2)
        assert mesh.n_elements == 6 * k * k and mesh.n_interfaces == 12 * k * k
        n = 3
        used = set()
        x = mesh.node_coordinates

        def node(code, kk):
            return {0: 0, 1: n - 1, 2: kk, 3: n - 1 - kk}[code]

        for I in range(mesh.n_interfaces):
            ep, es = mesh.interface_neighbor_ids[:, I]
            ip, isx = mesh.interface_node_indices[:, 0, I], mesh.interface_node_indices[:, 1, I]
            for side, e, c in ((0, ep, ip), (1, es, isx)):
                face = (int(e), int(c[0]) if c[0] <= 1 else 2 + int(c[1]))
                assert face not in used
                used.add(face)
            for kk in range(n):
                a = x[:, node(ip[0], kk), node(ip[1], kk), ep]
                b = x[:, node(isx[0], kk), node(isx[1], kk), es]
                np.testing.assert_allclose(a, b, atol=1e-12)
        assert len(used) == 4 * mesh.n_elements
