"""atmo_max_dt3d_noalloc (the allocation-free max_dt of the 3D Euler
equations, kernels with their partials in the library's fixed scratch array)
against atmo_max_dt on a handle: the same step size bit for bit, on two calls
in a row, the second on a smaller mesh, each from device arrays owned by the
caller; sizes beyond the scratch array are refused before any GPU work."""
import ctypes

import numpy as np
import pytest

import trixiatmo_hip as T
from trixiatmo_hip import _lib

torch = pytest.importorskip("torch")


def _config(mesh, eq, n_elements=None):
    return _lib.AtmoConfig(equations=1, polydeg=mesh.polydeg,
                           n_elements=mesh.n_elements if n_elements is None else n_elements,
                           c_p=eq.c_p, c_v=eq.c_v, gravity=eq.gravity)


def _device(a):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a).ravel(order="F"))).cuda()


def _case(trees, polydeg, seed):
    mesh, eq, solver, bcs, u0 = T.workloads.periodic_cube(trees=trees, polydeg=polydeg, seed=seed)
    semi = T.SemidiscretizationHyperbolic(mesh, eq, solver, bcs)
    semi.set_state(u0)
    return mesh, eq, semi, u0


def test_oversized_input_is_refused():
    mesh, eq, _, _, _ = T.workloads.periodic_cube(trees=2, polydeg=3)
    cfg = _config(mesh, eq, n_elements=262145)
    p = ctypes.c_void_p(256)  # never dereferenced: the size check comes first
    rc = _lib.lib().atmo_max_dt3d_noalloc(ctypes.byref(cfg), p, p, p, 0.5, p, None)
    assert rc == -1
    assert b"scratch" in _lib.lib().atmo_last_error(None)


@pytest.mark.gpu
def test_noalloc_max_dt_matches_the_handle_twice():
    cfl = 0.5
    dt_dev = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
    for trees, polydeg, seed in ((4, 3, 0), (2, 3, 1)):  # the second call on a smaller mesh
        print(f"case: noalloc max_dt, periodic cube {trees}^3, polydeg {polydeg}", flush=True)
        mesh, eq, semi, u0 = _case(trees, polydeg, seed)
        ref = semi.max_dt(cfl)
        u, cv, ij = _device(u0), _device(mesh.contravariant_vectors), _device(mesh.inverse_jacobian)
        assert u.numel() == 6 * mesh.nnodes ** 3 * mesh.n_elements
        assert cv.numel() == 9 * mesh.nnodes ** 3 * mesh.n_elements
        torch.cuda.synchronize()
        cfg = _config(mesh, eq)
        rc = _lib.lib().atmo_max_dt3d_noalloc(
            ctypes.byref(cfg), ctypes.c_void_p(u.data_ptr()), ctypes.c_void_p(cv.data_ptr()),
            ctypes.c_void_p(ij.data_ptr()), cfl, ctypes.c_void_p(dt_dev.data_ptr()), None)
        assert rc == 0, _lib.lib().atmo_last_error(None)
        torch.cuda.synchronize()
        got = float(dt_dev.cpu()[0])
        print(f"  dt handle {ref!r} noalloc {got!r}", flush=True)
        assert got == ref
        del semi
