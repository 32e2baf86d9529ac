"""The operation-counting run of the 3D oracle (oracle/counted.h,
ora_e3_count_flops) that gives the algorithmic flops per node: it computes
exactly the oracle's rhs!, and its counts are additive over elements (a
periodic brick twice as long, with the same element state repeated, counts
exactly twice the operations)."""
import ctypes

import numpy as np

import oracle_lib as OL
from trixiatmo_hip import workloads as W
from trixiatmo_hip.meshes import P4estMesh


def counts(oracle, mesh, eq, solver, u):
    p = OL.params_for((mesh, eq, solver, 0.0))
    c, keep = oracle._conn(mesh)
    cv, ij = oracle._geo(mesh)
    uf = np.ascontiguousarray(np.asarray(u).ravel(order="F"), dtype=float)
    du = np.zeros_like(uf)
    k = (ctypes.c_longlong * 4)()
    assert oracle.L.ora_e3_count_flops(ctypes.byref(p), ctypes.byref(c), OL._d(cv), OL._d(ij),
                                       OL._d(uf), OL._d(du), k) == 0
    return np.array(list(k)), du, p


def test_counting_run_computes_the_oracle_rhs(oracle):
    mesh, eq, solver, bcs, u0 = W.shell(2, 2, polydeg=3, seed=4)
    k, du, p = counts(oracle, mesh, eq, solver, u0)
    np.testing.assert_array_equal(du, oracle.rhs(p, mesh, u0))
    assert k.min() > 0


def test_counts_are_additive_over_elements(oracle):
    mesh1, eq, solver, _, _ = W.periodic_cube(trees=2, polydeg=3, seed=0)
    mesh2 = P4estMesh((4, 2, 2), polydeg=3, coordinates_min=(0.0, 0.0, 0.0),
                      coordinates_max=(2.0, 1.0, 1.0), periodicity=(True, True, True))
    # one smooth state per element, the same in every element (counts do not
    # depend on values except through the ln_mean branch: keep the states equal)
    n = mesh1.nnodes
    rng = np.random.default_rng(3)
    node = np.empty((6, n, n, n))
    node[0] = 1.0 + 0.05 * rng.random((n, n, n))
    node[1:4] = 0.1 * rng.standard_normal((3, n, n, n))
    node[4] = 2.5 + 0.1 * rng.random((n, n, n))
    node[5] = 0.0
    u1 = np.repeat(node[..., None], mesh1.n_elements, axis=-1)
    u2 = np.repeat(node[..., None], mesh2.n_elements, axis=-1)
    k1, _, _ = counts(oracle, mesh1, eq, solver, u1)
    k2, _, _ = counts(oracle, mesh2, eq, solver, u2)
    np.testing.assert_array_equal(k2, 2 * k1)
