"""Algorithmic floating-point work of one rhs! evaluation per node, counted by
running the oracle's own code with an operation-counting scalar (counted.h:
+, -, * one each; /, sqrt, log, exp, pow one each; comparisons, fabs,
min/max free).  The volume terms are counted as the reference evaluates
them: every unordered node pair of a line once, with the nonconservative
term per ordered pair.  Test infrastructure: it reads the oracle and the
package's host-side mesh builders, needs no GPU, and prints the table that
profiles/algorithmic_flops.txt records.

Usage (repository root): python oracle/count_flops.py
"""
import ctypes
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "trixiatmo.jl_amd"))

import numpy as np  # noqa: E402

import oracle_lib as OL  # noqa: E402
import trixiatmo_hip as T  # noqa: E402


def count(oracle, mesh, eq, solver, omega, u):
    p = OL.params_for((mesh, eq, solver, omega))
    c, keep = oracle._conn(mesh)
    cv, ij = oracle._geo(mesh)
    uf = np.ascontiguousarray(np.asarray(u).ravel(order="F"), dtype=float)
    du = np.zeros_like(uf)
    k = (ctypes.c_longlong * 4)()
    rc = oracle.L.ora_e3_count_flops(ctypes.byref(p), ctypes.byref(c), OL._d(cv), OL._d(ij),
                                     OL._d(uf), OL._d(du), k)
    assert rc == 0
    # the counting run computes the oracle's values
    assert np.array_equal(du, oracle.rhs(p, mesh, u))
    nodes = mesh.n_elements * mesh.nnodes ** 3
    return list(k), nodes


def main():
    W = T.workloads
    oracle = OL.Oracle()
    cases = [
        ("C1 periodic cube 4^3, polydeg 3, flux_ranocha / flux_lax_friedrichs",
         W.periodic_cube(trees=4, polydeg=3, seed=0), 0.0),
        ("C3 shell 6x16^2x8, polydeg 4, KG + Waruszewski / KG + LLF + Waruszewski, slip walls",
         W.shell(16, 8, polydeg=4, seed=0), 0.0),
        ("C4 shell 6x8^2x4, polydeg 7, flux_ranocha + Waruszewski volume and surface",
         W.precision_stress("fp64"), 0.0),
    ]
    print(f"{'workload':88s} {'add':>12s} {'mul':>12s} {'div':>11s} {'special':>10s} "
          f"{'nodes':>9s} {'flop/node':>9s}")
    for name, (mesh, eq, solver, bcs, u0), omega in cases:
        t0 = time.perf_counter()
        k, nodes = count(oracle, mesh, eq, solver, omega, u0)
        print(f"{name:88s} {k[0]:12d} {k[1]:12d} {k[2]:11d} {k[3]:10d} {nodes:9d} "
              f"{sum(k) / nodes:9.1f}   ({time.perf_counter() - t0:.0f} s)", flush=True)


if __name__ == "__main__":
    main()
