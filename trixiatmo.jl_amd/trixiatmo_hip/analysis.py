"""AnalysisCallback error norms on the host (Trixi's calc_error_norms for
P4estMesh{3} [ext], the 3D counterpart of
src/callbacks_step/analysis_covariant.jl:144-189).

The solution, the node coordinates and J are interpolated to the LGL nodes of
degree 2 * polydeg; l2 = sqrt(sum w J diff^2 / sum w J), linf = max |diff|.
"""
import numpy as np

from .meshes import lgl_basis


def interpolation_matrix(x_in, x_out):
    """Barycentric Lagrange interpolation from nodes x_in to points x_out; a
    point that coincides with a node gives a unit row."""
    n = len(x_in)
    bw = np.array([1.0 / np.prod([x_in[j] - x_in[k] for k in range(n) if k != j])
                   for j in range(n)])
    V = np.zeros((len(x_out), n))
    for i, xo in enumerate(x_out):
        hit = np.isclose(xo, x_in, rtol=0, atol=1e-14)
        if hit.any():
            V[i, np.argmax(hit)] = 1.0
            continue
        t = bw / (xo - x_in)
        V[i] = t / t.sum()
    return V


def _apply(V, a):
    # a [..., n, n, n, E] -> [..., m, m, m, E]
    a = np.tensordot(V, a, axes=([1], [-4]))      # [m, ..., n, n, E]
    a = np.moveaxis(a, 0, -4)
    a = np.tensordot(V, a, axes=([1], [-3]))
    a = np.moveaxis(a, 0, -3)
    a = np.tensordot(V, a, axes=([1], [-2]))
    a = np.moveaxis(a, 0, -2)
    return a


def calc_error_norms(u, mesh, exact, analysis_polydeg=None):
    """u: [V, n, n, n, E]; exact(x [3, ...]) -> [V, ...]. Returns (l2, linf)."""
    n = mesh.nnodes
    m = 2 * mesh.polydeg if analysis_polydeg is None else analysis_polydeg
    x_sol, _, _, _ = lgl_basis(mesh.polydeg)
    x_ana, w_ana, _, _ = lgl_basis(m)
    V = interpolation_matrix(x_sol, x_ana)
    u_i = _apply(V, u)
    x_i = _apply(V, mesh.node_coordinates)
    J_i = _apply(V, 1.0 / mesh.inverse_jacobian)
    diff = exact(x_i) - u_i
    W = w_ana[:, None, None] * w_ana[None, :, None] * w_ana[None, None, :]
    wJ = W[..., None] * np.abs(J_i)
    l2 = np.sqrt((diff ** 2 * wJ).sum(axis=(1, 2, 3, 4)) / wJ.sum())
    linf = np.abs(diff).max(axis=(1, 2, 3, 4))
    return l2, linf


def calc_error_norms_covariant(u, mesh, aux, exact, t):
    """src/callbacks_step/analysis_covariant.jl:144-189: errors of the
    conserved variables at the solution nodes, the l2 norm weighted by
    w_i w_j J with J the area element of the auxiliary variables.
    u [3,n,n,E], aux [26,n,n,E], exact(x, t, aux) -> [3,n,n,E]."""
    _, w, _, _ = lgl_basis(mesh.polydeg)
    diff = exact(mesh.node_coordinates, t, aux) - u
    wJ = w[:, None, None] * w[None, :, None] * aux[12]
    l2 = np.sqrt((diff ** 2 * wJ).sum(axis=(1, 2, 3)) / wJ.sum())
    linf = np.abs(diff).max(axis=(1, 2, 3))
    return l2, linf
