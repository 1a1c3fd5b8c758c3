"""Element matrices for structured meshes (host, once per mesh).

Q4 plane stress with 2x2 Gauss and H8 with 2x2x2 Gauss, DOFs interleaved
per node, the interface of EDGETO/physics/_linelas.py: auto_stiffness(x0s, E,
nu) -> (K, D, B, A) (:477-547), quadrilateral_element_stiffness (:195-309),
hexahedron_element_stiffness (:338-447).  Computed here in vectorised form:
all Gauss points at once, B assembled from the shape-function gradients.
"""
import numpy as np

_GP = np.array([-1.0, 1.0]) / np.sqrt(3.0)
# corner signs of the reference elements (node order of the structured meshes)
_Q4_S = np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]], dtype=np.float64)
_H8_S = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                  [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=np.float64)


def _gauss_points(dim):
    g = np.stack(np.meshgrid(*([_GP] * dim), indexing="ij"), axis=-1).reshape(-1, dim)
    return g  # unit weights


def _dN_dxi(signs, xi):
    """Shape-function derivatives w.r.t. the reference coordinates: (n_gp, dim, n_nodes)."""
    npts, dim = xi.shape
    f = 1.0 + xi[:, None, :] * signs[None, :, :]  # (n_gp, n_nodes, dim) factors (1 + s_a xi_a)
    out = np.empty((npts, dim, signs.shape[0]))
    scale = 0.5 ** dim
    for d in range(dim):
        prod = np.ones((npts, signs.shape[0]))
        for a in range(dim):
            prod *= signs[None, :, a] if a == d else f[:, :, a]
        out[:, d, :] = scale * prod
    return out


def _b_matrices(x0s, signs):
    dim = signs.shape[1]
    xi = _gauss_points(dim)
    dN = _dN_dxi(signs, xi)
    J = dN @ x0s  # (n_gp, dim, dim)
    detJ = np.linalg.det(J)
    G = np.linalg.solve(J, dN)  # physical gradients (n_gp, dim, n_nodes)
    n = signs.shape[0]
    if dim == 2:
        B = np.zeros((xi.shape[0], 3, 2 * n))
        B[:, 0, 0::2] = G[:, 0]
        B[:, 1, 1::2] = G[:, 1]
        B[:, 2, 0::2] = G[:, 1]
        B[:, 2, 1::2] = G[:, 0]
    else:
        B = np.zeros((xi.shape[0], 6, 3 * n))
        for i in range(3):
            B[:, i, i::3] = G[:, i]
        B[:, 3, 0::3] = G[:, 1]
        B[:, 3, 1::3] = G[:, 0]
        B[:, 4, 2::3] = G[:, 1]
        B[:, 4, 1::3] = G[:, 2]
        B[:, 5, 0::3] = G[:, 2]
        B[:, 5, 2::3] = G[:, 0]
    return B, detJ


def quadrilateral_element_stiffness(x0s, E=1.0, NU=0.33, e_type=0, t=1.0):
    """Q4 (K 8x8, D 3x3, averaged B 3x8); e_type 0 plane stress, 1 plane strain."""
    x0s = np.asarray(x0s, dtype=np.float64)
    if e_type == 0:
        c = E / (1.0 - NU * NU)
        D = np.array([[c, c * NU, 0.0], [c * NU, c, 0.0], [0.0, 0.0, c * (1.0 - NU) / 2.0]])
    else:
        c = E / ((1.0 + NU) * (1.0 - 2.0 * NU))
        D = np.array([[c * (1 - NU), c * NU, 0.0], [c * NU, c * (1 - NU), 0.0],
                      [0.0, 0.0, c * (1 - 2 * NU) / 2.0]])
    B, detJ = _b_matrices(x0s, _Q4_S)
    K = np.einsum("gij,jk,gkl,g->il", B.transpose(0, 2, 1), D, B, detJ * t)
    return K, D, B.mean(axis=0)


def hexahedron_element_stiffness(x0s, E=1.0, nu=0.33):
    """H8 (K 24x24, C 6x6, averaged B 6x24)."""
    x0s = np.asarray(x0s, dtype=np.float64)
    g = (1.0 - 2.0 * nu) / 2.0
    C = E / ((1.0 + nu) * (1.0 - 2.0 * nu)) * np.array(
        [[1 - nu, nu, nu, 0, 0, 0], [nu, 1 - nu, nu, 0, 0, 0], [nu, nu, 1 - nu, 0, 0, 0],
         [0, 0, 0, g, 0, 0], [0, 0, 0, 0, g, 0], [0, 0, 0, 0, 0, g]])
    B, detJ = _b_matrices(x0s, _H8_S)
    K = np.einsum("gij,jk,gkl,g->il", B.transpose(0, 2, 1), C, B, detJ)
    return K, C, B.mean(axis=0)


def _signed_measure(x0s):
    x0s = np.asarray(x0s, dtype=np.float64)
    if x0s.shape == (4, 2):  # shoelace area
        x, y = x0s[:, 0], x0s[:, 1]
        return 0.5 * float(np.dot(x, np.roll(y, -1)) - np.dot(y, np.roll(x, -1)))
    # hexahedron: six tetrahedra sharing the 1-7 diagonal
    V = 0.0
    for idx in ((0, 1, 3, 7), (1, 2, 3, 7), (0, 1, 7, 4), (1, 2, 7, 6), (1, 5, 7, 4), (1, 5, 6, 7)):
        p = x0s[list(idx)]
        V += float(np.linalg.det(p[1:] - p[0])) / 6.0
    return V


def auto_stiffness(x0s, E, Nu):
    """(K, D, B, A) for one Q4 (4x2 coordinates) or H8 (8x3) element."""
    x0s = np.asarray(x0s, dtype=np.float64)
    A = _signed_measure(x0s)
    if A < 0:
        A = -A
    if x0s.shape == (4, 2):
        K, D, B = quadrilateral_element_stiffness(x0s, E, Nu)
    elif x0s.shape == (8, 3):
        K, D, B = hexahedron_element_stiffness(x0s, E, Nu)
    else:
        raise ValueError("structured meshes use Q4 (4x2) or H8 (8x3) elements")
    return K, D, B, A
