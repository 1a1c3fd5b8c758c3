"""Element matrices (host, once per mesh).

T3 (constant-strain triangle), Q4 (2x2 Gauss), T4 (constant-strain
tetrahedron) and H8 (2x2x2 Gauss), DOFs interleaved per node, the interface
of EDGETO/physics/_linelas.py: auto_stiffness(x0s, E, nu) -> (K, D, B, A)
(:477-547), triangle_element_stiffness (:3-50), quadrilateral_element_stiffness
(:195-309), tetrahedron_element_stiffness (:70-176), hexahedron_element_stiffness
(:338-447).  Computed here in vectorised form: all Gauss points at once, B
assembled from the shape-function gradients (T3 / T4: the gradients of the
barycentric coordinates), and element_matrices() evaluates a whole mesh's
elements of one type in one batch.
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


def _d_plane(E, NU, e_type=0):
    if e_type == 0:
        c = E / (1.0 - NU * NU)
        return np.array([[c, c * NU, 0.0], [c * NU, c, 0.0], [0.0, 0.0, c * (1.0 - NU) / 2.0]])
    c = E / ((1.0 + NU) * (1.0 - 2.0 * NU))
    return np.array([[c * (1 - NU), c * NU, 0.0], [c * NU, c * (1 - NU), 0.0], [0.0, 0.0, c * (1 - 2 * NU) / 2.0]])


def _c_solid(E, nu):
    g = (1.0 - 2.0 * nu) / 2.0
    return E / ((1.0 + nu) * (1.0 - 2.0 * nu)) * np.array(
        [[1 - nu, nu, nu, 0, 0, 0], [nu, 1 - nu, nu, 0, 0, 0], [nu, nu, 1 - nu, 0, 0, 0],
         [0, 0, 0, g, 0, 0], [0, 0, 0, 0, g, 0], [0, 0, 0, 0, 0, g]])


def _b_from_grads(G):
    """Strain-displacement matrices from nodal gradients G (..., dim, n): rows xx, yy, xy
    (2-D) or xx, yy, zz, xy, yz, zx (3-D), engineering shears."""
    dim, n = G.shape[-2], G.shape[-1]
    if dim == 2:
        B = np.zeros(G.shape[:-2] + (3, 2 * n))
        B[..., 0, 0::2] = G[..., 0, :]
        B[..., 1, 1::2] = G[..., 1, :]
        B[..., 2, 0::2] = G[..., 1, :]
        B[..., 2, 1::2] = G[..., 0, :]
    else:
        B = np.zeros(G.shape[:-2] + (6, 3 * n))
        for i in range(3):
            B[..., i, i::3] = G[..., i, :]
        B[..., 3, 0::3] = G[..., 1, :]
        B[..., 3, 1::3] = G[..., 0, :]
        B[..., 4, 1::3] = G[..., 2, :]
        B[..., 4, 2::3] = G[..., 1, :]
        B[..., 5, 0::3] = G[..., 2, :]
        B[..., 5, 2::3] = G[..., 0, :]
    return B


def _simplex_grads(x):
    """Gradients of the barycentric coordinates of simplices x (n, dim+1, dim): (n, dim, dim+1),
    and the signed measures (area / volume).  P = [1 x] per vertex; column a of P^-1 holds
    (c_a, grad lambda_a) with lambda_a(p_b) = delta_ab."""
    n, k, dim = x.shape
    P = np.concatenate([np.ones((n, k, 1)), x], axis=2)
    measure = np.linalg.det(P) / (2.0 if dim == 2 else 6.0)
    Pinv = np.linalg.inv(P)
    return Pinv[:, 1:, :], measure


def triangle_element_stiffness(x0s, E=1.0, NU=0.33, e_type=0, t=1.0):
    """T3 (K 6x6, D 3x3, B 3x6): K = t A B^T D B, counter-clockwise nodes (the reference
    raises otherwise, physics/_linelas.py:23-24)."""
    x0s = np.asarray(x0s, dtype=np.float64)
    G, A = _simplex_grads(x0s[None])
    if A[0] < 0:
        raise ValueError("Node Order Is Not Correct")
    D = _d_plane(E, NU, e_type)
    B = _b_from_grads(G)[0]
    return t * A[0] * B.T @ D @ B, D, B


def tetrahedron_element_stiffness(x0s, E=1.0, NU=0.33):
    """T4 (K 12x12, C 6x6, B 6x12): K = V B^T C B, positively oriented nodes (the reference
    raises otherwise, physics/_linelas.py:85-86)."""
    x0s = np.asarray(x0s, dtype=np.float64)
    G, V = _simplex_grads(x0s[None])
    if V[0] < 0:
        raise ValueError("Node Order Is Not Correct")
    C = _c_solid(E, NU)
    B = _b_from_grads(G)[0]
    return V[0] * B.T @ C @ B, C, B


def triangle_element_area(x0s):
    return float(_simplex_grads(np.asarray(x0s, dtype=np.float64)[None])[1][0])


def tetrahedron_element_volume(x0s):
    return float(_simplex_grads(np.asarray(x0s, dtype=np.float64)[None])[1][0])


def element_matrices(x, E=1.0, nu=0.33):
    """(Ks, Ds, Bs, As) of a batch of elements of one type, x (n, nodes, dim): T3 (n, 3, 2),
    Q4 (n, 4, 2), T4 (n, 4, 3) or H8 (n, 8, 3), as auto_stiffness gives them one by one:
    simplices of either orientation (their matrices are in the given node order, the
    measure is |A|), Q4 / H8 with positive orientation only (physics/_linelas.py:477-547)."""
    x = np.asarray(x, dtype=np.float64)
    n, k, dim = x.shape
    if (k, dim) in ((3, 2), (4, 3)):
        G, meas = _simplex_grads(x)
        B = _b_from_grads(G)
        D = _d_plane(E, nu) if dim == 2 else _c_solid(E, nu)
        A = np.abs(meas)
        K = np.einsum("nji,jk,nkl,n->nil", B, D, B, A)
        return K, np.broadcast_to(D, (n,) + D.shape).copy(), B, A
    if (k, dim) == (4, 2):
        signs, D = _Q4_S, _d_plane(E, nu)
    elif (k, dim) == (8, 3):
        signs, D = _H8_S, _c_solid(E, nu)
    else:
        raise ValueError("element type not supported (T3, Q4, T4 or H8)")
    xi = _gauss_points(dim)
    dN = _dN_dxi(signs, xi)                          # (g, dim, k)
    J = np.einsum("gdk,nke->ngde", dN, x)            # (n, g, dim, dim)
    detJ = np.linalg.det(J)
    meas = _signed_measure_batch(x)
    if np.any(meas < 0):
        raise ValueError(f"Node Order Is Not Correct (element {int(np.argmax(meas < 0))})")
    Gp = np.linalg.solve(J, np.broadcast_to(dN, (n,) + dN.shape))  # (n, g, dim, k)
    Bg = _b_from_grads(Gp)                            # (n, g, nb, dim*k)
    K = np.einsum("ngji,jk,ngkl,ng->nil", Bg, D, Bg, detJ)
    return K, np.broadcast_to(D, (n,) + D.shape).copy(), Bg.mean(axis=1), meas


def _signed_measure_batch(x):
    n, k, dim = x.shape
    if dim == 2:
        X, Y = x[:, :, 0], x[:, :, 1]
        return 0.5 * (np.sum(X * np.roll(Y, -1, axis=1), axis=1) - np.sum(Y * np.roll(X, -1, axis=1), axis=1))
    V = np.zeros(n)
    for idx in ((0, 1, 3, 7), (1, 2, 3, 7), (0, 1, 7, 4), (1, 2, 7, 6), (1, 5, 7, 4), (1, 5, 6, 7)):
        p = x[:, list(idx)]
        V += np.linalg.det(p[:, 1:] - p[:, :1]) / 6.0
    return V


def auto_stiffness(x0s, E, Nu):
    """(K, D, B, A) for one element: T3 (3x2 coordinates), Q4 (4x2), T4 (4x3) or H8 (8x3).
    Simplices with negative orientation are evaluated in the reversed order and mapped
    back (physics/_linelas.py:492-512); Q4 and H8 must be positively oriented (the
    reference's final re-evaluation in the given order raises, :525-527, :543)."""
    x0s = np.asarray(x0s, dtype=np.float64)
    if x0s.shape in ((3, 2), (4, 3)):
        K, D, B, A = element_matrices(x0s[None], E, Nu)
        return K[0], D[0], B[0], float(A[0])
    A = _signed_measure(x0s)
    if A < 0:
        raise ValueError("Node Order Is Not Correct")
    if x0s.shape == (4, 2):
        K, D, B = quadrilateral_element_stiffness(x0s, E, Nu)
    elif x0s.shape == (8, 3):
        K, D, B = hexahedron_element_stiffness(x0s, E, Nu)
    else:
        raise ValueError("element type not supported (T3, Q4, T4 or H8)")
    return K, D, B, A


class AutoStiffness:
    """auto_stiffness(x0s, E, nu) as a callable with a batched form (element_matrices), the
    default stiffness of the general meshes (geom/_mesh.py:147)."""

    def __init__(self, E=1.0, nu=0.33):
        self.E, self.nu = float(E), float(nu)

    def __call__(self, x0s):
        return auto_stiffness(x0s, self.E, self.nu)

    def batch(self, x):
        return element_matrices(x, self.E, self.nu)
