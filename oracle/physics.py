"""Element matrices (oracle restatement; test infrastructure only).

Restates EDGETO/physics/_linelas.py:195-309 (Q4, plane stress, 2x2 Gauss),
:338-447 (H8, 2x2x2 Gauss), :3-50 (T3, beta / gamma cofactors), :70-176 (T4,
3x3 cofactor determinants) and auto_stiffness :477-547.
Pinned in tests by: the published top88 Q4 element matrix, the affine-field
energy identity u^T K u = V * eps:C:eps, symmetry, the rigid-body null space,
and an independent 3-point Gauss integration.
"""
import numpy as np

_G2 = (-1.0 / np.sqrt(3.0), 1.0 / np.sqrt(3.0))


def q4_stiffness(x0s, E=1.0, nu=0.33, t=1.0, n_gauss=2):
    """Q4 plane-stress K (8x8), D (3x3), averaged B (3x8).

    Follows physics/_linelas.py:195-309: node order (0,0),(1,0),(1,1),(0,1),
    DOFs interleaved [n0x, n0y, n1x, ...], B rows (xx, yy, xy-engineering).
    """
    x0s = np.asarray(x0s, dtype=np.float64)
    c = E / (1.0 - nu * nu)
    D = np.array([[c, c * nu, 0.0], [c * nu, c, 0.0], [0.0, 0.0, c * (1.0 - nu) / 2.0]])
    pts, wts = np.polynomial.legendre.leggauss(n_gauss)
    K = np.zeros((8, 8))
    Bavg = np.zeros((3, 8))
    for r, wr in zip(pts, wts):
        for s, ws in zip(pts, wts):
            dNdr = 0.25 * np.array([-(1 - s), (1 - s), (1 + s), -(1 + s)])
            dNds = 0.25 * np.array([-(1 - r), -(1 + r), (1 + r), (1 - r)])
            J = np.array([[dNdr @ x0s[:, 0], dNdr @ x0s[:, 1]],
                          [dNds @ x0s[:, 0], dNds @ x0s[:, 1]]])
            detJ = np.linalg.det(J)
            iJ = np.linalg.inv(J)
            dNdx = iJ[0, 0] * dNdr + iJ[0, 1] * dNds
            dNdy = iJ[1, 0] * dNdr + iJ[1, 1] * dNds
            B = np.zeros((3, 8))
            B[0, 0::2] = dNdx
            B[1, 1::2] = dNdy
            B[2, 0::2] = dNdy
            B[2, 1::2] = dNdx
            K += B.T @ D @ B * (t * wr * ws * detJ)
            Bavg += B / (n_gauss * n_gauss)
    return K, D, Bavg


def h8_stiffness(x0s, E=1.0, nu=0.33, n_gauss=2):
    """H8 K (24x24), C (6x6), averaged B (6x24).

    Follows physics/_linelas.py:338-447: bottom face CCW then top face CCW,
    strain order (xx, yy, zz, xy, yz, zx) with engineering shears, B rows
    3 = (d/dy u_x + d/dx u_y), 4 = (d/dy u_z + d/dz u_y), 5 = (d/dz u_x + d/dx u_z).
    """
    x0s = np.asarray(x0s, dtype=np.float64)
    f = E / ((1 + nu) * (1 - 2 * nu))
    g = (1 - 2 * nu) / 2
    C = f * np.array([[1 - nu, nu, nu, 0, 0, 0],
                      [nu, 1 - nu, nu, 0, 0, 0],
                      [nu, nu, 1 - nu, 0, 0, 0],
                      [0, 0, 0, g, 0, 0],
                      [0, 0, 0, 0, g, 0],
                      [0, 0, 0, 0, 0, g]])
    # reference corner signs (xi1, xi2, xi3) of the 8 nodes
    sx = np.array([-1, 1, 1, -1, -1, 1, 1, -1], dtype=np.float64)
    sy = np.array([-1, -1, 1, 1, -1, -1, 1, 1], dtype=np.float64)
    sz = np.array([-1, -1, -1, -1, 1, 1, 1, 1], dtype=np.float64)
    pts, wts = np.polynomial.legendre.leggauss(n_gauss)
    K = np.zeros((24, 24))
    Bavg = np.zeros((6, 24))
    for a, wa in zip(pts, wts):
        for b, wb in zip(pts, wts):
            for c, wc in zip(pts, wts):
                dS = 0.125 * np.array([sx * (1 + sy * b) * (1 + sz * c),
                                       sy * (1 + sx * a) * (1 + sz * c),
                                       sz * (1 + sx * a) * (1 + sy * b)])
                J = dS @ x0s
                aux = np.linalg.inv(J) @ dS
                B = np.zeros((6, 24))
                for i in range(3):
                    B[i, i::3] = aux[i]
                B[3, 0::3] = aux[1]
                B[3, 1::3] = aux[0]
                B[4, 2::3] = aux[1]
                B[4, 1::3] = aux[2]
                B[5, 0::3] = aux[2]
                B[5, 2::3] = aux[0]
                K += B.T @ C @ B * (np.linalg.det(J) * wa * wb * wc)
                Bavg += B / (n_gauss ** 3)
    return K, C, Bavg


def quad_area(x0s):
    """Two-triangle area, physics/_linelas.py:312-335."""
    def tri(p):
        return (p[0, 0] * (p[1, 1] - p[2, 1]) + p[1, 0] * (p[2, 1] - p[0, 1])
                + p[2, 0] * (p[0, 1] - p[1, 1])) / 2
    x0s = np.asarray(x0s, dtype=np.float64)
    return tri(x0s[[0, 1, 2]]) + tri(x0s[[0, 2, 3]])


def hex_volume(x0s):
    """Six-tetrahedron volume, physics/_linelas.py:450-474 (det/6 of [1 x y z])."""
    x0s = np.asarray(x0s, dtype=np.float64)
    V = 0.0
    for idx in ([0, 1, 3, 7], [1, 2, 3, 7], [0, 1, 7, 4], [1, 2, 7, 6], [1, 5, 7, 4], [1, 5, 6, 7]):
        p = np.concatenate([np.ones((4, 1)), x0s[idx]], axis=1)
        V += np.linalg.det(p) / 6.0
    return V


def tri_area(p):
    """physics/_linelas.py:53-67."""
    p = np.asarray(p, dtype=np.float64)
    return (p[0, 0] * (p[1, 1] - p[2, 1]) + p[1, 0] * (p[2, 1] - p[0, 1]) + p[2, 0] * (p[0, 1] - p[1, 1])) / 2


def t3_stiffness(x0s, E=1.0, nu=0.33, t=1.0):
    """T3 plane stress, physics/_linelas.py:3-50: B from the cofactors beta_i = y_j - y_m,
    gamma_i = x_m - x_j over 2A; K = t A B^T D B."""
    p = np.asarray(x0s, dtype=np.float64)
    A = tri_area(p)
    if A < 0:
        raise ValueError("Node Order Is Not Correct")
    bi, bj, bm = p[1, 1] - p[2, 1], p[2, 1] - p[0, 1], p[0, 1] - p[1, 1]
    gi, gj, gm = p[2, 0] - p[1, 0], p[0, 0] - p[2, 0], p[1, 0] - p[0, 0]
    B = np.array([[bi, 0, bj, 0, bm, 0], [0, gi, 0, gj, 0, gm], [gi, bi, gj, bj, gm, bm]]) / (2 * A)
    D = (E / (1 - nu * nu)) * np.array([[1, nu, 0], [nu, 1, 0], [0, 0, (1 - nu) / 2]])
    return t * A * B.T @ D @ B, D, B


def tet_volume(p):
    """physics/_linelas.py:179-192: det([1 x y z]) / 6."""
    p = np.asarray(p, dtype=np.float64)
    return np.linalg.det(np.concatenate([np.ones((4, 1)), p], axis=1)) / 6


def t4_stiffness(x0s, E=1.0, nu=0.33):
    """T4, physics/_linelas.py:70-176: beta / gamma / delta as signed 3x3 determinants of
    [1, two coordinates] over the other three vertices; B / (6V); K = V B^T C B."""
    p = np.asarray(x0s, dtype=np.float64)
    V = tet_volume(p)
    if V < 0:
        raise ValueError("Node Order Is Not Correct")
    others = ([1, 2, 3], [0, 2, 3], [0, 1, 3], [0, 1, 2])
    sgn = (-1, 1, -1, 1)
    beta = [sgn[a] * np.linalg.det(np.concatenate([np.ones((3, 1)), p[o][:, 1:]], axis=1))
            for a, o in enumerate(others)]
    gamma = [-sgn[a] * np.linalg.det(np.concatenate([np.ones((3, 1)), p[o][:, [0, 2]]], axis=1))
             for a, o in enumerate(others)]
    delta = [sgn[a] * np.linalg.det(np.concatenate([np.ones((3, 1)), p[o][:, 0:2]], axis=1))
             for a, o in enumerate(others)]
    B = np.zeros((6, 12))
    for a in range(4):
        B[0, 3 * a] = beta[a]
        B[1, 3 * a + 1] = gamma[a]
        B[2, 3 * a + 2] = delta[a]
        B[3, 3 * a], B[3, 3 * a + 1] = gamma[a], beta[a]
        B[4, 3 * a + 1], B[4, 3 * a + 2] = delta[a], gamma[a]
        B[5, 3 * a], B[5, 3 * a + 2] = delta[a], beta[a]
    B /= 6 * V
    C = np.zeros((6, 6))
    C[:3, :3] = [[1 - nu, nu, nu], [nu, 1 - nu, nu], [nu, nu, 1 - nu]]
    C[3:, 3:] = np.eye(3) * (1 - 2 * nu) / 2
    C *= E / (1 + nu) / (1 - 2 * nu)
    return V * B.T @ C @ B, C, B


def _reorder(K, B, order, dof):
    idx = np.array([[dof * o + d for d in range(dof)] for o in order]).reshape(-1)
    return K[np.ix_(idx, idx)], B[:, idx]


def auto_stiffness(x0s, E, nu):
    """(K, D, B, A) for a T3, Q4, T4 or H8 element.

    physics/_linelas.py:477-547. For these two element types the reference
    ends by recomputing K, D, B on the original node order (:527, :543), so
    the reorientation branch only changes the returned A.
    """
    x0s = np.asarray(x0s, dtype=np.float64)
    if x0s.shape == (3, 2):
        A = tri_area(x0s)
        if A < 0:
            K, D, B = t3_stiffness(x0s[::-1], E, nu)
            K, B = _reorder(K, B, [2, 1, 0], 2)
            return K, D, B, tri_area(x0s[::-1])
        K, D, B = t3_stiffness(x0s, E, nu)
        return K, D, B, A
    if x0s.shape == (4, 3):
        V = tet_volume(x0s)
        if V < 0:
            K, D, B = t4_stiffness(x0s[[0, 1, 3, 2]], E, nu)
            K, B = _reorder(K, B, [0, 1, 3, 2], 3)
            return K, D, B, tet_volume(x0s[[0, 1, 3, 2]])
        K, D, B = t4_stiffness(x0s, E, nu)
        return K, D, B, V
    if x0s.shape == (4, 2):
        A = quad_area(x0s)
        if A < 0:
            A = quad_area(x0s[::-1])
        K, D, B = q4_stiffness(x0s, E, nu)
        return K, D, B, A
    if x0s.shape[0] == 8:
        A = hex_volume(x0s)
        if A < 0:
            A = hex_volume(x0s[[4, 5, 6, 7, 0, 1, 2, 3]])
        K, D, B = h8_stiffness(x0s, E, nu)
        return K, D, B, A
    raise ValueError("oracle supports T3, Q4, T4 and H8 elements only")
