"""Small general meshes for the unstructured-mesh tests (host arrays, no GPU)."""
import numpy as np

# Kuhn split of the unit cube into 6 tetrahedra around the 0-6 diagonal (corner order of H8)
_KUHN = [(0, 1, 2, 6), (0, 2, 3, 6), (0, 3, 7, 6), (0, 7, 4, 6), (0, 4, 5, 6), (0, 5, 1, 6)]


def grid_nodes(nel, size=None):
    nel = tuple(nel)
    size = tuple(float(n) for n in (size or nel))
    axes = [np.linspace(0.0, s, n + 1) for s, n in zip(size, nel)]
    g = np.meshgrid(*axes, indexing="ij")
    return np.stack([a.reshape(-1) for a in g], axis=1)


def _nid(nel, *ijk):
    if len(nel) == 2:
        i, j = ijk
        return i * (nel[1] + 1) + j
    i, j, k = ijk
    return (i * (nel[1] + 1) + j) * (nel[2] + 1) + k


def quad_elements(nel):
    """Q4 node lists (counter-clockwise) of a 2-D grid."""
    out = []
    for i in range(nel[0]):
        for j in range(nel[1]):
            out.append([_nid(nel, i, j), _nid(nel, i + 1, j), _nid(nel, i + 1, j + 1), _nid(nel, i, j + 1)])
    return out


def hex_elements(nel):
    """H8 node lists (bottom face counter-clockwise, then top) of a 3-D grid."""
    out = []
    for i in range(nel[0]):
        for j in range(nel[1]):
            for k in range(nel[2]):
                c = [(i, j, k), (i + 1, j, k), (i + 1, j + 1, k), (i, j + 1, k),
                     (i, j, k + 1), (i + 1, j, k + 1), (i + 1, j + 1, k + 1), (i, j + 1, k + 1)]
                out.append([_nid(nel, *p) for p in c])
    return out


def tri_mesh(nel, size=None, flip_every=0):
    """Conforming T3 mesh: each grid cell split along the same diagonal; every flip_every-th
    triangle listed clockwise (auto_stiffness reorders it)."""
    nodes = grid_nodes(nel, size)
    els = []
    for q in quad_elements(nel):
        for t in ([q[0], q[1], q[2]], [q[0], q[2], q[3]]):
            if flip_every and len(els) % flip_every == flip_every - 1:
                t = t[::-1]
            els.append(t)
    return nodes, els


def tet_mesh(nel, size=None, swap_every=0):
    """Conforming T4 mesh (Kuhn split of every cell); every swap_every-th tetrahedron listed with
    negative orientation."""
    nodes = grid_nodes(nel, size)
    els = []
    for h in hex_elements(nel):
        for t in _KUHN:
            tt = [h[a] for a in t]
            if swap_every and len(els) % swap_every == swap_every - 1:
                tt = [tt[0], tt[1], tt[3], tt[2]]
            els.append(tt)
    return nodes, els


def mixed_2d(nel, seed=0):
    """Q4 cells with about a third of them split into two triangles (non-conforming mix is fine
    for operator tests), node coordinates jittered."""
    rng = np.random.default_rng(seed)
    nodes = grid_nodes(nel) + rng.uniform(-0.15, 0.15, (int(np.prod(np.array(nel) + 1)), 2))
    els = []
    for q in quad_elements(nel):
        if rng.uniform() < 0.35:
            els.append([q[0], q[1], q[2]])
            els.append([q[0], q[2], q[3]][::-1] if rng.uniform() < 0.5 else [q[0], q[2], q[3]])
        else:
            els.append(q)
    return nodes, els


def mixed_3d(nel, seed=0):
    """H8 cells with about a third of them split into Kuhn tetrahedra, coordinates jittered."""
    rng = np.random.default_rng(seed)
    nodes = grid_nodes(nel) + rng.uniform(-0.1, 0.1, (int(np.prod(np.array(nel) + 1)), 3))
    els = []
    for h in hex_elements(nel):
        if rng.uniform() < 0.35:
            for t in _KUHN:
                tt = [h[a] for a in t]
                els.append([tt[0], tt[1], tt[3], tt[2]] if rng.uniform() < 0.3 else tt)
        else:
            els.append(h)
    return nodes, els
