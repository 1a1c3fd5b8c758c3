"""Structured mesh restatement: connectivity, coordinates, node->element star.

Node numbering (node counts nx = nelx+1, ...):
  2-D  node = j*nx + i                     (core/_geom.py:73-76, meshgrid 'xy')
  3-D  node = j*nx*nz + i*nz + k           (core/_geom.py:86-90; y slowest, z fastest)
Element numbering:
  2-D  e = i*nely + j                      (core/_geom.py:5-29)
  3-D  e = i*nely*nelz + j*nelz + k        (core/_geom.py:31-59)
Pinned in tests by brute-force enumeration against the nested-loop order the
reference comments describe, and by the coordinate identity nodes[e][0] = (i, j, k).
"""
import numpy as np


def elements_2d(nelx, nely):
    """Q4 connectivity, core/_geom.py:5-29 and core/_cuda.py:1089-1107."""
    nx = nelx + 1
    e = np.arange(nelx * nely, dtype=np.int64)
    i = e // nely
    j = e % nely
    out = np.stack([j * nx + i, j * nx + i + 1, (j + 1) * nx + i + 1, (j + 1) * nx + i], axis=1)
    return out.astype(np.int32)


def elements_3d(nelx, nely, nelz):
    """H8 connectivity, core/_geom.py:31-59 and core/_cuda.py:1109-1135."""
    nx, nz = nelx + 1, nelz + 1
    e = np.arange(nelx * nely * nelz, dtype=np.int64)
    i = e // (nely * nelz)
    j = (e % (nely * nelz)) // nelz
    k = e % nelz
    b = nz * i + nz * nx * j + k
    out = np.stack([b, b + nz, b + nz + nz * nx, b + nz * nx,
                    b + 1, b + nz + 1, b + nz + nz * nx + 1, b + nz * nx + 1], axis=1)
    return out.astype(np.int32)


def nodes_2d(nelx, nely, lx, ly):
    """Node coordinates, core/_geom.py:70-76 (np.meshgrid default 'xy')."""
    x = np.linspace(0, lx, nelx + 1)
    y = np.linspace(0, ly, nely + 1)
    xx, yy = np.meshgrid(x, y)
    return np.stack([xx.ravel(), yy.ravel()], axis=-1)


def nodes_3d(nelx, nely, nelz, lx, ly, lz):
    """Node coordinates, core/_geom.py:82-90 (meshgrid 'xy': shape (ny, nx, nz))."""
    x = np.linspace(0, lx, nelx + 1)
    y = np.linspace(0, ly, nely + 1)
    z = np.linspace(0, lz, nelz + 1)
    xx, yy, zz = np.meshgrid(x, y, z)
    return np.stack([xx.ravel(), yy.ravel(), zz.ravel()], axis=-1)


def node_index(nel, ijk):
    """Linear node index of integer node coordinates (i, j[, k])."""
    ijk = np.asarray(ijk, dtype=np.int64)
    if len(nel) == 2:
        return ijk[..., 1] * (nel[0] + 1) + ijk[..., 0]
    nx, nz = nel[0] + 1, nel[2] + 1
    return ijk[..., 1] * nx * nz + ijk[..., 0] * nz + ijk[..., 2]


def node_star(elements, n_nodes):
    """Node->element CSR used by the node-gather kernels, kernels/_cuda.py:77-87.

    elements_flat, el_ids = repeat(arange(nel)), sorter = stable argsort of
    elements_flat, node_ids = searchsorted(..., side='left'). Returns
    (elements_flat, el_ids, sorter, node_ids); all int32.
    """
    flat = elements.reshape(-1).astype(np.int32)
    el_ids = np.repeat(np.arange(elements.shape[0], dtype=np.int32), elements.shape[1])
    sorter = np.argsort(flat, kind="stable").astype(np.int32)
    node_ids = np.searchsorted(flat, np.arange(n_nodes, dtype=np.int32), sorter=sorter,
                               side="left").astype(np.int32)
    return flat, el_ids, sorter, node_ids


def dof_map(elements, dof):
    """Element DOF table (nel, nE*dof): node*dof + d, interleaved (core/_templated_cuda.cu:818)."""
    e = elements.astype(np.int64)
    return (e[:, :, None] * dof + np.arange(dof)[None, None, :]).reshape(e.shape[0], -1)
