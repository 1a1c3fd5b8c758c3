"""General (unstructured) meshes: CPU restatement (test infrastructure only).

Follows EDGETO/geom/_mesh.py:146-229 (GeneralMesh: centroids, element pointers,
per-element matrices), kernels/_cuda.py:305-815 (Uniform / General stiffness
kernels: the node -> element star of :320-331, matvec with constraints,
diagonal, process_grad), core/_templated_cuda.cu:1722-1800 (FEA integrals of
general meshes) and core/_filter.py:126-182 (the KD-tree filter of general
meshes).  Nothing here is imported by the product.

Pinned in tests by: a structured grid written as a general mesh, whose
operator, energy, FEA and filter must equal the structured oracle's (pinned
independently); symmetry and the rigid-body null space of the assembled
operator; the constant-strain patch test of T3 / T4 meshes; and brute-force
pair enumeration for the filter neighbourhoods.
"""
import numpy as np
import scipy.sparse as sp
from scipy.spatial import cKDTree


def flat_layout(elements):
    """(elements_flat, elements_ptr) of a list of node lists, geom/_mesh.py:160-169."""
    sizes = np.array([len(e) for e in elements], dtype=np.int64)
    ptr = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    flat = np.concatenate([np.asarray(e, dtype=np.int64) for e in elements])
    return flat, ptr


def star(elements_flat, n_nodes):
    """Stable argsort of the flat node list, its element ids and the per-node offsets
    (sorter / el_ids / node_ids of kernels/_cuda.py:320-331, node_ids with the end appended)."""
    flat = np.asarray(elements_flat, dtype=np.int64)
    sorter = np.argsort(flat, kind="stable")
    node_ptr = np.searchsorted(flat[sorter], np.arange(n_nodes + 1), side="left")
    return sorter, node_ptr


def assemble(Ks, elements, E, cons, n_dof, dof):
    """A_hat = M K(E) M + I_c, K(E) = sum_e E_e Q_e^T K_e Q_e (constrained rows and columns
    replaced by the identity, core/_templated_cuda.cu:669-780)."""
    rows, cols, vals = [], [], []
    for e, (K, nodes) in enumerate(zip(Ks, elements)):
        idx = (np.asarray(nodes)[:, None] * dof + np.arange(dof)[None, :]).reshape(-1)
        rows.append(np.repeat(idx, idx.size))
        cols.append(np.tile(idx, idx.size))
        vals.append((E[e] * np.asarray(K)).reshape(-1))
    A = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(n_dof, n_dof)).tocsr()
    c = np.asarray(cons, dtype=bool)
    M = sp.diags((~c).astype(np.float64))
    return (M @ A @ M + sp.diags(c.astype(np.float64))).tocsr()


def process_dk(Ks, elements, U, dof):
    """dk_e = -u_e^T K_e u_e, process_dk_{full,flat}_kernel_cuda (:196-231)."""
    out = np.empty(len(Ks))
    for e, (K, nodes) in enumerate(zip(Ks, elements)):
        idx = (np.asarray(nodes)[:, None] * dof + np.arange(dof)[None, :]).reshape(-1)
        ue = U[idx]
        out[e] = -ue @ np.asarray(K) @ ue
    return out


def fea(Ks, Bs, Ds, elements, E, U, dof):
    """strain = B_e u_e, stress = D_e strain, energy = 1/2 u_e^T (E_e K_e) u_e, von Mises
    (FEA_Integrals_node_basis_parallel_{full,flat}, :1722-1800; TopOpt/_cuda.py:123-127)."""
    n = len(Ks)
    nb = np.asarray(Bs[0]).shape[0]
    strain, stress, energy = np.empty((n, nb)), np.empty((n, nb)), np.empty(n)
    for e, (K, B, D, nodes) in enumerate(zip(Ks, Bs, Ds, elements)):
        idx = (np.asarray(nodes)[:, None] * dof + np.arange(dof)[None, :]).reshape(-1)
        ue = U[idx]
        strain[e] = np.asarray(B) @ ue
        stress[e] = np.asarray(D) @ strain[e]
        energy[e] = 0.5 * ue @ (E[e] * np.asarray(K)) @ ue
    s = stress
    if nb == 3:
        vm = np.sqrt(s[:, 0] ** 2 + s[:, 1] ** 2 - s[:, 0] * s[:, 1] + 3 * s[:, 2] ** 2)
    else:
        vm = np.sqrt(0.5 * ((s[:, 0] - s[:, 1]) ** 2 + (s[:, 1] - s[:, 2]) ** 2 + (s[:, 2] - s[:, 0]) ** 2
                            + 6 * (s[:, 3] ** 2 + s[:, 4] ** 2 + s[:, 5] ** 2)))
    return strain, stress, energy, vm


def centroids(nodes, elements):
    """geom/_mesh.py:165: the mean of each element's node coordinates."""
    return np.array([np.asarray(nodes)[np.asarray(e)].mean(axis=0) for e in elements])


def general_filter(cents, r_min):
    """filter_kernel_{2D,3D}_general (core/_filter.py:126-182): neighbours within r_min of
    each centroid (KD-tree ball query), weights (r_min - d) / r_min normalised per row."""
    cents = np.asarray(cents, dtype=np.float64)
    tree = cKDTree(cents)
    nb = tree.query_ball_point(cents, r_min)
    rows, cols, vals = [], [], []
    for i, js in enumerate(nb):
        js = np.asarray(js, dtype=np.int64)
        w = (r_min - np.linalg.norm(cents[js] - cents[i], axis=-1)) / r_min
        w = w / w.sum()
        rows.append(np.full(js.size, i))
        cols.append(js)
        vals.append(w)
    n = cents.shape[0]
    return sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(n, n)).tocsr()


def general_filter_bruteforce(cents, r_min):
    """The same matrix by enumerating every pair (an independent check of the KD-tree query)."""
    cents = np.asarray(cents, dtype=np.float64)
    d = np.linalg.norm(cents[:, None, :] - cents[None, :, :], axis=-1)
    W = np.where(d <= r_min, (r_min - d) / r_min, 0.0)
    return W / W.sum(axis=1, keepdims=True)
