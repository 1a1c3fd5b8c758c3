"""Constrained matrix-free stiffness operator (oracle restatement).

The operator is  A_hat = M K(E) M + I_c,  K(E) = sum_e E_e Q_e^T K0 Q_e,
where M zeroes constrained DOFs and I_c is the identity on them: constrained
rows return the input entry, constrained columns are skipped.

  matvec            core/_templated_cuda.cu:782-834 (GPU), core/_ops.py:330-368 (CPU)
  matvec (no cons)  core/_templated_cuda.cu:137-178, core/_ops.py:212-244
  diagonal          core/_templated_cuda.cu:1068-1113, core/_ops.py:45-75
  process_dk        core/_templated_cuda.cu:180-194, core/_ops.py:5-13
  fea_integrals     TopOpt/_cuda.py:78-129, core/_templated_cuda.cu:1681-1719

Pinned in tests against a SciPy COO assembly of the same operator (built from
the element list independently of the gather) and, on tiny meshes, against a
literal per-node loop (matvec_node_loop) that follows the CUDA kernel line by line.
"""
import numpy as np
import scipy.sparse as sp

from .mesh import dof_map


def matvec(K0, elements, E, u, cons=None):
    """y = A_hat u (vectorised: gather u_e, contract with K0, scale by E_e, scatter)."""
    dof = K0.shape[0] // elements.shape[1]
    edof = dof_map(elements, dof)
    uu = np.array(u, dtype=np.float64, copy=True)
    if cons is not None:
        uu[cons] = 0.0
    ue = uu[edof]
    ve = (ue @ K0.T) * np.asarray(E, dtype=np.float64)[:, None]
    y = np.bincount(edof.ravel(), weights=ve.ravel(), minlength=u.shape[0])
    if cons is not None:
        y[cons] = u[cons]
    return y


def matvec_node_loop(K0, elements, E, u, cons, star):
    """Literal per-node restatement of mat_vec_node_basis_parallel_wcon (tiny meshes only)."""
    flat, el_ids, sorter, node_ids = star
    n_nodes = node_ids.shape[0]
    nE = elements.shape[1]
    dof = K0.shape[0] // nE
    Kf = K0.reshape(-1)
    out = np.zeros(n_nodes * dof)
    for i in range(n_nodes):
        st = node_ids[i]
        en = node_ids[i + 1] if i < n_nodes - 1 else flat.shape[0]
        for k in range(dof):
            out[i * dof + k] = u[i * dof + k] if cons[i * dof + k] else 0.0
        for j in range(en - st):
            e_ind = sorter[st + j]
            eid = el_ids[e_ind]
            start = eid * nE
            w = E[eid]
            rel = e_ind - start
            k_start = rel * dof * nE * dof
            for k in range(dof):
                if cons[i * dof + k]:
                    continue
                val = 0.0
                for l in range(nE * dof):
                    kk = flat[start + l // dof] * dof + l % dof
                    if cons[kk]:
                        continue
                    val += Kf[k_start + k * nE * dof + l] * u[kk]
                out[i * dof + k] += val * w
    return out


def assemble(K0, elements, E, cons=None, n_dof=None):
    """Explicit CSR of A_hat from a COO of element blocks (brute-force check)."""
    dof = K0.shape[0] // elements.shape[1]
    edof = dof_map(elements, dof)
    if n_dof is None:
        n_dof = int(elements.max() + 1) * dof
    nd = edof.shape[1]
    rows = np.repeat(edof, nd, axis=1).ravel()
    cols = np.tile(edof, (1, nd)).ravel()
    vals = (np.asarray(E, dtype=np.float64)[:, None] * K0.reshape(1, -1)).ravel()
    A = sp.coo_matrix((vals, (rows, cols)), shape=(n_dof, n_dof)).tocsr()
    if cons is not None and np.any(cons):
        m = sp.diags((~cons).astype(np.float64))
        A = (m @ A @ m + sp.diags(cons.astype(np.float64))).tocsr()
    A.sum_duplicates()
    return A


def diagonal(K0, elements, E, cons=None, n_dof=None):
    """diag(A_hat): sum_e E_e K0[l, l] on free DOFs, 1 on constrained ones."""
    dof = K0.shape[0] // elements.shape[1]
    edof = dof_map(elements, dof)
    if n_dof is None:
        n_dof = int(elements.max() + 1) * dof
    d = np.bincount(edof.ravel(), weights=(np.asarray(E, dtype=np.float64)[:, None]
                                           * np.diag(K0)[None, :]).ravel(), minlength=n_dof)
    if cons is not None:
        d[cons] = 1.0
    return d


def process_dk(K0, elements, U):
    """dk_e = -u_e^T K0 u_e on the full (unmasked) displacement."""
    dof = K0.shape[0] // elements.shape[1]
    ue = np.asarray(U, dtype=np.float64)[dof_map(elements, dof)]
    return -np.einsum("ei,ij,ej->e", ue, K0, ue)


def fea_integrals(K0, D, B, elements, E, U):
    """Per-element post-processing of a solved displacement (TopOpt/_cuda.py:78-129 with
    FEA_Integrals_node_basis_parallel_cuda_kernel, core/_templated_cuda.cu:1681-1719):
      strain_e = B u_e (B: the element's Gauss-averaged strain-displacement matrix),
      stress_e = D strain_e (not scaled by the density),
      energy_e = 1/2 u_e^T (E_e K0) u_e,
      von Mises from the stress (2-D plane stress / 3-D, TopOpt/_cuda.py:121-124).
    u_e is the full displacement (constrained DOFs included)."""
    dof = K0.shape[0] // elements.shape[1]
    ue = np.asarray(U, dtype=np.float64)[dof_map(elements, dof)]
    strain = ue @ B.T
    stress = strain @ D.T
    energy = 0.5 * np.asarray(E, dtype=np.float64) * np.einsum("ei,ij,ej->e", ue, K0, ue)
    s = stress
    if B.shape[0] == 3:
        vm = np.sqrt(s[:, 0] ** 2 + s[:, 1] ** 2 - s[:, 0] * s[:, 1] + 3 * s[:, 2] ** 2)
    else:
        vm = np.sqrt(0.5 * ((s[:, 0] - s[:, 1]) ** 2 + (s[:, 1] - s[:, 2]) ** 2 + (s[:, 2] - s[:, 0]) ** 2
                            + 6 * (s[:, 3] ** 2 + s[:, 4] ** 2 + s[:, 5] ** 2)))
    return strain, stress, energy, vm
