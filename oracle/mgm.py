"""Geometric multigrid transfer operators and Galerkin coarse operators (oracle).

Restriction  v_c[c] = sum_{d in {-1,0,1}^dim, in bounds} 2^-|d|_1 v_f[2c+d] / tw(c)
             tw(c) = prod_axis (1.5 if c on that axis' boundary else 2)
             core/_templated_cuda.cu:232-364, core/_mgm.py:10-92 (3-D)
CPU 2-D      core/_mgm.py:94-166 adds the in-bounds weights to tw during the
             d == 0 pass, so every coarse value is divided by 2*tw (R/2).
Prolongation v_f[f] = (1/count) sum_{c in parents(f)} v_c[c] / tw(c)
             core/_templated_cuda.cu:366-475, core/_mgm.py:175-303
Galerkin     A_{l+1} = R A_l R^T with R the (GPU) restriction matrix
             core/_mgm_cuda.cu:1-344 (3-D), :346-649 (2-D), core/_mgm.py:659-1295

Pinned in tests by R 1 = 1, P = R^T on even grids, agreement of the
kernel-style application with the explicit sparse R, and the Galerkin product
checked against dense brute force on tiny grids.
"""
import itertools

import numpy as np
import scipy.sparse as sp


def _grid_nodes(nel):
    return tuple(int(n) + 1 for n in nel)


def _lin(nel, ijk):
    """Linear node index from per-axis node coords (axis order x, y[, z])."""
    if len(nel) == 2:
        nx = nel[0] + 1
        return ijk[1] * nx + ijk[0]
    nx, nz = nel[0] + 1, nel[2] + 1
    return ijk[1] * nx * nz + ijk[0] * nz + ijk[2]


def _coarse_nel(nel):
    return tuple(int(n) // 2 for n in nel)


def _tw(coords, ncoarse):
    tw = np.ones(coords[0].shape)
    for c, n in zip(coords, ncoarse):
        tw *= np.where((c == 0) | (c == n - 1), 1.5, 2.0)
    return tw


def restriction_matrix(nel, dof):
    """Explicit R (n_coarse*dof x n_fine*dof), GPU / 3-D CPU normalisation."""
    nel = tuple(int(n) for n in nel)
    dim = len(nel)
    nf = _grid_nodes(nel)
    ncoarse = _grid_nodes(_coarse_nel(nel))
    cc = np.meshgrid(*[np.arange(n) for n in ncoarse], indexing="ij")
    cc = [c.ravel() for c in cc]
    tw = _tw(cc, ncoarse)
    cidx = _lin(_coarse_nel(nel), cc)
    rows, cols, vals = [], [], []
    for d in itertools.product((-1, 0, 1), repeat=dim):
        f = [2 * c + dd for c, dd in zip(cc, d)]
        ok = np.ones(cc[0].shape, dtype=bool)
        for fa, n in zip(f, nf):
            ok &= (fa >= 0) & (fa < n)
        w = 0.5 ** sum(abs(x) for x in d)
        fidx = _lin(nel, [np.where(ok, fa, 0) for fa in f])
        for k in range(dof):
            rows.append(cidx[ok] * dof + k)
            cols.append(fidx[ok] * dof + k)
            vals.append(w / tw[ok])
    n_c = int(np.prod(ncoarse)) * dof
    n_f = int(np.prod(nf)) * dof
    return sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(n_c, n_f)).tocsr()


def prolongation_matrix(nel, dof):
    """Explicit P of the prolongation kernels (n_fine*dof x n_coarse*dof)."""
    nel = tuple(int(n) for n in nel)
    dim = len(nel)
    nf = _grid_nodes(nel)
    ncoarse = _grid_nodes(_coarse_nel(nel))
    ff = np.meshgrid(*[np.arange(n) for n in nf], indexing="ij")
    ff = [f.ravel() for f in ff]
    fidx = _lin(nel, ff)
    # count = number of in-bounds parents
    count = np.ones(ff[0].shape)
    for fa, n in zip(ff, ncoarse):
        count *= np.where((fa % 2 == 1) & (fa // 2 + 1 < n), 2.0, 1.0)
    rows, cols, vals = [], [], []
    for s in itertools.product((0, 1), repeat=dim):
        c = [fa // 2 + sa * (fa % 2) for fa, sa in zip(ff, s)]
        ok = np.ones(ff[0].shape, dtype=bool)
        for fa, sa, ca, n in zip(ff, s, c, ncoarse):
            ok &= ca < n
            if sa == 1:
                ok &= fa % 2 == 1
        tw = _tw([np.where(ok, ca, 0) for ca in c], ncoarse)
        cidx = _lin(_coarse_nel(nel), [np.where(ok, ca, 0) for ca in c])
        for k in range(dof):
            rows.append(fidx[ok] * dof + k)
            cols.append(cidx[ok] * dof + k)
            vals.append(1.0 / tw[ok] / count[ok])
    n_c = int(np.prod(ncoarse)) * dof
    n_f = int(np.prod(nf)) * dof
    return sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(n_f, n_c)).tocsr()


def restrict(v, nel, dof, cpu2d_quirk=False):
    """Kernel-style restriction. cpu2d_quirk reproduces core/_mgm.py:158-160 (R/2)."""
    nel = tuple(int(n) for n in nel)
    dim = len(nel)
    nf = _grid_nodes(nel)
    ncoarse = _grid_nodes(_coarse_nel(nel))
    cc = np.meshgrid(*[np.arange(n) for n in ncoarse], indexing="ij")
    cc = [c.ravel() for c in cc]
    tw = _tw(cc, ncoarse)
    cidx = _lin(_coarse_nel(nel), cc)
    V = v.reshape(-1, dof)
    acc = np.zeros((cc[0].shape[0], dof))
    wsum = np.zeros(cc[0].shape[0])
    for d in itertools.product((-1, 0, 1), repeat=dim):
        f = [2 * c + dd for c, dd in zip(cc, d)]
        ok = np.ones(cc[0].shape, dtype=bool)
        for fa, n in zip(f, nf):
            ok &= (fa >= 0) & (fa < n)
        w = 0.5 ** sum(abs(x) for x in d)
        fidx = _lin(nel, [np.where(ok, fa, 0) for fa in f])
        acc += np.where(ok[:, None], w * V[fidx], 0.0)
        wsum += np.where(ok, w, 0.0)
    if cpu2d_quirk and dim == 2:
        tw = tw + wsum
    out = np.zeros((int(np.prod(ncoarse)), dof))
    out[cidx] = acc / tw[:, None]
    return out.reshape(-1)


def prolong(v, nel, dof):
    return prolongation_matrix(nel, dof) @ v


def galerkin(A, nel, dof):
    """A_c = R A R^T with the GPU restriction matrix (reference Galerkin kernels)."""
    R = restriction_matrix(nel, dof)
    Ac = (R @ A @ R.T).tocsr()
    Ac.sum_duplicates()
    return Ac
