"""Linear-hat density filter on structured grids (oracle restatement).

  offsets / weights   geom/_filters.py:86-99 (3-D), :132-146 (2-D), :175-193 (GPU)
  normaliser N_i      core/_templated_cuda.cu:1585-1630, core/_filter.py:6-39
  forward             x~_i = sum_k w_k x_{i+o_k} / sum_k w_k  (in-bounds k)
                      core/_templated_cuda.cu:1526-1583, core/_filter.py:78-124
  transpose           y_i = sum_k (w_k / N_{i+o_k}) v_{i+o_k}
                      core/_templated_cuda.cu:1632-1679, core/_filter.py:42-76

Element order is i*nely*nelz + j*nelz + k (2-D: i*nely + j). Pinned in tests by
an explicit sparse H built from element centroids with a KD-tree, the adjoint
identity <Hx, y> = <x, H^T y>, and row sums of H equal to one.
"""
import numpy as np
import scipy.sparse as sp


def offsets_weights(r_min, scales):
    """Integer offsets with strictly positive hat weight, weights normalised to sum 1."""
    scales = np.asarray(scales, dtype=np.float64)
    scales = scales / scales.min()
    n = int(np.ceil(r_min))
    rng = np.arange(-n, n + 1, dtype=np.int32)
    grids = np.meshgrid(*([rng] * scales.shape[0]), indexing="ij")
    offsets = np.stack([g.ravel() for g in grids], axis=1).astype(np.int32)
    d = np.linalg.norm(offsets * scales[None], axis=1)
    w = (r_min - d) / r_min
    keep = w > 0
    offsets, w = offsets[keep], w[keep]
    return offsets, w / w.sum()


def _shifted(shape, off):
    """For every element, the linear index of element+off and an in-bounds mask."""
    idx = np.indices(shape).reshape(len(shape), -1)
    nb = idx + np.asarray(off)[:, None]
    ok = np.all((nb >= 0) & (nb < np.asarray(shape)[:, None]), axis=0)
    lin = np.ravel_multi_index(np.where(ok, nb, 0), shape)
    return lin, ok


def normalizer(shape, offsets, weights):
    n = np.zeros(int(np.prod(shape)))
    for off, w in zip(offsets, weights):
        _, ok = _shifted(shape, off)
        n += np.where(ok, w, 0.0)
    return n


def forward(x, shape, offsets, weights):
    s = np.zeros(x.shape[0])
    n = np.zeros(x.shape[0])
    for off, w in zip(offsets, weights):
        lin, ok = _shifted(shape, off)
        s += np.where(ok, w * x[lin], 0.0)
        n += np.where(ok, w, 0.0)
    return np.where(n > 0, s / np.where(n > 0, n, 1.0), 0.0)


def transpose(v, shape, offsets, weights, norm):
    y = np.zeros(v.shape[0])
    for off, w in zip(offsets, weights):
        lin, ok = _shifted(shape, off)
        y += np.where(ok, (w / norm[lin]) * v[lin], 0.0)
    return y


def explicit_matrix(shape, r_min):
    """H built from element centroids with a KD-tree (independent construction)."""
    from scipy.spatial import cKDTree
    cen = np.indices(shape).reshape(len(shape), -1).T.astype(np.float64) + 0.5
    tree = cKDTree(cen)
    pairs = tree.query_pairs(r_min - 1e-12, output_type="ndarray")
    i = np.concatenate([pairs[:, 0], pairs[:, 1], np.arange(cen.shape[0])])
    j = np.concatenate([pairs[:, 1], pairs[:, 0], np.arange(cen.shape[0])])
    d = np.linalg.norm(cen[i] - cen[j], axis=1)
    w = np.maximum(r_min - d, 0.0)
    H = sp.coo_matrix((w, (i, j)), shape=(cen.shape[0],) * 2).tocsr()
    H.eliminate_zeros()
    rs = np.asarray(H.sum(axis=1)).ravel()
    return sp.diags(1.0 / rs) @ H
