"""Random problem configurations for structured 2-D / 3-D domains (reference:
randomizer/BoundaryConditions3D.py).

A configuration is the tuple (dims, loads, load_v, bcs, bc_c, E_mat, vf):
  dims    (dim,) int        elements per axis of a unit-element grid;
  loads   (n_l, dim)        node coordinates (integers in [0, dims]) of the loaded nodes;
  load_v  (n_l, dim)        load vector of each loaded node;
  bcs     (n_b, dim)        node coordinates of the constrained nodes;
  bc_c    (n_b, dim)        per-axis constraint flags (0/1) of each constrained node;
  E_mat   (n_material, 1)   material moduli, the largest 1;
  vf      (n_material, 1)   volume fraction per material.

Every draw goes through NumPy's global generator (np.random.*), in the reference's order,
so that np.random.seed(s) followed by the same call gives the reference's configuration:
the order of the draws is part of the output's definition here.  The probability
arguments are renormalised one after the other, each with the values already updated
(BoundaryConditions3D.py:107-114, kept as is).
"""
import numpy as np

__all__ = ["generate_random_configuration", "generate_random_boundary_condition", "generate_random_load",
           "generate_random_domain", "create_random_configuration_from_list"]


def _volume_fractions(n_material, min_vf, max_vf, min_vf_per_material):
    """E_mat (largest 1) and per-material volume fractions summing to U(min_vf, max_vf),
    redrawn until every material has at least min_vf_per_material (:27-41)."""
    E_mat = np.random.uniform(size=n_material)
    E_mat = (E_mat / E_mat.max()).reshape(-1, 1)
    total_vf = np.random.uniform(min_vf, max_vf)
    while True:
        vf = np.random.uniform(size=n_material)
        vf = vf / vf.sum() * total_vf
        if not np.any(vf < min_vf_per_material):
            return E_mat, vf.reshape(-1, 1)


def _shares(point, line, surface, dim):
    """The three type probabilities, renormalised in sequence as the reference does
    (each division uses the values already updated)."""
    if dim == 2:
        surface = 0
        line /= point + line
        point /= point + line
    else:
        surface /= point + line + surface
        point /= point + line + surface
        line /= point + line + surface
    return point, line, surface


def _kind(point, line):
    u = np.random.uniform()
    return 0 if u < point else (1 if u < point + line else 2)


def _boundary_point(dims, dim, inner):
    """A node inside the domain (inner) or on a random face of its boundary."""
    if inner:
        return np.random.randint(1, dims)
    loc = np.random.randint(0, dims + 1)
    axis = np.random.randint(0, dim)
    loc[axis] = np.random.choice([0, dims[axis]])
    return loc


def _segment(p1, p2, dim):
    """Nodes along p1 -> p2: 5 samples per unit length, p1 + the running sum of the step,
    truncated to integers, duplicates removed (rows sorted)."""
    n = int(np.linalg.norm(p2 - p1) * 5)
    step = (p2 - p1) / n
    pts = np.zeros((n, dim)) + p1[None]
    pts += np.cumsum(np.ones((n, dim)) * step[None], axis=0)
    return np.unique(pts.astype(int), axis=0)


def _distinct_pair(dims, lo, axis=None, side=None):
    """Two distinct grid points (low bound lo), optionally pinned to the face axis = side."""
    p1 = np.random.randint(lo, dims)
    if axis is not None:
        p1[axis] = side
    p2 = np.random.randint(lo, dims)
    if axis is not None:
        p2[axis] = side
    while np.all(p1 == p2):
        p2 = np.random.randint(lo, dims)
        if axis is not None:
            p2[axis] = side
    return p1, p2


def _line(dims, dim, inner):
    if inner:
        p1, p2 = _distinct_pair(dims, 1)
    else:
        axis = np.random.randint(0, dim)
        side = np.random.choice([0, dims[axis]])
        p1, p2 = _distinct_pair(dims, 0, axis, side)
    return _segment(p1, p2, dim)


def _surface(dims, dim):
    """Grid nodes of the rectangle spanned by two points of one outer face (3-D)."""
    axis = np.random.randint(0, dim)
    side = np.random.choice([0, dims[axis]])
    p1, p2 = _distinct_pair(dims, 0, axis, side)
    a0, a1 = [a for a in range(dim) if a != axis][:2]
    g0, g1 = np.meshgrid(np.arange(min(p1[a0], p2[a0]), max(p1[a0], p2[a0]) + 1),
                         np.arange(min(p1[a1], p2[a1]), max(p1[a1], p2[a1]) + 1))
    pts = np.zeros((g0.size, dim))
    pts[:, a0] = g0.reshape(-1)
    pts[:, a1] = g1.reshape(-1)
    pts[:, axis] = side
    return np.unique(pts, axis=0)


def _spread(pts, dim):
    """One random direction for a set of nodes, unit length per node, divided by the count."""
    v = np.ones_like(pts) * (np.random.uniform(size=dim) - 0.5)
    v /= np.linalg.norm(v, axis=1)[:, None]
    return v / v.shape[0]


def generate_random_boundary_condition(dims, dim, line_bc=1 / 3, point_bc=1 / 3, surface_bc=1 / 3, inner_bc=0.0):
    """One constraint group (BoundaryConditions3D.py:103-232): a point, a line or (3-D) a
    face rectangle, with one random 0/1 flag per axis shared by the group.  Returns
    (nodes (n, dim), flags (n, dim)); a point group keeps integer flags, the others bool."""
    point_bc, line_bc, surface_bc = _shares(point_bc, line_bc, surface_bc, dim)
    inner = np.random.uniform() < inner_bc
    kind = _kind(point_bc, line_bc)
    if kind == 0:
        loc = _boundary_point(dims, dim, inner)
        return loc[None], np.random.randint(0, 2, size=dim)[None]
    pts = _line(dims, dim, inner) if kind == 1 else _surface(dims, dim)
    flags = np.random.randint(0, 2, size=dim)[None].repeat(pts.shape[0], axis=0)
    return pts, flags.astype(bool)


def generate_random_load(dims, dim, point_load=1 / 3, line_loads=1 / 3, surface_loads=1 / 3, inner_loads=0.5):
    """One load group (BoundaryConditions3D.py:235-365): a point load of unit length in a
    random direction, or a line / face rectangle sharing one direction whose total is 1."""
    point_load, line_loads, surface_loads = _shares(point_load, line_loads, surface_loads, dim)
    inner = np.random.uniform() < inner_loads
    kind = _kind(point_load, line_loads)
    if kind == 0:
        loc = _boundary_point(dims, dim, inner)
        v = np.random.uniform(size=dim) - 0.5
        return loc[None], (v / np.linalg.norm(v))[None]
    pts = _line(dims, dim, inner) if kind == 1 else _surface(dims, dim)
    return pts, _spread(pts, dim)


def generate_random_domain(nel, min_nel_per_axis=10, dim=2):
    """Elements per axis for about nel elements with at least min_nel_per_axis on every axis
    but the last (BoundaryConditions3D.py:368-380)."""
    cap = int(np.floor(nel / min_nel_per_axis ** (dim - 1)))
    dims = np.zeros(dim).astype(int)
    for i in range(dim - 1):
        dims[i] = np.random.randint(min_nel_per_axis, cap + 1)
        cap = int(np.floor(nel / dims[i] / min_nel_per_axis ** (dim - 2 - i)))
    dims[-1] = nel // np.prod(dims[:-1])
    return dims


def _constraint_groups(n_bc, dims, dim, line_bc, point_bc, surface_bc, inner_bc):
    groups = [generate_random_boundary_condition(dims, dim, line_bc, point_bc, surface_bc, inner_bc)
              for _ in range(n_bc)]
    return np.concatenate([g[0] for g in groups]), np.concatenate([g[1] for g in groups])


def generate_random_configuration(min_element, max_element, dim=2, min_nel_per_axis=10, multi_load=True,
                                  line_loads=0.5, surface_loads=0.0, point_loads=0.5, inner_loads=0.0, max_loads=5,
                                  line_bc=0.5, point_bc=0.5, surface_bc=0.0, inner_bc=0.0, max_bc=10, n_material=1,
                                  min_vf=0.1, max_vf=0.75, min_vf_per_material=0.02, dims=None):
    """A random configuration (BoundaryConditions3D.py:4-100): materials and volume fractions,
    the domain (unless dims is given), 1..max_loads load groups scaled by U(0, 1) magnitudes
    normalised to a largest of 1, and dim..max_bc constraint groups, all redrawn until every
    axis is constrained somewhere."""
    E_mat, vf = _volume_fractions(n_material, min_vf, max_vf, min_vf_per_material)
    if dims is None:
        dims = generate_random_domain(np.random.randint(min_element, max_element + 1), min_nel_per_axis, dim)
    n_loads = np.random.randint(1, max_loads + 1) if multi_load else 1
    mags = np.random.uniform(size=n_loads)
    mags /= mags.max()
    groups = [generate_random_load(dims, dim, point_loads, line_loads, surface_loads, inner_loads)
              for _ in range(n_loads)]
    loads = np.concatenate([g[0] for g in groups])
    load_v = np.concatenate([g[1] * mags[i] for i, g in enumerate(groups)])
    n_bc = np.random.randint(dim, max_bc + 1)
    bcs, bc_c = _constraint_groups(n_bc, dims, dim, line_bc, point_bc, surface_bc, inner_bc)
    while not np.all(bc_c.sum(axis=0) > 0):
        bcs, bc_c = _constraint_groups(n_bc, dims, dim, line_bc, point_bc, surface_bc, inner_bc)
    return dims, loads, load_v, bcs, bc_c, E_mat, vf


def create_random_configuration_from_list(domains, BCs, dim_idx=None, bc_idx=None, multi_load=False, line_loads=0.0,
                                          surface_loads=0.0, point_loads=1.0, inner_loads=0.0, max_loads=1,
                                          n_material=4, min_vf=0.16, max_vf=0.6, min_vf_per_material=0.04):
    """A configuration on one of the given domains with one of its given constraint sets
    (BoundaryConditions3D.py:383-398): random loads, materials and volume fractions.  The
    inner call also draws (and discards) `dim` point constraints, as the reference does, so
    the generator's state advances identically.  Returns (config, (dim_idx, bc_idx))."""
    if dim_idx is None:
        dim_idx = np.random.choice(domains.shape[0])
    dims = domains[dim_idx]
    if bc_idx is None:
        bc_idx = np.random.choice(len(BCs[dim_idx]))
    bcs, bc_c = BCs[dim_idx][bc_idx]
    dim = dims.shape[0]
    if n_material > 1:
        n_material = np.random.randint(2, n_material + 1)
    _, loads, load_v, _, _, E_mat, vf = generate_random_configuration(
        0, 0, dim, 0, multi_load, line_loads, surface_loads, point_loads, inner_loads, max_loads,
        0, 1, 0, 0, dim, n_material, min_vf, max_vf, min_vf_per_material, dims=dims)
    return (dims, loads, load_v, bcs, bc_c, E_mat, vf), (dim_idx, bc_idx)
