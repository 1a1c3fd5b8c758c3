"""Random 2-D load / support configurations (reference: randomizer/BoundaryConditions2D.py).

generate_random_condition returns (dims, load_pos, load_vec, fixtures, fixture_dirs, E_mat, vf):
  dims          (2,) int       grid extent; positions use 0 .. dims-1 on each axis;
  load_pos      (n_l, 2) int   loaded positions (distinct);
  load_vec      (n_l, 2)       load vectors, the largest of unit length;
  fixtures      (n_f, 2)       supported positions (sorted rows, distinct);
  fixture_dirs  (n_f, 2) int   1 where the position is fixed in x / y;
  E_mat, vf     (n_material, 1).

All draws go through NumPy's global generator in the reference's order, so a seeded
np.random gives the reference's configuration; the order of the draws, of the option
tables and of the set-based de-duplication is part of the output's definition.  Reference
behaviours kept as they are: an "edge_point" on a y edge also takes y from {0, dims[1]-1};
rectangle loads are not clipped to the domain; the "all loads coincide with supports" test
uses ndarray membership (a coordinate match in either column counts).
"""
import numpy as np

__all__ = ["generate_random_condition", "find_duplicates_in_two_arrays", "get_load_mag_and_vec",
           "get_constraints", "random_load_or_cond", "random_vector"]

# option -> weight, in the reference's order; the distributed kinds are switched by flags
_POINT_KINDS = ("edge_point", "corner", "internal_point")
_OPTION_ORDER = ("edge_point", "full_edge", "partial_edge", "corner", "internal_distributed_orthogonal",
                 "internal_distributed", "internal_distributed_ellipse", "internal_distributed_rectangle",
                 "internal_point")
_BASE_WEIGHT = {"edge_point": 0.1, "full_edge": 0.1, "partial_edge": 0.1, "corner": 0.1,
                "internal_distributed_orthogonal": 0.025, "internal_distributed": 0.025,
                "internal_distributed_ellipse": 0.025, "internal_distributed_rectangle": 0.025,
                "internal_point": 0.50}


def _options(switch):
    """Option table with every distributed kind multiplied by `switch` (a bool or 0/1)."""
    return {k: _BASE_WEIGHT[k] * (1 if k in _POINT_KINDS else switch) for k in _OPTION_ORDER}


def _probabilities(options_dict):
    keys = list(options_dict.keys())
    p = np.array([options_dict[k] for k in keys])
    return keys, p / np.sum(p)


def _span(n):
    """A centre in range(n) and a half length in range(1, n // 2), clipped to [0, n-1]: the
    half-open range of the segment (half length drawn first)."""
    half = np.random.choice(range(1, n // 2))
    c = np.random.choice(range(n))
    return max(0, c - half), min(n - 1, c + half)


def _edge_segment(dims, along, fixed):
    """Positions of a segment along axis `along` at coordinate `fixed` of the other axis."""
    lo, hi = _span(dims[along])
    return [(v, fixed) if along == 0 else (fixed, v) for v in range(lo, hi)]


def random_load_or_cond(dims, options, probs, face=None):
    """Positions of one load or support group (BoundaryConditions2D.py:189-344)."""
    kind = np.random.choice(options, p=probs)
    X, Y = dims[0], dims[1]
    if kind == "internal_point":
        x = np.random.choice(range(0, X - 1))
        y = np.random.choice(range(0, Y - 1))
        pos = [(x, y)]
    elif kind == "face":
        if not face:
            face = np.random.choice(["x", "y"])
        if face == "x":
            x = np.random.choice([0, X - 1])
            y = np.random.choice(range(0, Y - 1))
        else:
            x = np.random.choice(range(0, X - 1))
            y = np.random.choice([0, Y - 1])
        pos = [(x, y)]
    elif kind == "edge_point":
        if np.random.choice(["x", "y"]) == "x":
            x = np.random.choice(range(X))
        else:
            x = np.random.choice([0, X - 1])
        y = np.random.choice([0, Y - 1])
        pos = [(x, y)]
    elif kind == "full_edge":
        if np.random.choice(["x", "y"]) == "x":
            y = np.random.choice([0, Y - 1])
            pos = [(x, y) for x in range(X)]
        else:
            x = np.random.choice([0, X - 1])
            pos = [(x, y) for y in range(Y)]
    elif kind == "partial_edge":
        if np.random.choice(["x", "y"]) == "x":
            y = np.random.choice([0, Y - 1])
            pos = _edge_segment(dims, 0, y)
        else:
            x = np.random.choice([0, X - 1])
            pos = _edge_segment(dims, 1, x)
    elif kind == "internal_distributed_orthogonal":
        if np.random.choice(["x", "y"]) == "x":
            y = np.random.choice(range(Y))
            pos = _edge_segment(dims, 0, y)
        else:
            x = np.random.choice(range(X))
            pos = _edge_segment(dims, 1, x)
    elif kind == "internal_distributed":
        # a straight segment across the box [xc +- xlen] x [yc +- ylen] (clipped), one position
        # per unit of its longer side
        xc = np.random.choice(range(X))
        yc = np.random.choice(range(Y))
        xlen = np.random.choice(range(1, X // 2))
        ylen = np.random.choice(range(1, Y // 2))
        xl, xr = max(0, xc - xlen), min(X - 1, xc + xlen)
        yl, yr = max(0, yc - ylen), min(Y - 1, yc + ylen)
        if xlen >= ylen:
            pos = [(x, int((x - xl) / (xr - xl) * (yr - yl) + yl)) for x in range(xl, xr)]
        else:
            pos = [(int((y - yl) / (yr - yl) * (xr - xl) + xl), y) for y in range(yl, yr)]
    elif kind == "internal_distributed_ellipse":
        xc = np.random.choice(range(X))
        yc = np.random.choice(range(Y))
        xlen = int(np.random.geometric(min(20 / float(X), 1)))
        ylen = int(np.random.geometric(min(20 / float(Y), 1)))
        angle = np.random.uniform(0, 2 * np.pi)
        c, s = np.cos(angle), np.sin(angle)
        # scalar arithmetic per grid point (NumPy scalar squares, as the reference evaluates
        # them), so points on the rim fall on the same side
        pos = [(x, y) for x in range(X) for y in range(Y)
               if ((x - xc) * c + (y - yc) * s) ** 2 / xlen ** 2 + ((x - xc) * s - (y - yc) * c) ** 2 / ylen ** 2 <= 1]
    elif kind == "internal_distributed_rectangle":
        xc = np.random.choice(range(X))
        yc = np.random.choice(range(Y))
        xlen = int(np.random.geometric(min(20 / float(X), 1)))
        ylen = int(np.random.geometric(min(20 / float(Y), 1)))
        pos = [(x, y) for x in range(xc - xlen, xc + xlen) for y in range(yc - ylen, yc + ylen)]
    elif kind == "corner":
        x = np.random.choice([0, X - 1])
        y = np.random.choice([0, Y - 1])
        pos = [(x, y)]
    return np.array(pos)


def random_vector(mag, type):
    """(mag, 0), (0, mag) or mag in a uniformly random direction."""
    if type == "x":
        return np.array((mag, 0))
    if type == "y":
        return np.array((0, mag))
    angle = np.random.uniform(0, 2 * np.pi)
    return np.array((mag * np.cos(angle), mag * np.sin(angle)))


def get_load_mag_and_vec(dims, load_options_dict, load_dir):
    """One load group: magnitude ~ Exp(1) shared equally by the group's positions."""
    options, probs = _probabilities(load_options_dict)
    mag = np.random.exponential(1)
    pos = random_load_or_cond(dims, options, probs)
    vec = random_vector(mag, load_dir)
    n = len(pos)
    return pos, np.array([vec / n] * n)


def get_constraints(dims, constraint_options_dict, face=None):
    options, probs = _probabilities(constraint_options_dict)
    return random_load_or_cond(dims, options, probs, face=face)


def find_duplicates_in_two_arrays(arr1, arr2):
    """Distinct rows of arr1 and arr2 (sorted) and, per row, whether it occurs in arr1 / arr2."""
    rows = np.unique(np.vstack((arr1, arr2)), axis=0)
    src = np.zeros((len(rows), 2), dtype=int)
    for i, r in enumerate(rows):
        src[i, 0] = 1 if np.any(np.all(arr1 == r, axis=1)) else 0
        src[i, 1] = 1 if np.any(np.all(arr2 == r, axis=1)) else 0
    return rows, src


def _volume_fractions(n_material, min_vf, max_vf, min_vf_per_material):
    E_mat = np.random.uniform(size=n_material)
    E_mat = (E_mat / E_mat.max()).reshape(-1, 1)
    total_vf = np.random.uniform(min_vf, max_vf)
    while True:
        vf = np.random.uniform(size=n_material)
        vf = vf / vf.sum() * total_vf
        if not np.any(vf < min_vf_per_material):
            return E_mat, vf.reshape(-1, 1)


def generate_random_condition(min_element_num=None, max_element_num=None, dimensions=None, distributed=True,
                              multi_load=True, n_material=1, max_vf=0.75, min_vf=0.15, min_vf_per_material=0.02):
    """A random 2-D configuration (BoundaryConditions2D.py:4-134): Geometric(0.3) load groups
    (one if not multi_load), Geometric(0.2) + 1 support groups with x / y / xy directions,
    redrawn from the start while the supports cannot hold the domain (fewer than one x and
    two y supports, or the other way round) or every load sits on a support."""
    args = (min_element_num, max_element_num, dimensions, distributed, multi_load, n_material, max_vf, min_vf,
            min_vf_per_material)
    while True:
        out = _draw_condition(*args)
        if out is not None:
            return out


def _draw_condition(min_element_num, max_element_num, dimensions, distributed, multi_load, n_material, max_vf,
                    min_vf, min_vf_per_material):
    if dimensions:
        dims = dimensions
    elif min_element_num and max_element_num:
        n = np.random.uniform(min_element_num, max_element_num)
        ar = np.exp(np.random.normal(0, 1, 1))
        a = (n * ar) ** (1 / 2)
        dims = np.rint((a, a / ar)).astype(int).flatten()
    else:
        raise Exception("Either dimensions or min_element_num and max_element_num must be specified.")

    num_loads = np.random.geometric(0.3) if multi_load else 1
    num_constraints = np.random.geometric(0.2) + 1

    load_pos, load_vec = [], []
    load_table = _options(distributed * multi_load)
    for _ in range(num_loads):
        direction = np.random.choice(["x", "y", "xy"], p=[0.2, 0.2, 0.6])
        pos_s, vec_s = get_load_mag_and_vec(dims, load_table, direction)
        for i in range(len(pos_s)):
            p = tuple(pos_s[i, :])
            if p not in load_pos:  # a position keeps its first load
                load_pos.append(p)
                load_vec.append(tuple(vec_s[i, :]))
    load_pos, load_vec = np.array(load_pos), np.array(load_vec)

    directions = [np.random.choice(["x", "y", "xy"], p=[0.3, 0.3, 0.4]) for _ in range(num_constraints)]
    support_table = _options(distributed)
    fx, fy = [], []
    for d in directions:
        pos_s = get_constraints(dims, support_table)
        for i in range(len(pos_s)):
            p = tuple(pos_s[i, :])
            if d in ("x", "xy"):
                fx.append(p)
            if d in ("y", "xy"):
                fy.append(p)
    fx, fy = np.array(list(set(fx))), np.array(list(set(fy)))

    lo, hi = sorted([len(fx), len(fy)])
    if lo < 1 or hi < 2:
        return None
    if all((p in fx) or (p in fy) for p in load_pos):
        return None

    load_vec = load_vec / np.max(np.linalg.norm(load_vec, axis=1))
    if n_material > 1:
        n_material = np.random.randint(2, n_material + 1)
    E_mat, vf = _volume_fractions(n_material, min_vf, max_vf, min_vf_per_material)
    fixtures, dirs = find_duplicates_in_two_arrays(fx, fy)
    return dims, load_pos, load_vec, fixtures, dirs, E_mat, vf
