"""Write tests/golden/randomizer_vectors.npz: outputs of the reference's random configuration
generators (randomizer/BoundaryConditions2D.py, randomizer/BoundaryConditions3D.py) for fixed
seeds and arguments.  Run in a checkout that has the reference at REF (the two files import
only NumPy and are loaded on their own, not through their package):

    python tests/golden/make_randomizer_vectors.py [REF]

Each record stores the generator's outputs and `tail`, the next np.random.randint(2**31)
after the call, which pins the number of draws the call consumed.
"""
import importlib.util
import os
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/EDGETO/randomizer"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "randomizer_vectors.npz")


def _load(name):
    spec = importlib.util.spec_from_file_location("ref_" + name, os.path.join(REF, name + ".py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def cases():
    """(key, module, function, args, kwargs, seed): shared with the test that replays them."""
    out = []
    for s in range(8):
        out.append((f"cfg3_{s}", "BoundaryConditions3D", "generate_random_configuration", (800, 3000),
                    dict(dim=3, min_nel_per_axis=6, surface_loads=0.3, surface_bc=0.3, inner_loads=0.25,
                         inner_bc=0.25, n_material=3), 100 + s))
    for s in range(6):
        out.append((f"cfg2d_{s}", "BoundaryConditions3D", "generate_random_configuration", (200, 1500), {}, 200 + s))
    for s in range(4):
        out.append((f"dom_{s}", "BoundaryConditions3D", "generate_random_domain", (5000, 10, 3), {}, 300 + s))
    dims3 = np.array([14, 9, 7])
    for kind, kw in (("pt", dict(point_bc=1.0, line_bc=0.0, surface_bc=0.0)),
                     ("ln", dict(point_bc=0.0, line_bc=1.0, surface_bc=0.0)),
                     ("sf", dict(point_bc=0.0, line_bc=0.0, surface_bc=1.0))):
        for inner in (0.0, 1.0):
            for s in range(2):
                out.append((f"bc_{kind}_{int(inner)}_{s}", "BoundaryConditions3D", "generate_random_boundary_condition",
                            (dims3, 3), dict(kw, inner_bc=inner), 400 + s))
                lk = dict(point_load=kw["point_bc"], line_loads=kw["line_bc"], surface_loads=kw["surface_bc"],
                          inner_loads=inner)
                out.append((f"ld_{kind}_{int(inner)}_{s}", "BoundaryConditions3D", "generate_random_load",
                            (dims3, 3), lk, 500 + s))
    for s in range(12):
        out.append((f"cond_{s}", "BoundaryConditions2D", "generate_random_condition", (),
                    dict(min_element_num=400, max_element_num=2500, n_material=1 + s % 3), 600 + s))
    for s in range(3):
        out.append((f"condfix_{s}", "BoundaryConditions2D", "generate_random_condition", (),
                    dict(dimensions=(40, 20), multi_load=False, distributed=False, n_material=3), 700 + s))
    kinds = ["edge_point", "full_edge", "partial_edge", "corner", "internal_distributed_orthogonal",
             "internal_distributed", "internal_distributed_ellipse", "internal_distributed_rectangle",
             "internal_point", "face"]
    for k in kinds:
        for s in range(3):
            probs = np.array([1.0 if kk == k else 0.0 for kk in kinds])
            out.append((f"lc_{k}_{s}", "BoundaryConditions2D", "random_load_or_cond", (np.array([30, 17]), kinds, probs),
                        {}, 800 + s))
    for s in range(3):
        out.append((f"vec_{s}", "BoundaryConditions2D", "random_vector", (2.5, "xy"), {}, 900 + s))
    return out


def flatten(key, result):
    """Result -> dict of arrays (tuples flattened by position)."""
    if isinstance(result, tuple):
        items = {}
        for i, r in enumerate(result):
            items.update(flatten(f"{key}.{i}", r))
        return items
    return {key: np.asarray(result)}


def main():
    mods = {m: _load(m) for m in ("BoundaryConditions2D", "BoundaryConditions3D")}
    data = {}
    for key, mod, fn, args, kwargs, seed in cases():
        np.random.seed(seed)
        res = getattr(mods[mod], fn)(*args, **kwargs)
        data.update(flatten(key, res))
        data[key + ".tail"] = np.asarray(np.random.randint(2 ** 31))
    domains = np.array([[20, 10, 10], [16, 16, 8]])
    BCs = [[(np.array([[0, 0, 0], [0, 10, 10]]), np.array([[1, 1, 1], [1, 0, 1]]))],
           [(np.array([[0, 0, 0]]), np.array([[1, 1, 1]])), (np.array([[16, 16, 8]]), np.array([[0, 1, 1]]))]]
    for s in range(4):
        np.random.seed(1000 + s)
        res = mods["BoundaryConditions3D"].create_random_configuration_from_list(domains, BCs)
        data.update(flatten(f"list_{s}", res))
        data[f"list_{s}.tail"] = np.asarray(np.random.randint(2 ** 31))
    np.savez_compressed(OUT, **data)
    print(f"wrote {len(data)} arrays to {OUT}")


if __name__ == "__main__":
    main()
