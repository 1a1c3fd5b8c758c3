"""Random problem configurations and their optimisation (reference: randomizer/*.py).

BoundaryConditions2D / BoundaryConditions3D draw loads, supports, domains, materials and
volume fractions with NumPy's global generator, in the reference's order (a seeded
np.random gives the reference's configurations).  ConfigSolve optimises one configuration
(generate_topology) or many at once on the batched engine (generate_topologies).
"""
from . import BoundaryConditions2D, BoundaryConditions3D  # noqa: F401
