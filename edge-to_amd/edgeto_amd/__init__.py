"""edgeto_amd -- MI355X-native drop-in for EDGE-TO's SIMP hot path.

The public names are those of the reference package (EDGETO/__init__.py),
restricted to the GPU path:

    CuStructuredMesh2D / CuStructuredMesh3D     geom
    CuStructuredFilter2D / CuStructuredFilter3D geom
    CuStructuredStiffnessKernel                 kernels
    CuGeneralMesh, CuGeneralFilter,
    CuUniformStiffnessKernel,
    CuGeneralStiffnessKernel                    general (unstructured meshes)
    CuMultiGrid, CuCG, CuGMRES                  solvers
    SingleMaterial                              MaterialModels
    CuTopOpt                                    TopOpt
    auto_stiffness                              physics

plus compliance(), a torch.autograd.Function with a hand-written adjoint.
All numerical work runs in libedgeto_amd.so (HIP, gfx950); importing this
package loads it and raises if it is missing.
"""
from . import _lib
from .geom import (CuStructuredFilter2D, CuStructuredFilter3D, CuStructuredMesh2D,
                   CuStructuredMesh3D, FilterKernel, StructuredFilter2D, StructuredFilter3D,
                   StructuredMesh2D, StructuredMesh3D)
from .kernels import CuStructuredStiffnessKernel, StiffnessKernel, StructuredStiffnessKernel
from .MaterialModels import (MaterialModel, PenalizedMultiMaterial, PenalizedMultiMaterialRP,
                             SingleMaterial)
from .physics import (auto_stiffness, hexahedron_element_stiffness,
                      quadrilateral_element_stiffness)
from .solvers import CG, GMRES, CuCG, CuGMRES, CuMultiGrid, MultiGrid, Solver
from .TopOpt import CuTopOpt, TopOpt
from .autograd import ComplianceFunction, compliance, energy_compliance
from .batch import BatchTopOpt, gmg_levels, random_problem
from .general import (CuGeneralFilter, CuGeneralMesh, CuGeneralStiffnessKernel, CuUniformStiffnessKernel,
                      GeneralFilter, GeneralMesh, GeneralStiffnessKernel, UniformStiffnessKernel)
from .physics import (AutoStiffness, element_matrices, tetrahedron_element_stiffness,
                      triangle_element_stiffness)
from . import torch_ops
from . import randomizer  # noqa: F401  (generators; ConfigSolve is imported on use)

_lib.load(require_gpu=False)

__all__ = ["CuStructuredMesh2D", "CuStructuredMesh3D", "CuStructuredFilter2D",
           "CuStructuredFilter3D", "CuStructuredStiffnessKernel", "CuMultiGrid", "SingleMaterial",
           "CuTopOpt", "auto_stiffness", "StructuredMesh2D", "StructuredMesh3D",
           "StructuredFilter2D", "StructuredFilter3D", "StructuredStiffnessKernel", "MultiGrid",
           "TopOpt", "FilterKernel", "StiffnessKernel", "MaterialModel", "Solver",
           "quadrilateral_element_stiffness", "hexahedron_element_stiffness", "compliance",
           "ComplianceFunction", "energy_compliance", "CuCG", "CuGMRES", "CG", "GMRES", "BatchTopOpt",
           "gmg_levels", "random_problem", "PenalizedMultiMaterial", "PenalizedMultiMaterialRP",
           "CuGeneralMesh", "GeneralMesh", "CuUniformStiffnessKernel", "CuGeneralStiffnessKernel",
           "UniformStiffnessKernel", "GeneralStiffnessKernel", "CuGeneralFilter", "GeneralFilter",
           "AutoStiffness", "element_matrices", "triangle_element_stiffness", "tetrahedron_element_stiffness"]
