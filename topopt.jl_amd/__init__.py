"""topopt.jl_amd — MI355X-native SIMP compliance inner loop with TopOpt.jl's API.

Host-side mirror of the TopOpt.jl (v0.14.0) interface for the matrix-free
displacement solve, the compliance sensitivity and the density filter, over
the C ABI of ``libtopopt_hip.so`` (include/topopt_hip.h).  The directory name
contains a dot, so import it through ``load_package()`` in the repository
root helpers, which registers it as ``topopt_jl_amd``.
"""
from ._lib import TohipError, load_library
from .beso import BESO, BESOResult
from .functions import (ComplianceFun, DensityFilterFun, ProjectedDensityFilterFun,
                        PseudoDensities, VolumeFun, oc_update)
from .io import nodal_field, save_mesh
from .problems import PointLoadCantilever, ferrite_numbering
from .simp import MMA87, SIMPResult, simp_run
from .heat import (HeatConductionProblem, HeatFEASolver, HeatTree, TemperatureFun,
                   ThermalComplianceFun, cell_temperature)
from .solver import (ConvergenceCriteria, DefaultCriteria, EnergyCriteria, FEASolver,
                     GenericFEASolver, HeavisideProjectionFun, HIPCGAssemblySolver,
                     HIPMatrixFreeSolver,
                     PowerPenaltyFun, ProjectedPenaltyFun, RationalPenaltyFun,
                     SigmoidProjectionFun, SinhPenaltyFun)

__all__ = [
    "TohipError", "load_library", "ComplianceFun", "DensityFilterFun",
    "ProjectedDensityFilterFun", "PseudoDensities",
    "VolumeFun", "oc_update", "PointLoadCantilever", "ferrite_numbering", "SIMPResult",
    "simp_run", "FEASolver", "GenericFEASolver", "HIPMatrixFreeSolver", "PowerPenaltyFun",
    "RationalPenaltyFun", "SinhPenaltyFun", "ProjectedPenaltyFun", "HeavisideProjectionFun",
    "SigmoidProjectionFun", "save_mesh", "nodal_field", "BESO", "BESOResult",
    "ConvergenceCriteria", "DefaultCriteria", "EnergyCriteria", "HeatConductionProblem",
    "HeatTree", "HeatFEASolver", "ThermalComplianceFun", "TemperatureFun", "cell_temperature",
    "HIPCGAssemblySolver", "MMA87",
]
