"""MI355X backend of insem's explicit time-stepping hot path.

The object model mirrors the reference package (python/insem/__init__.py);
the time loop runs on the GPU through the C-ABI declared in
include/insem_hip.h (libinsem_hip.so), fed by the host setup in
libinsem_host.so.
"""

from ._lib import InsemError
from .core import (
    Discretization,
    MacroMesh,
    Material,
    Receiver,
    SetupImage,
    Simulation,
    BatchSolver,
    Solver,
    Source,
    SourceTimeFunction,
    TimeConfig,
    derivative_matrix,
    gauss_lobatto_jacobi,
    gaussian,
    glj01,
    gll,
    gaussian_derivative,
    real_dtype,
    ricker,
    sampled_stf,
)
from . import mesh
from .many import plan_batches, solve_many

__version__ = "0.1.0"

__all__ = [
    "Discretization",
    "InsemError",
    "MacroMesh",
    "Material",
    "Receiver",
    "SetupImage",
    "Simulation",
    "BatchSolver",
    "Solver",
    "Source",
    "SourceTimeFunction",
    "TimeConfig",
    "__version__",
    "gaussian",
    "gaussian_derivative",
    "mesh",
    "real_dtype",
    "ricker",
    "plan_batches",
    "sampled_stf",
    "solve_many",
]
