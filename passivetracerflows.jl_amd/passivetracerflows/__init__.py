"""passivetracerflows: the PassiveTracerFlows.jl TracerAdvectionDiffusion time step on MI355X.

The physics API mirrors the Julia package (``TracerAdvectionDiffusion`` module,
FourierFlows ``Problem``/``stepforward!``/grids) with Julia's ``!`` written as
a trailing ``_``.  Every transform and time step runs in the HIP library
libptf_hip (gfx950); this package is the host-side binding.
"""
from . import _lib
from . import fourierflows as FourierFlows
from . import traceradvectiondiffusion as TracerAdvectionDiffusion
from . import flows
from .ensemble import Ensemble
from .output import Diagnostic, Output, increment_, jldopen, record_, save_records, saveoutput, saveproblem
from .fourierflows import (GPU, Clock, DeviceArray, OneDGrid, ROCm, ThreeDGrid, TwoDGrid, dealias_,
                           gridpoints, ldiv_, mul_, step_until_, stepforward_)
from .traceradvectiondiffusion import (OneDAdvectingFlow, SpectralStreamfunctionFlow, ThreeDAdvectingFlow,
                                       TwoDAdvectingFlow, set_c_, updatevars_)

stepforward = stepforward_

__all__ = [
    "flows", "FourierFlows", "TracerAdvectionDiffusion", "Ensemble", "ROCm", "GPU", "Clock", "DeviceArray",
    "OneDGrid", "TwoDGrid", "ThreeDGrid", "gridpoints", "mul_", "ldiv_", "dealias_", "stepforward_", "step_until_",
    "OneDAdvectingFlow", "TwoDAdvectingFlow", "ThreeDAdvectingFlow", "SpectralStreamfunctionFlow",
    "set_c_", "updatevars_", "Diagnostic", "increment_", "Output", "saveproblem", "saveoutput",
    "save_records", "record_", "jldopen",
]
