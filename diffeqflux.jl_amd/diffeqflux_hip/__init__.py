"""MI355X (gfx950) NeuralODE hot path: DiffEqFlux's ``NeuralODE`` forward
(Tsit5) and ``InterpolatingAdjoint`` gradient on hand-written HIP kernels.

Host-side mirror of the reference's layer API (``src/neural_de.jl:35-61``)
over the C ABI declared in ``include/node_hip.h``.
"""
from .neural_ode import (NeuralODE, Chain, Dense, Cube, Tsit5, InterpolatingAdjoint,
                         ZygoteVJP, BacksolveAdjoint, TrackerAdjoint, QuadratureAdjoint,
                         GaussAdjoint, ODESolution, save_plan, AugmentedNDELayer, DimMover)
from .handle import Handle
from ._lib import load, NodeHipError
from .ffjord import FFJORD, FFJORDDistribution
from .multiple_shooting import (group_ranges, multiple_shoot, EnsembleProblem, EnsembleThreads,
                                EnsembleSerial, EnsembleGPUArray)

__all__ = ["NeuralODE", "Chain", "Dense", "Cube", "Tsit5", "InterpolatingAdjoint",
           "ZygoteVJP", "BacksolveAdjoint", "TrackerAdjoint", "QuadratureAdjoint",
           "GaussAdjoint", "ODESolution", "save_plan", "AugmentedNDELayer", "DimMover",
           "Handle", "load", "NodeHipError", "FFJORD", "FFJORDDistribution", "group_ranges", "multiple_shoot",
           "EnsembleProblem", "EnsembleThreads", "EnsembleSerial", "EnsembleGPUArray"]
