"""misamplers — MI355X-native Monte Carlo sweep engine with the dwave-samplers API.

``SimulatedAnnealingSampler`` (alias ``Neal``) mirrors
``dwave.samplers.SimulatedAnnealingSampler``; ``PathIntegralAnnealingSampler`` and
``RotorModelAnnealingSampler`` mirror their ``dwave.samplers.sqa`` namesakes and
the ``SteepestDescent*`` classes mirror ``dwave.samplers.greedy``, and
``TreeDecompositionSampler`` and ``TreeDecompositionSolver`` mirror ``dwave.samplers.tree``;
the sweeps, descents and bucket eliminations run in hand-written HIP kernels for gfx950 behind the C ABI declared
in include/misamplers.h.
"""
from . import dimodlite
from .dimodlite import BINARY, SPIN, BinaryQuadraticModel, SampleSet, Vartype
from .sa import Neal, SimulatedAnnealingSampler, default_beta_range, simulated_annealing
from .sqa import PathIntegralAnnealingSampler, RotorModelAnnealingSampler
from .greedy import SteepestDescentComposite, SteepestDescentSampler, SteepestDescentSolver
from .tree import TreeDecompositionSampler, TreeDecompositionSolver

__version__ = "0.1.0"

__all__ = ["SimulatedAnnealingSampler", "Neal", "PathIntegralAnnealingSampler", "RotorModelAnnealingSampler",
           "SteepestDescentSolver", "TreeDecompositionSampler", "TreeDecompositionSolver",
           "SteepestDescentSampler", "SteepestDescentComposite", "simulated_annealing", "default_beta_range",
           "BinaryQuadraticModel", "SampleSet", "Vartype", "SPIN", "BINARY"]
