from .basesolver import BaseSolver
from .static_implicit import StaticImplicitSolver
from .dynamic_implicit import DynamicImplicitSolver
from .dynamic_explicit import DynamicExplicitSolver
from ._linear_solver import conjugate_gradient, PCG, BatchedPCG, PcgVectors
