"""MI355X (gfx950) implementation of RandLinearAlgebra.jl's data-parallel hot
path: compressor application (SparseSign / CountSketch, SRHT / FJLT,
Gaussian), the RangeFinder / RandSVD loop and block Kaczmarz, behind the
reference's complete_compressor / update_compressor! / mul! / rapproximate /
rsolve! surface.

The directory name contains a dot, so load it with
`__graft_entry__.load_package()` (it registers the module as
`randlinearalgebra_amd`).
"""
from ._lib import (AdmissionError, ArgumentError, DeviceError, DimensionMismatch,  # noqa: F401
                   DomainError)
from .compressors import (Cardinality, Compressor, CompressorAdjoint,  # noqa: F401
                          CompressorRecipe, CountSketch, CountSketchRecipe, FJLT, FJLTRecipe,
                          Gaussian, GaussianRecipe, Left, Right, SRHT,
                          SRHTRecipe, SparseSign, SparseSignRecipe, Undef, View, adjoint,
                          as_view, complete_compressor, fwht, left_mul_dimcheck, mul, mul_,
                          pack_bits, padded_size, right_mul_dimcheck, seed, times,
                          update_compressor, update_compressor_)
from .approximators import (Approximator, ApproximatorAdjoint, ApproximatorRecipe,  # noqa: F401
                            RandSVD, RandSVDRecipe, RangeFinder, RangeFinderRecipe,
                            complete_approximator, rapproximate, rapproximate_)
from .solvers import (Agmon, AgmonRecipe, BasicLogger, BasicLoggerRecipe,  # noqa: F401
                      ColumnProjection, sample_distribution,
                      ColumnProjectionRecipe, DistributionRecipe,
                      FullResidual, FullResidualRecipe, IHS, IHSRecipe, Kaczmarz,
                      LSGradient, LSGradientRecipe,
                      KaczmarzRecipe, L2Norm,
                      LQSolver, LQSolverRecipe, Sampling, SamplingRecipe, Uniform,
                      complete_distribution, complete_logger, complete_solver, compute_error,
                      rsolve, rsolve_, threshold_stop, update_distribution)
from .batched import (QRSolver, QRSolverRecipe, SketchSolveBatch,  # noqa: F401
                      complete_sub_solver, ldiv, update_sub_solver)
from .serving import KaczmarzRequest, RangeFinderRequest, Server, poisson_arrivals  # noqa: F401
from . import _dense, _lib  # noqa: F401
