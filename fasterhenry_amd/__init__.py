"""fasterhenry_amd -- MI355X-native (gfx950) PEEC extraction hot path.

A drop-in for the solve paths of fasterhenry 0.2.0: the same problem
types (Geometry, Port, Discretization, SolverChoice,
DENSE_PATH_MAX_FILAMENTS) and entry points (solve, MeshSystem), computed by
hand-written HIP kernels in libfasterhenry_gpu.so through its C ABI
(include/fasterhenry_gpu.h).  Importing the package needs no GPU; the first
solve loads the native library and raises if it or the device is missing.
"""
from .errors import *  # noqa: F401,F403
from .coupling import (DEFAULT_GROUP, Coupling, CouplingError, NonTransitiveCoupling,  # noqa: F401
                       PairMask, SegmentCountMismatch, TruncationWarning, UnknownGroup)
from .model import *  # noqa: F401,F403
from .model import DENSE_PATH_MAX_FILAMENTS, MU0  # noqa: F401
from .api import (ConductorPeak, Counts, CurrentSolution, FilamentCurrent,  # noqa: F401
                  IterativeSystem, MeshSystem, PfftOperator, PowerBalance, SegmentCurrent, SweepResult,
                  unit_excitation,
                  Timing, device_math, discretize_geometry, discretize_packed, discretize_tangent,
                  gmres_dense, iterative_params, partial_inductance_derivative,
                  partial_inductance_matrix, partial_inductance_matrix_masked, pfft_params, solve,
                  solve_batch, spice_subckt, tune_load, tune_save,
                  subckt_name_from_path, version, zc_mat)
from . import deck  # noqa: E402,F401
