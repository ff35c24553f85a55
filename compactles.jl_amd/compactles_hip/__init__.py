"""compactles_hip: host-side mirror of the CompactLES frontend and operator
plugin interface, driving the gfx950 kernels of libcompactles_hip.so through
its C ABI."""
from ._lib import ClesError, SolverFailure, load  # noqa: F401
from .device import Context, DeviceArray  # noqa: F401
from .operators import Decomp, DirPlan, apply_along_, apply_along_timed, field, filter_field_, plan_direction  # noqa: F401
from .schemes import (BandedClosureRow, BandedCompactScheme, ClosureRow, CompactScheme,  # noqa: F401
                      compact_filter, lele_d1_10, lele_d1_6, pade_d1_4)
from .problem import (ArtParams, AxisBC, ConstantBodyForce, DirichletBC, ExtrapolationBC, NoSlipWallBC, OriginBC, PoleBC, SwitchableBC,
                      switch_, switched, IdealMixture, IdealSpecies, NSCBCInflowBC,  # noqa: F401
                      NSCBCOutflowBC, Numerics, PeriodicBC, Prim, Problem, SlipWallBC, Solver, SolverSpec, StepControl,
                      StiffenedGas, Stretch, sine_cluster, CartesianMetric, CylindricalMetric, SphericalMetric,
                      Transport, conserved_from_prim, initialize_, run_, setup, single_species, tanh_blend)
from .batch import Batch, Instance  # noqa: F401
from .stack import Stack  # noqa: F401
from .server import Server  # noqa: F401
from . import tune  # noqa: F401
from .callbacks import AtTime, Callback, EveryStep, EveryTime, ProgressLog, WhenState  # noqa: F401
from .nasa9 import Nasa9Interval, Nasa9Mixture, Nasa9Species, nasa9_constant_cp, read_nasa9  # noqa: F401
