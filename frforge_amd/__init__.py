"""frforge_amd -- MI355X-native FRForge 2D FR residual and SSP-RK3 time step.

Host-side mirror of FRForge's Julia API for SolutionState2D / Euler2D; the residual, sensor,
CFL step and time stepping run in hand-written gfx950 HIP kernels behind the C ABI in
include/frforge_hip.h (libfrforge_hip.so).
"""
from .capturing import (AbstractCapturingMethod, ElementArtificialViscosity, NullCapturing,
                        PerssonAVMethod, PerssonSensor, ScaledPerssonMethod, get_capturing_method,
                        list_methods, method_params, register_method)
from .learned import LearnedSensorMethod   # registers "learned_persson"
from .equations import (Euler2D, conserved_to_primitives, pressure, primitives_to_conserved,
                        sound_speed)
from .mesh import Mesh2D, PeriodicBC, TransmissiveBC, build_mesh_metrics
from .operators import FROperators, build_operators
from .solver import (compute_dt, element_viscosities, integrate, residual, sense, ssp_rk3,
                     ssp_rk3_batch, ssp_rk3_step)
from .state import (SchemeConfig, SolutionState2D, allocate_state, discrete_mass, has_nonfinite,
                    l2_error, positivity_ok, set_initial_condition)

__all__ = [n for n in dir() if not n.startswith("_")]
