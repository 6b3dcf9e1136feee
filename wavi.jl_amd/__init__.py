"""wavi.jl_amd - MI355X-native velocity solve of WAVI.jl.

Host-side mirror of the reference's API (Grid, Params, SolverParams,
InitialConditions, Model, Simulation, update_velocities!, update_state!,
timestep!, run_simulation!) over libwavi_hip.so, whose HIP kernels for gfx950
do all of the solve's arithmetic.  Julia names ending in "!" drop the "!".
"""
from .core import (ArgumentError, BoundsError, DimensionMismatch, Grid, InitialConditions, Params,
                   SolverParams, TimesteppingParams, OutputParams, UniformMeltRate, clip, get_c_mask,
                   get_glx, get_u_mask, get_v_mask, glen_b, height_above_floatation, mismip_plus_bed,
                   orientations2bc, volume_above_floatation)
from .model import (BasicParallelSpec, Fields, HIPSpec, Model, get_op_diag, get_rhs, haar,
                    inner_update, misfit_gradient, misfit_gradient_batch, op_apply, pull_fields,
                    push_fields, update_geometry, update_preconditioner, update_state,
                    update_velocities, update_velocities_batch, update_wavelets, vcycle)
from .simulation import (Clock, Simulation, load_checkpoint, pickup, run_simulation, save_checkpoint, timestep,
                         update_thickness, write_vel)
from .schwarz import (SharedMemorySpec, schwarz_model, schwarz_partition_of_unity, schwarz_prolong_velocities,
                      schwarz_restrict_velocities)
from . import _lib, workloads

__all__ = [
    "ArgumentError", "DimensionMismatch", "Grid", "InitialConditions", "Params", "SolverParams",
    "TimesteppingParams", "UniformMeltRate", "Model", "HIPSpec", "BasicParallelSpec", "Fields",
    "Simulation", "Clock", "update_velocities", "update_velocities_batch", "update_state",
    "inner_update", "update_wavelets", "update_preconditioner", "update_geometry",
    "misfit_gradient", "misfit_gradient_batch", "timestep", "run_simulation",
    "update_thickness", "op_apply", "get_rhs", "get_op_diag", "vcycle", "haar",
    "push_fields", "pull_fields", "glen_b", "mismip_plus_bed", "orientations2bc",
    "get_u_mask", "get_v_mask", "get_c_mask", "get_glx", "clip", "height_above_floatation",
    "volume_above_floatation", "workloads", "BoundsError", "SharedMemorySpec", "schwarz_model",
    "schwarz_partition_of_unity", "schwarz_prolong_velocities", "schwarz_restrict_velocities",
    "OutputParams", "save_checkpoint", "load_checkpoint", "pickup", "write_vel",
]
