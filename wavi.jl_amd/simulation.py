"""Simulation driver: Simulation, timestep!, update_thickness!, run_simulation!.

Mirrors src/Simulations/Simulation.jl:24-59 and run_simulation.jl:6-100.
The whole time loop runs on the device (wavi_run_steps): fields are copied to
the device once before the loop and back once after it.  Checkpoint, output
and pickup files are not part of this package.
"""
import ctypes
from dataclasses import dataclass

from . import _lib
from .core import ArgumentError
from .model import _check, _report_solve, _schwarz, pull_fields, push_fields, update_preconditioner, update_state


@dataclass
class Clock:
    """Clock (src/Clock.jl:2-12)."""
    n_iter: int = 0
    time: float = 0.0


class Simulation:
    """Simulation(; model, timestepping_params) (Simulation.jl:24-51)."""

    def __init__(self, model=None, timestepping_params=None, output_params=None,
                 pickup_output_update_flag=False):
        if timestepping_params is None:
            raise ArgumentError("You must specify a timestepping parameters")
        if timestepping_params.niter0 > 0:
            raise ArgumentError("pickup from checkpoints is not supported")
        self.model = model
        self.timestepping_params = timestepping_params
        self.output_params = output_params
        self.clock = Clock(0, 0.0)
        # set_dt_in_model! (Simulation.jl:56-59)
        model.params.dt = timestepping_params.dt


def update_clock(simulation):
    """update_clock! (run_simulation.jl:49-54)."""
    simulation.clock.n_iter += 1
    simulation.clock.time += simulation.timestepping_params.dt
    return simulation


def _run_steps(simulation, nsteps, verbose=False):
    model = simulation.model
    if _schwarz(model):
        # restricted additive Schwarz: the steps are driven from the host
        for _ in range(nsteps):
            update_state(model, verbose=verbose)
            if simulation.timestepping_params.step_thickness:
                update_thickness(simulation)
            update_clock(simulation)
        return simulation
    update_preconditioner(model)
    spec = model.parallel_spec
    push_fields(model)
    rep = _lib.SolveReport()
    if simulation.timestepping_params.step_thickness:
        _check(spec._lib, spec.ctx, spec._lib.wavi_run_steps(spec.ctx, nsteps, ctypes.byref(rep)))
    else:
        for _ in range(nsteps):
            _check(spec._lib, spec.ctx, spec._lib.wavi_update_state(spec.ctx, ctypes.byref(rep)))
    pull_fields(model)
    _report_solve(model, rep, verbose)
    for _ in range(nsteps):
        update_clock(simulation)
    return simulation


def timestep(simulation, verbose=False):
    """timestep!(simulation) (run_simulation.jl:6-17)."""
    return _run_steps(simulation, 1, verbose)


def update_thickness(simulation):
    """update_thickness!(simulation) (run_simulation.jl:24-42)."""
    model = simulation.model
    update_preconditioner(model)
    spec = model.parallel_spec
    push_fields(model)
    _check(spec._lib, spec.ctx, spec._lib.wavi_update_thickness(spec.ctx, simulation.timestepping_params.dt))
    pull_fields(model)
    return simulation


def run_simulation(simulation, verbose=False):
    """run_simulation!(simulation) (run_simulation.jl:61-100) without
    checkpoints or outputs."""
    n = simulation.timestepping_params.n_iter_total - simulation.clock.n_iter
    if n > 0:
        _run_steps(simulation, n, verbose)
    return simulation
