"""Simulation driver: Simulation, timestep!, update_thickness!, run_simulation!.

Mirrors src/Simulations/Simulation.jl:24-86 and run_simulation.jl:6-125.
The time loop runs on the device (wavi_run_steps) between the steps at
which run_simulation writes something: fields travel to the device once per
stretch and back at its end.  Checkpoints (temporary A/B and permanent,
with pickup from a permanent one) hold every field of the model and the
clock in numpy .npz files under the reference's names (ChkptA, ChkptB,
PChkpt_<10-digit step>), since no JLD2 writer is available here; the
velocity dump at the final step (dump_vel) is the reference's big-endian
<prefix>_U.bin / <prefix>_V.bin.
"""
import ctypes
import math
import os
from dataclasses import dataclass

import numpy as np

from . import _lib
from .core import ArgumentError, OutputParams
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
        self.model = model
        self.timestepping_params = timestepping_params
        self.output_params = OutputParams() if output_params is None else output_params
        self.clock = Clock(0, 0.0)
        # set_dt_in_model! (Simulation.jl:56-59)
        model.params.dt = timestepping_params.dt
        pickup(self, pickup_output_update_flag)


CHECKPOINT_EXT = ".npz"


def save_checkpoint(simulation, fname):
    """Every field of the model (the device state of a solve) and the clock."""
    from .model import FIELD_MAP
    m = simulation.model
    arrays = {fid: np.asarray(m._array(c, a), dtype=np.float64) for c, a, fid in FIELD_MAP}
    arrays["clock"] = np.array([simulation.clock.n_iter, simulation.clock.time], dtype=np.float64)
    arrays["dt"] = np.array([m.params.dt])
    np.savez(fname, **arrays)


def load_checkpoint(simulation, fname):
    """Restore the fields, the model's dt and the clock written by
    save_checkpoint into the simulation's model (built with the same grid and
    parameters).  pickup! replaces the model with the saved one
    (Simulation.jl:77-79), so the model's dt, which the velocity solve's
    implicit terms use, is the saved dt even where the new
    TimesteppingParams has another; the thickness steps use the new
    TimesteppingParams.dt, as in the reference (run_simulation.jl:28)."""
    from .model import FIELD_MAP
    m = simulation.model
    with np.load(fname, allow_pickle=False) as d:
        for c, a, fid in FIELD_MAP:
            m._set_array(c, a, d[fid])
        n_iter, time = d["clock"]
        m.params.dt = float(d["dt"][0])
    simulation.clock = Clock(int(n_iter), float(time))
    return simulation


def pickup(simulation, pickup_output_update_flag=False):
    """pickup! (Simulation.jl:68-86): with niter0 > 0, resume from
    PChkpt_<niter0, 10 digits> in the working directory."""
    niter0 = simulation.timestepping_params.niter0
    if niter0 > 0:
        fname = "PChkpt_%010d%s" % (niter0, CHECKPOINT_EXT)
        print("detected niter0 > 0 (niter0 = %d). Looking for pickup..." % niter0)
        if not os.path.exists(fname):
            raise RuntimeError("Pickup error, terminating run: %s not found" % fname)
        load_checkpoint(simulation, fname)
        print("Pickup successful")
    return simulation


def write_vel(simulation):
    """write_vel (run_simulation.jl:107-125): u without its last x row and v
    without its last y column, big-endian Float64, column-major."""
    m = simulation.model
    prefix = simulation.output_params.prefix
    np.asfortranarray(m.fields.gu.u[:-1, :]).astype(">f8").ravel(order="F").tofile(prefix + "_U.bin")
    np.asfortranarray(m.fields.gv.v[:, :-1]).astype(">f8").ravel(order="F").tofile(prefix + "_V.bin")


def update_clock(simulation):
    """update_clock! (run_simulation.jl:49-54)."""
    simulation.clock.n_iter += 1
    simulation.clock.time += simulation.timestepping_params.dt
    return simulation


def _run_steps(simulation, nsteps, verbose=False):
    model = simulation.model
    if _schwarz(model):
        # restricted additive Schwarz: each step's pieces are launched from the
        # host, on fields that stay on the device for the whole stretch
        from .schwarz import state_steps_on_device
        update_preconditioner(model)
        push_fields(model)
        state_steps_on_device(model, nsteps, simulation.timestepping_params.dt,
                              simulation.timestepping_params.step_thickness, verbose)
        pull_fields(model)
        for _ in range(nsteps):
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
    """run_simulation!(simulation) (run_simulation.jl:61-100): the steps run
    on the device in stretches that end at the checkpoint steps and the final
    step, where the checkpoints and the velocity dump are written."""
    tp = simulation.timestepping_params
    events = [tp.n_iter_chkpt, tp.n_iter_pchkpt]
    chkpt_tag = "A"
    i = simulation.clock.n_iter
    while i < tp.n_iter_total:
        # the next step at which something is written (or the last step)
        nxt = tp.n_iter_total
        for e in events:
            if e != math.inf:
                nxt = min(nxt, (i // e + 1) * e)
        _run_steps(simulation, nxt - i, verbose)
        i = nxt
        if tp.n_iter_chkpt != math.inf and i % tp.n_iter_chkpt == 0:
            save_checkpoint(simulation, "%sChkpt%s%s" % (tp.chkpt_path, chkpt_tag, CHECKPOINT_EXT))
            chkpt_tag = "B" if chkpt_tag == "A" else "A"
            print("making temporary checkpoint at timestep number %d" % simulation.clock.n_iter)
        if tp.n_iter_pchkpt != math.inf and i % tp.n_iter_pchkpt == 0:
            save_checkpoint(simulation, "%sPChkpt_%010d%s" % (tp.chkpt_path, simulation.clock.n_iter,
                                                              CHECKPOINT_EXT))
            print("making permanent checkpoint at timestep number %d" % simulation.clock.n_iter)
        if i == tp.n_iter_total and simulation.output_params.dump_vel:
            write_vel(simulation)
    return simulation
