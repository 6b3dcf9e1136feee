"""Checkpoints, pickup and the velocity dump of run_simulation
(run_simulation.jl:61-125, Simulation.jl:68-86) with the device loop: a run
interrupted at a permanent checkpoint and picked up from it ends bit-equal
to the uninterrupted run; the dump holds u and v in the reference's layout."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _step(name):
    print("[gpu] %s" % name, flush=True)
    sys.stdout.flush()


def _sim(W, path, niter0=0, n=20, dump=False, precision="fp64"):
    m = W.workloads.mismip_plus_fixture(precision)
    tp = W.TimesteppingParams(niter0=niter0, dt=0.1, n_iter_total=n, chkpt_freq=0.5, pchkpt_freq=1.0,
                              chkpt_path=str(path))
    op = W.OutputParams(prefix=os.path.join(str(path), "run"), dump_vel=dump)
    return W.Simulation(model=m, timestepping_params=tp, output_params=op)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pickup_continues_bit_equal(W, tmp_path, monkeypatch, precision):
    _step("checkpoint / pickup, %s" % precision)
    monkeypatch.chdir(tmp_path)
    full = _sim(W, tmp_path, dump=True, precision=precision)
    W.run_simulation(full)
    names = sorted(os.listdir(tmp_path))
    assert "PChkpt_0000000010.npz" in names and "PChkpt_0000000020.npz" in names
    assert "ChkptA.npz" in names and "ChkptB.npz" in names
    # the velocity dump: u without its last x row, v without its last y column, big-endian
    u = np.fromfile(os.path.join(str(tmp_path), "run_U.bin"), dtype=">f8")
    v = np.fromfile(os.path.join(str(tmp_path), "run_V.bin"), dtype=">f8")
    np.testing.assert_array_equal(u, full.model.fields.gu.u[:-1, :].ravel(order="F"))
    np.testing.assert_array_equal(v, full.model.fields.gv.v[:, :-1].ravel(order="F"))
    resumed = _sim(W, tmp_path, niter0=10, precision=precision)
    assert resumed.clock.n_iter == 10
    W.run_simulation(resumed)
    assert resumed.clock.n_iter == 20
    for a, b in ((full.model.fields.gh.h, resumed.model.fields.gh.h),
                 (full.model.fields.gu.u, resumed.model.fields.gu.u),
                 (full.model.fields.gv.v, resumed.model.fields.gv.v),
                 (full.model.fields.g3d.η, resumed.model.fields.g3d.η)):
        np.testing.assert_array_equal(a, b)


def test_missing_pickup_raises(W, tmp_path, monkeypatch):
    _step("missing pickup")
    monkeypatch.chdir(tmp_path)
    with pytest.raises(RuntimeError):
        _sim(W, tmp_path, niter0=30)
