"""wavi_update_velocities_lockstep_noalloc: the lockstep ensemble solve with
its coarse-CG tables in the library's static device scratch instead of
per-call allocations.  Through the C ABI, its results must equal those of
wavi_update_velocities_lockstep bit for bit (velocities, reports), on two
calls in a row, the second with fewer and smaller members than the first
(so stale tables of the first call would show).  The size limit is checked
before any device work, so its rejection runs without a GPU."""
import ctypes
import sys

import numpy as np
import pytest


def _step(name):
    print("[gpu] %s" % name, flush=True)
    sys.stdout.flush()


def _call(W, fn, models):
    """Run the named lockstep entry point on the models' contexts; returns the
    velocities and reports."""
    lib = W._lib.load()
    n = len(models)
    ctxs = (ctypes.c_void_p * n)(*[m.parallel_spec.ctx for m in models])
    reps = (W._lib.SolveReport * n)()
    codes = (ctypes.c_int * n)()
    rc = getattr(lib, fn)(ctxs, n, reps, codes)
    assert rc == W._lib.WAVI_OK, (fn, rc, lib.wavi_last_error(None))
    assert list(codes) == [0] * n
    out = []
    for m, rep in zip(models, reps):
        W.pull_fields(m)
        out.append((m.fields.gu.u.copy(), m.fields.gv.v.copy(), m.fields.gh.ηav.copy(), rep.as_dict()))
    return out


def _members(W, specs):
    ms = []
    for sp in specs:
        m = W.workloads.ensemble_member(sp, "mixed", maxiter_picard=6)
        W.update_geometry(m)
        ms.append(m)
    return ms


@pytest.mark.gpu
def test_noalloc_lockstep_equals_lockstep_on_two_calls(W):
    small = [sp for sp in W.workloads.ensemble_specs(64, seed=5) if sp["nx"] <= 320]
    first = small[:8]
    second = [sp for sp in small[8:] if sp["nx"] == 160][:3]
    assert len({(s["nx"], s["nσ"]) for s in first}) > 1 and len(second) == 3
    assert sum(s["nx"] * s["ny"] for s in second) < sum(s["nx"] * s["ny"] for s in first)
    for call, specs in enumerate((first, second)):
        _step("no-alloc lockstep, call %d: %d members %s" % (call + 1, len(specs),
                                                             sorted({(s["nx"], s["ny"]) for s in specs})))
        ref = _call(W, "wavi_update_velocities_lockstep", _members(W, specs))
        got = _call(W, "wavi_update_velocities_lockstep_noalloc", _members(W, specs))
        for k, (a, b) in enumerate(zip(ref, got)):
            for x, y in zip(a[:3], b[:3]):
                np.testing.assert_array_equal(x, y, err_msg="call %d member %d" % (call + 1, k))
            assert a[3] == b[3], (call + 1, k)
            assert b[3]["picard_iters"] > 0 and b[3]["coarse_cg_total"] > 0


def test_noalloc_lockstep_rejects_more_members_than_its_scratch_holds(W):
    """More than WAVI_NOALLOC_MAX_MEMBERS members: WAVI_EARG with a message,
    before any context is touched (so NULL contexts do not matter)."""
    lib_path = W._lib.LIB_PATH
    import os
    if not os.path.exists(lib_path):
        pytest.skip("libwavi_hip.so not built")
    lib = W._lib.load()
    n = 65
    ctxs = (ctypes.c_void_p * n)()
    assert lib.wavi_update_velocities_lockstep_noalloc(ctxs, n, None, None) == W._lib.WAVI_EARG
    assert b"at most 64 members" in lib.wavi_last_error(None)
    assert lib.wavi_update_velocities_lockstep_noalloc(ctxs, 0, None, None) == W._lib.WAVI_OK
