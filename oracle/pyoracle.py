"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, the package smoke check and the
benchmark's CPU-baseline leg as the checker.  An OracleModel is built from a
host-side Model of the package (same grid, parameters and fields) and
restates the reference algorithm on the CPU.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_dp = ctypes.POINTER(ctypes.c_double)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    lib = ctypes.CDLL(LIB)
    vp = ctypes.c_void_p
    sig = {
        "om_create": (vp, [vp, vp, vp]),
        "om_destroy": (None, [vp]),
        "om_field": (_dp, [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
        "om_set_dt": (None, [vp, ctypes.c_double]),
        "om_init_quad_f2": (None, [vp]),
        "om_masks": (None, [vp, _u8p, _u8p, _u8p, _u8p, _u8p]),
        "om_op_apply": (None, [vp, _dp, _dp, _dp, _dp]),
        "om_inner_update": (None, [vp]),
        "om_rhs": (None, [vp, _dp, _dp]),
        "om_op_diag": (None, [vp, _dp, _dp]),
        "om_vcycle": (ctypes.c_int, [vp, _dp, _dp, _dp, _dp]),
        "om_coarse_op": (None, [vp, _dp, _dp, _dp, _dp]),
        "om_coarse_cg": (ctypes.c_int, [vp, _dp, _dp, _dp, _dp, ctypes.c_double, ctypes.c_int]),
        "om_precondition": (ctypes.c_int, [vp, _dp, ctypes.POINTER(ctypes.c_int)]),
        "om_update_velocities": (ctypes.c_int, [vp, vp]),
        "om_update_state": (None, [vp, vp]),
        "om_update_geometry": (None, [vp]),
        "om_update_thickness": (None, [vp, ctypes.c_double]),
        "om_run_steps": (None, [vp, ctypes.c_int, vp]),
        "om_update_wavelets": (None, [vp]),
        "om_haar": (None, [vp, ctypes.c_int, ctypes.c_int, _dp, _dp]),
        "om_wavelet_matrix": (None, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]),
        "om_canonical_dot": (ctypes.c_double, [vp, _dp, _dp, _dp, _dp]),
        "om_det_pow": (ctypes.c_double, [ctypes.c_double, ctypes.c_double]),
        "om_pos_fraction": (None, [ctypes.c_int, ctypes.c_int, _dp, _u8p, _dp, _dp, _dp]),
        "om_misfit_gradient": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp, _dp,
                                              ctypes.POINTER(ctypes.c_double), _dp, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _d(a):
    return a.ctypes.data_as(_dp)


def det_pow(x, y):
    return load().om_det_pow(float(x), float(y))


def wavelet_matrix(n, levels, forward):
    """Dense wavelet_matrix(n, levels, "forward"/"reverse")."""
    W = np.empty((n, n))
    load().om_wavelet_matrix(n, levels, 1 if forward else 0, _d(W))
    return W


def pos_fraction(z, mask=None):
    """pos_fraction(z1; mask) (src/utilities.jl:255-396)."""
    z = np.asfortranarray(z, dtype=float)
    M, N = z.shape
    mask = np.ones((M, N), bool) if mask is None else np.asarray(mask, bool)
    zb = np.ascontiguousarray(z.ravel(order="F"))
    mb = np.ascontiguousarray(mask.ravel(order="F").astype(np.uint8))
    af = np.empty(M * N)
    au = np.empty((M + 1) * N)
    av = np.empty(M * (N + 1))
    load().om_pos_fraction(M, N, _d(zb), mb.ctypes.data_as(_u8p), _d(af), _d(au), _d(av))
    return (af.reshape((M, N), order="F"), au.reshape((M + 1, N), order="F"),
            av.reshape((M, N + 1), order="F"))


class OracleModel:
    """CPU oracle state built from a host Model of the package."""

    def __init__(self, model):
        import importlib
        pkg = importlib.import_module(type(model).__module__.rsplit(".", 1)[0])
        self._pkg = pkg
        self._lib = load()
        self.model = model
        self.nx, self.ny = model.grid.nx, model.grid.ny
        keep, desc, phys, solver = pkg.model.grid_desc(model)
        self._keep = (keep, desc, phys, solver)
        self.h = self._lib.om_create(ctypes.addressof(desc), ctypes.addressof(phys),
                                     ctypes.addressof(solver))
        self.load_fields(model)

    def __del__(self):
        if getattr(self, "h", None):
            self._lib.om_destroy(self.h)
            self.h = None

    def _view(self, fid):
        n = ctypes.c_int(0)
        p = self._lib.om_field(self.h, fid, ctypes.byref(n))
        return np.ctypeslib.as_array(p, shape=(n.value,))

    def load_fields(self, model):
        """Copy every field of a host Model into the oracle."""
        _l = self._pkg._lib
        for cont, attr, fid in self._pkg.model.FIELD_MAP:
            a = model._array(cont, attr)
            self._view(_l.FIELD[fid])[:] = np.asarray(a, dtype=float).ravel(order="F")
        self._lib.om_set_dt(self.h, model.params.dt)

    def store_fields(self, model):
        """Copy every oracle field into a host Model."""
        _l = self._pkg._lib
        for cont, attr, fid in self._pkg.model.FIELD_MAP:
            model._set_array(cont, attr, self._view(_l.FIELD[fid]).copy())

    def field(self, name):
        """A copy of field `name` (a library field name such as "U_U") in its
        natural shape."""
        _l = self._pkg._lib
        fid = _l.FIELD[name]
        v = self._view(fid).copy()
        nx, ny = self.nx, self.ny
        if name.startswith(("H_", "P_")):
            return v.reshape((nx, ny), order="F")
        if name.startswith(("U_", "WU_")):
            return v.reshape((nx + 1, ny), order="F")
        if name.startswith(("V_", "WV_")):
            return v.reshape((nx, ny + 1), order="F")
        return v.reshape((nx, ny, -1), order="F")

    def set_dt(self, dt):
        self._lib.om_set_dt(self.h, dt)

    def _u(self, a):
        return np.ascontiguousarray(np.asarray(a, dtype=float).ravel(order="F"))

    def _shape(self, b, comp):
        nx, ny = self.nx, self.ny
        return b.reshape((nx + 1, ny) if comp == "u" else (nx, ny + 1), order="F")

    def op_apply(self, xu, xv):
        bu, bv = self._u(xu), self._u(xv)
        yu, yv = np.empty_like(bu), np.empty_like(bv)
        self._lib.om_op_apply(self.h, _d(bu), _d(bv), _d(yu), _d(yv))
        return self._shape(yu, "u"), self._shape(yv, "v")

    def inner_update(self):
        self._lib.om_inner_update(self.h)

    def rhs(self):
        bu = np.empty((self.nx + 1) * self.ny)
        bv = np.empty(self.nx * (self.ny + 1))
        self._lib.om_rhs(self.h, _d(bu), _d(bv))
        return self._shape(bu, "u"), self._shape(bv, "v")

    def coarse_op(self, yu, yv):
        au, av = self._u(yu), self._u(yv)
        ou, ov = np.empty_like(au), np.empty_like(av)
        self._lib.om_coarse_op(self.h, _d(au), _d(av), _d(ou), _d(ov))
        return self._shape(ou, "u"), self._shape(ov, "v")

    def coarse_cg(self, bu, bv, xu, xv, tol_coarse, maxiter):
        """The coarse CG alone from the warm start (xu, xv): returns
        (iterations, xu, xv)."""
        b1, b2 = self._u(bu), self._u(bv)
        x1, x2 = self._u(xu).copy(), self._u(xv).copy()
        it = self._lib.om_coarse_cg(self.h, _d(b1), _d(b2), _d(x1), _d(x2), float(tol_coarse), int(maxiter))
        return it, self._shape(x1, "u"), self._shape(x2, "v")

    def op_diag(self):
        du = np.empty((self.nx + 1) * self.ny)
        dv = np.empty(self.nx * (self.ny + 1))
        self._lib.om_op_diag(self.h, _d(du), _d(dv))
        return self._shape(du, "u"), self._shape(dv, "v")

    def vcycle(self, ru, rv):
        bu, bv = self._u(ru), self._u(rv)
        eu, ev = np.empty_like(bu), np.empty_like(bv)
        it = self._lib.om_vcycle(self.h, _d(bu), _d(bv), _d(eu), _d(ev))
        return self._shape(eu, "u"), self._shape(ev, "v"), it

    def haar(self, comp, direction, x):
        d = {"forward": 0, "inverse": 1, "inverse_t": 2}[direction]
        b = self._u(x)
        out = np.empty_like(b)
        self._lib.om_haar(self.h, 0 if comp == "u" else 1, d, _d(b), _d(out))
        return self._shape(out, comp)

    def canonical_dot(self, au, av, bu=None, bv=None):
        a1, a2 = self._u(au), self._u(av)
        if bu is None:
            return self._lib.om_canonical_dot(self.h, _d(a1), _d(a2), None, None)
        b1, b2 = self._u(bu), self._u(bv)
        return self._lib.om_canonical_dot(self.h, _d(a1), _d(a2), _d(b1), _d(b2))

    def update_velocities(self):
        rep = self._pkg._lib.SolveReport()
        self._lib.om_update_velocities(self.h, ctypes.addressof(rep))
        return rep.as_dict()

    def update_state(self):
        rep = self._pkg._lib.SolveReport()
        self._lib.om_update_state(self.h, ctypes.addressof(rep))
        return rep.as_dict()

    def update_geometry(self):
        self._lib.om_update_geometry(self.h)

    def misfit_gradient(self, n_picard, tol_linear, maxit_linear, u_obs, v_obs, gradient=True):
        """J and (gradient=True) dJ/dlog C of the adjoint map; see
        wavi_misfit_gradient.  Returns (J, grad (nx, ny) or None, report)."""
        uo, vo = self._u(u_obs), self._u(v_obs)
        J = ctypes.c_double(0.0)
        g = np.zeros(self.nx * self.ny) if gradient else None
        rep = self._pkg._lib.AdjointReport()
        rc = self._lib.om_misfit_gradient(self.h, n_picard, tol_linear, maxit_linear, _d(uo), _d(vo),
                                          ctypes.byref(J), _d(g) if gradient else None, ctypes.addressof(rep))
        if rc != 0:
            raise ValueError("om_misfit_gradient: code %d" % rc)
        return J.value, (g.reshape((self.nx, self.ny), order="F") if gradient else None), rep.as_dict()

    def update_thickness(self, dt):
        self._lib.om_update_thickness(self.h, dt)

    def run_steps(self, nsteps):
        rep = self._pkg._lib.SolveReport()
        self._lib.om_run_steps(self.h, nsteps, ctypes.addressof(rep))
        return rep.as_dict()

    def update_wavelets(self):
        self._lib.om_update_wavelets(self.h)

    def masks(self):
        nx, ny = self.nx, self.ny
        um = np.empty((nx + 1) * ny, np.uint8)
        vm = np.empty(nx * (ny + 1), np.uint8)
        cm = np.empty(max((nx - 1) * (ny - 1), 1), np.uint8)
        ui = np.empty_like(um)
        vi = np.empty_like(vm)
        p = lambda a: a.ctypes.data_as(_u8p)  # noqa: E731
        self._lib.om_masks(self.h, p(um), p(vm), p(cm), p(ui), p(vi))
        return {"u_mask": self._shape(um, "u").astype(bool), "v_mask": self._shape(vm, "v").astype(bool),
                "c_mask": cm[:(nx - 1) * (ny - 1)].reshape((nx - 1, ny - 1), order="F").astype(bool),
                "u_inner": self._shape(ui, "u").astype(bool), "v_inner": self._shape(vi, "v").astype(bool)}
