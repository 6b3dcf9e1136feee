"""torch.autograd binding of the adjoint basal-drag gradient.

MisfitProblem fixes a model's starting state (the fields as they are when
the problem is made), the observations and the Picard count; its misfit is
then a pure function of log C, evaluated by wavi_misfit_gradient: the
forward Picard iterations with converged solves and, in the same call, the
reverse sweep whose gradient backward() returns.  Import this module
explicitly (it needs torch; the package itself does not).
No reference counterpart: WAVI.jl has no adjoint.
"""
import numpy as np
import torch

from . import model as _model


class MisfitProblem:
    """J(log C) = 1/2 ||u - u_obs||^2 for one model (see misfit_gradient)."""

    def __init__(self, model, n_picard, u_obs, v_obs, tol_linear=1e-13, maxit_linear=60):
        self.model = model
        self.n_picard = int(n_picard)
        self.u_obs, self.v_obs = u_obs, v_obs
        self.tol_linear, self.maxit_linear = float(tol_linear), int(maxit_linear)
        self._start = [(c, a, np.array(model._array(c, a), copy=True)) for c, a, _ in _model.FIELD_MAP]

    def evaluate(self, logc, gradient=True):
        """(J, dJ/dlog C or None, report) at log C (an nx x ny array)."""
        m = self.model
        for c, a, v in self._start:
            m._set_array(c, a, v.copy())
        m.params.weertman_c = np.asfortranarray(np.exp(np.asarray(logc, dtype=np.float64)))
        _model.update_geometry(m)
        return _model.misfit_gradient(m, self.n_picard, self.u_obs, self.v_obs, self.tol_linear,
                                      self.maxit_linear, gradient=gradient)


class MisfitLogC(torch.autograd.Function):
    """J = MisfitLogC.apply(logc, problem): the misfit as a differentiable
    function of a float64 tensor of shape (nx, ny)."""

    @staticmethod
    def forward(ctx, logc, problem):
        J, g, _ = problem.evaluate(logc.detach().cpu().numpy(), gradient=logc.requires_grad)
        if g is not None:
            ctx.save_for_backward(torch.from_numpy(np.ascontiguousarray(g)).to(logc.device, logc.dtype))
        return logc.new_tensor(J)

    @staticmethod
    def backward(ctx, J_bar):
        (g,) = ctx.saved_tensors
        return J_bar * g, None
