"""Drag of a rotating, blowing cylinder and its gradient by the discrete adjoint.

The body of a LUO case (a circle of Lagrangian points) is given the surface velocity

    u_k = Vt t_k + b_k n_k,   Vt = alpha (D / 2) omega_ref,

with t_k, n_k the tangent and outward normal at point k (angle theta_k about the centre),
``alpha`` the rotation rate in units of ``omega_ref`` and ``b_k`` a normal blowing
velocity per point; the solid nodes inside turn rigidly at Omega = alpha omega_ref. The
objective is the mean drag coefficient over steps n0..n1 of an nsteps run from the
impulsive start. ``cuibm_adjoint_run`` returns dJ/du_k, dJ/dv_k and dJ/dOmega; this
module applies the chain rule to (alpha, b) and puts the pair behind
``torch.autograd.Function`` so that ``J.backward()`` fills ``alpha.grad`` and
``b.grad``. The reference has no adjoint: this is a build extension over its LUO step
(src/solvers/NavierStokes/luoIBM.cu:60-109).
"""
import numpy as np
import torch

import cuibm_amd

__all__ = ["RotatingCylinder", "MeanDrag", "mean_drag"]


class RotatingCylinder:
    """One objective evaluation per call: a fresh solver of `case_folder` (with the
    command-line `overrides`), the surface velocities of (alpha, blow), and the forward
    run, or the forward run and the adjoint sweep, of `nsteps` steps."""

    def __init__(self, case_folder, nsteps, n0, n1, every, overrides=(), omega_ref=2.0, diameter=1.0):
        self.case_folder = str(case_folder)
        self.overrides = [str(a) for a in overrides]
        self.nsteps, self.n0, self.n1, self.every = int(nsteps), int(n0), int(n1), int(every)
        self.omega_ref, self.diameter = float(omega_ref), float(diameter)
        s = cuibm_amd.Solver(self.case_folder, *self.overrides, diagnostics=False)
        bx, by = s.case_array("body_X"), s.case_array("body_Y")
        s.close()
        self.theta = np.arctan2(by - by.mean(), bx - bx.mean())
        self.npts = bx.size
        self.last = None

    def surface(self, alpha, blow):
        """(ub, vb, Omega) of rotation rate alpha and normal blowing `blow`"""
        blow = np.asarray(blow, dtype=np.float64)
        if blow.shape != (self.npts,):
            raise ValueError(f"blow needs {self.npts} values, one per body point")
        vt = alpha * 0.5 * self.diameter * self.omega_ref
        c, s = np.cos(self.theta), np.sin(self.theta)
        return -vt * s + blow * c, vt * c + blow * s, alpha * self.omega_ref

    def evaluate(self, alpha, blow, backward=True):
        """J and, with backward, (dJ/dalpha, dJ/dblow); the run's statistics in .last"""
        ub, vb, om = self.surface(alpha, blow)
        with cuibm_amd.Solver(self.case_folder, *self.overrides, diagnostics=False) as s:
            s.initialise()
            s.set_body_velocity(ub, vb, om)
            r = s.adjoint_run(self.nsteps, self.n0, self.n1, self.every, backward=backward)
        self.last = r
        if not backward:
            return r["J"], None, None
        c, sn = np.cos(self.theta), np.sin(self.theta)
        gu, gv = r["grad_ub"], r["grad_vb"]
        dvt = 0.5 * self.diameter * self.omega_ref
        galpha = float(np.sum(gu * (-sn) + gv * c) * dvt + r["grad_omega"] * self.omega_ref)
        gblow = gu * c + gv * sn
        return r["J"], galpha, gblow


class MeanDrag(torch.autograd.Function):
    """J = problem's mean drag at (alpha, blow); the adjoint sweep runs in the forward
    call when a gradient is wanted (one forward and one backward pass of the flow)."""

    @staticmethod
    def forward(ctx, alpha, blow, problem):
        want = bool(ctx.needs_input_grad[0] or ctx.needs_input_grad[1])
        J, ga, gb = problem.evaluate(float(alpha.detach()), blow.detach().cpu().numpy(), backward=want)
        if want:
            ctx.save_for_backward(torch.tensor(ga, dtype=torch.float64), torch.from_numpy(np.array(gb)))
        return torch.tensor(J, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gJ):
        ga, gb = ctx.saved_tensors
        return gJ * ga, gJ * gb, None


def mean_drag(alpha, blow, problem):
    """differentiable mean drag: alpha a 0-d and blow a 1-d float64 tensor"""
    return MeanDrag.apply(alpha, blow, problem)
