"""Discrete adjoint of the steady convection-diffusion path on the GPU.

Primal map (parameters: face diffusivity Gamma[F], cell source s[n], values
phi_b of the Dirichlet boundary faces):

    phi_0 = 0 in the cells; boundary entries = phi_b on Dirichlet faces,
            the owner's value on every other boundary face
    for k = 1..K (deferred-correction passes):
        g_{k-1} = Gauss gradient of phi_{k-1}             (fcp_grad_gauss)
        A, b    = facefluxsc assembly with Gamma, phi_{k-1}, g_{k-1},
                  b = s V + face and boundary terms        (fcp_convdiff, urf 1)
        phi_k   = A^{-1} b         (FP64 DILU-BiCGStab, ||b-Ax||/||b|| <= 1e-13,
                                    then one FP64 refinement pass)
    J = sum_i w_i phi_K,i V_i

A depends only on Gamma and the mass flux, so every pass shares it.  The
reverse sweep solves A^T lam_k = phibar_k (A^T by the face-based coefficient
swap, fcp_matrix_transpose), then accumulates with the backward kernels
(fcp_convdiff_backward, fcp_grad_gauss_backward):

    Gammabar += -lam^T (dA/dGamma) phi_k + lam^T db/dGamma
    sbar     += lam V
    phibar_{k-1} = (db/dphi_{k-1})^T lam + G^T (db/dg_{k-1})^T lam

and routes the boundary entries of phibar_{k-1} to phi_b (Dirichlet faces)
or to the owner cell (copied faces).  ``ConvDiffSolve`` wraps the map as a
torch.autograd.Function.
"""
from __future__ import annotations

import numpy as np
import torch

from .api import Context, FCP_CRIT_TRUE_L2, FCP_PREC_FP64


class ConvDiffProblem:
    """One mesh on one device with a fixed mass flux and boundary kinds."""

    def __init__(self, ctx: Context, mesh, geo, flux, bc_kind, K: int = 2, gam: float = 1.0,
                 scheme: int = 1, tol: float = 1e-13, itr_max: int = 5000,
                 device: str = "cuda:0", log=None, refine: int = 1):
        self.ctx, self.mesh = ctx, mesh
        self.K, self.gam, self.scheme, self.tol, self.itr_max = K, gam, scheme, tol, itr_max
        self.refine = refine
        self.log = log  # callable taking one line per linear solve, or None
        dev = torch.device(device)
        self.dev = dev
        T = lambda a, dt=torch.float64: torch.from_numpy(np.ascontiguousarray(a)).to(dev, dt)  # noqa: E731
        self.n, self.nt = mesh.n_cells, mesh.n_total
        self.flux = T(flux)
        self.bc_kind = T(np.asarray(bc_kind, np.uint8), torch.uint8)
        self.dirichlet = T(np.asarray(bc_kind) == 0, torch.bool)
        self.own_b = T(mesh.owner[mesh.n_inner:].astype(np.int64), torch.int64)
        self.vol = T(geo["vol"])

    def _full(self, phi_int, phib):
        phi = torch.empty(self.nt, dtype=torch.float64, device=self.dev)
        phi[: self.n] = phi_int
        phi[self.n:] = torch.where(self.dirichlet, phib, phi_int[self.own_b])
        return phi

    def _solve_once(self, rhs, tol):
        """e = A^{-1} rhs to ||rhs - A e||/||rhs|| <= tol.  The reference's
        BiCGStab guards its denominators with an absolute 1e-20
        (linear_solvers.f90:652,747), which stalls it once the squared norms
        of its vectors approach that value; rhs is therefore scaled by a power
        of two to a norm of about 2^20 (exact in binary arithmetic) and the
        solution scaled back."""
        nrm = float(torch.linalg.vector_norm(rhs))
        k = 20 - int(np.floor(np.log2(nrm))) if nrm > 0 else 0
        x = torch.zeros(self.n, dtype=torch.float64, device=self.dev)
        rc, info = self.ctx.csrsolve("bicgstab", x, rhs * (2.0 ** k),
                                     itr_max=self.itr_max, tol_abs=0.0, tol_rel=tol,
                                     precision=FCP_PREC_FP64, criterion=FCP_CRIT_TRUE_L2)
        if self.log is not None:
            self.log(f"solve: {info['iterations']} iterations, rel. residual "
                     f"{info['rel_res_l2']:.2e}, {info['time_ms']:.0f} ms")
        if rc != 0:
            raise RuntimeError(f"adjoint path: solve did not converge ({info})")
        return x * (2.0 ** -k)

    def _solve(self, rhs):
        """x = A^{-1} rhs (the resident matrix: A, or A^T during backward):
        one solve to tol, then `refine` FP64 refinement passes that solve for
        the correction of the true residual to 1e-6 of it.  Central
        differences with step h resolve the sensitivities only to about
        (solution error)/h, so the solutions must be accurate far below the
        Krylov tolerance."""
        x = self._solve_once(rhs, self.tol)
        y = torch.empty_like(x)
        for _ in range(self.refine):
            self.ctx.spmv(x, y)
            r = rhs - y
            if float(torch.linalg.vector_norm(r)) == 0.0:
                break
            x = x + self._solve_once(r, 1e-6)
        return x

    def _assemble(self, gamma, phi_old, g, su):
        """A (resident in the context) and b (into su) of one pass."""
        sp = torch.zeros(self.n, dtype=torch.float64, device=self.dev)
        work = phi_old.clone()  # fcp_convdiff rewrites zero-gradient boundary entries
        self.ctx.convdiff(self.flux, gamma, work, g, self.gam, self.scheme, self.bc_kind, su, sp,
                          urf=1.0)

    def run(self, gamma, s, phib):
        """(phi_K, tape): the cells' solution for parameters gamma[F], s[n],
        phib[F_b] (device) and the per-call record the reverse sweep needs
        (gamma and every pass's input field, gradient and solution)."""
        n, nt = self.n, self.nt
        gamma = gamma.detach().contiguous().clone()
        phi_int = torch.zeros(n, dtype=torch.float64, device=self.dev)
        states = []
        for _ in range(self.K):
            phi_old = self._full(phi_int, phib)
            g = torch.zeros((nt, 3), dtype=torch.float64, device=self.dev)
            self.ctx.grad_gauss(phi_old, out=g)
            su = s * self.vol
            self._assemble(gamma, phi_old, g, su)
            phi_int = self._solve(su)
            states.append((phi_old, g, phi_int))
        return phi_int, {"gamma": gamma, "states": states}

    def forward(self, gamma, s, phib):
        """phi_K (cells) for parameters gamma[F], s[n], phib[F_b] (device)."""
        return self.run(gamma, s, phib)[0]

    def backward(self, phibar, tape):
        """(gamma_bar[F], s_bar[n], phib_bar[F_b]) for dJ/dphi_K = phibar at
        the forward call that recorded `tape`.  The context's matrix may have
        changed since (other forward calls, other assemblies), so A is first
        re-assembled from the taped gamma, then transposed in place."""
        n, nt = self.n, self.nt
        F = self.mesh.n_faces
        gamma, states = tape["gamma"], tape["states"]
        gamma_bar = torch.zeros(F, dtype=torch.float64, device=self.dev)
        s_bar = torch.zeros(n, dtype=torch.float64, device=self.dev)
        phib_bar = torch.zeros(nt - n, dtype=torch.float64, device=self.dev)
        src = phibar.contiguous()
        self._assemble(gamma, states[0][0], states[0][1],
                       torch.zeros(n, dtype=torch.float64, device=self.dev))
        self.ctx.matrix_transpose()
        try:
            for k in reversed(range(self.K)):
                phi_old, g, phi_new = states[k]
                lam = self._solve(src)
                s_bar += lam * self.vol
                pob = torch.zeros(nt, dtype=torch.float64, device=self.dev)
                gob = torch.zeros((nt, 3), dtype=torch.float64, device=self.dev)
                self.ctx.convdiff_backward(self.flux, gamma, phi_old, g, self.gam,
                                           self.scheme, self.bc_kind, lam, phi_new, gamma_bar,
                                           pob, gob)
                self.ctx.grad_gauss_backward(gob, pob)
                bnd = pob[n:]
                phib_bar += torch.where(self.dirichlet, bnd, torch.zeros_like(bnd))
                cells = pob[:n].clone()
                cells.index_add_(0, self.own_b, torch.where(self.dirichlet, torch.zeros_like(bnd), bnd))
                src = cells
        finally:
            self.ctx.matrix_transpose()
        return gamma_bar, s_bar, phib_bar


class ConvDiffSolve(torch.autograd.Function):
    """phi_K = ConvDiffSolve.apply(gamma, s, phib, problem)."""

    @staticmethod
    def forward(ctx, gamma, s, phib, problem):
        phi, tape = problem.run(gamma, s, phib)
        ctx.problem = problem
        ctx.tape = tape
        return phi

    @staticmethod
    def backward(ctx, phibar):
        g, s, b = ctx.problem.backward(phibar, ctx.tape)
        return g, s, b, None
