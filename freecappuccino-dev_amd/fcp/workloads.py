"""Synthetic workloads of the assemble-and-solve path.

Every input comes from ``fcp.rng`` (counter-based SplitMix64) so the oracle,
the reference driver and the GPU path receive identical arrays.

poisson_uniform (32^3 ICCG case)
    Unit cube, N^3 uniform hexahedra.  mu = -1 per cell passed to
    ``laplacian`` (SPD -div(grad) operator, laplacian.f90:45-96 with the sign
    convention of applications/Poisson/poisson.f90:75), Dirichlet phi_b = 0 on
    all six patches, cell source s ~ U[-1,1] (seed 0, stream 0), right-hand side
    b_i = s_i * V_i (poisson.f90:63).  ICCG, tol_rel 1e-10, tol_abs 0.

poisson_jittered (128^3 pressure-correction-style case)
    N^3 hexahedra, interior nodes jittered by U(-0.2h, 0.2h) per coordinate
    (seed 2).  Face diffusivity Gamma_f = 10^U(-1,1) (seed 2, stream 10) on all
    faces; source s ~ U[-1,1] made zero-mean (stream 11); Dirichlet phi_b = 0
    on the x- patch ('wall'), zero-gradient elsewhere ('symmetry').  One step
    = Laplacian assembly with face diffusivity (operator -div(Gamma grad),
    i.e. mu_f = -Gamma_f in the reference's sign convention) + Gauss gradient
    of the previous iterate phi_prev + explicit non-orthogonal correction
    Gamma_f (grad phi)_f . (S_f - D_f d_PN) (scalar_fluxes.f90:99-109) added to
    b = s*V + boundary terms, then a solve to ||b-Ax||_2/||b||_2 <= 1e-8 from
    x0 = 0.  phi_prev is the smooth field sin(2 pi x) cos(2 pi y) exp(z) at
    cell centres with boundary values from the Dirichlet value (x-) and
    zero-gradient copies elsewhere (updateBoundary.f90:31-44).
"""
from __future__ import annotations

import numpy as np

from .mesh import hex_mesh
from .rng import uniform


def poisson_uniform_mesh(N: int = 32):
    return hex_mesh(N, jitter=0.0, types=("wall",) * 6)


def poisson_uniform_inputs(mesh, vol):
    n = mesh.n_cells
    s = uniform(0, 0, n, -1.0, 1.0)
    return dict(mu=-np.ones(n), phi=np.zeros(mesh.n_total), source=s, rhs=s * vol)


def poisson_jittered_mesh(N: int = 128):
    return hex_mesh(N, jitter=0.2, seed=2,
                    types=("wall", "symmetry", "symmetry", "symmetry", "symmetry", "symmetry"))


def poisson_jittered_inputs(mesh, vol, xc, yc, zc):
    n, F = mesh.n_cells, mesh.n_faces
    gamma = 10.0 ** uniform(2, 10, F, -1.0, 1.0)
    s = uniform(2, 11, n, -1.0, 1.0)
    s = s - s.mean()
    bc_kind = np.where(mesh.patch_of_bface() == 0, 0, 1).astype(np.uint8)
    phi_prev = np.zeros(mesh.n_total)
    phi_prev[:n] = np.sin(2 * np.pi * xc) * np.cos(2 * np.pi * yc) * np.exp(zc)
    own_b = mesh.owner[mesh.n_inner:]
    phi_prev[n:] = np.where(bc_kind == 0, 0.0, phi_prev[own_b])
    return dict(gamma=gamma, mu_f=-gamma, source=s, bc_kind=bc_kind, phi_prev=phi_prev,
                phi_bc=np.zeros(mesh.n_total))
