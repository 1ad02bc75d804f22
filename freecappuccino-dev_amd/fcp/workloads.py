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


# ---------------------------------------------------------------------------
# Momentum-like convection-diffusion (192^3; 128^3 jittered for sensitivities)
# ---------------------------------------------------------------------------
# Unit cube of N^3 hexahedra; x+ is an outlet, the other five patches are
# no-slip walls.  Velocity u(x) = sum_{m<8} a_m e_m sin(2 pi k_m.x + theta_m)
# with a_m ~ N(0,1), k_m in {1..4}^3, theta_m ~ U[0,2pi), e_m a random unit
# vector (seed 5, streams 30-35).  Face fluxes F_f = u(x_f).S_f on inner and
# outlet faces (0 on walls) are made discretely divergence-free by one
# Poisson projection (unit face diffusivity, p = 0 on the outlet, zero normal
# gradient on walls): F_f <- F_f - L_PN (p_N - p_P), where L is the assembled
# Laplacian, so the corrected cell divergence equals the projection solve's
# residual.  Face diffusivity Gamma_f = max(|F_f|, F_min) / (Pe_f D_f) with
# Pe_f = 10^U(-1,2) (stream 40), F_min = 1e-3 mean |F_inner|.  Components
# u, v, w of the field are the three transported quantities; their sources
# s ~ U[-1,1] (streams 50-52) times V, linear-upwind deferred correction
# (gamma = 1) with Gauss gradients, under-relaxation 0.8, walls Dirichlet 0,
# outlet zero-gradient.  All three share one matrix.
CONVDIFF_SEED = 5
CONVDIFF_URF = 0.8
CONVDIFF_TYPES = ("wall", "outlet", "wall", "wall", "wall", "wall")


def convdiff_mesh(N: int = 192, jitter: float = 0.0):
    return hex_mesh(N, jitter=jitter, seed=2, types=CONVDIFF_TYPES)


def _normal01(seed, s1, s2, n):
    from .rng import uniform01
    u1 = uniform01(seed, s1, np.arange(n, dtype=np.uint64))
    u2 = uniform01(seed, s2, np.arange(n, dtype=np.uint64))
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)


def fourier_velocity(x, y, z, seed: int = CONVDIFF_SEED, modes: int = 8):
    """u(x) of the convection-diffusion workloads at points (x, y, z)."""
    from .rng import uniform01
    a = _normal01(seed, 30, 31, modes)
    k = 1.0 + np.floor(4.0 * uniform01(seed, 32, np.arange(3 * modes, dtype=np.uint64)))
    k = k.reshape(modes, 3)
    th = 2.0 * np.pi * uniform01(seed, 33, np.arange(modes, dtype=np.uint64))
    e = _normal01(seed, 34, 35, 3 * modes).reshape(modes, 3)
    e /= np.linalg.norm(e, axis=1, keepdims=True)
    u = np.zeros((np.size(x), 3))
    for m in range(modes):
        w = a[m] * np.sin(2.0 * np.pi * (k[m, 0] * x + k[m, 1] * y + k[m, 2] * z) + th[m])
        u += w[:, None] * e[m][None, :]
    return u


def convdiff_flux(mesh, geo, csr, ctx):
    """Divergence-free face fluxes by one Poisson projection on ``ctx`` (a
    Context holding ``mesh`` and its CSR arrays ``csr``, 0-based; the resident
    matrix is overwritten).  Returns (flux[F], cell divergence[n])."""
    n, Fi, F = mesh.n_cells, mesh.n_inner, mesh.n_faces
    uf = fourier_velocity(geo["xf"], geo["yf"], geo["zf"])
    flux = uf[:, 0] * geo["arx"] + uf[:, 1] * geo["ary"] + uf[:, 2] * geo["arz"]
    patch = mesh.patch_of_bface()
    outlet = np.asarray(mesh.bctype_codes[patch] == 4)
    flux[Fi:][~outlet] = 0.0
    own, nb = mesh.owner, mesh.neighbour
    div = np.bincount(own, weights=flux, minlength=n) - np.bincount(nb, weights=flux[:Fi],
                                                                    minlength=n)
    bc_kind = np.where(outlet, 0, 1).astype(np.uint8)
    a = np.zeros(ctx.nnz)
    rhs = -div
    ctx.laplacian_face(-np.ones(F), bc_kind, np.zeros(mesh.n_total), rhs, a_out=a)
    p = np.zeros(n)
    rc, info = ctx.csrsolve("iccg", p, rhs, itr_max=20000, tol_abs=0.0, tol_rel=1e-13,
                            criterion=1)
    # L = -a: F_f -= L_PN (p_N - p_P) = F_f + a_PN (p_N - p_P)
    flux[:Fi] += a[csr["slot_pn"]] * (p[nb] - p[own[:Fi]])
    rowsum = np.add.reduceat(a, csr["ia"][:-1])
    ob = np.nonzero(outlet)[0]
    cells = own[Fi + ob]
    flux[Fi + ob] += rowsum[cells] * p[cells]
    div2 = np.bincount(own, weights=flux, minlength=n) - np.bincount(nb, weights=flux[:Fi],
                                                                     minlength=n)
    return flux, div2


def convdiff_inputs(mesh, geo, flux):
    """Face diffusivity, boundary kinds, sources and current fields of the
    three transported components (see the block comment above)."""
    n, Fi, F = mesh.n_cells, mesh.n_inner, mesh.n_faces
    patch = mesh.patch_of_bface()
    outlet = np.asarray(mesh.bctype_codes[patch] == 4)
    D = np.empty(F)
    D[:Fi] = geo["Df"]
    D[Fi:] = geo["srdb"]
    pe = 10.0 ** uniform(CONVDIFF_SEED, 40, F, -1.0, 2.0)
    fmin = 1e-3 * np.mean(np.abs(flux[:Fi]))
    gamma = np.maximum(np.abs(flux), fmin) / (pe * D)
    bc_kind = np.where(outlet, 2, 0).astype(np.uint8)   # kind 2 in/outlet, 0 Dirichlet wall
    uc = fourier_velocity(geo["xc"], geo["yc"], geo["zc"])
    own_b = mesh.owner[Fi:]
    phi = np.zeros((3, mesh.n_total))
    src = np.zeros((3, n))
    for r in range(3):
        phi[r, :n] = uc[:, r]
        phi[r, n:] = np.where(outlet, uc[own_b, r], 0.0)
        src[r] = uniform(CONVDIFF_SEED, 50 + r, n, -1.0, 1.0)
    return dict(gamma=gamma, bc_kind=bc_kind, phi=phi, source=src, urf=CONVDIFF_URF)


# ---------------------------------------------------------------------------
# Unstructured tetrahedra (96^3 cubes x 6 Kuhn tets, cells permuted)
# ---------------------------------------------------------------------------
# Mesh: ``tet_mesh(N, jitter=0.15, seed=3, perm_seed=3)`` -- interior nodes
# jittered by +-0.15h, cell order scrambled by Fisher-Yates (seed 3); the GPU
# then renumbers it by reverse Cuthill-McKee.  Field for the gradients:
# f = sin(2 pi x) cos(2 pi y) exp(z) + 1e-3 U[-1,1] at cell centres and
# boundary-face centres (seed 3, streams 300/301).  Poisson: face diffusivity
# 10^U(-2,2) (stream 302), zero-mean source U[-1,1] (stream 303), Dirichlet 0
# on x-, zero normal gradient elsewhere.
def tet_mesh_workload(N: int = 96):
    from .mesh import tet_mesh
    return tet_mesh(N, jitter=0.15, seed=3, perm_seed=3,
                    types=("wall", "symmetry", "symmetry", "symmetry", "symmetry", "symmetry"))


def analytic_field(x, y, z):
    return np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y) * np.exp(z)


def analytic_grad(x, y, z):
    e = np.exp(z)
    return np.stack([2 * np.pi * np.cos(2 * np.pi * x) * np.cos(2 * np.pi * y) * e,
                     -2 * np.pi * np.sin(2 * np.pi * x) * np.sin(2 * np.pi * y) * e,
                     np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y) * e], axis=1)


def tet_inputs(mesh, geo):
    n, Fi, F = mesh.n_cells, mesh.n_inner, mesh.n_faces
    f = np.empty(mesh.n_total)
    f[:n] = analytic_field(geo["xc"], geo["yc"], geo["zc"]) + 1e-3 * uniform(3, 300, n, -1.0, 1.0)
    f[n:] = (analytic_field(geo["xf"][Fi:], geo["yf"][Fi:], geo["zf"][Fi:]) +
             1e-3 * uniform(3, 301, F - Fi, -1.0, 1.0))
    gamma = 10.0 ** uniform(3, 302, F, -2.0, 2.0)
    s = uniform(3, 303, n, -1.0, 1.0)
    s = s - s.mean()
    bc_kind = np.where(mesh.patch_of_bface() == 0, 0, 1).astype(np.uint8)
    return dict(f=f, gamma=gamma, mu_f=-gamma, source=s, bc_kind=bc_kind,
                phi_bc=np.zeros(mesh.n_total))


# ---------------------------------------------------------------------------
# Adjoint sensitivities (128^3 jittered, one transported scalar)
# ---------------------------------------------------------------------------
# Mesh ``convdiff_mesh(N, jitter=0.2)`` with the projected flux and face
# diffusivity of the convection-diffusion workload; source s = component 0
# of its sources; Dirichlet wall values phi_b = 0.1 U[-1,1] (seed 4,
# stream 61); objective J = sum_i w_i phi_i V_i with w ~ U[0,1] (seed 4,
# stream 60); K = 2 deferred-correction passes of the linear-upwind scheme.
ADJOINT_K = 2


def adjoint_inputs(mesh, geo, flux):
    base = convdiff_inputs(mesh, geo, flux)
    n, Fb = mesh.n_cells, mesh.n_bnd
    phib = np.where(base["bc_kind"] == 0, 0.1 * uniform(4, 61, Fb, -1.0, 1.0), 0.0)
    return dict(gamma=base["gamma"], source=base["source"][0], phib=phib,
                bc_kind=base["bc_kind"], w=uniform(4, 60, n, 0.0, 1.0), K=ADJOINT_K)
