"""Batched (multi right-hand-side) solves, mixed-precision DILU-BiCGStab and
the convection-diffusion workload on the GPU.

A batched solve runs every column's own recurrence with the single-vector
per-row arithmetic and reduction structure, so each column must be
bit-equal to fcp_csrsolve on that column alone; the single-vector solves are
themselves pinned to the reference trajectories in test_gpu_parity.py.
"""
import numpy as np
import pytest

from fcp import workloads
from fcp.mesh import hex_mesh
from oracle import Oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api():
    from fcp import api as _api
    return _api


def _setup(api, mesh):
    ctx = api.Context(0)
    ctx.read_mesh(mesh)
    geo = ctx.geometry()
    csr = ctx.create_CSR_matrix()
    return ctx, geo, csr


def _convdiff_system(api, N, jitter=0.2):
    """Small convection-diffusion workload: divergence-free flux by Poisson
    projection, linear-upwind assembly of three components sharing A."""
    mesh = workloads.convdiff_mesh(N, jitter=jitter)
    ctx, geo, csr = _setup(api, mesh)
    flux, div = workloads.convdiff_flux(mesh, geo, csr, ctx)
    inp = workloads.convdiff_inputs(mesh, geo, flux)
    n = mesh.n_cells
    rhs = np.zeros((3, n))
    a = np.zeros(csr["ja"].size)
    for r in range(3):
        phi = inp["phi"][r].copy()
        gr = ctx.grad_gauss(phi)
        su = inp["source"][r] * geo["vol"]
        sp = np.zeros(n)
        ctx.convdiff(flux, inp["gamma"], phi, gr, 1.0, 1, inp["bc_kind"], su, sp,
                     urf=inp["urf"], a_out=a)
        rhs[r] = su
    return mesh, ctx, geo, csr, flux, div, a, rhs


def test_convdiff_flux_divergence_free(api):
    print("[case] convdiff flux projection 16^3", flush=True)
    mesh, ctx, geo, csr, flux, div, a, rhs = _convdiff_system(api, 16)
    n, Fi = mesh.n_cells, mesh.n_inner
    # the projection removes the divergence down to its solve tolerance
    raw = workloads.fourier_velocity(geo["xf"], geo["yf"], geo["zf"])
    f0 = raw[:, 0] * geo["arx"] + raw[:, 1] * geo["ary"] + raw[:, 2] * geo["arz"]
    div0 = np.bincount(mesh.owner[:Fi], weights=f0[:Fi], minlength=n) - np.bincount(
        mesh.neighbour, weights=f0[:Fi], minlength=n)
    assert np.abs(div).max() < 1e-9 * np.abs(div0).max()
    # walls carry no flux; Peclet numbers span the stated range
    bcode = mesh.bctype_codes[mesh.patch_of_bface()]
    assert np.all(flux[Fi:][bcode == 1] == 0.0)
    inp = workloads.convdiff_inputs(mesh, geo, flux)
    pe = np.abs(flux[:Fi]) / (inp["gamma"][:Fi] * geo["Df"])
    big = np.abs(flux[:Fi]) > 1e-2 * np.abs(flux[:Fi]).max()
    assert pe[big].min() >= 0.1 * (1 - 1e-12) and pe[big].max() <= 100.0 * (1 + 1e-12)
    ctx.close()


@pytest.mark.parametrize("precision,inner,N", [(0, 1, 14), (1, 1, 14), (1, 1, 26), (1, 2, 26)])
def test_multi_bicgstab_columns_bitequal_single(api, precision, inner, N):
    """3-column DILU-BiCGStab (FP64, and FP32 inner + FP64 refinement with the
    tile-local preconditioner: inner iterations in tile order (1) or with the
    row-gather sweeps (2)) equals three single-column solves bit for bit, and
    reaches the tolerance.  At 26^3 the preconditioner has several tiles."""
    print(f"[case] multi bicgstab {N}^3 precision={precision} inner={inner}", flush=True)
    mesh, ctx, geo, csr, flux, div, a, rhs = _convdiff_system(api, N)
    n = mesh.n_cells
    ctx.tune(api.FCP_TUNE_INNER_PRECOND, inner)
    if N == 26:
        assert ctx.sweep_tiles().max() >= 3
    ctx.matrix_set(a)
    X = np.zeros((3, n))
    rc, info, res0 = ctx.csrsolve_multi("bicgstab", X, rhs, itr_max=3000, tol_rel=1e-9,
                                        precision=precision, criterion=1)
    assert rc == 0, info
    assert info["rel_res_l2"] <= 1e-9
    for r in range(3):
        x = np.zeros(n)
        rc1, info1 = ctx.csrsolve("bicgstab", x, rhs[r].copy(), itr_max=3000, tol_rel=1e-9,
                                  precision=precision, criterion=1)
        assert rc1 == 0
        assert np.array_equal(X[r], x), (r, np.abs(X[r] - x).max())
    # against the oracle's tight FP64 solve
    o = Oracle(mesh)
    for r in range(3):
        ref = o.solve("bicgstab", a, rhs[r], itr_max=20000, tol_abs=0.0, tol_rel=1e-14)["fi"]
        err = np.linalg.norm(X[r] - ref) / np.linalg.norm(ref)
        assert err < 1e-7, (r, err)
    ctx.close()


def test_multi_iccg_columns_bitequal_single(api):
    print("[case] multi iccg 16^3 two columns", flush=True)
    mesh = hex_mesh(16, jitter=0.2, seed=9)
    ctx, geo, csr = _setup(api, mesh)
    n = mesh.n_cells
    o = Oracle(mesh)
    a, _ = o.laplacian(-(10.0 ** np.sin(np.arange(n))), np.zeros(mesh.n_total))
    ctx.matrix_set(a)
    B = np.stack([np.sin(np.arange(n) * 0.37), np.cos(np.arange(n) * 0.11)])
    for precision in (0, 1):
        X = np.zeros((2, n))
        rc, info, _ = ctx.csrsolve_multi("iccg", X, B, itr_max=3000, tol_rel=1e-10,
                                         precision=precision, criterion=1)
        assert rc == 0
        for r in range(2):
            x = np.zeros(n)
            ctx.csrsolve("iccg", x, B[r].copy(), itr_max=3000, tol_rel=1e-10,
                         precision=precision, criterion=1)
            assert np.array_equal(X[r], x), (precision, r)
    ctx.close()


def test_multi_reference_criterion_matches_reference_trajectory(api):
    """Batched FP64 BiCGStab with the reference L1 criterion: each column
    follows the oracle trajectory of the same fixed iteration count."""
    print("[case] multi bicgstab trajectories 10^3", flush=True)
    mesh, ctx, geo, csr, flux, div, a, rhs = _convdiff_system(api, 10)
    ctx.matrix_set(a)
    o = Oracle(mesh)
    for it in (1, 3, 10):
        X = np.zeros((3, mesh.n_cells))
        rc, info, _ = ctx.csrsolve_multi("bicgstab", X, rhs, itr_max=it, tol_rel=1e-30)
        assert rc == it  # tol_rel 1e-30 is never met: not converged, iterations used
        for r in range(3):
            ref = o.solve("bicgstab", a, rhs[r], itr_max=it, tol_abs=0.0, tol_rel=1e-30)["fi"]
            assert np.allclose(X[r], ref, rtol=1e-10, atol=1e-12 * np.abs(ref).max()), (it, r)
    ctx.close()
