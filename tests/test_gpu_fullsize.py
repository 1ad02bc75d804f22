"""The 128^3 jittered-hex pressure-correction Poisson at full size (2,097,152
cells, the benchmark workload), through size-independent properties:

- the mixed-precision, the FP64 and the FP64-with-tile-preconditioner ICCG
  reach the requested true relative residual, and the mixed and FP64-tile
  solutions' errors against a tight FP64 solve stay within three times the
  FP64 solution's own error at the same tolerance;
- the tile-local DIC of the FP32 inner solves is a symmetric operator
  (r2 . M^-1 r1 = r1 . M^-1 r2) in FP64 and, within FP32 rounding, in the
  FP32 tile-order kernel;
- three batched mixed columns equal three single-column solves bit for bit.

The oracle pins the same kernels at sizes it finishes in seconds
(test_gpu_parity.py); here only properties that do not need it are checked.
"""
import numpy as np
import pytest

from fcp import workloads

pytestmark = pytest.mark.gpu

N = 128


@pytest.fixture(scope="module")
def system():
    from fcp import api
    print(f"[case] full size {N}^3 setup", flush=True)
    mesh = workloads.poisson_jittered_mesh(N)
    ctx = api.Context(0)
    ctx.read_mesh(mesh)
    geo = ctx.geometry()
    ctx.create_CSR_matrix(want_arrays=False)
    inp = workloads.poisson_jittered_inputs(mesh, geo["vol"], geo["xc"], geo["yc"], geo["zc"])
    su = inp["source"] * geo["vol"]
    ctx.laplacian_face(inp["mu_f"], inp["bc_kind"], inp["phi_bc"], su)
    grad = ctx.grad_gauss(inp["phi_prev"])
    ctx.nonorth_correction(inp["gamma"], grad, su)
    yield api, mesh, ctx, su
    ctx.close()


def test_fullsize_mixed_error_within_fp64_error(system):
    """Mixed and FP64 ICCG reach the requested true relative residual, and the
    mixed solution's error against a tight FP64 solve (1e-13) is at most three
    times the FP64 solution's own error at the same tolerance (1e-8 and
    1e-12): the FP32 inner iterations cost no accuracy beyond what the
    stopping test allows."""
    api, mesh, ctx, su = system
    print(f"[case] full size {N}^3 mixed vs fp64 errors", flush=True)

    def solve(prec, tol):
        x = np.zeros(mesh.n_cells)
        rc, info = ctx.csrsolve("iccg", x, su.copy(), itr_max=50000, tol_abs=0.0, tol_rel=tol,
                                precision=prec, criterion=api.FCP_CRIT_TRUE_L2, check_every=16)
        assert rc == 0 and info["rel_res_l2"] <= tol, (prec, tol, info)
        return x

    xr = solve(api.FCP_PREC_FP64, 1e-13)
    nr = np.linalg.norm(xr)
    err = {}
    for tol in (1e-8, 1e-12):
        for prec in (api.FCP_PREC_MIXED, api.FCP_PREC_FP64, api.FCP_PREC_FP64_TILE):
            err[prec, tol] = np.linalg.norm(solve(prec, tol) - xr) / nr
        print(f"[case] tol {tol:g}: err mixed {err[api.FCP_PREC_MIXED, tol]:.3e} "
              f"fp64 {err[api.FCP_PREC_FP64, tol]:.3e} "
              f"fp64-tile {err[api.FCP_PREC_FP64_TILE, tol]:.3e}", flush=True)
        assert err[api.FCP_PREC_MIXED, tol] <= 3.0 * err[api.FCP_PREC_FP64, tol], (tol, err)
        assert err[api.FCP_PREC_FP64_TILE, tol] <= 3.0 * err[api.FCP_PREC_FP64, tol], (tol, err)
    assert err[api.FCP_PREC_MIXED, 1e-12] <= 1e-9, err


@pytest.mark.parametrize("kind,tol", [(3, 1e-12), (5, 1e-5)])
def test_fullsize_tile_local_dic_is_symmetric(system, kind, tol):
    api, mesh, ctx, su = system
    print(f"[case] full size {N}^3 precond symmetry kind {kind}", flush=True)
    n = mesh.n_cells
    r1 = np.sin(np.arange(n) * 0.37)
    r2 = np.cos(np.arange(n) * 0.11) * (1.0 + (np.arange(n) % 7))
    z1, _ = ctx.precond_apply(r1, kind=kind)
    z2, _ = ctx.precond_apply(r2, kind=kind)
    a, b = float(r2 @ z1), float(r1 @ z2)
    assert abs(a - b) <= tol * max(abs(a), abs(b)), (a, b)
    # positive definite on these vectors
    assert r1 @ z1 > 0 and r2 @ z2 > 0


def test_fullsize_batched_columns_bitequal_single(system):
    api, mesh, ctx, su = system
    print(f"[case] full size {N}^3 batched vs single", flush=True)
    n = mesh.n_cells
    B = np.stack([su, np.cos(np.arange(n) * 0.013) * 1e-4, np.sin(np.arange(n) * 0.029) * 1e-4])
    X = np.zeros((3, n))
    rc, info, _ = ctx.csrsolve_multi("iccg", X, B, itr_max=20000, tol_rel=1e-8,
                                     precision=api.FCP_PREC_MIXED,
                                     criterion=api.FCP_CRIT_TRUE_L2, check_every=16)
    assert rc == 0 and info["rel_res_l2"] <= 1e-8, info
    for r in range(3):
        x = np.zeros(n)
        rc1, info1 = ctx.csrsolve("iccg", x, B[r].copy(), itr_max=20000, tol_rel=1e-8,
                                  precision=api.FCP_PREC_MIXED, criterion=api.FCP_CRIT_TRUE_L2,
                                  check_every=16)
        assert rc1 == 0
        assert np.array_equal(X[r], x), r
