"""The LIS interface (LIS_linear_solvers.f90:38-200) without the LIS library:
lis_create_solver's option string mapped onto this library's solvers and
lis_spsolve's report values, on a 24^3 jittered pressure-correction system
(SPD) and a convection-diffusion system (nonsymmetric).

LIS itself is absent from the reference and from this image, so the
iterates are not compared with LIS's: each mapped solve is compared with
the same solver through csrsolve, the solution with the oracle's residual,
and the reported initial residual with its definition (:135-143)."""
import numpy as np
import pytest

from fcp import workloads
from oracle import Oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def system():
    from fcp import api
    mesh = workloads.poisson_jittered_mesh(24)
    ctx = api.Context(0)
    ctx.read_mesh(mesh)
    geo = ctx.geometry()
    csr = ctx.create_CSR_matrix()
    inp = workloads.poisson_jittered_inputs(mesh, geo["vol"], geo["xc"], geo["yc"], geo["zc"])
    su = inp["source"] * geo["vol"]
    a = np.zeros(csr["ja"].size)
    ctx.laplacian_face(inp["mu_f"], inp["bc_kind"], inp["phi_bc"], su, a_out=a)
    o = Oracle(mesh)
    yield api, mesh, ctx, o, a, su
    ctx.close()


def _l1_res(o, a, b, x):
    return float(np.abs(b - o.spmv(a, x)).sum())


@pytest.mark.parametrize("opts,solver", [("-i cg -p ilu", "iccg"), ("-i 1 -p 1", "dpcg"),
                                         ("-i gmres -restart 20 -p ilu", "pmgmres")])
def test_lis_spd_solvers(system, opts, solver):
    api, mesh, ctx, o, a, su = system
    print(f"[case] lis '{opts}' 24^3", flush=True)
    n = mesh.n_cells
    ctx.matrix_set(a)
    ctx.lis_create_solver(opts + " -tol 1.0e-10 -maxiter 4000 -print 0")
    fi = np.sin(np.arange(n) * 0.3)  # the current field: only the report's res0 uses it
    rc, info = ctx.lis_spsolve(fi, su, "p")
    assert rc == 0, info
    assert abs(info["res0"] - _l1_res(o, a, su, np.sin(np.arange(n) * 0.3))) <= 1e-12 * info["res0"]
    rel = np.linalg.norm(su - o.spmv(a, fi)) / np.linalg.norm(su)
    assert rel <= 1.01e-10 and info["resid"] <= 1e-10, (rel, info)
    assert abs(info["resid"] - rel) <= 1e-3 * rel
    if solver != "pmgmres":  # the same solver through csrsolve: the same iterations
        x = np.zeros(n)
        rc2, inf2 = ctx.csrsolve(solver, x, su, itr_max=4000, tol_abs=0.0, tol_rel=1e-10,
                                 criterion=api.FCP_CRIT_TRUE_L2)
        assert inf2["iterations"] == info["iter"] and np.array_equal(x, fi)
    ctx.lis_destroy_solver()


def test_lis_warm_start_and_conv_cond(system):
    """-initx_zeros false starts from the field passed in; convergence
    condition 0 is relative to the initial residual, 1 to ||b||."""
    api, mesh, ctx, o, a, su = system
    print("[case] lis warm start 24^3", flush=True)
    n = mesh.n_cells
    ctx.matrix_set(a)
    ctx.lis_create_solver("-i cg -p ilu -tol 1e-6")
    xs = np.zeros(n)
    ctx.lis_spsolve(xs, su)
    r_s = np.linalg.norm(su - o.spmv(a, xs))
    for cond in (0, 1):
        ctx.lis_create_solver(f"-i cg -p ilu -tol 1e-4 -initx_zeros false -conv_cond {cond}")
        x = xs.copy()
        rc, info = ctx.lis_spsolve(x, su)
        assert rc == 0
        r = np.linalg.norm(su - o.spmv(a, x))
        ref = r_s if cond == 0 else np.linalg.norm(su)
        assert r <= 1.01e-4 * ref and abs(info["resid"] - r / ref) <= 1e-3 * (r / ref)
    # warm start from near the solution takes fewer iterations than from zero
    ctx.lis_create_solver("-i cg -p ilu -tol 1e-8 -conv_cond 1")
    x0 = np.zeros(n)
    _, cold = ctx.lis_spsolve(x0, su)
    ctx.lis_create_solver("-i cg -p ilu -tol 1e-8 -conv_cond 1 -initx_zeros false")
    x1 = xs.copy()
    _, warm = ctx.lis_spsolve(x1, su)
    assert warm["iter"] < cold["iter"]


def test_lis_bicgstab_and_gs_nonsymmetric(system):
    api, mesh, ctx, o, a, su = system
    print("[case] lis bicgstab / gs convection-diffusion 24^3", flush=True)
    n, F = mesh.n_cells, mesh.n_faces
    bk = np.full(mesh.n_bnd, 2, np.uint8)
    a_cd, _, _, _ = o.convdiff(np.cos(np.arange(F)) * 1e-2, 10.0 ** np.sin(np.arange(F)),
                               np.zeros(mesh.n_total), np.zeros((mesh.n_total, 3)), 1.0, 0, bk)
    su_cd = np.sin(np.arange(n) * 0.7)
    ctx.matrix_set(a_cd)
    ctx.lis_create_solver("-i bicgstab -p ilu -tol 1e-10 -maxiter 500")
    x = np.zeros(n)
    rc, info = ctx.lis_spsolve(x, su_cd, "T")
    assert rc == 0
    assert np.linalg.norm(su_cd - o.spmv(a_cd, x)) <= 1.01e-10 * np.linalg.norm(su_cd)
    # 30 Gauss-Seidel sweeps (an unreachable tolerance) equal csrsolve's
    ctx.lis_create_solver("-i gs -p none -tol 1e-30 -maxiter 30")
    x = np.zeros(n)
    rc, info = ctx.lis_spsolve(x, su_cd, "T")
    assert rc > 0 and info["iter"] == 30
    x2 = np.zeros(n)
    ctx.csrsolve("gauss-seidel", x2, su_cd, itr_max=30, tol_abs=0.0, tol_rel=1e-30,
                 criterion=api.FCP_CRIT_TRUE_L2)
    assert np.array_equal(x, x2)


@pytest.mark.parametrize("opts", ["-i cgs -p ilu", "-i cg -p ssor", "-i cg -p ilu -f quad",
                                  "-i cg -p ilu -tol", "-i cg -p ilu -bogus 3"])
def test_lis_rejects_what_it_cannot_map(system, opts):
    api, mesh, ctx, o, a, su = system
    print(f"[case] lis rejects '{opts}'", flush=True)
    with pytest.raises(api.FcpError):
        ctx.lis_create_solver(opts)
