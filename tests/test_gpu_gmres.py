"""GMRES(m) with ILU(0) on the GPU (pmgmres_ilu, ilu_cr, lus_cr, atx_cr of
linear_solvers.f90:789-1555), through the C-ABI.

ILU(0) factors, one preconditioner application and the transposed product
keep the reference's per-row arithmetic and must be bit-exact with the
reference's golden vectors (tests/golden/gmres.npz) and the oracle.  GMRES
trajectories differ from the reference only by the summation order of the
inner products and norms, so solutions are compared within stated tolerances
and iteration counts must agree.
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

from fcp.mesh import hex_mesh
from fcp import workloads
from oracle import Oracle, atx, ilu0, ilu_apply

pytestmark = pytest.mark.gpu

HEX5J = dict(jitter=0.2, seed=2, types=("wall", "symmetry", "outlet", "wall", "inlet", "pressure"))


@pytest.fixture(scope="module")
def api():
    from fcp import api as _api
    return _api


@pytest.fixture(scope="module")
def gm(golden_dir):
    return np.load(os.path.join(golden_dir, "gmres.npz"))


def _setup(api, mesh):
    ctx = api.Context(0)
    ctx.read_mesh(mesh)
    geo = ctx.geometry()
    csr = ctx.create_CSR_matrix()
    return ctx, geo, csr


def test_ilu_lus_atx_bitexact_vs_reference(api, golden_dir, gm):
    print("[case] ilu/lus/atx hex5j", flush=True)
    hx = np.load(os.path.join(golden_dir, "hex5j.npz"))
    ctx, geo, csr = _setup(api, hex_mesh(5, **HEX5J))
    for key, a in (("lap", hx["lap_a"]), ("cd", hx["cd_linearUpwind_a"])):
        ctx.matrix_set(a)
        assert np.array_equal(ctx.ilu_factor(a.size), gm[f"{key}_ilu"]), key
        z, _ = ctx.precond_apply(hx["in_b"], kind=2)
        assert np.array_equal(z, gm[f"{key}_lus_b"]), key
        assert np.array_equal(ctx.spmv_transpose(hx["in_b"]), gm[f"{key}_atx_b"]), key
    ctx.close()


def test_ilu_lus_atx_bitexact_vs_oracle_20cube(api):
    """382-level-deep factorisation order on a jittered 20^3 mesh, SPD and
    non-symmetric matrices."""
    print("[case] ilu/lus/atx 20^3", flush=True)
    mesh = hex_mesh(20, jitter=0.2, seed=4)
    ctx, geo, csr = _setup(api, mesh)
    o = Oracle(mesh)
    mu = -(10.0 ** np.sin(np.arange(mesh.n_cells)))
    a_lap, _ = o.laplacian(mu, np.zeros(mesh.n_total))
    F = mesh.n_faces
    bk = np.full(mesh.n_bnd, 2, np.uint8)
    a_cd, _, _, _ = o.convdiff(np.cos(np.arange(F)), 10.0 ** np.sin(np.arange(F)),
                               np.zeros(mesh.n_total), np.zeros((mesh.n_total, 3)), 1.0, 0, bk)
    r = np.sin(np.arange(mesh.n_cells) * 0.1)
    for a in (a_lap, a_cd):
        ctx.matrix_set(a)
        l_o = ilu0(o.ia, o.ja, o.diag, a)
        assert np.array_equal(ctx.ilu_factor(a.size), l_o)
        z, _ = ctx.precond_apply(r, kind=2)
        assert np.array_equal(z, ilu_apply(o.ia, o.ja, o.diag, l_o, r))
        assert np.array_equal(ctx.spmv_transpose(r), atx(o.ia, o.ja, o.diag, a, r))
    ctx.close()


@pytest.mark.parametrize("key", ["lap", "cd"])
def test_pmgmres_trajectories_vs_reference(api, golden_dir, gm, key):
    print(f"[case] pmgmres trajectories hex5j {key}", flush=True)
    hx = np.load(os.path.join(golden_dir, "hex5j.npz"))
    ctx, geo, csr = _setup(api, hex_mesh(5, **HEX5J))
    ctx.matrix_set(hx["lap_a"] if key == "lap" else hx["cd_linearUpwind_a"])
    for it in (1, 2, 5, 20, 500):
        x = np.zeros(125)
        rc, info = ctx.csrsolve("pmgmres", x, hx["in_b"], itr_max=it, tol_abs=0.0, tol_rel=1e-10)
        ref = gm[f"{key}_gm_{it}"]
        # only the reduction order of dots and norms differs from the reference
        assert np.allclose(x, ref, rtol=1e-10, atol=1e-12 * np.abs(ref).max()), it
        assert info["iterations"] == int(gm[f"{key}_gm_{it}_iters"]), it
        assert info["resor"] == pytest.approx(float(gm[f"{key}_gm_{it}_resor"]), rel=1e-13)
        assert info["resl"] == pytest.approx(float(gm[f"{key}_gm_{it}_resl"]), rel=6e-4)
        assert rc == (0 if it >= 20 else info["iterations"]), it
    ctx.close()


def test_pmgmres_explicit_entry(api, golden_dir, gm):
    """fcp_pmgmres (explicit arrays, pmgmres_ilu's own signature):
    - the hex5j matrices with mr = 4 against the reference's explicit calls;
    - mr = 7 against the oracle;
    - the dense 5x5 systems with mr = 1 against the oracle.  With mr = 4
      the reference's answer on them is set by rounding noise: ILU(0) is the
      exact LU there, so the second Krylov vector is pure rounding error and
      only the sequential summation order reproduces it (the oracle does,
      bit for bit).  With mr = 1 that noise enters at second order only."""
    print("[case] pmgmres explicit entry", flush=True)
    from oracle import solve_explicit
    hx = np.load(os.path.join(golden_dir, "hex5j.npz"))
    kat = np.load(os.path.join(golden_dir, "kat5x5.npz"))
    for key, a in (("lap", hx["lap_a"]), ("cd", hx["cd_linearUpwind_a"])):
        for it in (2, 20):
            x = np.zeros(125)
            api.explicit_solve("pmgmres", hx["c_ia"], hx["c_ja"], a, hx["c_diag"], x, hx["in_b"],
                               it, 0.0, 1e-10)
            ref = gm[f"{key}_gm_{it}"]
            assert np.allclose(x, ref, rtol=1e-10, atol=1e-12 * np.abs(ref).max()), (key, it)
        x = np.zeros(125)
        api.explicit_solve("pmgmres", hx["c_ia"], hx["c_ja"], a, hx["c_diag"], x, hx["in_b"], 10,
                           0.0, 1e-10, mr=7)
        ref = solve_explicit("pmgmres", hx["c_ia"], hx["c_ja"], hx["c_diag"], a, hx["in_b"],
                             itr_max=10, tol_rel=1e-10, mr=7)["fi"]
        assert np.allclose(x, ref, rtol=1e-10, atol=1e-12 * np.abs(ref).max()), (key, "mr7")
    for which in ("ns", "sp"):
        A, b = kat["A_" + which].reshape(-1), kat["b_" + which]
        x = np.zeros(5)
        api.explicit_solve("pmgmres", kat["ia"], kat["ja"], A, kat["diag"], x, b, 5, 0.0, 1e-10,
                           mr=1)
        ref = solve_explicit("pmgmres", kat["ia"], kat["ja"], kat["diag"], A, b, itr_max=5,
                             tol_rel=1e-10, mr=1)["fi"]
        assert np.allclose(x, ref, rtol=1e-12, atol=1e-14), which


def test_pmgmres_poisson32_vs_reference(api, gm):
    """csrsolve('pmgmres') on the 32^3 Poisson: 3 cycles, and 3000 cycles in
    which the reference's stopping test is never met (12000 inner steps).

    Here the reference's normalisation (the preconditioned residual divided by
    the unpreconditioned norm) makes the second Gram-Schmidt step cancel by a
    factor ~300 (av/h(k+1,k) = 305), so the result depends on the summation
    order of the inner products well above rounding: redoing the reference's
    loop on the CPU with only pairwise or blocked dot products moves the
    solution by 2.2e-8 / 1.9e-8 (3 cycles) and 3.2e-9 (3000 cycles).  The
    tolerances below sit just above that."""
    print("[case] pmgmres poisson32", flush=True)
    mesh = workloads.poisson_uniform_mesh(32)
    ctx, geo, csr = _setup(api, mesh)
    inp = workloads.poisson_uniform_inputs(mesh, geo["vol"])
    su = inp["rhs"].copy()
    ctx.laplacian(inp["mu"], inp["phi"], su)
    assert np.array_equal(su, gm["p32_rhs"])
    for it, rtol in ((3, 1e-7), (3000, 5e-8)):
        x = np.zeros(mesh.n_cells)
        rc, info = ctx.csrsolve("pmgmres", x, su, itr_max=it, tol_abs=0.0, tol_rel=1e-10,
                                check_every=64)
        ref = gm[f"p32_gm_{it}"]
        assert np.linalg.norm(x - ref) / np.linalg.norm(ref) < rtol, it
        assert info["iterations"] == int(gm[f"p32_gm_{it}_iters"]), it
        assert rc == info["iterations"]  # not converged, as in the reference
    ctx.close()


@pytest.mark.parametrize("case", ["poisson32", "convdiff20"])
def test_pmgmres_true_l2_converges_to_direct_solution(api, case):
    """FCP_CRIT_TRUE_L2: preconditioned residual normalised by its own norm,
    stop on ||b - A x||_2 / ||b||_2 <= tol at a restart; the result matches
    a direct sparse solve."""
    print(f"[case] pmgmres true-L2 {case}", flush=True)
    if case == "poisson32":
        mesh = workloads.poisson_uniform_mesh(32)
        ctx, geo, csr = _setup(api, mesh)
        inp = workloads.poisson_uniform_inputs(mesh, geo["vol"])
        b = inp["rhs"].copy()
        a = np.zeros(csr["ja"].size)
        ctx.laplacian(inp["mu"], inp["phi"], b, a_out=a)
    else:
        mesh = hex_mesh(20, jitter=0.2, seed=4)
        ctx, geo, csr = _setup(api, mesh)
        o = Oracle(mesh)
        F = mesh.n_faces
        bk = np.full(mesh.n_bnd, 2, np.uint8)
        a, _, _, _ = o.convdiff(np.cos(np.arange(F)), 10.0 ** np.sin(np.arange(F)),
                                np.zeros(mesh.n_total), np.zeros((mesh.n_total, 3)), 1.0, 0, bk)
        ctx.matrix_set(a)
        b = np.sin(np.arange(mesh.n_cells) * 0.37)
    x = np.zeros(mesh.n_cells)
    rc, info = ctx.csrsolve("pmgmres", x, b, itr_max=5000, tol_abs=0.0, tol_rel=1e-9,
                            criterion=api.FCP_CRIT_TRUE_L2)
    assert rc == 0, info
    assert info["rel_res_l2"] <= 1e-9
    A = sp.csr_matrix((a, csr["ja"], csr["ia"]), shape=(mesh.n_cells,) * 2)
    xd = sla.spsolve(A.tocsc(), b)
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-6
    ctx.close()
