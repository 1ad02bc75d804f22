"""Pins the oracle's GMRES(m)/ILU(0) restatement (linear_solvers.f90:789-1555).

Golden vectors come from the flang-built reference (tests/golden/gmres.npz,
written by make_golden.py): ilu_cr factors, one lus_cr application, atx_cr,
and pmgmres_ilu after 1, 2, 5, 20 and 500 restart cycles on two mesh
matrices (SPD Laplacian, non-symmetric convection-diffusion) and the two
dense 5x5 systems, plus the 32^3 Poisson through csrsolve('pmgmres').
Comparisons against them are bit-exact.  The mathematics-fixed pins use the
dense systems, where ILU(0) drops nothing and is the exact LU factorisation.
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp

from fcp import workloads
from oracle import Oracle, atx, ilu0, ilu_apply, solve_explicit


@pytest.fixture(scope="module")
def gm(golden_dir):
    return np.load(os.path.join(golden_dir, "gmres.npz"))


def _systems(golden_dir):
    hx = np.load(os.path.join(golden_dir, "hex5j.npz"))
    kat = np.load(os.path.join(golden_dir, "kat5x5.npz"))
    return {
        "lap": (hx["c_ia"], hx["c_ja"], hx["c_diag"], hx["lap_a"], hx["in_b"]),
        "cd": (hx["c_ia"], hx["c_ja"], hx["c_diag"], hx["cd_linearUpwind_a"], hx["in_b"]),
        "ns": (kat["ia"], kat["ja"], kat["diag"], kat["A_ns"].reshape(-1), kat["b_ns"]),
        "sp": (kat["ia"], kat["ja"], kat["diag"], kat["A_sp"].reshape(-1), kat["b_sp"]),
    }


@pytest.mark.parametrize("key", ["lap", "cd", "ns", "sp"])
def test_ilu_lus_atx_bitexact(golden_dir, gm, key):
    ia, ja, dg, a, b = _systems(golden_dir)[key]
    l = ilu0(ia, ja, dg, a)
    assert np.array_equal(l, gm[f"{key}_ilu"])
    assert np.array_equal(ilu_apply(ia, ja, dg, l, b), gm[f"{key}_lus_b"])
    assert np.array_equal(atx(ia, ja, dg, a, b), gm[f"{key}_atx_b"])


@pytest.mark.parametrize("key", ["lap", "cd", "ns", "sp"])
def test_pmgmres_trajectories_bitexact(golden_dir, gm, key):
    ia, ja, dg, a, b = _systems(golden_dir)[key]
    for itmax in (1, 2, 5, 20, 500):
        out = solve_explicit("pmgmres", ia, ja, dg, a, b, itr_max=itmax, tol_abs=0.0,
                             tol_rel=1e-10)
        assert np.array_equal(out["fi"], gm[f"{key}_gm_{itmax}"]), itmax
        assert out["iters"] == int(gm[f"{key}_gm_{itmax}_iters"]), itmax
        assert out["resor"] == float(gm[f"{key}_gm_{itmax}_resor"]), itmax
        # the report line prints resl with 4 significant digits (1PE10.3)
        assert out["resl"] == pytest.approx(float(gm[f"{key}_gm_{itmax}_resl"]), rel=6e-4)


def test_pmgmres_poisson32_csrsolve(gm):
    mesh = workloads.poisson_uniform_mesh(32)
    o = Oracle(mesh)
    inp = workloads.poisson_uniform_inputs(mesh, o.geom["vol"])
    a, su = o.laplacian(inp["mu"], inp["phi"], inp["rhs"])
    assert np.array_equal(su, gm["p32_rhs"])
    for itmax in (3, 3000):
        out = o.solve("pmgmres", a, su, itr_max=itmax, tol_abs=0.0, tol_rel=1e-10)
        assert np.array_equal(out["fi"], gm[f"p32_gm_{itmax}"]), itmax
        assert out["iters"] == int(gm[f"p32_gm_{itmax}_iters"])
        assert out["resor"] == float(gm[f"p32_gm_{itmax}_resor"])


@pytest.mark.parametrize("which", ["ns", "sp"])
def test_ilu_of_dense_matrix_is_exact_lu(golden_dir, which):
    """No entry is dropped for a dense pattern: L (unit lower, the stored
    multipliers) times U (the rest) reproduces A, and the preconditioner
    application is a direct solve."""
    ia, ja, dg, a, b = _systems(golden_dir)[which]
    l = ilu0(ia, ja, dg, a).reshape(5, 5)
    L = np.tril(l, -1) + np.eye(5)
    U = np.triu(l)
    A = a.reshape(5, 5)
    assert np.allclose(L @ U, A, rtol=0, atol=1e-13 * np.abs(A).max())
    assert np.allclose(ilu_apply(ia, ja, dg, l.reshape(-1), b), np.linalg.solve(A, b),
                       rtol=1e-12)


def test_atx_is_transpose_product(golden_dir):
    for key, (ia, ja, dg, a, b) in _systems(golden_dir).items():
        A = sp.csr_matrix((a, ja, ia), shape=(len(ia) - 1,) * 2)
        assert np.allclose(atx(ia, ja, dg, a, b), A.T @ b, rtol=1e-14, atol=1e-14), key


def test_ilu_on_tridiagonal_closed_form():
    """Tridiagonal A: ILU(0) = LU with pivots u_i = a_ii - a_i,i-1 a_i-1,i / u_i-1
    and multipliers l_i = a_i,i-1 / u_i-1 (Thomas algorithm)."""
    n = 9
    rng = np.random.default_rng(3)
    lo, up = rng.uniform(-1, 0, n), rng.uniform(-1, 0, n)
    dgv = 3.0 + rng.uniform(0, 1, n)
    A = np.diag(dgv) + np.diag(lo[1:], -1) + np.diag(up[:-1], 1)
    S = sp.csr_matrix(A)
    S.sort_indices()
    ia, ja, a = S.indptr.astype(np.int32), S.indices.astype(np.int32), S.data.copy()
    dg = np.array([ia[i] + list(ja[ia[i]:ia[i + 1]]).index(i) for i in range(n)], np.int32)
    l = ilu0(ia, ja, dg, a)
    u = np.zeros(n)
    u[0] = dgv[0]
    for i in range(1, n):
        u[i] = dgv[i] - lo[i] * up[i - 1] / u[i - 1]
        assert l[dg[i] - 1] == pytest.approx(lo[i] / u[i - 1], rel=1e-14)
    assert np.allclose(l[dg], u, rtol=1e-14)
    b = rng.uniform(-1, 1, n)
    assert np.allclose(ilu_apply(ia, ja, dg, l, b), np.linalg.solve(A, b), rtol=1e-12)
