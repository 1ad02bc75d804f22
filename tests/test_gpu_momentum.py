"""The momentum components' solves (velocity.f90:611-748): u, v and w share
the convection-diffusion off-diagonals while each has its own diagonal
(sp_u / sp_v / sp_w minus the off-diagonal sum, under-relaxed).
fcp_csrsolve_multi_diag solves them together -- one batched SpMV and one
batched sweep per step read the off-diagonals once -- and each column must
equal, bit for bit and in iterations, fcp_csrsolve on the matrix with that
column's diagonal (the reference's three separate csrsolve calls), and the
oracle's solve of the same system within the solver's tolerance."""
import numpy as np
import pytest

from fcp.mesh import hex_mesh
from oracle import Oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api():
    from fcp import api as _api
    return _api


def _momentum_like(o, mesh, sym):
    """Off-diagonals of a convection-diffusion (or, sym, a diffusion)
    operator and three diagonals: the off-diagonal sum plus a per-component
    source, under-relaxed by 0.7, 0.8, 0.9 (velocity.f90:617-627)."""
    F, n = mesh.n_faces, mesh.n_cells
    bk = np.full(mesh.n_bnd, 0, np.uint8)
    flm = np.zeros(F) if sym else 0.02 * np.cos(np.arange(F) * 0.7)
    a, _, _, _ = o.convdiff(flm, 10.0 ** np.sin(np.arange(F)), np.zeros(mesh.n_total),
                            np.zeros((mesh.n_total, 3)), 1.0, 0, bk)
    rows = np.repeat(np.arange(n), np.diff(o.ia))
    off = np.where(np.arange(a.size) == o.diag[rows], 0.0, a)
    offsum = np.bincount(rows, off, n)
    D = []
    for c, urf in enumerate((0.7, 0.8, 0.9)):
        sp = 0.05 * (1.0 + np.sin(np.arange(n) * (0.3 + 0.2 * c)) ** 2)
        D.append((sp - offsum) / urf)
    return a, np.array(D)


@pytest.mark.parametrize("solver,sym", [("bicgstab", False), ("iccg", True)])
def test_momentum_columns_equal_single_solves(api, solver, sym):
    print(f"[case] momentum {solver} 20^3 own diagonals", flush=True)
    mesh = hex_mesh(20, jitter=0.2, seed=5)
    o = Oracle(mesh)
    a, D = _momentum_like(o, mesh, sym)
    n = mesh.n_cells
    B = np.stack([np.sin(np.arange(n) * (0.11 + 0.07 * c)) for c in range(3)])
    ctx = api.Context(0)
    ctx.read_mesh(mesh)
    ctx.geometry()
    ctx.create_CSR_matrix()
    ctx.matrix_set(a)
    X = np.zeros((3, n))
    rc, info, res0 = ctx.csrsolve_multi(solver, X, B, itr_max=2000, tol_rel=1e-10,
                                        criterion=api.FCP_CRIT_TRUE_L2, diag=D)
    assert rc == 0 and info["rel_res_l2"] <= 1e-10, info
    for c in range(3):
        ac = a.copy()
        ac[o.diag] = D[c]
        ctx.matrix_set(ac)
        x = np.zeros(n)
        rc1, i1 = ctx.csrsolve(solver, x, B[c].copy(), itr_max=2000, tol_rel=1e-10,
                               criterion=api.FCP_CRIT_TRUE_L2)
        assert rc1 == 0
        assert np.array_equal(X[c], x), c
        ref = o.solve(solver, ac, B[c], itr_max=5000, tol_abs=0.0, tol_rel=1e-13)["fi"]
        assert np.linalg.norm(X[c] - ref) <= 1e-7 * np.linalg.norm(ref), c
    # the resident matrix's own diagonal is not used: columns with the
    # matrix's diagonal reproduce the shared-matrix batched solve
    ctx.matrix_set(a)
    X2, X3 = np.zeros((3, n)), np.zeros((3, n))
    ctx.csrsolve_multi(solver, X2, B, itr_max=2000, tol_rel=1e-10, criterion=api.FCP_CRIT_TRUE_L2,
                       diag=np.tile(a[o.diag], (3, 1)))
    ctx.csrsolve_multi(solver, X3, B, itr_max=2000, tol_rel=1e-10, criterion=api.FCP_CRIT_TRUE_L2)
    assert np.array_equal(X2, X3)
    with pytest.raises(api.FcpError):  # FP64 only
        ctx.csrsolve_multi(solver, np.zeros((3, n)), B, precision=api.FCP_PREC_MIXED, diag=D)
    ctx.close()


@pytest.fixture(scope="module")
def mom(golden_dir):
    import os
    import sys
    if golden_dir not in sys.path:
        sys.path.insert(0, golden_dir)
    import make_golden as mg
    return mg, np.load(os.path.join(golden_dir, "momentum.npz"))


@pytest.mark.parametrize("key", ["hex", "per"])
@pytest.mark.parametrize("ic", [0, 1, 2, 3])
def test_calcuvw_assembly_and_solves_vs_reference(api, mom, key, ic):
    """fcp_momentum_assemble against the reference's own calcuvw (fixture
    momentum.npz, velocity.f90:50-750): on a jittered 6^3 hex with inlet /
    outlet / wall / symmetry / empty patches and on a periodic 6^3 mesh,
    steady / BDF2 / CN / BDF1, the resident matrix starting from the same
    previous content.  The w system (a and its right-hand side), the v and w
    sources, the diagonal sources and 1/a_P of all three components, the
    updated boundary values and the gradients are bit-exact; with five
    BiCGStab iterations the three components solved together with their own
    diagonals equal the reference's separate solves to 1e-9 (reduction
    order)."""
    mg, g = mom
    print(f"[case] calcuvw {key} case {ic}", flush=True)
    mesh = (hex_mesh(6, jitter=0.2, seed=19, types=mg.PCORR_TYPES) if key == "hex"
            else mg.periodic_mesh())
    ctx = api.Context(0)
    ctx.read_mesh(mesh)
    ctx.geometry()
    csr = ctx.create_CSR_matrix()
    n, nt = mesh.n_cells, mesh.n_total
    I = {k: np.array(g[f"{key}_in_{k}"]) for k in ("u", "v", "w", "uo", "vo", "wo", "uoo", "voo",
                                                  "p", "den", "vis", "visw", "flmass", "astale")}
    I["woo"] = 0.5 * I["uoo"]
    ctx.matrix_set(I["astale"])
    ts, sch, gds, urf, its = mg.MOMENTUM_CASES[ic]
    dp = np.zeros(3 * nt)
    grad = np.zeros(9 * nt)
    ap, spv, src = np.zeros((3, n)), np.zeros((3, n)), np.zeros((3, n))
    a = np.zeros(csr["ja"].size)
    rhs, diag = ctx.momentum_assemble(I["u"], I["v"], I["w"], I["p"], I["den"], I["vis"],
                                      I["flmass"], 1e-3, urf, uo=I["uo"], vo=I["vo"], wo=I["wo"],
                                      uoo=I["uoo"], voo=I["voo"], woo=I["woo"], visw=I["visw"],
                                      dt=0.05, time_scheme=ts, conv_scheme=sch, gds=gds,
                                      dpdxi=dp, grad=grad, ap=ap, spv=spv, src=src, a_out=a)
    t = f"{key}_m{ic}"
    assert np.array_equal(grad[:3 * nt].reshape(-1, 3), g[f"{t}_dudxi"])
    assert np.array_equal(dp.reshape(-1, 3), g[f"{t}_dpdxi"])
    assert np.array_equal(I["p"], g[f"{t}_p"])
    for r, k in enumerate(("spu", "spv", "sp")):
        assert np.array_equal(spv[r], g[f"{t}_{k}"]), k
    for r, k in enumerate(("apu", "apv", "apw")):
        assert np.array_equal(ap[r], g[f"{t}_{k}"]), k
    if its == 0:
        assert np.array_equal(a, g[f"{t}_a"])
        assert np.array_equal(rhs[2], g[f"{t}_su"])
        assert np.array_equal(src[1], g[f"{t}_sv"]) and np.array_equal(src[2], g[f"{t}_sw"])
        for k in ("u", "v", "w"):
            assert np.array_equal(I[k], g[f"{t}_{k}"]), k
    else:
        X = np.ascontiguousarray(np.stack([I["u"][:n], I["v"][:n], I["w"][:n]]))
        rc, info, _ = ctx.csrsolve_multi("bicgstab", X, rhs, itr_max=its, tol_abs=0.0,
                                         tol_rel=1e-3, diag=diag)
        for r, k in enumerate(("u", "v", "w")):
            ref = g[f"{t}_{k}"][:n]
            assert np.allclose(X[r], ref, rtol=0, atol=1e-9 * np.abs(ref).max()), k
    ctx.close()
