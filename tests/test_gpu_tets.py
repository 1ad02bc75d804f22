"""Tetrahedral meshes with randomly permuted cells: GPU reverse Cuthill-McKee
against the CPU rule, geometry/CSR/gradients after renumbering, and a
Poisson solve with strongly varying diffusivity."""
import numpy as np
import pytest

from fcp import workloads
from fcp.mesh import permute_cells, tet_mesh
from oracle import Oracle

pytestmark = pytest.mark.gpu

GEOM = ("xc", "yc", "zc", "vol", "arx", "ary", "arz", "xf", "yf", "zf", "facint", "Df", "dnb", "srdb")


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


@pytest.mark.parametrize("N", [4, 12])
def test_rcm_bitexact_vs_cpu(api, N):
    print(f"[case] rcm tets {N}^3x6", flush=True)
    mesh = tet_mesh(N, jitter=0.15, perm_seed=3)
    ctx, geo, csr = _setup(api, mesh)
    o = Oracle(mesh)
    for pp in (0, 4):
        assert np.array_equal(ctx.rcm(pp), o.rcm(pp)), pp
    ctx.close()


def test_tets_renumbered_geometry_gradients_poisson(api):
    """Permuted 10^3x6 tets -> GPU RCM -> renumbered mesh: geometry and CSR
    bit-exact vs the oracle, every gradient variant vs the oracle, LSQ exact
    for a linear field, and a Poisson solve (diffusivity 10^U(-2,2))
    against the oracle's tight solve."""
    print("[case] tets 10^3x6 rcm + gradients + poisson", flush=True)
    mesh0 = tet_mesh(10, jitter=0.15, perm_seed=3)
    ctx0, _, _ = _setup(api, mesh0)
    perm = ctx0.rcm(4)
    ctx0.close()
    mesh = permute_cells(mesh0, perm)
    ctx, geo, csr = _setup(api, mesh)
    o = Oracle(mesh)
    for k in GEOM:
        assert np.array_equal(geo[k], o.geom[k]), k
    for k in ("ia", "ja", "diag", "slot_pn", "slot_np"):
        assert np.array_equal(csr[k], getattr(o, k)), k
    n, nt, Fi = mesh.n_cells, mesh.n_total, mesh.n_inner
    inp = workloads.tet_inputs(mesh, geo)
    f = inp["f"]
    assert np.array_equal(ctx.grad_gauss(f), o.grad_gauss(f))
    for method in (1, 2, 3):
        ctx.create_lsq_grad_matrix(method, compat=False)
        g = ctx.grad(method, f)
        g_o = o.grad_lsq(method - 1, f, compat=0)
        if method < 3:
            assert np.array_equal(g, g_o), method
        else:
            assert np.allclose(g, g_o, rtol=1e-12, atol=1e-12 * np.abs(g_o).max())
    xs = np.concatenate([geo["xc"], geo["xf"][Fi:]])
    ys = np.concatenate([geo["yc"], geo["yf"][Fi:]])
    zs = np.concatenate([geo["zc"], geo["zf"][Fi:]])
    lin = 0.5 * xs + 3.0 * ys - zs
    for method in (1, 2, 3):
        ctx.create_lsq_grad_matrix(method, compat=False)
        assert np.allclose(ctx.grad(method, lin)[:n], [0.5, 3.0, -1.0], rtol=0, atol=1e-9)
    # Poisson, FP64 and mixed precision
    su = inp["source"] * geo["vol"]
    a = np.zeros(csr["ja"].size)
    ctx.laplacian_face(inp["mu_f"], inp["bc_kind"], inp["phi_bc"], su, a_out=a)
    a_o, su_o = o.laplacian_face(inp["mu_f"], inp["bc_kind"], inp["phi_bc"],
                                 inp["source"] * o.geom["vol"])
    assert np.array_equal(a, a_o) and np.array_equal(su, su_o)
    ref = o.solve("iccg", a_o, su_o, itr_max=20000, tol_abs=0.0, tol_rel=1e-14)["fi"]
    for precision in (0, 1):
        x = np.zeros(n)
        rc, info = ctx.csrsolve("iccg", x, su, itr_max=20000, tol_rel=1e-10, precision=precision,
                                criterion=1)
        assert rc == 0 and info["rel_res_l2"] <= 1e-10
        err = np.linalg.norm(x - ref) / np.linalg.norm(ref)
        assert err < 1e-7, (precision, err)
    ctx.close()
