"""Polyhedral cells (pairs of jittered hexes merged into ten-face cells,
rows of up to eleven CSR entries): geometry, CSR, assembly, gradients and
every standalone SpMV variant bit-exact vs the oracle, then FP64 and mixed
ICCG solves on cells with more in-tile dependencies per row than the
tile-order kernels take."""
import numpy as np
import pytest

from fcp import workloads
from fcp.mesh import hex_mesh, merge_cells, staggered_pairs
from oracle import Oracle

pytestmark = pytest.mark.gpu

GEOM = ("xc", "yc", "zc", "vol", "arx", "ary", "arz", "xf", "yf", "zf", "facint", "Df", "dnb", "srdb")


@pytest.fixture(scope="module")
def api():
    from fcp import api as _api
    return _api


def _poly(N):
    return merge_cells(hex_mesh(N, jitter=0.2), staggered_pairs(N))


def test_polyhedra_bitexact_vs_oracle(api):
    print("[case] polyhedra 16^3 merged pairs: geometry, CSR, assembly, SpMV", flush=True)
    mesh = _poly(16)
    ctx = api.Context(0)
    ctx.read_mesh(mesh)
    geo = ctx.geometry()
    csr = ctx.create_CSR_matrix()
    o = Oracle(mesh)
    assert np.diff(o.ia).max() == 11
    for k in GEOM:
        assert np.array_equal(geo[k], o.geom[k]), k
    for k in ("ia", "ja", "diag", "slot_pn", "slot_np"):
        assert np.array_equal(csr[k], getattr(o, k)), k
    inp = workloads.poisson_jittered_inputs(mesh, geo["vol"], geo["xc"], geo["yc"], geo["zc"])
    su = inp["source"] * geo["vol"]
    a = np.zeros(csr["ja"].size)
    ctx.laplacian_face(inp["mu_f"], inp["bc_kind"], inp["phi_bc"], su, a_out=a)
    gr = ctx.grad_gauss(inp["phi_prev"])
    ctx.nonorth_correction(inp["gamma"], gr, su)
    a_o, su_o = o.laplacian_face(inp["mu_f"], inp["bc_kind"], inp["phi_bc"],
                                 inp["source"] * o.geom["vol"])
    gr_o = o.grad_gauss(inp["phi_prev"])
    su_o = o.nonorth_correction(inp["gamma"], gr_o, su_o)
    assert np.array_equal(gr, gr_o)
    assert np.array_equal(a, a_o) and np.array_equal(su, su_o)
    for method in (1, 2):
        ctx.create_lsq_grad_matrix(method, compat=False)
        assert np.array_equal(ctx.grad(method, inp["phi_prev"]),
                              o.grad_lsq(method - 1, inp["phi_prev"], compat=0)), method
    # every variant of the standalone SpMV, rows past the unrolled eight
    # entries included, equals the reference row loop bit for bit
    x = np.cos(np.arange(mesh.n_cells) * 0.37) * 10.0 ** np.sin(np.arange(mesh.n_cells))
    y_o = o.spmv(a, x)
    for v in (-1, 0, 1, 2, 3, 4):
        ctx.tune(api.FCP_TUNE_SPMV, v)
        assert np.array_equal(ctx.spmv(x), y_o), v
    ctx.tune(api.FCP_TUNE_SPMV, -1)
    ctx.close()


def test_polyhedra_solves(api):
    print("[case] polyhedra 40^3 merged pairs: FP64 / mixed / FP64-tile ICCG", flush=True)
    mesh = _poly(40)
    ctx = api.Context(0)
    ctx.read_mesh(mesh)
    geo = ctx.geometry()
    csr = ctx.create_CSR_matrix()
    n = mesh.n_cells
    inp = workloads.poisson_jittered_inputs(mesh, geo["vol"], geo["xc"], geo["yc"], geo["zc"])
    su = inp["source"] * geo["vol"]
    a = np.zeros(csr["ja"].size)
    ctx.laplacian_face(inp["mu_f"], inp["bc_kind"], inp["phi_bc"], su, a_out=a)
    o = Oracle(mesh)
    ref = o.solve("iccg", a, su, itr_max=20000, tol_abs=0.0, tol_rel=1e-14)["fi"]
    for precision in (api.FCP_PREC_FP64, api.FCP_PREC_MIXED):
        x = np.zeros(n)
        rc, info = ctx.csrsolve("iccg", x, su, itr_max=20000, tol_rel=1e-10, precision=precision,
                                criterion=api.FCP_CRIT_TRUE_L2)
        assert rc == 0 and info["rel_res_l2"] <= 1e-10, (precision, rc, info)
        print(f"precision {precision}: {info['iterations']} iterations", flush=True)
        assert np.linalg.norm(su - o.spmv(a, x)) <= 1.01e-10 * np.linalg.norm(su), precision
        err = np.linalg.norm(x - ref) / np.linalg.norm(ref)
        assert err < 1e-7, (precision, err)
    # the tile-order kernels take at most four in-tile dependencies per row;
    # the merged cells have up to five lower neighbours, so the mixed solve
    # above ran the row-gather inner PCG and the FP64-tile mode, which has no
    # other form, refuses the mesh
    with pytest.raises(api.FcpError, match="does not apply"):
        ctx.csrsolve("iccg", np.zeros(n), su, itr_max=20000, tol_rel=1e-10,
                     precision=api.FCP_PREC_FP64_TILE, criterion=api.FCP_CRIT_TRUE_L2)
    ctx.close()
