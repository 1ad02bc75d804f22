"""Periodic boundaries on the GPU against the reference's own output (golden
fixture periodic.npz, written by the flang-built reference on a jittered
6^3 hex whose x+ and z+ patches are periodic with shuffled twin faces) and
against the oracle on a larger mesh:

- face_mapping (geometry.f90:1848-1997) as fcp_geometry runs it: the twin
  faces' centres and area vectors permuted as the reference permutes them;
- the CSR pattern with the periodic pairs and the slot maps over inner +
  periodic faces (sparse_matrix.f90:141-171, 262-293);
- the Laplacian and the Gauss gradient on the mapped owners, and
  updateBoundary's periodic average (updateBoundary.f90:46-68);
- fcp_transport with calculate_temperature's periodic branch
  (temperature.f90:262-313, facefluxsc_periodic scalar_fluxes.f90:147-251);
- the same mesh read from disk by the library (the twin start face on the
  periodic patches' boundary lines).

Integer, geometry and assembly outputs are bit-exact; the solved field
carries the solver's summation order and is compared within 1e-11.
"""
import os
import sys

import numpy as np
import pytest

from fcp.mesh import hex_mesh, write_native
from oracle import Oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api():
    from fcp import api as _api
    return _api


@pytest.fixture(scope="module")
def per(golden_dir):
    if golden_dir not in sys.path:
        sys.path.insert(0, golden_dir)
    import make_golden as mg
    return mg, np.load(os.path.join(golden_dir, "periodic.npz"))


def _ctx(api, mesh):
    ctx = api.Context(0)
    ctx.read_mesh(mesh)
    geo = ctx.geometry()
    csr = ctx.create_CSR_matrix()
    return ctx, geo, csr


def test_periodic_geometry_csr_bitexact(api, per):
    mg, g = per
    print("[case] periodic geometry + CSR 6^3", flush=True)
    mesh = mg.periodic_mesh()
    ctx, geo, csr = _ctx(api, mesh)
    for k in ("xc", "yc", "zc", "vol", "arx", "ary", "arz", "xf", "yf", "zf", "facint", "Df"):
        assert np.array_equal(geo[k], g[f"g_{k}"]), k
    for k in ("ia", "ja", "diag"):
        assert np.array_equal(csr[k], g[f"c_{k}"]), k
    assert np.array_equal(csr["slot_pn"], g["p_slot_pn"])
    assert np.array_equal(csr["slot_np"], g["p_slot_np"])
    ctx.close()


def test_periodic_operators_bitexact(api, per):
    mg, g = per
    print("[case] periodic laplacian / grad / updateBoundary 6^3", flush=True)
    mesh = mg.periodic_mesh()
    ctx, geo, csr = _ctx(api, mesh)
    a = np.full(csr["ja"].size, 7.0)  # stale values: the periodic entries must come back zero
    ctx.matrix_set(a)
    su = g["in_su0"].copy()
    ctx.laplacian(g["in_mu"], g["in_phi"], su, a_out=a)
    assert np.array_equal(a, g["lap_a"]) and np.array_equal(su, g["lap_su"])
    assert np.array_equal(ctx.grad_gauss(g["in_u"]), g["grad_gauss"])
    phi = g["in_phi"].copy()
    ctx.update_boundary(phi)
    assert np.array_equal(phi, g["update_boundary"])
    ctx.close()


@pytest.mark.parametrize("ic", [0, 1])
def test_periodic_transport_vs_calculate_temperature(api, per, ic):
    mg, g = per
    ts, sch, slv, itm, trel, gm, urf, gds = mg.PERIODIC_TCASES[ic]
    print(f"[case] periodic transport {ic} ({ts}, {sch}) 6^3", flush=True)
    mesh = mg.periodic_mesh()
    ctx, geo, csr = _ctx(api, mesh)
    n, nt = mesh.n_cells, mesh.n_total
    I = {k: g[f"in_{k}"] for k in ("T", "To", "Too", "den", "vis", "flm", "wf")}
    T = I["T"].copy()
    out = dict(grad=np.zeros((nt, 3)), a=np.zeros(csr["ja"].size), su=np.zeros(n), sp=np.zeros(n))
    rc, info = ctx.transport(T, I["vis"], I["flm"], I["wf"], 1e-3, 0.7, 1.0, urf, gds, slv, itm,
                             1e-13, trel, time_scheme=ts, den=I["den"], dt=1e-2, phi_o=I["To"],
                             phi_oo=I["Too"], conv_scheme=sch, grad_out=out["grad"],
                             a_out=out["a"], su_out=out["su"], sp_out=out["sp"])
    ctx.close()
    assert rc >= 0, info
    for k in ("grad", "a", "su", "sp"):
        assert np.array_equal(out[k], g[f"t{ic}_{k}"]), k
    Tr = g[f"t{ic}_T"]
    assert np.abs(T - Tr).max() <= 1e-11 * np.abs(Tr).max()
    r = float(g[f"t{ic}_resor"])
    assert abs(info["resor"] - r) <= 1e-12 * abs(r)


def test_periodic_read_from_disk_same_as_upload(api, per, tmp_path):
    """fcp_read_polymesh takes the twin start face from the periodic lines
    of polyMesh/boundary: geometry and CSR equal the upload path's."""
    mg, g = per
    print("[case] periodic mesh from disk 6^3", flush=True)
    mesh = mg.periodic_mesh()
    write_native(mesh, str(tmp_path))
    ctx = api.Context(0)
    sizes, patches = ctx.read_mesh_file(str(tmp_path), "native")
    ctx.mesh = mesh  # same arrays, for the output sizes
    geo = ctx.geometry()
    csr = ctx.create_CSR_matrix()
    ctx.close()
    for k in ("xf", "arx", "facint"):
        assert np.array_equal(geo[k], g[f"g_{k}"]), k
    assert np.array_equal(csr["ja"], g["c_ja"]) and np.array_equal(csr["slot_np"], g["p_slot_np"])


@pytest.mark.parametrize("ts", ["cn", "bdf2"])
def test_periodic_transport_vs_oracle_16cube(api, ts):
    """A 16^3 jittered hex periodic in x and y (shuffled twins), with an
    isothermal wall and an inlet on the z patches: the GPU assembly against
    the oracle's, bit for bit."""
    from fcp.mesh import fisher_yates
    print(f"[case] periodic transport vs oracle 16^3 {ts}", flush=True)
    types = ("empty", "periodic", "empty", "periodic", "wall", "inlet")
    mesh = hex_mesh(16, jitter=0.2, seed=21, types=types)
    for k, t in enumerate(mesh.twin_start):
        nf = 256
        perm = fisher_yates(nf, 31 + k)
        sl = slice(int(t), int(t) + nf)
        mesh.node[sl] = mesh.node[sl][perm]
        mesh.nnodes[sl] = mesh.nnodes[sl][perm]
        mesh.owner[sl] = mesh.owner[sl][perm]
    o = Oracle(mesh)
    from fcp.rng import uniform
    n, nt, F = mesh.n_cells, mesh.n_total, mesh.n_faces
    T = uniform(41, 1, nt, 0.0, 1.0)
    To = uniform(41, 2, n, 0.0, 1.0)
    Too = uniform(41, 3, n, 0.0, 1.0)
    den = 1.0 + 0.1 * uniform(41, 4, n, 0.0, 1.0)
    vis = 1e-3 * (1.0 + uniform(41, 5, n, 0.0, 1.0))
    flm = 1e-2 * uniform(41, 6, F, -1.0, 1.0)
    wf = np.array([0, 0, 0, 0, 1, 0], np.float64)
    Tg = T.copy()
    ctx, geo, csr = _ctx(api, mesh)
    out = dict(a=np.zeros(csr["ja"].size), su=np.zeros(n), sp=np.zeros(n))
    rc, info = ctx.transport(Tg, vis, flm, wf, 1e-3, 0.7, 1.0, 0.9, 1.0, "bicgstab", 30, 1e-13,
                             1e-8, time_scheme=ts, den=den, dt=1e-2, phi_o=To, phi_oo=Too,
                             conv_scheme="linearUpwind", a_out=out["a"], su_out=out["su"],
                             sp_out=out["sp"])
    ctx.close()
    assert rc >= 0, info
    To_, d = o.temperature(T, To, Too, den, vis, flm, wf, 1e-3, 0.7, 1.0, 0.9, 1.0, 1e-2, ts,
                           "linearUpwind", "bicgstab", 30, 1e-13, 1e-8)
    for k in ("a", "su", "sp"):
        assert np.array_equal(out[k], d[k]), k
    assert np.abs(Tg - To_).max() <= 1e-9 * np.abs(To_).max()
