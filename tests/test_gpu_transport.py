"""The scalar transport step on the GPU (fcp_transport, the replacement of
calculate_temperature, temperature.f90:35-385) against the reference
routine's own output (golden fixture transport.npz, written by the
flang-built reference) and against the oracle on a larger mesh.

The gradient and the assembled system (a, su, sp) are bit-exact; the updated
field carries the linear solver's trajectory, whose dot products are summed
in another order than the reference's loops, so it is compared within 1e-11
of its largest value.
"""
import os
import sys

import numpy as np
import pytest

from fcp.mesh import hex_mesh
from oracle import Oracle, orc_small

pytestmark = pytest.mark.gpu

GRAD = {"gauss": 0, "lsq": 1, "lsq_dm": 2, "lsq_qr": 3}


@pytest.fixture(scope="module")
def cases(golden_dir):
    if golden_dir not in sys.path:
        sys.path.insert(0, golden_dir)
    import make_golden as mg
    return mg, np.load(os.path.join(golden_dir, "transport.npz"))


def _run(api, mesh, inp, setting, a_size):
    ts, sch, slv, itm, trel, gm, urf, gds = setting
    ctx = api.Context(0)
    ctx.read_mesh(mesh)
    ctx.geometry()
    ctx.create_CSR_matrix()
    if gm != "gauss":
        ctx.create_lsq_grad_matrix(GRAD[gm], compat=True)
    n, nt = mesh.n_cells, mesh.n_total
    T = inp["T"].copy()
    out = dict(grad=np.zeros((nt, 3)), a=np.zeros(a_size), su=np.zeros(n), sp=np.zeros(n))
    rc, info = ctx.transport(T, inp["vis"], inp["flm"], inp["wf"], 1e-3, 0.7, 1.0, urf, gds, slv,
                             itm, 1e-13, trel, time_scheme=ts, den=inp["den"], dt=1e-2,
                             phi_o=inp["To"], phi_oo=inp["Too"], conv_scheme=sch,
                             grad_method=GRAD[gm], grad_out=out["grad"], a_out=out["a"],
                             su_out=out["su"], sp_out=out["sp"])
    ctx.close()
    assert rc >= 0, info
    return T, out, info


@pytest.mark.parametrize("ic", range(5))
def test_transport_step_vs_calculate_temperature(cases, ic):
    from fcp import api
    mg, g = cases
    mk, *setting = mg.TRANSPORT_CASES[ic]
    print(f"[case] transport {ic}: mesh {mk} {setting}", flush=True)
    types, _ = mg.TRANSPORT_MESHES[mk]
    mesh = hex_mesh(5, jitter=0.2, seed=9, types=types)
    inp = {k: g[f"{mk}_in_{k}"] for k in ("T", "To", "Too", "den", "vis", "flm", "wf")}
    T, out, info = _run(api, mesh, inp, setting, g[f"c{ic}_a"].size)
    n = mesh.n_cells
    assert np.array_equal(out["grad"][:n], g[f"c{ic}_grad"][:n])
    for k in ("a", "su", "sp"):
        assert np.array_equal(out[k], g[f"c{ic}_{k}"]), k
    Tr = g[f"c{ic}_T"]
    assert np.abs(T - Tr).max() <= 1e-11 * np.abs(Tr).max()
    r = float(g[f"c{ic}_resor"])
    assert abs(info["resor"] - r) <= 1e-12 * abs(r)
    assert abs(info["fimax"] - Tr[:n].max()) <= 1e-11 * np.abs(Tr).max()
    if info["fimin"] < 0.0:  # the clip ran
        assert T[:n].min() == orc_small()


@pytest.mark.parametrize("ts", ["cn", "bdf2"])
def test_transport_step_vs_oracle_16cube(ts):
    """A 16^3 jittered hex with every non-periodic boundary treatment: the
    GPU step against the oracle's restatement (bit-exact assembly)."""
    from fcp import api
    from fcp.rng import uniform
    print(f"[case] transport 16^3 {ts}", flush=True)
    types = ("inlet", "wall", "wall", "outlet", "wallQFlux", "pressure")
    mesh = hex_mesh(16, jitter=0.2, seed=11, types=types)
    n, nt, F = mesh.n_cells, mesh.n_total, mesh.n_faces
    inp = dict(T=uniform(8, 1, nt, 0.0, 1.0), To=uniform(8, 2, n, 0.0, 1.0),
               Too=uniform(8, 3, n, 0.0, 1.0), den=1.0 + 0.1 * uniform(8, 4, n, 0.0, 1.0),
               vis=1e-3 * (1.0 + uniform(8, 5, n, 0.0, 1.0)),
               flm=1e-3 * uniform(8, 6, F, -1.0, 1.0), wf=np.array([0, 0, 1, 0, 0, 0], float))
    setting = (ts, "linearUpwind", "bicgstab", 30, 1e-9, "gauss", 0.8, 1.0)
    o = Oracle(mesh)
    T, out, info = _run(api, mesh, inp, setting, o.nnz)
    To_, d = o.temperature(inp["T"], inp["To"], inp["Too"], inp["den"], inp["vis"], inp["flm"],
                           inp["wf"], 1e-3, 0.7, 1.0, 0.8, 1.0, 1e-2, ts, "linearUpwind",
                           "bicgstab", 30, 1e-13, 1e-9, "gauss")
    for k in ("a", "su", "sp"):
        assert np.array_equal(out[k], d[k]), k
    assert np.abs(T - To_).max() <= 1e-11 * np.abs(To_).max()
