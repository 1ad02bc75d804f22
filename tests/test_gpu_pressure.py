"""The pressure-correction step of SIMPLE on the GPU (fcp_pressure_assemble,
fcp_pressure_correct; calcp_simple.f90:69-419 for an incompressible case
without pressure patches):

- against the reference's own routines (golden fixture pcorr.npz, composed
  from facefluxmass2, facefluxmass2_periodic, adjustMassFlow,
  gradp_and_sources/bpres and updateVelocityAtBoundary in calcp_simple's
  order) on a jittered 6^3 hex with inlet/outlet/wall/symmetry/empty
  patches and on a periodic 6^3 mesh, host buffers, bit-exact;
- against the oracle on a jittered 20^3 hex with device buffers, bit-exact;
- the whole step (assembly, FP64 PCG on the resident matrix, correction):
  the corrected mass fluxes are divergence-free to the solve's tolerance.
"""
import os
import sys

import numpy as np
import pytest

from fcp.mesh import hex_mesh
from oracle import Oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api():
    from fcp import api as _api
    return _api


@pytest.fixture(scope="module")
def pc(golden_dir):
    if golden_dir not in sys.path:
        sys.path.insert(0, golden_dir)
    import make_golden as mg
    return mg, np.load(os.path.join(golden_dir, "pcorr.npz"))


def _mesh(mg, key):
    return (hex_mesh(6, jitter=0.2, seed=13, types=mg.PCORR_TYPES) if key == "hex"
            else mg.periodic_mesh())


def _setup(api, mesh):
    ctx = api.Context(0)
    ctx.read_mesh(mesh)
    ctx.geometry()
    csr = ctx.create_CSR_matrix()
    return ctx, csr


@pytest.mark.parametrize("key,cmf", [("hex", 0), ("hex", 1), ("per", 1)])
def test_pressure_assemble_bitexact_vs_reference(api, pc, key, cmf):
    mg, g = pc
    print(f"[case] pressure assemble {key} const_mflux={cmf}", flush=True)
    mesh = _mesh(mg, key)
    ctx, csr = _setup(api, mesh)
    I = {k: np.array(g[f"{key}_in_{k}"]) for k in ("u", "v", "w", "p", "dpdxi", "den", "apu",
                                                  "apv", "apw", "flmass")}
    su = np.zeros(mesh.n_cells)
    a = np.zeros(csr["ja"].size)
    ctx.pressure_assemble(I["u"], I["v"], I["w"], I["p"], I["dpdxi"], I["den"], I["apu"],
                          I["apv"], I["apw"], I["flmass"], su, const_mflux=cmf,
                          flomas=float(g[f"{key}_flomas"]), a_out=a)
    t = f"{key}_c{cmf}"
    assert np.array_equal(a, g[f"{t}_a"])
    assert np.array_equal(su, g[f"{t}_su"])
    assert np.array_equal(I["flmass"], g[f"{t}_flmass"])
    for k in ("u", "v", "w"):
        assert np.array_equal(I[k], g[f"{t}_{k}"]), k
    ctx.close()


@pytest.mark.parametrize("key", ["hex", "per"])
def test_pressure_correct_bitexact_vs_reference(api, pc, key):
    mg, g = pc
    print(f"[case] pressure correct {key}", flush=True)
    mesh = _mesh(mg, key)
    ctx, csr = _setup(api, mesh)
    n, nt = mesh.n_cells, mesh.n_total
    c1 = f"{key}_c1"
    ctx.matrix_set(g[f"{c1}_a"])
    pp = np.array(g[f"{key}_in_pp"])
    fm = np.array(g[f"{c1}_flmass"])
    u, v, w = (np.array(g[f"{c1}_{k}"]) for k in ("u", "v", "w"))
    p = np.array(g[f"{key}_in_p"])
    dp = np.zeros(3 * nt)
    su, sv, sw = np.zeros(n), np.zeros(n), np.zeros(n)
    ctx.pressure_correct(pp, fm, u, v, w, p, dp, g[f"{key}_in_apu"], g[f"{key}_in_apv"],
                         g[f"{key}_in_apw"], 0.3, 7, su, sv, sw)
    for k, x in (("pp", pp), ("flmass", fm), ("u", u), ("v", v), ("w", w), ("p", p),
                 ("dpdxi", dp), ("su", su), ("sv", sv), ("sw", sw)):
        assert np.array_equal(x, g[f"{key}_corr_{k}"]), k
    ctx.close()


def test_pressure_step_device_buffers_vs_oracle(api, pc):
    import torch
    mg, _ = pc
    print("[case] pressure step 20^3 device buffers", flush=True)
    mesh = hex_mesh(20, jitter=0.15, seed=21, types=mg.PCORR_TYPES)
    o = Oracle(mesh)
    I = mg.pcorr_inputs(mesh, 23)
    n, nt, Fi = mesh.n_cells, mesh.n_total, mesh.n_inner
    types = [mesh.types[pt] for pt in mesh.patch_of_bface()]
    inlet = np.array([Fi + b for b, t in enumerate(types) if t == "inlet"])
    flomas = -float(I["flmass"][inlet].sum())  # inflow: inlet fluxes are negative
    ctx, csr = _setup(api, mesh)
    dev = torch.device("cuda:0")
    t = lambda x: torch.from_numpy(np.ascontiguousarray(x, np.float64)).to(dev)  # noqa: E731
    D = {k: t(v) for k, v in I.items()}
    su = torch.zeros(n, dtype=torch.float64, device=dev)
    a = torch.zeros(csr["ja"].size, dtype=torch.float64, device=dev)
    ctx.pressure_assemble(D["u"], D["v"], D["w"], D["p"], D["dpdxi"], D["den"], D["apu"],
                          D["apv"], D["apw"], D["flmass"], su, const_mflux=False, flomas=flomas,
                          a_out=a)
    torch.cuda.synchronize()
    a_o, su_o, fm_o, u_o, v_o, w_o = o.pcorr_assemble(I["u"], I["v"], I["w"], I["p"], I["dpdxi"],
                                                      I["den"], I["apu"], I["apv"], I["apw"],
                                                      I["flmass"], 0, flomas)
    assert np.array_equal(a.cpu().numpy(), a_o)
    assert np.array_equal(su.cpu().numpy(), su_o)
    assert np.array_equal(D["flmass"].cpu().numpy(), fm_o)
    for k, ref in (("u", u_o), ("v", v_o), ("w", w_o)):
        assert np.array_equal(D[k].cpu().numpy(), ref), k
    # solve A pp = su on the resident matrix, then correct
    pp = np.zeros(nt)
    rc, info = ctx.csrsolve("dpcg", pp, su.cpu().numpy(), itr_max=2000, tol_rel=1e-12)
    fm_before = D["flmass"].cpu().numpy().copy()
    pp_d = t(pp)
    dp = torch.zeros(3 * nt, dtype=torch.float64, device=dev)
    S = [torch.zeros(n, dtype=torch.float64, device=dev) for _ in range(3)]
    ctx.pressure_correct(pp_d, D["flmass"], D["u"], D["v"], D["w"], D["p"], dp, D["apu"],
                         D["apv"], D["apw"], 0.7, 1, *S)
    torch.cuda.synchronize()
    r = o.pcorr_correct(a_o, pp, fm_o, u_o, v_o, w_o, I["p"], I["apu"], I["apv"], I["apw"], 0.7, 0)
    for k, x in (("pp", pp_d), ("flmass", D["flmass"]), ("u", D["u"]), ("v", D["v"]),
                 ("w", D["w"]), ("p", D["p"]), ("dpdxi", dp), ("su", S[0]), ("sv", S[1]),
                 ("sw", S[2])):
        assert np.array_equal(x.cpu().numpy(), r[k]), k
    # continuity: the corrected fluxes' net outflow per cell is A pp - su,
    # small against the uncorrected imbalance
    fm1 = D["flmass"].cpu().numpy()

    def net_out(fm):
        d = np.zeros(n)
        np.add.at(d, o.owner[:Fi], fm[:Fi])
        np.subtract.at(d, mesh.neighbour, fm[:Fi])
        for b, t_ in enumerate(types):
            if t_ in ("inlet", "outlet"):
                d[o.owner[Fi + b]] += fm[Fi + b]
        return d
    e0 = np.abs(net_out(fm_before)).sum()
    e1 = np.abs(net_out(fm1)).sum()
    assert e1 < 1e-9 * e0, (e0, e1)
    ctx.close()
