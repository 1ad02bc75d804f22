"""facefluxscalar (scalar_fluxes.f90:344-464) on the GPU: fcp_convdiff_cell,
the convection-diffusion assembly with the diffusivity given per cell and
interpolated to each inner face (:384-392):

- against the flang-built reference's own assembly (golden fixture
  fluxscalar.npz: jittered 5^3 hex with every boundary kind and a jittered
  3^3 tet mesh, 'cds' and 'linearUpwind'), host buffers;
- against the oracle on a permuted 6^3 tet mesh with device buffers and
  under-relaxation.

Bit-exact throughout.
"""
import os

import numpy as np
import pytest

from fcp.mesh import hex_mesh, tet_mesh
from oracle import Oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api():
    from fcp import api as _api
    return _api


def _setup(api, mesh):
    ctx = api.Context(0)
    ctx.read_mesh(mesh)
    ctx.geometry()
    csr = ctx.create_CSR_matrix()
    return ctx, csr


@pytest.mark.parametrize("key", ["hex", "tet"])
@pytest.mark.parametrize("scheme,code", [("cds", 0), ("linearUpwind", 1)])
def test_convdiff_cell_bitexact_vs_reference(api, golden_dir, key, scheme, code):
    print(f"[case] facefluxscalar {key} {scheme}", flush=True)
    g = np.load(os.path.join(golden_dir, "fluxscalar.npz"))
    mesh = (hex_mesh(5, jitter=0.2, seed=7, types=("inlet", "outlet", "wall", "symmetry",
                                                   "pressure", "empty"))
            if key == "hex" else tet_mesh(3, jitter=0.2, seed=3))
    ctx, csr = _setup(api, mesh)
    n = mesh.n_cells
    I = {k: g[f"{key}_in_{k}"] for k in ("flmass", "vis", "phi", "grad", "bckind")}
    su, sp = np.zeros(n), np.zeros(n)
    a = np.zeros(csr["ja"].size)
    phi = I["phi"].copy()
    ctx.convdiff_cell(I["flmass"], I["vis"], phi, I["grad"], 0.7, code, I["bckind"], su, sp,
                      urf=1.0, a_out=a)
    assert np.array_equal(a, g[f"{key}_{scheme}_a"])
    assert np.array_equal(su, g[f"{key}_{scheme}_su"])
    assert np.array_equal(sp, g[f"{key}_{scheme}_sp"])
    ctx.close()


def test_convdiff_cell_device_buffers_vs_oracle(api):
    import torch
    print("[case] facefluxscalar device buffers tet6p urf 0.7", flush=True)
    mesh = tet_mesh(6, jitter=0.2, seed=4, perm_seed=9)
    o = Oracle(mesh)
    n, F, nt = mesh.n_cells, mesh.n_faces, mesh.n_total
    rng = np.random.default_rng(44)
    flm = rng.uniform(-1, 1, F)
    vis = 10.0 ** rng.uniform(-1, 1, n)
    phi = rng.uniform(-1, 1, nt)
    gr = rng.uniform(-1, 1, (nt, 3))
    bk = (np.arange(mesh.n_bnd) % 3).astype(np.uint8)
    ctx, csr = _setup(api, mesh)
    dev = torch.device("cuda:0")
    t = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to(dev)  # noqa: E731
    su, sp = torch.zeros(n, dtype=torch.float64, device=dev), torch.zeros(n, dtype=torch.float64,
                                                                           device=dev)
    a = torch.zeros(csr["ja"].size, dtype=torch.float64, device=dev)
    ph = t(phi)
    ctx.convdiff_cell(t(flm), t(vis), ph, t(gr), 0.9, 1, t(bk), su, sp, urf=0.7, a_out=a)
    torch.cuda.synchronize()
    a_o, su_o, sp_o, ph_o = o.convdiff_cell(flm, vis, phi, gr, 0.9, 1, bk, urf=0.7)
    assert np.array_equal(a.cpu().numpy(), a_o)
    assert np.array_equal(su.cpu().numpy(), su_o)
    assert np.array_equal(sp.cpu().numpy(), sp_o)
    assert np.array_equal(ph.cpu().numpy(), ph_o)
    ctx.close()
