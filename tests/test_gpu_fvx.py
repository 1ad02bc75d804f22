"""fvxGradient.f90's Gauss gradients on the GPU (fcp_grad with
FCP_GRAD_GAUSS_FVX / FCP_GRAD_GAUSS_CORRECTED):

- against the flang-built reference's own output (golden fixture
  gradfvx.npz: the two-sweep grad_gauss, :1549-1645, and
  grad_gauss_corrected, :1664-1758, on a jittered hex mesh with every
  boundary type and a jittered tet mesh), host buffers;
- against the oracle on a larger permuted tet mesh and a 48^3 jittered hex,
  with device buffers, the corrected form updating its argument in place.

Every comparison is bit-exact: each cell gathers its faces in the order the
reference's face loops visit them, with the reference's arithmetic.
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


def _ctx(api, mesh):
    ctx = api.Context(0)
    ctx.read_mesh(mesh)
    ctx.geometry()
    return ctx


@pytest.mark.parametrize("key", ["hex", "tet"])
def test_grad_fvx_bitexact_vs_reference(api, golden_dir, key):
    print(f"[case] fvx gauss vs reference fixture ({key})", flush=True)
    g = np.load(os.path.join(golden_dir, "gradfvx.npz"))
    mesh = (hex_mesh(5, jitter=0.2, seed=7, types=("inlet", "outlet", "wall", "symmetry",
                                                   "pressure", "empty"))
            if key == "hex" else tet_mesh(3, jitter=0.2, seed=3))
    ctx = _ctx(api, mesh)
    n = mesh.n_cells
    phi, gin = g[f"{key}_in_phi"], g[f"{key}_in_gin"]
    g2 = ctx.grad(api.FCP_GRAD_GAUSS_FVX, phi)
    assert np.array_equal(g2[:n], g[f"{key}_gauss2"]) and not g2[n:].any()
    gc = ctx.grad(api.FCP_GRAD_GAUSS_CORRECTED, phi, gold=gin)
    assert np.array_equal(gc[:n], g[f"{key}_corrected"]) and not gc[n:].any()
    ctx.close()


@pytest.mark.parametrize("which", ["tet6p", "hex48"])
def test_grad_fvx_device_buffers_vs_oracle(api, which):
    import torch
    print(f"[case] fvx gauss device buffers vs oracle ({which})", flush=True)
    mesh = (tet_mesh(6, jitter=0.2, seed=4, perm_seed=9) if which == "tet6p"
            else hex_mesh(48, jitter=0.15, seed=2))
    o = Oracle(mesh)
    n, nt = mesh.n_cells, mesh.n_total
    rng = np.random.default_rng(31)
    phi = rng.uniform(-1, 1, nt)
    gin = rng.uniform(-1, 1, (n, 3))
    ctx = _ctx(api, mesh)
    dev = torch.device("cuda:0")
    phi_d = torch.from_numpy(phi).to(dev)
    out = torch.full((nt, 3), 7.0, dtype=torch.float64, device=dev)
    ctx.grad(api.FCP_GRAD_GAUSS_FVX, phi_d, out=out)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), o.grad_gauss_fvx(phi))
    io = torch.zeros((nt, 3), dtype=torch.float64, device=dev)
    io[:n] = torch.from_numpy(gin).to(dev)
    ctx.grad(api.FCP_GRAD_GAUSS_CORRECTED, phi_d, out=io)
    torch.cuda.synchronize()
    assert np.array_equal(io.cpu().numpy(), o.grad_gauss_fvx(phi, gin))
    ctx.close()
