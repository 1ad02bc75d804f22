"""Discrete adjoint of the convection-diffusion path: GPU gradients of
J = sum w phi V with respect to face diffusivity, cell source and Dirichlet
boundary values, checked against central finite differences of the GPU
primal (step 1e-6) and against a CPU reference gradient (exhaustive central
differences of the oracle's primal) on a small mesh."""
import numpy as np
import pytest

from fcp import workloads
from oracle import Oracle

pytestmark = pytest.mark.gpu


def _problem(N, K=2):
    import torch
    from fcp.adjoint import ConvDiffProblem
    from fcp.api import Context
    mesh = workloads.convdiff_mesh(N, jitter=0.2)
    ctx = Context(0)
    ctx.read_mesh(mesh)
    geo = ctx.geometry()
    csr = ctx.create_CSR_matrix()
    flux, _ = workloads.convdiff_flux(mesh, geo, csr, ctx)
    inp = workloads.adjoint_inputs(mesh, geo, flux)
    prob = ConvDiffProblem(ctx, mesh, geo, flux, inp["bc_kind"], K=K)
    dev = torch.device("cuda:0")
    P = {k: torch.from_numpy(np.ascontiguousarray(inp[k])).to(dev) for k in ("gamma", "source", "phib", "w")}
    vol = torch.from_numpy(geo["vol"]).to(dev)
    return mesh, ctx, geo, csr, flux, inp, prob, P, vol


def _J(prob, P, vol, gamma, s, phib):
    phi = prob.forward(gamma, s, phib)
    return float((P["w"] * phi * vol).sum())


def test_adjoint_vs_central_differences_gpu():
    import torch
    from fcp.adjoint import ConvDiffSolve
    print("[case] adjoint 8^3 jittered vs FD", flush=True)
    mesh, ctx, geo, csr, flux, inp, prob, P, vol = _problem(8)
    g = P["gamma"].clone().requires_grad_(True)
    s = P["source"].clone().requires_grad_(True)
    b = P["phib"].clone().requires_grad_(True)
    phi = ConvDiffSolve.apply(g, s, b, prob)
    J = (P["w"] * phi * vol).sum()
    J.backward()
    gg, gs, gb = g.grad, s.grad, b.grad
    assert torch.isfinite(gg).all() and torch.isfinite(gs).all() and torch.isfinite(gb).all()
    rng = np.random.default_rng(4)
    dirichlet = torch.from_numpy(inp["bc_kind"] == 0).to(gb.device)
    h = 1e-6
    fds, ads = [], []
    for _ in range(16):
        vg = torch.from_numpy(rng.uniform(-1, 1, g.numel())).to(g.device) * P["gamma"]
        vs = torch.from_numpy(rng.uniform(-1, 1, s.numel())).to(g.device)
        vb = torch.from_numpy(rng.uniform(-1, 1, b.numel())).to(g.device) * dirichlet
        Jp = _J(prob, P, vol, P["gamma"] + h * vg, P["source"] + h * vs, P["phib"] + h * vb)
        Jm = _J(prob, P, vol, P["gamma"] - h * vg, P["source"] - h * vs, P["phib"] - h * vb)
        fds.append((Jp - Jm) / (2 * h))
        ads.append(float((gg * vg).sum() + (gs * vs).sum() + (gb * vb).sum()))
    fds, ads = np.array(fds), np.array(ads)
    # FD carries the primal solves' rounding (tolerance 1e-13) divided by 2h,
    # so the comparison is relative to the largest directional derivative
    err = np.abs(fds - ads).max() / np.abs(ads).max()
    print(f"max |FD - adjoint| / max |adjoint| over 16 directions: {err:.2e}; "
          f"per-direction worst {np.max(np.abs(fds - ads) / np.abs(ads)):.2e}", flush=True)
    assert err < 1e-6
    ctx.close()


def test_adjoint_vs_cpu_reference_gradient():
    """Every component of the GPU gradient against central differences of the
    oracle's primal (CPU, FP64) on a 4^3 mesh."""
    import torch
    from fcp.adjoint import ConvDiffSolve
    print("[case] adjoint 4^3 vs CPU reference gradient", flush=True)
    mesh, ctx, geo, csr, flux, inp, prob, P, vol = _problem(4)
    g = P["gamma"].clone().requires_grad_(True)
    s = P["source"].clone().requires_grad_(True)
    b = P["phib"].clone().requires_grad_(True)
    phi = ConvDiffSolve.apply(g, s, b, prob)
    (P["w"] * phi * vol).sum().backward()
    o = Oracle(mesh)
    n, nt, Fi = mesh.n_cells, mesh.n_total, mesh.n_inner
    bk = inp["bc_kind"]
    own_b = mesh.owner[Fi:]
    V = o.geom["vol"]

    def J_cpu(gamma, src, phib):
        x = np.zeros(n)
        for _ in range(inp["K"]):
            ph = np.empty(nt)
            ph[:n] = x
            ph[n:] = np.where(bk == 0, phib, x[own_b])
            gr = o.grad_gauss(ph)
            a, su, sp, _ = o.convdiff(flux, gamma, ph, gr, 1.0, 1, bk, urf=1.0)
            x = o.solve("bicgstab", a, src * V + su, itr_max=5000, tol_abs=0.0, tol_rel=1e-15)["fi"]
        return float(np.sum(inp["w"] * x * V))

    # relative step 1e-4 for the diffusivities: wall faces carry tiny Gamma,
    # where a 1e-6 relative step drowns in the primal's rounding
    h = 1e-4
    ref_g = np.zeros(mesh.n_faces)
    for f in range(mesh.n_faces):
        e = np.zeros(mesh.n_faces)
        e[f] = h * inp["gamma"][f]
        ref_g[f] = (J_cpu(inp["gamma"] + e, inp["source"], inp["phib"]) -
                    J_cpu(inp["gamma"] - e, inp["source"], inp["phib"])) / (2 * h * inp["gamma"][f])
    h = 1e-6
    ref_s = np.zeros(n)
    for i in range(n):
        e = np.zeros(n)
        e[i] = h
        ref_s[i] = (J_cpu(inp["gamma"], inp["source"] + e, inp["phib"]) -
                    J_cpu(inp["gamma"], inp["source"] - e, inp["phib"])) / (2 * h)
    ref_b = np.zeros(mesh.n_bnd)
    for j in np.nonzero(bk == 0)[0]:
        e = np.zeros(mesh.n_bnd)
        e[j] = h
        ref_b[j] = (J_cpu(inp["gamma"], inp["source"], inp["phib"] + e) -
                    J_cpu(inp["gamma"], inp["source"], inp["phib"] - e)) / (2 * h)
    for name, gpu, ref in (("gamma", g.grad, ref_g), ("source", s.grad, ref_s), ("phib", b.grad, ref_b)):
        gpu = gpu.cpu().numpy()
        err = np.abs(gpu - ref).max() / np.abs(ref).max()
        print(f"{name}: max |gpu-cpu|/max|cpu| = {err:.2e}", flush=True)
        assert err < 1e-5, name
    ctx.close()


def test_adjoint_two_calls_before_backward():
    """Two forward calls with different face diffusivity before either
    backward: each backward must return the gradients of its own call (the
    taped gamma is re-assembled before the transpose), equal to a forward and
    backward of that call alone."""
    import torch
    from fcp.adjoint import ConvDiffSolve
    print("[case] adjoint two calls before backward 6^3", flush=True)
    mesh, ctx, geo, csr, flux, inp, prob, P, vol = _problem(6)
    gammas = [P["gamma"], P["gamma"] * (1.5 + 0.5 * torch.cos(torch.arange(
        P["gamma"].numel(), device=P["gamma"].device, dtype=torch.float64)))]

    alone = []
    for gm in gammas:
        g = gm.clone().requires_grad_(True)
        s = P["source"].clone().requires_grad_(True)
        b = P["phib"].clone().requires_grad_(True)
        (P["w"] * ConvDiffSolve.apply(g, s, b, prob) * vol).sum().backward()
        alone.append((g.grad.clone(), s.grad.clone(), b.grad.clone()))
    leaves = [[t.clone().requires_grad_(True) for t in (gm, P["source"], P["phib"])]
              for gm in gammas]
    Js = [(P["w"] * ConvDiffSolve.apply(*lv, prob) * vol).sum() for lv in leaves]
    # backward of the earlier call after the later call replaced the matrix
    for k in (0, 1):
        Js[k].backward()
        for t, ref in zip(leaves[k], alone[k]):
            assert torch.allclose(t.grad, ref, rtol=1e-9, atol=1e-12 * float(ref.abs().max())), k
    # the two calls' gradients differ (the test would not see a mix-up otherwise)
    assert not torch.allclose(alone[0][0], alone[1][0], rtol=1e-3)
    ctx.close()
