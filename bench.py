"""Benchmark: cells/s to 1e-8 relative residual, mixed vs FP64, on one GPU.

Workload (defaults): pressure-correction-style Poisson on 128^3 hexahedra with
interior nodes jittered +-20% h, face diffusivity 10^U(-1,1), zero-mean
source, Dirichlet on the x- patch and zero-gradient elsewhere.  One step =
Laplacian assembly with face diffusivity + Gauss gradient of the previous
iterate + explicit non-orthogonal correction + DIC-PCG solve from x0 = 0 to
||b - A x||_2 / ||b||_2 <= 1e-8 (FP64 true residual).  Inputs are resident
in HBM before the timed region; every step is a full assemble+solve.

value = cells / (seconds per step) of the mixed-precision path (FP32 inner
PCG inside FP64 iterative refinement); the plain FP64 path (the reference's
global DIC) and FP64 iterations with the mixed path's own two-level tile
preconditioner (fp64_tile, the like-for-like precision baseline) are timed
in the same run and reported beside it.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--size N]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "freecappuccino-dev_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def log(*a):
    print(*a, file=sys.stderr, flush=True)


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "sweeps_spmv_128cube_hbm_traffic.txt")


def measured_traffic(kernel, N):
    """HBM bytes per launch of `kernel` from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (calibrated against a copy of known size),
    which covers the 128^3 workload only; None otherwise."""
    if N != 128 or not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as f:
        for line in f:
            t = line.split()
            if len(t) >= 3 and t[0] == "traffic" and t[1] == kernel:
                return float(t[2])
    return None


def alg_bytes(name, n, nnz, Fi, Fb, local=None):
    """Algorithmic (compulsory) HBM bytes of one launch of a named kernel;
    `local`: strict-lower entries inside the tile-local preconditioner's tiles
    (inner faces whose two cells share a tile)."""
    lower = Fi  # strict-lower entries = inner faces
    tab = {
        "spmv_pq_f32": nnz * 8 + 2 * n * 4,
        "spmv_pq_f64": nnz * 12 + 2 * n * 8,
        "spmv_sell_f64": nnz * 12 + 2 * n * 8,
        "sweep_fwd_f32": lower * 8 + 4 * n + 3 * n * 4,
        "sweep_bwd_f32": lower * 8 + 4 * n + 3 * n * 4,
        "sweep_fwd_f64": lower * 12 + 4 * n + 3 * n * 8,
        "sweep_bwd_f64": lower * 12 + 4 * n + 3 * n * 8,
        "dic_setup": lower * 12 + 4 * n + 2 * n * 8,
        # inner PCG in tile order: per position x, p, r, q read and x, r, z
        # written, the pivot, 2 x (4-byte coefficient + 2-byte slot) per
        # in-tile dependency; SpMV values + columns, z and p read, p and q
        # written
        "tin_pre": 32 * n + 12 * local if local is not None else None,
        "tin_spmv": nnz * 8 + 4 * n * 4,
        # coarse correction: z read and written (the dense coarse rows stay
        # in L2)
        "tin_coarse": 2 * n * 4,
        "bj_sweep_fwd_f32": local * 8 + 4 * n + 3 * n * 4 if local is not None else None,
        "bj_sweep_bwd_f32": local * 8 + 4 * n + 3 * n * 4 if local is not None else None,
        "residual_f64": nnz * 12 + 3 * n * 8,
        "dot_rz": 2 * n * 8,
        "pupdate": 3 * n * 8,
        "xr_update": 6 * n * 8,
    }
    return tab.get(name)


def run_gpu(args):
    import torch

    from fcp import workloads
    from fcp.api import Context, FCP_CRIT_TRUE_L2, FCP_PREC_FP64, FCP_PREC_FP64_TILE, FCP_PREC_MIXED

    N = args.size
    t0 = time.time()
    mesh = workloads.poisson_jittered_mesh(N)
    log(f"[bench] mesh {N}^3: n={mesh.n_cells} Fi={mesh.n_inner} ({time.time()-t0:.1f}s)")
    ctx = Context(0)
    ctx.read_mesh(mesh)
    geo = ctx.geometry()
    ctx.create_CSR_matrix(want_arrays=False)
    n, nnz, Fi, Fb = mesh.n_cells, ctx.nnz, mesh.n_inner, mesh.n_bnd
    log(f"[bench] geometry+CSR+levels on GPU done ({time.time()-t0:.1f}s) nnz={nnz}")
    tiles = ctx.sweep_tiles()
    local = int(np.count_nonzero(tiles[mesh.owner[:Fi]] == tiles[mesh.neighbour]))
    inp = workloads.poisson_jittered_inputs(mesh, geo["vol"], geo["xc"], geo["yc"], geo["zc"])
    dev = torch.device("cuda:0")
    T = lambda a, dt=torch.float64: torch.from_numpy(np.ascontiguousarray(a)).to(dev, dt)  # noqa: E731
    mu_f = T(inp["mu_f"])
    gamma = T(inp["gamma"])
    bc_kind = T(inp["bc_kind"], torch.uint8)
    phi_bc = T(inp["phi_bc"])
    phi_prev = T(inp["phi_prev"])
    sV = T(inp["source"] * geo["vol"])
    su = torch.empty(n, dtype=torch.float64, device=dev)
    x = torch.empty(n, dtype=torch.float64, device=dev)
    grad = torch.empty((mesh.n_total, 3), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()

    def step(precision, tol=1e-8):
        su.copy_(sV)
        x.zero_()
        torch.cuda.synchronize()
        ctx.laplacian_face(mu_f, bc_kind, phi_bc, su)
        ctx.grad_gauss(phi_prev, grad)
        ctx.nonorth_correction(gamma, grad, su)
        rc, info = ctx.csrsolve("iccg", x, su, itr_max=20000, tol_abs=0.0, tol_rel=tol,
                                precision=precision, criterion=FCP_CRIT_TRUE_L2,
                                inner_tol=args.inner_tol, check_every=args.check_every)
        if rc != 0 or not info["rel_res_l2"] <= tol:
            raise RuntimeError(f"solve did not reach {tol:g}: rc={rc} info={info}")
        return info

    # fp64: the reference's global DIC; fp64_tile: FP64 iterations with the
    # mixed path's two-level tile preconditioner (the like-for-like baseline)
    precs = (("fp64", FCP_PREC_FP64), ("fp64_tile", FCP_PREC_FP64_TILE), ("mixed", FCP_PREC_MIXED))
    try:  # the FP64 tile kernels need the tiles' FP64 operands within LDS
        step(FCP_PREC_FP64_TILE)
    except Exception as e:  # noqa: BLE001
        log(f"[bench] fp64_tile not run: {e}")
        precs = tuple(p for p in precs if p[0] != "fp64_tile")
    results, xs = {}, {"fp64_tile": None}
    for name, prec in precs:
        for _ in range(args.warmup):
            info = step(prec)
            log(f"[bench] {name} warmup done: iters={info['iterations']} solve {info['time_ms']:.1f} ms")
        torch.cuda.synchronize()
        ts = time.perf_counter()
        infos = []
        for k in range(args.steps):
            infos.append(step(prec))
            log(f"[bench] {name} step {k} done: iters={infos[-1]['iterations']} "
                f"solve {infos[-1]['time_ms']:.1f} ms")
        torch.cuda.synchronize()
        te = time.perf_counter()
        ms = (te - ts) * 1e3 / args.steps
        results[name] = dict(ms_per_step=ms, cells_per_s=n / (ms / 1e3),
                             iterations=infos[-1]["iterations"],
                             outer_iterations=infos[-1]["outer_iterations"],
                             rel_res_l2=infos[-1]["rel_res_l2"], solve_ms=infos[-1]["time_ms"])
        log(f"[bench] {name}: {ms:.2f} ms/step, {n/(ms/1e3)/1e6:.2f} Mcells/s, "
            f"iters={infos[-1]['iterations']} outer={infos[-1]['outer_iterations']} "
            f"relres={infos[-1]['rel_res_l2']:.2e}")
        xs[name] = x.cpu().numpy().copy()

    # error of the mixed solution against a tight FP64 GPU solve
    su.copy_(sV)
    torch.cuda.synchronize()
    ctx.laplacian_face(mu_f, bc_kind, phi_bc, su)
    ctx.grad_gauss(phi_prev, grad)
    ctx.nonorth_correction(gamma, grad, su)
    x_ref = torch.zeros(n, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    ctx.csrsolve("iccg", x_ref, su, itr_max=50000, tol_rel=1e-13, precision=FCP_PREC_FP64,
                 criterion=FCP_CRIT_TRUE_L2)
    xr = x_ref.cpu().numpy()
    nxr = np.linalg.norm(xr)
    errs = {k: (float(np.linalg.norm(v - xr) / nxr) if v is not None else None)
            for k, v in xs.items()}
    log("[bench] error vs the FP64 solve to 1e-13: " +
        " ".join(f"{k} {v:.2e}" for k, v in errs.items() if v is not None))

    # the same step solved to ||b-Ax||/||b|| <= 1e-12
    tight = {}
    for name, prec in precs:
        step(prec, 1e-12)
        torch.cuda.synchronize()
        ts = time.perf_counter()
        for _ in range(args.tight_steps):
            info = step(prec, 1e-12)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - ts) * 1e3 / args.tight_steps
        err = float(np.linalg.norm(x.cpu().numpy() - xr) / nxr)
        tight[name] = dict(ms_per_step=ms, cells_per_s=n / (ms / 1e3), iterations=info["iterations"],
                           outer_iterations=info["outer_iterations"],
                           rel_res_l2=info["rel_res_l2"], err_vs_fp64_1e13=err)
        log(f"[bench] {name} to 1e-12: {ms:.2f} ms/step iters={info['iterations']} "
            f"relres={info['rel_res_l2']:.2e} err={err:.2e}")
    tight["mixed_over_fp64"] = tight["fp64"]["ms_per_step"] / tight["mixed"]["ms_per_step"]
    tight["mixed_over_fp64_same_precond"] = (
        tight["fp64_tile"]["ms_per_step"] / tight["mixed"]["ms_per_step"]
        if "fp64_tile" in tight else None)

    # per-kernel device time (HIP events on the library stream) over profiled steps
    ctx.profile(True)
    its = outer = 0
    for _ in range(args.prof_steps):
        info = step(FCP_PREC_MIXED)
        its += info["iterations"]
        outer += info["outer_iterations"]
    prof = ctx.profile_report()
    ctx.profile(False)
    # Launches that did work: every inner solve applies the preconditioner once
    # before its first iteration, then once per iteration (tk_pre, tc_zc); the
    # SpMV runs once per iteration.  Launches queued after the device-side
    # stopping test exit at once; their few microseconds stay in the time
    # divided by the working count (a conservative per-launch figure).
    working = {"tin_pre": its + outer, "tin_coarse": its + outer, "tin_spmv": its}
    total = sum(v[1] for v in prof.values())
    log("[bench] kernel profile (mixed, %d steps): " % args.prof_steps +
        ", ".join(f"{k}={v[1]/args.prof_steps:.2f}ms/{v[0]//args.prof_steps}x" for k, v in
                  sorted(prof.items(), key=lambda kv: -kv[1][1])[:8]))
    kernels = {}
    for k, (cnt, tms) in prof.items():
        b = alg_bytes(k, n, nnz, Fi, Fb, local)
        wk = min(cnt, working.get(k, cnt))
        avg_us = tms / wk * 1e3
        kernels[k] = dict(launches_per_step=cnt / args.prof_steps,
                          working_launches_per_step=wk / args.prof_steps, avg_us=avg_us,
                          share=tms / total if total else 0.0)
        if b:
            kernels[k]["GBs"] = b / (avg_us * 1e-6) / 1e9
    name, (cnt, tms) = max(prof.items(), key=lambda kv: kv[1][1])
    wk = min(cnt, working.get(name, cnt))
    avg_s = tms / wk / 1e3
    b = alg_bytes(name, n, nnz, Fi, Fb, local)
    roof = dict(bound="hbm", kernel=name, achieved=(b / avg_s / 1e9) if b else None,
                peak=HBM_PEAK_GBS, unit="GB/s", frac=(b / avg_s / 1e9 / HBM_PEAK_GBS) if b else None,
                traffic=measured_traffic(name, N), alg_bytes_per_launch=b,
                avg_launch_us=avg_s * 1e6, launches=cnt, working_launches=wk)

    # standalone SpMV FP64 (reference row order) bandwidth; the 128^3 operator
    # and vectors (209 MB) fit in the 256 MB Infinity Cache
    spmv = spmv_rate(ctx, n, nnz, Fi, Fb)
    if spmv:
        spmv["note"] = "operator + vectors fit in the 256 MB Infinity Cache"
    ctx.close()
    out = dict(n=n, nnz=nnz, Fi=Fi, results=results, errs=errs, tight=tight,
               roofline=roof, kernels=kernels, spmv=spmv)
    if args.spmv_size:
        out["spmv_big"] = run_spmv_big(args)
    return out


def spmv_rate(ctx, n, nnz, Fi, Fb, reps=20):
    """Average device time and algorithmic bandwidth of the FP64 SELL SpMV
    (reference row order), two input vectors alternated."""
    import torch
    dev = torch.device("cuda:0")
    xs = [torch.rand(n, dtype=torch.float64, device=dev) for _ in range(2)]
    ys = torch.empty(n, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    ctx.profile(True)
    for k in range(reps):
        ctx.spmv(xs[k & 1], ys)
    sp = ctx.profile_report().get("spmv_sell_f64")
    ctx.profile(False)
    if not sp:
        return None
    avg = sp[1] / sp[0] / 1e3
    bb = alg_bytes("spmv_sell_f64", n, nnz, Fi, Fb)
    return dict(kernel="spmv_sell_f64", avg_us=avg * 1e6, GBs=bb / avg / 1e9,
                frac=bb / avg / 1e9 / HBM_PEAK_GBS, alg_bytes=bb, launches=sp[0])


def run_spmv_big(args):
    """The FP64 SpMV on the same kind of operator at a size whose operator
    alone exceeds the Infinity Cache (192^3: 733 MB of algorithmic bytes)."""
    import torch

    from fcp import workloads
    from fcp.api import Context

    N = args.spmv_size
    mesh = workloads.poisson_jittered_mesh(N)
    ctx = Context(0)
    ctx.read_mesh(mesh)
    geo = ctx.geometry()
    ctx.create_CSR_matrix(want_arrays=False)
    inp = workloads.poisson_jittered_inputs(mesh, geo["vol"], geo["xc"], geo["yc"], geo["zc"])
    dev = torch.device("cuda:0")
    T = lambda a, dt=torch.float64: torch.from_numpy(np.ascontiguousarray(a)).to(dev, dt)  # noqa: E731
    su = T(inp["source"] * geo["vol"])
    ctx.laplacian_face(T(inp["mu_f"]), T(inp["bc_kind"], torch.uint8), T(inp["phi_bc"]), su)
    out = spmv_rate(ctx, mesh.n_cells, ctx.nnz, mesh.n_inner, mesh.n_bnd)
    if out:
        out.update(cells=mesh.n_cells, nnz=ctx.nnz, size=f"{N}^3")
        log(f"[bench] SpMV FP64 {N}^3: {out['avg_us']:.1f} us, {out['GBs']:.0f} GB/s "
            f"({100 * out['frac']:.1f}% of peak)")
    ctx.close()
    return out


def _timed(fn, reps=1):
    import torch
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        out = fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t) / reps, out


def run_convdiff(args):
    """Momentum-like convection-diffusion, three components sharing one
    matrix: Gauss gradients + facefluxsc assembly (linear upwind) for u, v, w,
    then one batched 3-column DILU-BiCGStab solve to ||b-Ax||/||b|| <= 1e-8
    per column (FP64, and FP32 inner + FP64 refinement); the same columns
    solved one at a time for comparison."""
    import torch

    from fcp import workloads
    from fcp.api import Context, FCP_CRIT_TRUE_L2, FCP_PREC_FP64, FCP_PREC_MIXED

    N = args.size
    t0 = time.time()
    mesh = workloads.convdiff_mesh(N)
    ctx = Context(0)
    ctx.read_mesh(mesh)
    geo = ctx.geometry()
    csr = ctx.create_CSR_matrix()
    log(f"[convdiff] mesh+geometry+CSR {N}^3 n={mesh.n_cells} ({time.time()-t0:.1f}s)")
    flux, div = workloads.convdiff_flux(mesh, geo, csr, ctx)
    inp = workloads.convdiff_inputs(mesh, geo, flux)
    log(f"[convdiff] projected flux, max|div|={np.abs(div).max():.1e} ({time.time()-t0:.1f}s)")
    dev = torch.device("cuda:0")
    T = lambda a, dt=torch.float64: torch.from_numpy(np.ascontiguousarray(a)).to(dev, dt)  # noqa: E731
    n, nt = mesh.n_cells, mesh.n_total
    F_, G_ = T(flux), T(inp["gamma"])
    bk = T(inp["bc_kind"], torch.uint8)
    phi0 = T(inp["phi"])
    sV = T(inp["source"] * geo["vol"][None, :])
    phi = torch.empty_like(phi0)
    grad = torch.empty((nt, 3), dtype=torch.float64, device=dev)
    rhs = torch.empty((3, n), dtype=torch.float64, device=dev)
    sp = torch.empty(n, dtype=torch.float64, device=dev)
    X = torch.zeros((3, n), dtype=torch.float64, device=dev)

    def assemble():
        phi.copy_(phi0)
        rhs.copy_(sV)
        for r in range(3):
            sp.zero_()
            ctx.grad_gauss(phi[r], grad)
            ctx.convdiff(F_, G_, phi[r], grad, 1.0, 1, bk, rhs[r], sp, urf=inp["urf"])

    out = {}
    for name, prec in (("fp64", FCP_PREC_FP64), ("mixed", FCP_PREC_MIXED)):
        def step():
            assemble()
            X.zero_()
            rc, info, _ = ctx.csrsolve_multi("bicgstab", X, rhs, itr_max=20000, tol_rel=1e-8,
                                             precision=prec, criterion=FCP_CRIT_TRUE_L2,
                                             inner_tol=args.inner_tol, check_every=args.check_every)
            if rc != 0:
                raise RuntimeError(f"batched solve failed: {info}")
            return info
        step()
        dt, info = _timed(step, args.steps)
        out[name] = dict(ms_per_step=dt * 1e3, cells_per_s=n / dt, iterations=info["iterations"],
                         outer=info["outer_iterations"], worst_rel_res=info["rel_res_l2"])
        log(f"[convdiff] batched {name}: {dt*1e3:.1f} ms/step  iters={info['iterations']} "
            f"rel={info['rel_res_l2']:.1e}")
        # the same three columns one at a time
        def seq():
            assemble()
            it = 0
            for r in range(3):
                x = torch.zeros(n, dtype=torch.float64, device=dev)
                rc, inf1 = ctx.csrsolve("bicgstab", x, rhs[r].clone(), itr_max=20000, tol_rel=1e-8,
                                        precision=prec, criterion=FCP_CRIT_TRUE_L2,
                                        inner_tol=args.inner_tol, check_every=args.check_every)
                it = max(it, inf1["iterations"])
            return it
        dts, it = _timed(seq, 1)
        out[name]["sequential_ms"] = dts * 1e3
        out[name]["batched_speedup"] = dts / dt
        log(f"[convdiff] sequential {name}: {dts*1e3:.1f} ms  (batched speedup {dts/dt:.2f}x)")
    ctx.close()
    return {"metric": "cells/s to 1e-8 rel-res (3 components, batched DILU-BiCGStab)",
            "value": out["mixed"]["cells_per_s"], "unit": "cells/s", "n_gpus": 1,
            "steps": args.steps, "higher_is_better": True, "dtype": "fp32+fp64",
            "data": "synthetic",
            "config": {"workload": f"momentum-like convection-diffusion {N}^3 hex, 8 Fourier modes "
                                   "projected divergence-free, Pe 0.1..100, linear upwind, "
                                   "3 RHS sharing A", "cells": n},
            "fp64": out["fp64"], "mixed": out["mixed"],
            "mixed_over_fp64": out["fp64"]["ms_per_step"] / out["mixed"]["ms_per_step"]}


def run_tets(args):
    """Tetrahedra with scrambled cell order: GPU reverse Cuthill-McKee, the
    renumbered mesh's gradients (Gauss, lsq, lsq_dm, lsq_qr) of the analytic
    field against its exact gradient, and an FP64 Poisson solve."""
    import torch

    from fcp import workloads
    from fcp.api import Context, FCP_CRIT_TRUE_L2, FCP_PREC_FP64, FCP_PREC_MIXED
    from fcp.mesh import permute_cells

    N = args.size
    t0 = time.time()
    mesh0 = workloads.tet_mesh_workload(N)
    log(f"[tets] mesh {N}^3x6 n={mesh0.n_cells} ({time.time()-t0:.1f}s)")
    ctx = Context(0)
    ctx.read_mesh(mesh0)
    ctx.compute_geometry()
    ctx.create_CSR_matrix(want_arrays=False)
    dt_rcm, perm = _timed(lambda: ctx.rcm(4))
    own, nb = mesh0.owner[:mesh0.n_inner].astype(np.int64), mesh0.neighbour.astype(np.int64)
    bw0 = int(np.abs(own - nb).max())
    bw1 = int(np.abs(perm[own] - perm[nb]).max())
    log(f"[tets] rcm {dt_rcm*1e3:.1f} ms, bandwidth {bw0} -> {bw1}")
    ctx.close()
    mesh = permute_cells(mesh0, perm)
    del mesh0
    ctx = Context(0)
    ctx.read_mesh(mesh)
    geo = ctx.geometry()
    ctx.create_CSR_matrix(want_arrays=False)
    n, nt, Fi = mesh.n_cells, mesh.n_total, mesh.n_inner
    inp = workloads.tet_inputs(mesh, geo)
    dev = torch.device("cuda:0")
    f = torch.from_numpy(inp["f"]).to(dev)
    gout = torch.empty((nt, 3), dtype=torch.float64, device=dev)
    exact = workloads.analytic_grad(geo["xc"], geo["yc"], geo["zc"])
    grads = {}
    for method, name in ((0, "gauss"), (1, "lsq"), (2, "lsq_dm"), (3, "lsq_qr")):
        if method:
            ctx.create_lsq_grad_matrix(method, compat=False)
        dtg, _ = _timed(lambda: ctx.grad(method, f, gout), 5)
        g = gout[:n].cpu().numpy()
        err = float(np.sqrt(np.mean(np.sum((g - exact) ** 2, axis=1))) /
                    np.sqrt(np.mean(np.sum(exact ** 2, axis=1))))
        grads[name] = dict(ms=dtg * 1e3, rel_rms_error=err)
        log(f"[tets] grad {name}: {dtg*1e3:.2f} ms, rel rms error vs analytic {err:.2e}")
    T = lambda a, dt=torch.float64: torch.from_numpy(np.ascontiguousarray(a)).to(dev, dt)  # noqa: E731
    mu_f, bk, pb = T(inp["mu_f"]), T(inp["bc_kind"], torch.uint8), T(inp["phi_bc"])
    sV = T(inp["source"] * geo["vol"])
    su = torch.empty(n, dtype=torch.float64, device=dev)
    x = torch.zeros(n, dtype=torch.float64, device=dev)
    # mesh setup: tile size of the FP32 inner solves' preconditioner
    su.copy_(sV)
    ctx.laplacian_face(mu_f, bk, pb, su)
    t0 = time.time()
    tile, _ = ctx.autotune("iccg")
    log(f"[tets] autotune: inner tiles of {tile} cells ({time.time()-t0:.1f}s)")
    res = {}
    for name, prec in (("fp64", FCP_PREC_FP64), ("mixed", FCP_PREC_MIXED)):
        def step():
            su.copy_(sV)
            ctx.laplacian_face(mu_f, bk, pb, su)
            x.zero_()
            rc, info = ctx.csrsolve("iccg", x, su, itr_max=50000, tol_rel=1e-8, precision=prec,
                                    criterion=FCP_CRIT_TRUE_L2, inner_tol=args.inner_tol,
                                    check_every=args.check_every)
            return info
        step()
        dts, info = _timed(step, args.steps)
        res[name] = dict(ms_per_step=dts * 1e3, cells_per_s=n / dts, iterations=info["iterations"])
        log(f"[tets] poisson {name}: {dts*1e3:.1f} ms/step iters={info['iterations']}")
    ctx.close()
    return {"metric": "tets: RCM ms, gradient errors, Poisson cells/s to 1e-8",
            "value": res["mixed"]["cells_per_s"], "unit": "cells/s", "n_gpus": 1,
            "higher_is_better": True, "dtype": "fp32+fp64", "data": "synthetic",
            "config": {"workload": f"{N}^3x6 Kuhn tets jittered +-15%, cells permuted (seed 3), "
                                   "GPU RCM, Gamma 10^U(-2,2)", "cells": n},
            "rcm_ms": dt_rcm * 1e3, "bandwidth_before": bw0, "bandwidth_after": bw1,
            "gradients": grads, "poisson": res, "inner_tile": tile}


def run_adjoint(args):
    """Sensitivities of J = sum w phi V on the jittered convection-diffusion
    problem: primal (2 deferred-correction passes) + adjoint on the GPU, and
    central finite differences (step 1e-6) along random directions."""
    import torch

    from fcp import workloads
    from fcp.adjoint import ConvDiffProblem, ConvDiffSolve
    from fcp.api import Context

    N = args.size
    mesh = workloads.convdiff_mesh(N, jitter=0.2)
    ctx = Context(0)
    ctx.read_mesh(mesh)
    geo = ctx.geometry()
    csr = ctx.create_CSR_matrix()
    flux, _ = workloads.convdiff_flux(mesh, geo, csr, ctx)
    inp = workloads.adjoint_inputs(mesh, geo, flux)
    dev = torch.device("cuda:0")
    prob = ConvDiffProblem(ctx, mesh, geo, flux, inp["bc_kind"], K=inp["K"],
                           log=lambda line: log(f"[adjoint] {line}"))
    P = {k: torch.from_numpy(np.ascontiguousarray(inp[k])).to(dev) for k in ("gamma", "source", "phib", "w")}
    vol = torch.from_numpy(geo["vol"]).to(dev)
    g = P["gamma"].clone().requires_grad_(True)
    s = P["source"].clone().requires_grad_(True)
    b = P["phib"].clone().requires_grad_(True)
    torch.cuda.synchronize()
    t = time.perf_counter()
    phi = ConvDiffSolve.apply(g, s, b, prob)
    J = (P["w"] * phi * vol).sum()
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t
    t = time.perf_counter()
    J.backward()
    torch.cuda.synchronize()
    t_bwd = time.perf_counter() - t
    log(f"[adjoint] forward {t_fwd*1e3:.0f} ms, backward {t_bwd*1e3:.0f} ms")
    rng = np.random.default_rng(4)
    dirichlet = torch.from_numpy(inp["bc_kind"] == 0).to(dev)
    h = 1e-6
    fds, ads = [], []
    for k in range(args.fd_dirs):
        vg = torch.from_numpy(rng.uniform(-1, 1, g.numel())).to(dev) * P["gamma"]
        vs = torch.from_numpy(rng.uniform(-1, 1, s.numel())).to(dev)
        vb = torch.from_numpy(rng.uniform(-1, 1, b.numel())).to(dev) * dirichlet
        Jp = float((P["w"] * prob.forward(P["gamma"] + h * vg, P["source"] + h * vs,
                                          P["phib"] + h * vb) * vol).sum())
        Jm = float((P["w"] * prob.forward(P["gamma"] - h * vg, P["source"] - h * vs,
                                          P["phib"] - h * vb) * vol).sum())
        fd = (Jp - Jm) / (2 * h)
        ad = float((g.grad * vg).sum() + (s.grad * vs).sum() + (b.grad * vb).sum())
        fds.append(fd)
        ads.append(ad)
        log(f"[adjoint] direction {k}: fd {fd:.10e} adjoint {ad:.10e} rel {abs(fd-ad)/abs(ad):.1e}")
    fds, ads = np.array(fds), np.array(ads)
    worst = float(np.abs(fds - ads).max() / np.abs(ads).max())
    worst_dir = float(np.max(np.abs(fds - ads) / np.abs(ads)))
    ctx.close()
    return {"metric": "adjoint gradient time and FD agreement", "value": (t_fwd + t_bwd) * 1e3,
            "unit": "ms (primal + adjoint)", "n_gpus": 1, "higher_is_better": False,
            "dtype": "fp64", "data": "synthetic",
            "config": {"workload": f"{N}^3 jittered convection-diffusion, J = sum w phi V, "
                                   f"K={inp['K']} deferred-correction passes", "cells": mesh.n_cells},
            "forward_ms": t_fwd * 1e3, "backward_ms": t_bwd * 1e3,
            "fd_directions": args.fd_dirs, "fd_step": h, "max_fd_mismatch_rel_to_max": worst,
            "worst_per_direction_rel": worst_dir}


def run_probe(args):
    """A short fixed kernel sequence for hardware-counter runs: a calibration
    copy of known size (256 MiB read + 256 MiB written), one assembly of the
    benchmark step (face Laplacian, Gauss gradient, non-orthogonal
    correction), then 10 mixed and 10 FP64 PCG iterations and 5 SELL SpMVs
    on the benchmark system."""
    import torch

    from fcp import workloads
    from fcp.api import Context, FCP_CRIT_TRUE_L2, FCP_PREC_FP64, FCP_PREC_MIXED

    N = args.size
    mesh = workloads.poisson_jittered_mesh(N)
    ctx = Context(0)
    ctx.read_mesh(mesh)
    geo = ctx.geometry()
    ctx.create_CSR_matrix(want_arrays=False)
    inp = workloads.poisson_jittered_inputs(mesh, geo["vol"], geo["xc"], geo["yc"], geo["zc"])
    dev = torch.device("cuda:0")
    T = lambda a, dt=torch.float64: torch.from_numpy(np.ascontiguousarray(a)).to(dev, dt)  # noqa: E731
    su = T(inp["source"] * geo["vol"])
    grad = torch.empty((mesh.n_total, 3), dtype=torch.float64, device=dev)
    src = torch.ones(64 << 20, dtype=torch.float32, device=dev)  # 256 MiB
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    ctx.laplacian_face(T(inp["mu_f"]), T(inp["bc_kind"], torch.uint8), T(inp["phi_bc"]), su)
    ctx.grad_gauss(T(inp["phi_prev"]), grad)
    ctx.nonorth_correction(T(inp["gamma"]), grad, su)
    torch.cuda.synchronize()
    x = torch.zeros(mesh.n_cells, dtype=torch.float64, device=dev)
    y = torch.empty_like(x)
    for prec in (FCP_PREC_MIXED, FCP_PREC_FP64):  # 10 PCG iterations each
        x.zero_()
        ctx.csrsolve("iccg", x, su, itr_max=10, tol_abs=0.0, tol_rel=1e-30, precision=prec,
                     criterion=FCP_CRIT_TRUE_L2, max_outer=1, check_every=10)
    for _ in range(5):
        ctx.spmv(x, y)
    torch.cuda.synchronize()
    ctx.close()
    log(f"[probe] done: n={mesh.n_cells} copy=256MiB")


def run_cpu_baseline(args, n_gpu_cells):
    """The oracle (serial C restatement of the reference path) on the same
    workload at a bounded size, 1 host core."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from fcp import workloads
    from oracle import Oracle

    N = args.cpu_size
    mesh = workloads.poisson_jittered_mesh(N)
    o = Oracle(mesh)
    g = o.geom
    inp = workloads.poisson_jittered_inputs(mesh, g["vol"], g["xc"], g["yc"], g["zc"])
    t0 = time.perf_counter()
    a, su = o.laplacian_face(inp["mu_f"], inp["bc_kind"], inp["phi_bc"], inp["source"] * g["vol"])
    gr = o.grad_gauss(inp["phi_prev"])
    su = o.nonorth_correction(inp["gamma"], gr, su)
    # the reference's DIC-PCG iterations, stopped on the GPU legs' test
    # ||b - A x||_2 <= 1e-8 ||b||_2 (x0 = 0) through the recursive residual
    out = o.solve("iccg-l2", a, su, itr_max=20000, tol_abs=0.0, tol_rel=1e-8)
    t1 = time.perf_counter()
    rel = float(np.linalg.norm(su - o.spmv(a, out["fi"])) / np.linalg.norm(su))
    return dict(value=mesh.n_cells / (t1 - t0), unit="cells/s", cores=1, kind="port",
                sample=f"same workload at {N}^3 ({mesh.n_cells} cells): oracle laplacian_face + "
                       f"grad_gauss + non-orthogonal correction + ICCG stopped at "
                       f"||r||_2/||b||_2 <= 1e-8 (true ||b-Ax||/||b|| = {rel:.1e}) in "
                       f"{t1-t0:.1f} s, {out['iters']} iterations, one thread",
                host=host_cpu())


def host_cpu():
    """CPU model and logical CPU count of this host."""
    model = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "nproc": os.cpu_count(),
            "affinity": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=None,
                    help="cells per edge (default: 128, tets 96, convdiff 192)")
    ap.add_argument("--cpu-size", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--prof-steps", type=int, default=1)
    ap.add_argument("--inner-tol", type=float, default=1e-4)
    ap.add_argument("--check-every", type=int, default=16)
    ap.add_argument("--probe", action="store_true",
                    help="run only the fixed kernel sequence for counter collection")
    ap.add_argument("--workload", default="poisson",
                    choices=["poisson", "convdiff", "tets", "adjoint"],
                    help="poisson (default, the headline) or one of the secondary workloads")
    ap.add_argument("--fd-dirs", type=int, default=8)
    ap.add_argument("--tight-steps", type=int, default=2,
                    help="timed steps of the 1e-12 measurement point")
    ap.add_argument("--spmv-size", type=int, default=192,
                    help="mesh size of the SpMV measured beyond the 256 MB Infinity Cache (0: skip)")
    args = ap.parse_args()
    if args.size is None:
        args.size = {"tets": 96, "convdiff": 192}.get(args.workload, 128)
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only; multi-GPU runs are not part of it "
              f"(asked for --gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)
    if args.probe:
        run_probe(args)
        return
    if args.workload != "poisson":
        fn = {"convdiff": run_convdiff, "tets": run_tets, "adjoint": run_adjoint}[args.workload]
        print(json.dumps(fn(args)))
        return
    g = run_gpu(args)
    cpu = None if args.no_cpu else run_cpu_baseline(args, g["n"])
    mixed, fp64 = g["results"]["mixed"], g["results"]["fp64"]
    f64t = g["results"].get("fp64_tile") or {k: None for k in ("cells_per_s", "ms_per_step",
                                                                 "iterations", "outer_iterations",
                                                                 "rel_res_l2")}
    line = {
        "metric": "cells/s to 1e-8 rel-res: mixed vs FP64 on GPU; SpMV HBM GB/s % of peak",
        "value": mixed["cells_per_s"],
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64",
        "data": "synthetic",
        "config": {"workload": f"pressure-correction Poisson, {args.size}^3 hex jittered +-20%, "
                               "face Gamma 10^U(-1,1), Dirichlet x-, Gauss grad + non-orth "
                               "correction + DIC-PCG to ||b-Ax||/||b||<=1e-8",
                   "cells": g["n"], "nnz": g["nnz"], "inner_faces": g["Fi"]},
        "fp64": {"cells_per_s": fp64["cells_per_s"], "ms_per_step": fp64["ms_per_step"],
                 "iterations": fp64["iterations"], "rel_res_l2": fp64["rel_res_l2"],
                 "err_vs_fp64_1e-13": g["errs"]["fp64"]},
        "mixed": {"iterations": mixed["iterations"], "outer_iterations": mixed["outer_iterations"],
                  "rel_res_l2": mixed["rel_res_l2"],
                  "err_vs_fp64_1e-13": g["errs"]["mixed"]},
        "mixed_over_fp64": fp64["ms_per_step"] / mixed["ms_per_step"],
        "fp64_tile": {"cells_per_s": f64t["cells_per_s"], "ms_per_step": f64t["ms_per_step"],
                      "iterations": f64t["iterations"],
                      "outer_iterations": f64t["outer_iterations"],
                      "rel_res_l2": f64t["rel_res_l2"],
                      "err_vs_fp64_1e-13": g["errs"]["fp64_tile"]},
        "mixed_over_fp64_same_precond": (f64t["ms_per_step"] / mixed["ms_per_step"]
                                         if f64t["ms_per_step"] else None),
        "tol_1e-12": g["tight"],
        "spmv": g["spmv"],
        "spmv_beyond_mall": g.get("spmv_big"),
        "roofline": g["roofline"],
        "cpu_baseline": cpu,
        "kernels": g["kernels"],
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
