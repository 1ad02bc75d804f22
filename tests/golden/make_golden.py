"""Writes the golden fixtures in tests/golden/ from the flang-built reference.

Run in the build container only (needs /root/reference and flang):
    python tests/golden/make_golden.py [case ...]     (default: every case)
The reference routines are compiled from the reference sources by
oracle/ref/Makefile into oracle/_ref/ and driven by oracle/ref/ref_driver.f90.
Fixtures hold inputs and the reference's outputs; large integer arrays are
stored as SHA-256 digests of their exact bytes.
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, os.path.join(ROOT, "freecappuccino-dev_amd"))
sys.path.insert(0, HERE)

from fcp.mesh import hex_mesh, read_foam  # noqa: E402
from fcp.rng import uniform  # noqa: E402
from fcp import workloads  # noqa: E402
from refdriver import RefCase  # noqa: E402


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def small_case(mesh, tag):
    """Geometry, CSR, operators and solver trajectories on a small mesh."""
    rc = RefCase(mesh)
    g, c, sizes = rc.dump()
    n, nt, F = mesh.n_cells, mesh.n_total, mesh.n_faces
    out = {f"g_{k}": v for k, v in g.items()}
    out.update({f"c_{k}": v for k, v in c.items()})
    out["sizes"] = np.array(sizes, np.int64)
    # inputs (seed 5 streams 1..9)
    mu = -(10.0 ** uniform(5, 1, n, -1.0, 1.0))
    phi = np.zeros(nt)
    phi[n:] = uniform(5, 2, nt - n, -1.0, 1.0)
    su0 = uniform(5, 3, n, -1.0, 1.0)
    u = uniform(5, 4, nt, -1.0, 1.0)
    b = uniform(5, 5, n, -1.0, 1.0)
    out.update(in_mu=mu, in_phi=phi, in_su0=su0, in_u=u, in_b=b)
    for k, v in (("mu", mu), ("phi", phi), ("su0", su0), ("u", u), ("b", b), ("fi0", np.zeros(n))):
        rc.put(k + ".bin", v)
    rc.run(["laplacian mu.bin phi.bin su0.bin a.bin su.bin", "grad_gauss u.bin gg.bin"])
    out["lap_a"] = rc.get("a.bin")
    out["lap_su"] = rc.get("su.bin")
    out["grad_gauss"] = rc.get("gg.bin").reshape(-1, 3)
    for name in ("lsq", "lsq_dm", "lsq_qr"):
        rc.run([f"grad_lsq {name} u.bin gl.bin"])
        out[f"grad_{name}"] = rc.get("gl.bin").reshape(-1, 3)
    # solver trajectories on the assembled SPD matrix
    for solver in ("iccg", "dpcg", "bicgstab", "gauss-seidel"):
        for it in (1, 2, 5, 20, 2000):
            rc.run([f"solve {solver} {it} 0.0 1.0d-10 a.bin fi0.bin b.bin fi.bin info.txt"])
            key = solver.replace("-", "_")
            out[f"sol_{key}_{it}"] = rc.get("fi.bin")
            out[f"sol_{key}_{it}_resor"] = np.array(
                float(open(os.path.join(rc.dir, "info.txt")).read()))
    # convection-diffusion assembly (both schemes, all three boundary kinds)
    flm = uniform(5, 6, F, -1.0, 1.0)
    dco = 10.0 ** uniform(5, 7, F, -1.0, 1.0)
    gr = uniform(5, 8, 3 * nt, -1.0, 1.0)
    bk = (np.arange(F - mesh.n_inner) % 3).astype(np.float64)
    out.update(in_flmass=flm, in_dcoef=dco, in_grad=gr.reshape(-1, 3), in_bckind=bk.astype(np.uint8))
    for k, v in (("flm", flm), ("dco", dco), ("gr", gr), ("bk", bk)):
        rc.put(k + ".bin", v)
    for scheme in ("cds", "linearUpwind"):
        rc.run([f"convdiff flm.bin dco.bin u.bin gr.bin 1.0d0 {scheme} bk.bin ca.bin csu.bin csp.bin"])
        out[f"cd_{scheme}_a"] = rc.get("ca.bin")
        out[f"cd_{scheme}_su"] = rc.get("csu.bin")
        out[f"cd_{scheme}_sp"] = rc.get("csp.bin")
    rc.close()
    np.savez_compressed(os.path.join(HERE, f"{tag}.npz"), **out)
    print("wrote", tag)


def bc_case():
    """Boundary handling and the explicit non-orthogonal correction on a
    jittered 5^3 hex whose six patches carry six different boundary types:
    updateBoundary.f90 on a random field; the inner-face non-orthogonal
    correction taken from facefluxsc with zero mass flux; and the
    calculate_temperature-style assembly with zero mass flux, Dirichlet on the
    inlet and wall patches and zero-gradient elsewhere (the face-diffusivity
    Laplacian with a Neumann mask, up to sign)."""
    types = ("inlet", "outlet", "wall", "symmetry", "pressure", "empty")
    mesh = hex_mesh(5, jitter=0.2, seed=7, types=types)
    rc = RefCase(mesh)
    g, c, sizes = rc.dump()
    n, nt, F, Fi = mesh.n_cells, mesh.n_total, mesh.n_faces, mesh.n_inner
    out = {f"g_{k}": v for k, v in g.items()}
    out.update({f"c_{k}": v for k, v in c.items()})
    out["sizes"] = np.array(sizes, np.int64)
    gam = 10.0 ** uniform(6, 1, F, -1.0, 1.0)
    gr = uniform(6, 2, 3 * nt, -1.0, 1.0)
    su0 = uniform(6, 3, n, -1.0, 1.0)
    phi = uniform(6, 4, nt, -1.0, 1.0)
    patch = mesh.patch_of_bface()
    bk = np.array([0 if types[p] in ("inlet", "wall") else 1 for p in patch], np.float64)
    out.update(in_gamma=gam, in_grad=gr.reshape(-1, 3), in_su0=su0, in_phi=phi,
               in_bckind=bk.astype(np.uint8))
    for k, v in (("gam", gam), ("gr", gr), ("su0", su0), ("phi", phi), ("bk", bk),
                 ("zero", np.zeros(F))):
        rc.put(k + ".bin", v)
    rc.run(["nonorth gam.bin gr.bin su0.bin no_su.bin",
            "update_boundary phi.bin ub.bin",
            "convdiff zero.bin gam.bin phi.bin gr.bin 1.0d0 cds bk.bin ca.bin csu.bin csp.bin"])
    out["nonorth_su"] = rc.get("no_su.bin")
    out["update_boundary"] = rc.get("ub.bin")
    out["cd0_a"] = rc.get("ca.bin")
    out["cd0_su"] = rc.get("csu.bin")
    out["cd0_sp"] = rc.get("csp.bin")
    rc.close()
    np.savez_compressed(os.path.join(HERE, "hexbc.npz"), **out)
    print("wrote hexbc")


TRANSPORT_MESHES = {
    # adiabatic wall before an isothermal one: the reference's srdw counter
    # then reads the adiabatic faces' |S|/dn for the isothermal ones
    "A": (("inlet", "outlet", "wall", "wall", "symmetry", "pressure"), (0, 0, 0, 1, 0, 0)),
    "B": (("wall", "wall", "wallQFlux", "outlet", "inlet", "empty"), (1, 0, 0, 0, 0, 0)),
}
# (mesh, time scheme, convection scheme, solver, itr_max, tol_rel, gradient, urf, gds)
TRANSPORT_CASES = [
    ("A", "bdf2", "linearUpwind", "bicgstab", 10, 0.02, "gauss", 0.9, 1.0),
    ("A", "cn", "cds", "bicgstab", 20, 1e-6, "gauss", 0.7, 0.5),
    ("A", "steady", "linearUpwind", "gauss-seidel", 5, 1e-6, "lsq", 0.8, 1.0),
    ("A", "bdf", "linearUpwind", "bicgstab", 10, 1e-4, "lsq_dm", 1.0, 1.0),
    ("B", "cn", "linearUpwind", "bicgstab", 10, 1e-8, "gauss", 0.9, 1.0),
]


def fd(x: float) -> str:
    """A REAL(8) literal of x for list-directed input."""
    return f"{x:.17e}".replace("e", "d")


def transport_case():
    """calculate_temperature (temperature.f90:35-385) itself, on jittered 5^3
    hexes with every non-periodic boundary treatment it has: the reference's
    module globals are set from the inputs, the routine runs once, and the
    assembled system, the updated field, its gradient and the solver's resor
    (what csrsolve returns in res0) are kept."""
    out = {}
    for key, (types, wf) in TRANSPORT_MESHES.items():
        mesh = hex_mesh(5, jitter=0.2, seed=9, types=types)
        rc = RefCase(mesh)
        rc.dump()
        n, nt, F = mesh.n_cells, mesh.n_total, mesh.n_faces
        T = uniform(7, 1, nt, 0.0, 1.0)
        To = uniform(7, 2, n, 0.0, 1.0)
        Too = uniform(7, 3, n, 0.0, 1.0)
        den = 1.0 + 0.1 * uniform(7, 4, n, 0.0, 1.0)
        vis = 1e-3 * (1.0 + uniform(7, 5, n, 0.0, 1.0))
        flm = 1e-2 * uniform(7, 6, F, -1.0, 1.0)
        ins = dict(T=T, To=To, Too=Too, den=den, vis=vis, flm=flm, wf=np.array(wf, np.float64))
        for k, v in ins.items():
            rc.put(k + ".bin", v)
            out[f"{key}_in_{k}"] = v
        for ic, (mk, ts, sch, slv, itm, trel, gm, urf, gds) in enumerate(TRANSPORT_CASES):
            if mk != key:
                continue
            rc.run([f"temperature T.bin To.bin Too.bin den.bin vis.bin flm.bin wf.bin 1.0d-3 0.7d0 "
                    f"1.0d0 {fd(urf)} {fd(gds)} 1.0d-2 {ts} {sch} {slv} {itm} 1.0d-13 {fd(trel)} "
                    f"{gm} Tn.bin a.bin su.bin sp.bin gr.bin info.txt"])
            out[f"c{ic}_T"] = rc.get("Tn.bin")
            out[f"c{ic}_a"] = rc.get("a.bin")
            out[f"c{ic}_su"] = rc.get("su.bin")
            out[f"c{ic}_sp"] = rc.get("sp.bin")
            out[f"c{ic}_grad"] = rc.get("gr.bin").reshape(-1, 3)
            out[f"c{ic}_resor"] = np.array(float(open(os.path.join(rc.dir, "info.txt")).read()))
        rc.close()
    np.savez_compressed(os.path.join(HERE, "transport.npz"), **out)
    print("wrote transport")


def grad_fvx_case():
    """fvxGradient.f90's Gauss gradients on a jittered 5^3 hex and a jittered
    3^3 tet mesh: the two-sweep corrected grad_gauss (:1549-1645, face
    values through gradco :1761-1817 with the previous sweep's gradient) and
    grad_gauss_corrected (:1664-1758) started from a given gradient."""
    from fcp.mesh import tet_mesh
    out = {}
    meshes = {"hex": hex_mesh(5, jitter=0.2, seed=7,
                              types=("inlet", "outlet", "wall", "symmetry", "pressure", "empty")),
              "tet": tet_mesh(3, jitter=0.2, seed=3)}
    for key, mesh in meshes.items():
        rc = RefCase(mesh)
        rc.dump()
        n, nt = mesh.n_cells, mesh.n_total
        phi = uniform(15, 1, nt, -1.0, 1.0)
        gin = uniform(15, 2, 3 * n, -1.0, 1.0)
        rc.put("phi.bin", phi)
        rc.put("gin.bin", gin)
        rc.run(["grad_fvx phi.bin gin.bin g2.bin gc.bin"])
        out[f"{key}_in_phi"] = phi
        out[f"{key}_in_gin"] = gin.reshape(-1, 3)
        out[f"{key}_gauss2"] = rc.get("g2.bin").reshape(-1, 3)
        out[f"{key}_corrected"] = rc.get("gc.bin").reshape(-1, 3)
        rc.close()
    np.savez_compressed(os.path.join(HERE, "gradfvx.npz"), **out)
    print("wrote gradfvx")


def flux_scalar_case():
    """facefluxscalar (scalar_fluxes.f90:344-464): the convection-diffusion
    assembly of the convdiff job with the diffusivity given per cell and
    interpolated to each inner face, on a jittered 5^3 hex (every boundary
    kind) and a jittered 3^3 tet mesh, under both face-value schemes."""
    from fcp.mesh import tet_mesh
    out = {}
    meshes = {"hex": hex_mesh(5, jitter=0.2, seed=7,
                              types=("inlet", "outlet", "wall", "symmetry", "pressure", "empty")),
              "tet": tet_mesh(3, jitter=0.2, seed=3)}
    for key, mesh in meshes.items():
        rc = RefCase(mesh)
        rc.dump()
        n, nt, F, Fi = mesh.n_cells, mesh.n_total, mesh.n_faces, mesh.n_inner
        flm = uniform(16, 1, F, -1.0, 1.0)
        vis = 10.0 ** uniform(16, 2, n, -1.0, 1.0)
        phi = uniform(16, 3, nt, -1.0, 1.0)
        gr = uniform(16, 4, 3 * nt, -1.0, 1.0)
        bk = (np.arange(F - Fi) % 3).astype(np.float64)
        for k, v in (("flm", flm), ("vis", vis), ("phi", phi), ("gr", gr), ("bk", bk)):
            rc.put(k + ".bin", v)
        out.update({f"{key}_in_flmass": flm, f"{key}_in_vis": vis, f"{key}_in_phi": phi,
                    f"{key}_in_grad": gr.reshape(-1, 3), f"{key}_in_bckind": bk.astype(np.uint8)})
        for scheme in ("cds", "linearUpwind"):
            rc.run([f"convdiff_cell flm.bin vis.bin phi.bin gr.bin 0.7d0 {scheme} bk.bin "
                    "ca.bin csu.bin csp.bin"])
            out[f"{key}_{scheme}_a"] = rc.get("ca.bin")
            out[f"{key}_{scheme}_su"] = rc.get("csu.bin")
            out[f"{key}_{scheme}_sp"] = rc.get("csp.bin")
        rc.close()
    np.savez_compressed(os.path.join(HERE, "fluxscalar.npz"), **out)
    print("wrote fluxscalar")


PCORR_TYPES = ("inlet", "outlet", "wall", "symmetry", "empty", "wall")


def pcorr_inputs(mesh, seed):
    """Inputs of the pressure-correction fixture: velocities, pressure and its
    gradient, density, 1/a_P of the momentum rows and a mass flux whose
    inlet part is what bcin would prescribe (den u.S of the inlet values)."""
    n, nt, F = mesh.n_cells, mesh.n_total, mesh.n_faces
    I = dict(u=uniform(seed, 1, nt, -1.0, 1.0), v=uniform(seed, 2, nt, -1.0, 1.0),
             w=uniform(seed, 3, nt, -1.0, 1.0), p=uniform(seed, 4, nt, -1.0, 1.0),
             dpdxi=uniform(seed, 5, 3 * nt, -1.0, 1.0),
             den=1.0 + 0.2 * uniform(seed, 6, nt, -1.0, 1.0),
             apu=10.0 ** uniform(seed, 7, n, -2.0, 0.0), apv=10.0 ** uniform(seed, 8, n, -2.0, 0.0),
             apw=10.0 ** uniform(seed, 9, n, -2.0, 0.0), flmass=uniform(seed, 10, F, -1.0, 1.0),
             pp=uniform(seed, 11, nt, -1.0, 1.0))
    return I


def pcorr_case():
    """The pressure-correction step of SIMPLE (calcp_simple.f90:69-419,
    composed from the reference's facefluxmass2, facefluxmass2_periodic,
    adjustMassFlow, gradp_and_sources, bpres and updateVelocityAtBoundary):
    the assembly on a jittered 6^3 hex with inlet/outlet/wall/symmetry/empty
    patches (with and without adjustMassFlow) and on the periodic 6^3 mesh,
    and the correction from a given pressure correction on both."""
    out = {}
    cases = {"hex": (hex_mesh(6, jitter=0.2, seed=13, types=PCORR_TYPES), 17),
             "per": (periodic_mesh(), 18)}
    for key, (mesh, seed) in cases.items():
        rc = RefCase(mesh)
        rc.dump()
        I = pcorr_inputs(mesh, seed)
        for k, v in I.items():
            rc.put(k + ".bin", v)
            out[f"{key}_in_{k}"] = v
        Fi = mesh.n_inner
        inlet = [i for i, pt in enumerate(mesh.patch_of_bface()) if mesh.types[pt] == "inlet"]
        # the mass flow entering through the inlets (inlet fluxes are
        # negative: S points out of the domain), to six decimals so that its
        # text reads back as the same double on both sides
        flomas = float(f"{-np.sum(I['flmass'][Fi + np.array(inlet, int)]):.6f}") if inlet else 0.0
        out[f"{key}_flomas"] = np.array(flomas)
        for cmf in ((0, 1) if key == "hex" else (1,)):
            rc.run([f"pcorr_assemble u.bin v.bin w.bin p.bin dpdxi.bin den.bin apu.bin apv.bin "
                    f"apw.bin flmass.bin {cmf} {flomas!r} a.bin su.bin fm.bin uo.bin vo.bin wo.bin"])
            t = f"{key}_c{cmf}"
            for k, f in (("a", "a.bin"), ("su", "su.bin"), ("flmass", "fm.bin"), ("u", "uo.bin"),
                         ("v", "vo.bin"), ("w", "wo.bin")):
                out[f"{t}_{k}"] = rc.get(f)
        cmf = 1
        t = f"{key}_c{cmf}"
        rc.run([f"pcorr_correct pp.bin fm.bin a.bin uo.bin vo.bin wo.bin p.bin apu.bin apv.bin "
                f"apw.bin 0.3d0 7 ppo.bin fmo.bin uo2.bin vo2.bin wo2.bin po.bin dpo.bin "
                f"csu.bin csv.bin csw.bin"])
        for k, f in (("pp", "ppo.bin"), ("flmass", "fmo.bin"), ("u", "uo2.bin"), ("v", "vo2.bin"),
                     ("w", "wo2.bin"), ("p", "po.bin"), ("dpdxi", "dpo.bin"), ("su", "csu.bin"),
                     ("sv", "csv.bin"), ("sw", "csw.bin")):
            out[f"{key}_corr_{k}"] = rc.get(f)
        rc.close()
    np.savez_compressed(os.path.join(HERE, "pcorr.npz"), **out)
    print("wrote pcorr")


# (time scheme, convection scheme, gds, urf u/v/w, bicgstab iterations)
MOMENTUM_CASES = [
    ("steady", "linearUpwind", 1.0, (0.7, 0.6, 0.5), 0),
    ("bdf2", "cds", 0.8, (0.8, 0.8, 0.8), 0),
    ("cn", "linearUpwind", 1.0, (0.7, 0.7, 0.7), 0),
    ("bdf", "linearUpwind", 1.0, (0.7, 0.7, 0.7), 5),
]


def momentum_inputs(mesh, seed):
    """Inputs of the momentum fixture: current, old and older velocities,
    pressure, density, viscosity (cells and boundary faces), the wall faces'
    effective viscosity, mass fluxes and the matrix's previous content."""
    n, nt, F = mesh.n_cells, mesh.n_total, mesh.n_faces
    nwal = int(sum(int(mesh.nfaces[ib]) for ib, t in enumerate(mesh.types) if t == "wall"))
    return dict(u=uniform(seed, 1, nt, -1.0, 1.0), v=uniform(seed, 2, nt, -1.0, 1.0),
                w=uniform(seed, 3, nt, -1.0, 1.0), uo=uniform(seed, 4, n, -1.0, 1.0),
                vo=uniform(seed, 5, n, -1.0, 1.0), wo=uniform(seed, 6, n, -1.0, 1.0),
                uoo=uniform(seed, 7, n, -1.0, 1.0), voo=uniform(seed, 8, n, -1.0, 1.0),
                p=uniform(seed, 9, nt, -1.0, 1.0), den=1.0 + 0.2 * uniform(seed, 10, nt, -1.0, 1.0),
                vis=10.0 ** uniform(seed, 11, nt, -3.0, -1.0),
                visw=10.0 ** uniform(seed, 12, max(nwal, 1), -3.0, -1.0),
                flmass=uniform(seed, 13, F, -1.0, 1.0) * 1e-2,
                astale=uniform(seed, 14, 7 * n + 2 * F, -1.0, 1.0))


def momentum_case():
    """calcuvw (velocity.f90:50-750) itself on a jittered 6^3 hex with inlet,
    outlet, wall, symmetry and empty patches and on the periodic 6^3 mesh:
    with zero solver iterations the assembled systems (a with w's diagonal,
    su = w's right-hand side, sv, sw, the diagonal sources and 1/a_P of all
    three components, gradients and pressure boundary values); with five
    BiCGStab iterations the solved velocities."""
    out = {}
    cases = {"hex": (hex_mesh(6, jitter=0.2, seed=19, types=PCORR_TYPES), 25),
             "per": (periodic_mesh(), 26)}
    for key, (mesh, seed) in cases.items():
        rc = RefCase(mesh)
        g, c, sizes = rc.dump()
        nnz = int(c["ja"].size)
        I = momentum_inputs(mesh, seed)
        I["astale"] = I["astale"][:nnz]
        for k, v in I.items():
            rc.put(k + ".bin", v)
            out[f"{key}_in_{k}"] = v
        for ic, (ts, sch, gds, urf, its) in enumerate(MOMENTUM_CASES):
            rc.run([f"momentum u.bin v.bin w.bin uo.bin vo.bin wo.bin uoo.bin voo.bin p.bin den.bin "
                    f"vis.bin visw.bin flmass.bin astale.bin 1.0d-3 0.05d0 {ts} {sch} {gds!r} "
                    f"{urf[0]!r} {urf[1]!r} {urf[2]!r} {its} 1.0d-3 a.bin su.bin sv.bin sw.bin "
                    f"spu.bin spv.bin sp.bin apu.bin apv.bin apw.bin uo2.bin vo2.bin wo2.bin "
                    f"po.bin dpo.bin duo.bin"])
            t = f"{key}_m{ic}"
            for k in ("a", "su", "sv", "sw", "spu", "spv", "sp", "apu", "apv", "apw"):
                out[f"{t}_{k}"] = rc.get(k + ".bin")
            for k, f in (("u", "uo2"), ("v", "vo2"), ("w", "wo2"), ("p", "po")):
                out[f"{t}_{k}"] = rc.get(f + ".bin")
            out[f"{t}_dpdxi"] = rc.get("dpo.bin").reshape(-1, 3)
            out[f"{t}_dudxi"] = rc.get("duo.bin").reshape(-1, 3)
        rc.close()
    np.savez_compressed(os.path.join(HERE, "momentum.npz"), **out)
    print("wrote momentum")


PERIODIC_TYPES = ("empty", "periodic", "wall", "inlet", "empty", "periodic")
PERIODIC_WF = (0, 0, 1, 0, 0, 0)
# (time scheme, convection scheme, solver, itr_max, tol_rel, gradient, urf, gds)
PERIODIC_TCASES = [
    ("bdf2", "linearUpwind", "bicgstab", 10, 1e-6, "gauss", 0.9, 1.0),
    ("cn", "cds", "bicgstab", 20, 1e-8, "gauss", 0.8, 0.5),
]


def periodic_mesh():
    """hex_mesh(6) with the faces of each twin patch in a seeded shuffled
    order, so that face_mapping has a permutation to find."""
    from fcp.mesh import fisher_yates
    mesh = hex_mesh(6, jitter=0.2, seed=11, types=PERIODIC_TYPES)
    for k, t in enumerate(mesh.twin_start):
        nf = int(mesh.nfaces[list(mesh.start_face).index(t)])
        perm = fisher_yates(nf, 17 + k)
        sl = slice(int(t), int(t) + nf)
        mesh.node[sl] = mesh.node[sl][perm]
        mesh.nnodes[sl] = mesh.nnodes[sl][perm]
        mesh.owner[sl] = mesh.owner[sl][perm]
    mesh.check()
    return mesh


def periodic_case():
    """Periodic boundaries on a jittered 6^3 hex: the x+ and z+ patches are
    'periodic' with the x- and z- patches (typed 'empty', as in the
    reference's examples) as twins.  The reference's read_mesh maps the twin
    faces (face_mapping, geometry.f90:1848-1997); kept are the geometry and
    owners after the mapping, the CSR pattern with the periodic pairs and the
    slot maps over inner + periodic faces (sparse_matrix.f90:141-171,
    262-293), the Laplacian and Gauss gradient on the mapped owners,
    updateBoundary's periodic average (updateBoundary.f90:46-68) and
    calculate_temperature with its periodic branch (temperature.f90:262-313)."""
    mesh = periodic_mesh()
    rc = RefCase(mesh)
    g, c, sizes = rc.dump()
    rc.run(["dump_periodic"])
    n, nt, F = mesh.n_cells, mesh.n_total, mesh.n_faces
    out = {f"g_{k}": v for k, v in g.items()}
    out.update({f"c_{k}": v for k, v in c.items()})
    out["sizes"] = np.array(sizes, np.int64)
    out["p_owner"] = rc.get("p_owner.bin", np.int32) - 1
    out["p_slot_pn"] = rc.get("p_slot_pn.bin", np.int32) - 1
    out["p_slot_np"] = rc.get("p_slot_np.bin", np.int32) - 1
    with open(os.path.join(rc.dir, "psizes.txt")) as f:
        out["psizes"] = np.array([int(t) for t in f.read().split()], np.int64)
    mu = -(10.0 ** uniform(13, 1, n, -1.0, 1.0))
    phi = uniform(13, 2, nt, -1.0, 1.0)
    su0 = uniform(13, 3, n, -1.0, 1.0)
    u = uniform(13, 4, nt, -1.0, 1.0)
    out.update(in_mu=mu, in_phi=phi, in_su0=su0, in_u=u)
    for k, v in (("mu", mu), ("phi", phi), ("su0", su0), ("u", u)):
        rc.put(k + ".bin", v)
    rc.run(["laplacian mu.bin phi.bin su0.bin a.bin su.bin", "grad_gauss u.bin gg.bin",
            "update_boundary phi.bin ub.bin"])
    out["lap_a"] = rc.get("a.bin")
    out["lap_su"] = rc.get("su.bin")
    out["grad_gauss"] = rc.get("gg.bin").reshape(-1, 3)
    out["update_boundary"] = rc.get("ub.bin")
    T = uniform(14, 1, nt, 0.0, 1.0)
    To = uniform(14, 2, n, 0.0, 1.0)
    Too = uniform(14, 3, n, 0.0, 1.0)
    den = 1.0 + 0.1 * uniform(14, 4, n, 0.0, 1.0)
    vis = 1e-3 * (1.0 + uniform(14, 5, n, 0.0, 1.0))
    flm = 1e-2 * uniform(14, 6, F, -1.0, 1.0)
    ins = dict(T=T, To=To, Too=Too, den=den, vis=vis, flm=flm,
               wf=np.array(PERIODIC_WF, np.float64))
    for k, v in ins.items():
        rc.put(k + ".bin", v)
        out[f"in_{k}"] = v
    for ic, (ts, sch, slv, itm, trel, gm, urf, gds) in enumerate(PERIODIC_TCASES):
        rc.run([f"temperature T.bin To.bin Too.bin den.bin vis.bin flm.bin wf.bin 1.0d-3 0.7d0 "
                f"1.0d0 {fd(urf)} {fd(gds)} 1.0d-2 {ts} {sch} {slv} {itm} 1.0d-13 {fd(trel)} "
                f"{gm} Tn.bin ta.bin tsu.bin tsp.bin tgr.bin info.txt"])
        out[f"t{ic}_T"] = rc.get("Tn.bin")
        out[f"t{ic}_a"] = rc.get("ta.bin")
        out[f"t{ic}_su"] = rc.get("tsu.bin")
        out[f"t{ic}_sp"] = rc.get("tsp.bin")
        out[f"t{ic}_grad"] = rc.get("tgr.bin").reshape(-1, 3)
        out[f"t{ic}_resor"] = np.array(float(open(os.path.join(rc.dir, "info.txt")).read()))
    rc.close()
    np.savez_compressed(os.path.join(HERE, "periodic.npz"), **out)
    print("wrote periodic")


def cavity_case():
    """The reference's own test mesh (test/testFieldOperations/polyMesh).
    The reference's OpenFOAM reader aborts under the flang runtime
    (list-directed CHARACTER read at geometry.f90:1072), so the same mesh is
    handed to the reference's native reader (identical connectivity, 1-based)."""
    cav = os.path.join(HERE, "cavity")
    mesh = read_foam(cav)
    rc = RefCase(mesh)
    g, c, sizes = rc.dump()
    out = {f"g_{k}": v for k, v in g.items()}
    out.update({f"c_{k}": v for k, v in c.items()})
    out["sizes"] = np.array(sizes, np.int64)
    n, nt = mesh.n_cells, mesh.n_total
    # Linear field psi = x+y+z with boundary values at face centroids
    # (testFieldOperations.f90:140-141).
    psi = np.zeros(nt)
    psi[:n] = g["xc"] + g["yc"] + g["zc"]
    fb = np.arange(mesh.n_inner, mesh.n_faces)
    psi[n:] = g["xf"][fb] + g["yf"][fb] + g["zf"][fb]
    out["in_psi"] = psi
    rc.put("psi.bin", psi)
    rc.put("mu.bin", -np.ones(n))
    rc.put("su0.bin", np.zeros(n))
    rc.run(["grad_gauss psi.bin gg.bin", "laplacian mu.bin psi.bin su0.bin a.bin su.bin"])
    out["grad_gauss"] = rc.get("gg.bin").reshape(-1, 3)
    out["lap_a"] = rc.get("a.bin")
    out["lap_su"] = rc.get("su.bin")
    rc.close()
    np.savez_compressed(os.path.join(HERE, "cavity.npz"), **out)
    print("wrote cavity")


def poisson32_case():
    """32^3 uniform Poisson, mu=-1, Dirichlet 0, ICCG to 1e-10."""
    mesh = workloads.poisson_uniform_mesh(32)
    rc = RefCase(mesh)
    g, c, sizes = rc.dump()
    n = mesh.n_cells
    inp = workloads.poisson_uniform_inputs(mesh, g["vol"])
    rc.put("mu.bin", inp["mu"])
    rc.put("phi.bin", inp["phi"])
    rc.put("su0.bin", inp["rhs"])
    rc.put("fi0.bin", np.zeros(n))
    rc.put("zero.bin", np.zeros(n))
    rc.run(["laplacian mu.bin phi.bin su0.bin a.bin su.bin"])
    a = rc.get("a.bin")
    su = rc.get("su.bin")
    out = dict(sizes=np.array(sizes, np.int64))
    for k, v in c.items():
        out[f"sha_c_{k}"] = np.array(sha(v))
    for k in ("vol", "xc", "facint", "Df"):
        out[f"sha_g_{k}"] = np.array(sha(g[k]))
    out["sha_lap_a"] = np.array(sha(a))
    out["sha_lap_su"] = np.array(sha(su))
    out["rhs"] = su
    for it in (1, 10, 50):
        rc.run([f"solve iccg {it} 0.0 1.0d-10 a.bin fi0.bin su.bin fi.bin info.txt"])
        out[f"sha_iccg_{it}"] = np.array(sha(rc.get("fi.bin")))
    log = rc.run(["solve iccg 10000 0.0 1.0d-10 a.bin fi0.bin su.bin fi.bin info.txt"])
    out["iccg_fi"] = rc.get("fi.bin")
    out["iccg_resor"] = np.array(float(open(os.path.join(rc.dir, "info.txt")).read()))
    iters = [int(line.split("No Iterations")[-1]) for line in log.splitlines() if "No Iterations" in line]
    out["iccg_iters"] = np.array(iters[-1])
    out["iccg_log"] = np.array([line for line in log.splitlines() if "PCG" in line][-1])
    rc.close()
    np.savez_compressed(os.path.join(HERE, "poisson32.npz"), **out)
    print("wrote poisson32", iters[-1])


def kat_5x5_case():
    """The 5x5 systems of test/test_linear_solvers_spsolve.f90:77-98,129-138.
    Entries are default-REAL literals, i.e. float32 values widened to REAL(8)."""
    ns = np.array([[6.80, -6.05, -0.45, 8.32, -9.67],
                   [-2.11, -3.30, 2.58, 2.71, -5.14],
                   [5.66, 5.36, -2.70, 4.35, -7.26],
                   [5.97, -4.44, 0.27, -7.17, 6.08],
                   [8.23, 1.08, 9.04, 2.14, -6.87]], dtype=np.float32).astype(np.float64)
    b_ns = np.array([4.02, 6.19, -8.22, -7.57, -3.03], dtype=np.float32).astype(np.float64)
    sp = np.array([[3.14, 0.17, -0.90, 1.65, -0.72],
                   [0.17, 0.79, 0.83, -0.65, 0.28],
                   [-0.90, 0.83, 4.53, -3.70, 1.60],
                   [1.65, -0.65, -3.70, 5.32, -1.37],
                   [-0.72, 0.28, 1.60, -1.37, 1.98]], dtype=np.float32).astype(np.float64)
    b_sp = np.array([-7.29, 9.25, 5.99, -1.94, -8.30], dtype=np.float32).astype(np.float64)
    ia = np.arange(0, 26, 5, dtype=np.int32)
    ja = np.tile(np.arange(5, dtype=np.int32), 5)
    dg = np.array([0, 6, 12, 18, 24], np.int32)
    out = dict(A_ns=ns, b_ns=b_ns, A_sp=sp, b_sp=b_sp, ia=ia, ja=ja, diag=dg,
               xa_ns=np.array([-0.80, -0.70, 0.59, 1.32, 0.57]),
               xa_sp=np.array([-6.02, 15.62, 3.02, 3.25, -8.78]))
    rc = RefCase(hex_mesh(2))
    (ia + 1).astype(np.int32).tofile(os.path.join(rc.dir, "eia.bin"))
    (ja + 1).astype(np.int32).tofile(os.path.join(rc.dir, "eja.bin"))
    (dg + 1).astype(np.int32).tofile(os.path.join(rc.dir, "edg.bin"))
    rc.put("x0.bin", np.zeros(5))
    runs = [("ns", "bicgstab", 5), ("ns", "bicgstab", 50), ("sp", "iccg", 5), ("sp", "dpcg", 5),
            ("sp", "dpcg", 50), ("sp", "gauss-seidel", 100), ("sp", "bicgstab", 5)]
    for which, solver, itmax in runs:
        A = out["A_" + which]
        rc.put("ea.bin", A.reshape(-1))
        rc.put("eb.bin", out["b_" + which])
        rc.run([f"solve_explicit {solver} 5 25 {itmax} 1.0d-10 1.0d-7 eia.bin eja.bin edg.bin "
                f"ea.bin x0.bin eb.bin ex.bin einfo.txt"])
        key = f"x_{which}_{solver.replace('-', '_')}_{itmax}"
        out[key] = rc.get("ex.bin")
    rc.close()
    np.savez_compressed(os.path.join(HERE, "kat5x5.npz"), **out)
    print("wrote kat5x5")


def _report(log, tag):
    """(iterations, final residual) of the last solver report line with tag."""
    line = [ln for ln in log.splitlines() if tag in ln][-1]
    its = int(line.split("No Iterations")[-1])
    resl = float(line.split("Final residual =")[-1].split(",")[0])
    return its, resl


def gmres_case():
    """pmgmres_ilu, ilu_cr, lus_cr and atx_cr (linear_solvers.f90:789-1555) on
    the hex5j Laplacian (SPD) and linearUpwind convection-diffusion (non-
    symmetric) matrices, the two dense 5x5 systems, and the 32^3 Poisson."""
    hx = np.load(os.path.join(HERE, "hex5j.npz"))
    kat = np.load(os.path.join(HERE, "kat5x5.npz"))
    out = {}
    rc = RefCase(hex_mesh(2))
    systems = {
        "lap": (hx["c_ia"], hx["c_ja"], hx["c_diag"], hx["lap_a"], hx["in_b"]),
        "cd": (hx["c_ia"], hx["c_ja"], hx["c_diag"], hx["cd_linearUpwind_a"], hx["in_b"]),
        "ns": (kat["ia"], kat["ja"], kat["diag"], kat["A_ns"].reshape(-1), kat["b_ns"]),
        "sp": (kat["ia"], kat["ja"], kat["diag"], kat["A_sp"].reshape(-1), kat["b_sp"]),
    }
    for key, (ia, ja, dg, a, b) in systems.items():
        n, nz = len(ia) - 1, len(ja)
        (ia + 1).astype(np.int32).tofile(os.path.join(rc.dir, "eia.bin"))
        (ja + 1).astype(np.int32).tofile(os.path.join(rc.dir, "eja.bin"))
        (dg + 1).astype(np.int32).tofile(os.path.join(rc.dir, "edg.bin"))
        rc.put("ea.bin", a)
        rc.put("eb.bin", b)
        rc.put("x0.bin", np.zeros(n))
        rc.run([f"ilu {n} {nz} eia.bin eja.bin edg.bin ea.bin eb.bin el.bin ez.bin ew.bin"])
        out[f"{key}_ilu"] = rc.get("el.bin")
        out[f"{key}_lus_b"] = rc.get("ez.bin")
        out[f"{key}_atx_b"] = rc.get("ew.bin")
        for itmax in (1, 2, 5, 20, 500):
            log = rc.run([f"solve_explicit pmgmres {n} {nz} {itmax} 0.0 1.0d-10 eia.bin eja.bin "
                          f"edg.bin ea.bin x0.bin eb.bin ex.bin einfo.txt"])
            out[f"{key}_gm_{itmax}"] = rc.get("ex.bin")
            out[f"{key}_gm_{itmax}_resor"] = np.array(
                float(open(os.path.join(rc.dir, "einfo.txt")).read()))
            its, resl = _report(log, "PMGMRES_ILU")
            out[f"{key}_gm_{itmax}_iters"] = np.array(its)
            out[f"{key}_gm_{itmax}_resl"] = np.array(resl)
    rc.close()
    # 32^3 Poisson through csrsolve('pmgmres'), mr = 4 (:85)
    mesh = workloads.poisson_uniform_mesh(32)
    rc = RefCase(mesh)
    g, c, sizes = rc.dump()
    n = mesh.n_cells
    inp = workloads.poisson_uniform_inputs(mesh, g["vol"])
    rc.put("mu.bin", inp["mu"])
    rc.put("phi.bin", inp["phi"])
    rc.put("su0.bin", inp["rhs"])
    rc.put("fi0.bin", np.zeros(n))
    rc.run(["laplacian mu.bin phi.bin su0.bin a.bin su.bin"])
    out["p32_rhs"] = rc.get("su.bin")
    out["p32_sha_a"] = np.array(sha(rc.get("a.bin")))
    for itmax in (3, 3000):
        log = rc.run([f"solve pmgmres {itmax} 0.0 1.0d-10 a.bin fi0.bin su.bin fi.bin info.txt"])
        out[f"p32_gm_{itmax}"] = rc.get("fi.bin")
        out[f"p32_gm_{itmax}_resor"] = np.array(
            float(open(os.path.join(rc.dir, "info.txt")).read()))
        its, resl = _report(log, "PMGMRES_ILU")
        out[f"p32_gm_{itmax}_iters"] = np.array(its)
        out[f"p32_gm_{itmax}_resl"] = np.array(resl)
    rc.close()
    np.savez_compressed(os.path.join(HERE, "gmres.npz"), **out)
    print("wrote gmres", int(out["p32_gm_3000_iters"]))


CASES = {
    "hex5j": lambda: small_case(
        hex_mesh(5, jitter=0.2, seed=2,
                 types=("wall", "symmetry", "outlet", "wall", "inlet", "pressure")), "hex5j"),
    "hexbc": bc_case,
    "periodic": periodic_case,
    "gradfvx": grad_fvx_case,
    "fluxscalar": flux_scalar_case,
    "pcorr": pcorr_case,
    "momentum": momentum_case,
    "transport": transport_case,
    "cavity": cavity_case,
    "kat5x5": kat_5x5_case,
    "poisson32": poisson32_case,
    "gmres": gmres_case,  # reads hex5j.npz and kat5x5.npz
}

if __name__ == "__main__":
    for name in (sys.argv[1:] or list(CASES)):
        CASES[name]()
