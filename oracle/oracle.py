"""ctypes wrapper of the CPU restatement (oracle/libfcp_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
the CPU-baseline leg of bench.py as the checker.  The product never imports
this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfcp_oracle.so")

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


class _Mesh(C.Structure):
    _fields_ = [("n_nodes", C.c_int), ("n_cells", C.c_int), ("n_inner", C.c_int),
                ("n_faces", C.c_int), ("n_patches", C.c_int),
                ("points", C.c_void_p), ("nnodes", C.c_void_p), ("node", C.c_void_p),
                ("owner", C.c_void_p), ("neighbour", C.c_void_p), ("nfaces", C.c_void_p),
                ("start_face", C.c_void_p), ("bctype", C.c_void_p), ("twin_start", C.c_void_p)]


class _Geom(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in
                ("xc", "yc", "zc", "vol", "arx", "ary", "arz", "xf", "yf", "zf",
                 "facint", "Df", "dnb", "srdb")]


class _Csr(C.Structure):
    _fields_ = [("n", C.c_int), ("nnz", C.c_int), ("ia", C.c_void_p), ("ja", C.c_void_p),
                ("diag", C.c_void_p)]


class _Info(C.Structure):
    _fields_ = [("resor", C.c_double), ("res0", C.c_double), ("resl", C.c_double),
                ("factor", C.c_double), ("iters", C.c_int)]


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class Oracle:
    """Holds one mesh plus its geometry and CSR pattern computed by the oracle."""

    GEOM_KEYS = ("xc", "yc", "zc", "vol", "arx", "ary", "arz", "xf", "yf", "zf",
                 "facint", "Df", "dnb", "srdb")

    def __init__(self, mesh):
        L = lib()
        self.mesh = mesh
        self._keep = dict(
            points=np.ascontiguousarray(mesh.points, dtype=np.float64),
            nnodes=np.ascontiguousarray(mesh.nnodes, dtype=np.int32),
            node=np.ascontiguousarray(mesh.node, dtype=np.int32),
            owner=np.ascontiguousarray(mesh.owner, dtype=np.int32),
            neighbour=np.ascontiguousarray(mesh.neighbour, dtype=np.int32),
            nfaces=np.ascontiguousarray(mesh.nfaces, dtype=np.int32),
            start_face=np.ascontiguousarray(mesh.start_face, dtype=np.int32),
            bctype=np.ascontiguousarray(mesh.bctype_codes, dtype=np.int32),
            twin_start=np.ascontiguousarray(list(mesh.twin_start) or [0], dtype=np.int32))
        k = self._keep
        twin = _p(k["twin_start"]) if mesh.twin_start else None
        self.cm = _Mesh(mesh.n_nodes, mesh.n_cells, mesh.n_inner, mesh.n_faces, mesh.n_patches,
                        _p(k["points"]), _p(k["nnodes"]), _p(k["node"]), _p(k["owner"]),
                        _p(k["neighbour"]), _p(k["nfaces"]), _p(k["start_face"]), _p(k["bctype"]),
                        twin)
        n, F, Fi = mesh.n_cells, mesh.n_faces, mesh.n_inner
        sizes = dict(xc=n, yc=n, zc=n, vol=n, arx=F, ary=F, arz=F, xf=F, yf=F, zf=F,
                     facint=Fi, Df=Fi, dnb=F - Fi, srdb=F - Fi)
        self.geom = {key: np.zeros(sizes[key]) for key in self.GEOM_KEYS}
        self.cg = _Geom(*[_p(self.geom[key]) for key in self.GEOM_KEYS])
        rc = L.orc_geometry(C.byref(self.cm), C.byref(self.cg))
        assert rc == 0, rc
        # owners after face_mapping (periodic twins), used from here on
        self.owner = k["owner"].copy()
        if mesh.twin_start:
            L.orc_face_mapping(C.byref(self.cm), C.byref(self.cg), _p(self.owner))
            k["owner_mapped"] = self.owner
            self.cm.owner = _p(self.owner)
        self.n_periodic = L.orc_n_periodic(C.byref(self.cm))
        nnz = L.orc_csr_nnz(C.byref(self.cm))
        assert nnz > 0, nnz
        self.ia = np.zeros(n + 1, np.int32)
        self.ja = np.zeros(nnz, np.int32)
        self.diag = np.zeros(n, np.int32)
        self.slot_pn = np.zeros(Fi + self.n_periodic, np.int32)
        self.slot_np = np.zeros(Fi + self.n_periodic, np.int32)
        rc = L.orc_csr_build(C.byref(self.cm), _p(self.ia), _p(self.ja), _p(self.diag),
                             _p(self.slot_pn), _p(self.slot_np))
        assert rc == nnz, rc
        self.nnz = nnz
        self.cc = _Csr(n, nnz, _p(self.ia), _p(self.ja), _p(self.diag))

    # -- operators ---------------------------------------------------------
    def laplacian(self, mu, phi, su0=None):
        m = self.mesh
        a = np.zeros(self.nnz)
        su = np.zeros(m.n_cells) if su0 is None else np.array(su0, dtype=np.float64, copy=True)
        lib().orc_laplacian(C.byref(self.cm), C.byref(self.cg), C.byref(self.cc),
                            _p(self.slot_pn), _p(self.slot_np),
                            _p(np.ascontiguousarray(mu, np.float64)),
                            _p(np.ascontiguousarray(phi, np.float64)), _p(a), _p(su))
        return a, su

    def laplacian_face(self, mu_f, bc_kind, phi, su0=None):
        m = self.mesh
        a = np.zeros(self.nnz)
        su = np.zeros(m.n_cells) if su0 is None else np.array(su0, dtype=np.float64, copy=True)
        lib().orc_laplacian_face(C.byref(self.cm), C.byref(self.cg), C.byref(self.cc),
                                 _p(self.slot_pn), _p(self.slot_np),
                                 _p(np.ascontiguousarray(mu_f, np.float64)),
                                 _p(np.ascontiguousarray(bc_kind, np.uint8)),
                                 _p(np.ascontiguousarray(phi, np.float64)), _p(a), _p(su))
        return a, su

    def grad_gauss(self, u):
        m = self.mesh
        out = np.zeros(3 * m.n_total)
        lib().orc_grad_gauss(C.byref(self.cm), C.byref(self.cg),
                             _p(np.ascontiguousarray(u, np.float64)), _p(out))
        return out.reshape(-1, 3)

    def grad_gauss_fvx(self, u, gold=None):
        """fvxGradient.f90 grad_gauss (:1549-1645, two sweeps) when gold is
        None, else grad_gauss_corrected (:1664-1758) from gold (n, 3)."""
        m = self.mesh
        out = np.zeros(3 * m.n_total)
        g = None if gold is None else np.ascontiguousarray(gold, np.float64).reshape(-1)
        rc = lib().orc_grad_gauss_fvx(C.byref(self.cm), C.byref(self.cg),
                                      _p(np.ascontiguousarray(u, np.float64)),
                                      None if g is None else _p(g), _p(out))
        assert rc == 0, rc
        return out.reshape(-1, 3)

    def lsq_matrix(self, variant: int, compat: int = 1):
        n = self.mesh.n_cells
        dmat = np.zeros(n * (18 if variant == 2 else 9))
        rc = lib().orc_lsq_matrix(C.byref(self.cm), C.byref(self.cg), variant, compat, _p(dmat))
        assert rc == 0, rc
        return dmat

    def grad_lsq(self, variant: int, phi, compat: int = 1, dmat=None):
        if dmat is None:
            dmat = self.lsq_matrix(variant, compat)
        out = np.zeros(3 * self.mesh.n_total)
        rc = lib().orc_grad_lsq(C.byref(self.cm), C.byref(self.cg), variant, compat, _p(dmat),
                                _p(np.ascontiguousarray(phi, np.float64)), _p(out))
        assert rc == 0, rc
        return out.reshape(-1, 3)

    def nonorth_correction(self, gamma_f, grad, su):
        su = np.array(su, dtype=np.float64, copy=True)
        lib().orc_nonorth_correction(C.byref(self.cm), C.byref(self.cg),
                                     _p(np.ascontiguousarray(gamma_f, np.float64)),
                                     _p(np.ascontiguousarray(grad, np.float64).reshape(-1)), _p(su))
        return su

    def update_boundary(self, phi):
        phi = np.array(phi, dtype=np.float64, copy=True)
        lib().orc_update_boundary(C.byref(self.cm), _p(phi))
        return phi

    def convdiff(self, flmass, dcoef, phi, grad, gam, scheme, bc_kind, urf=1.0):
        m = self.mesh
        a = np.zeros(self.nnz)
        su = np.zeros(m.n_cells)
        sp = np.zeros(m.n_cells)
        phi = np.array(phi, dtype=np.float64, copy=True)
        lib().orc_convdiff(C.byref(self.cm), C.byref(self.cg), C.byref(self.cc),
                           _p(self.slot_pn), _p(self.slot_np),
                           _p(np.ascontiguousarray(flmass, np.float64)),
                           _p(np.ascontiguousarray(dcoef, np.float64)), _p(phi),
                           _p(np.ascontiguousarray(grad, np.float64).reshape(-1)),
                           C.c_double(gam), int(scheme),
                           _p(np.ascontiguousarray(bc_kind, np.uint8)), C.c_double(urf),
                           _p(a), _p(su), _p(sp))
        return a, su, sp, phi

    def face_gamma_cells(self, diffcoef):
        """facefluxscalar's face diffusivity from a cell array
        (scalar_fluxes.f90:384-392); boundary faces take the owner's."""
        d = np.zeros(self.mesh.n_faces)
        lib().orc_face_gamma_cells(C.byref(self.cm), C.byref(self.cg),
                                   _p(np.ascontiguousarray(diffcoef, np.float64)), _p(d))
        return d

    def convdiff_cell(self, flmass, diffcoef, phi, grad, gam, scheme, bc_kind, urf=1.0):
        """The convdiff assembly with facefluxscalar on the inner faces."""
        return self.convdiff(flmass, self.face_gamma_cells(diffcoef), phi, grad, gam, scheme,
                             bc_kind, urf)

    def pcorr_assemble(self, u, v, w, p, dpdxi, den, apu, apv, apw, flmass, const_mflux, flomas):
        """calcp_simple.f90:69-228: returns (a, su, flmass, u, v, w)."""
        m = self.mesh
        f64 = lambda x: np.array(x, dtype=np.float64, copy=True).reshape(-1)  # noqa: E731
        u, v, w, flmass = f64(u), f64(v), f64(w), f64(flmass)
        keep = [f64(x) for x in (p, dpdxi, den, apu, apv, apw)]
        a = np.zeros(self.nnz)
        su = np.zeros(m.n_cells)
        rc = lib().orc_pcorr_assemble(C.byref(self.cm), C.byref(self.cg), C.byref(self.cc),
                                      _p(self.slot_pn), _p(self.slot_np), _p(u), _p(v), _p(w),
                                      *[_p(x) for x in keep], _p(flmass), int(const_mflux),
                                      C.c_double(flomas), _p(a), _p(su))
        assert rc == 0, rc
        return a, su, flmass, u, v, w

    def pcorr_correct(self, a, pp, flmass, u, v, w, p, apu, apv, apw, urfp, pref, dpdxi=None):
        """calcp_simple.f90:333-419 from the solved pp; pref 0-based.
        Returns dict of pp, flmass, u, v, w, p, dpdxi, su, sv, sw."""
        m = self.mesh
        f64 = lambda x: np.array(x, dtype=np.float64, copy=True).reshape(-1)  # noqa: E731
        o = dict(pp=f64(pp), flmass=f64(flmass), u=f64(u), v=f64(v), w=f64(w), p=f64(p),
                 dpdxi=np.zeros(3 * m.n_total) if dpdxi is None else f64(dpdxi),
                 su=np.zeros(m.n_cells), sv=np.zeros(m.n_cells), sw=np.zeros(m.n_cells))
        keep = [f64(x) for x in (a, apu, apv, apw)]
        lib().orc_pcorr_correct(C.byref(self.cm), C.byref(self.cg), _p(self.slot_pn),
                                _p(keep[0]), _p(o["pp"]), _p(o["flmass"]), _p(o["u"]), _p(o["v"]),
                                _p(o["w"]), _p(o["p"]), _p(keep[1]), _p(keep[2]), _p(keep[3]),
                                C.c_double(urfp), int(pref), _p(o["dpdxi"]), _p(o["su"]),
                                _p(o["sv"]), _p(o["sw"]))
        return o

    TIME_SCHEMES = {"steady": 0, "bdf": 1, "bdf2": 2, "cn": 3}

    def momentum(self, I, astale, viscos, dt, tscheme, scheme, gds, urf):
        """calcuvw's assembly (velocity.f90:50-750) before its solves.
        I: dict of u, v, w (numTotal), uo, vo, wo, uoo, voo, woo (cells), p, den,
        vis (numTotal), visw (wall faces), flmass.  Returns a dict."""
        m = self.mesh
        n, nt = m.n_cells, m.n_total
        f64 = lambda x: np.array(x, dtype=np.float64, copy=True).reshape(-1)  # noqa: E731
        o = {k: f64(I[k]) for k in ("u", "v", "w", "p")}
        keep = {k: f64(I[k]) for k in ("uo", "vo", "wo", "uoo", "voo", "woo", "den", "vis", "visw",
                                        "flmass")}
        o["a"] = f64(astale)
        for k in ("dpdxi", "dudxi", "dvdxi", "dwdxi"):
            o[k] = np.zeros(3 * nt)
        for k in ("rhs", "diag", "ap", "sp"):
            o[k] = np.zeros(3 * n)
        o["sv"], o["sw"] = np.zeros(n), np.zeros(n)
        sch = {"cds": 0, "linearUpwind": 1}[scheme] if isinstance(scheme, str) else int(scheme)
        ts = self.TIME_SCHEMES[tscheme] if isinstance(tscheme, str) else int(tscheme)
        urf = np.ascontiguousarray(urf, np.float64)
        rc = lib().orc_momentum(C.byref(self.cm), C.byref(self.cg), C.byref(self.cc),
                                _p(self.slot_pn), _p(self.slot_np), _p(o["u"]), _p(o["v"]),
                                _p(o["w"]), *[_p(keep[k]) for k in ("uo", "vo", "wo", "uoo", "voo",
                                                                     "woo")],
                                _p(o["p"]), _p(keep["den"]), _p(keep["vis"]), _p(keep["visw"]),
                                _p(keep["flmass"]), C.c_double(viscos), C.c_double(dt), ts, sch,
                                C.c_double(gds), _p(urf), _p(o["a"]), _p(o["dpdxi"]),
                                _p(o["dudxi"]), _p(o["dvdxi"]), _p(o["dwdxi"]), _p(o["rhs"]),
                                _p(o["diag"]), _p(o["ap"]), _p(o["sp"]), _p(o["sv"]), _p(o["sw"]))
        assert rc == 0, rc
        for k in ("rhs", "diag", "ap", "sp"):
            o[k] = o[k].reshape(3, n)
        return o

    def transport_dcoef(self, vis, bc_kind, viscos, pranl, sigt):
        d = np.zeros(self.mesh.n_faces)
        lib().orc_transport_dcoef(C.byref(self.cm), C.byref(self.cg),
                                  _p(np.ascontiguousarray(vis, np.float64)),
                                  _p(np.ascontiguousarray(bc_kind, np.uint8)), C.c_double(viscos),
                                  C.c_double(pranl), C.c_double(sigt), _p(d))
        return d

    def transport_assemble(self, flmass, dcoef, phi, grad, gam, scheme, bc_kind, srd_map=None,
                           urf=1.0, time_scheme=0, den=None, dt=1.0, phi_o=None, phi_oo=None):
        m = self.mesh
        n = m.n_cells
        a = np.zeros(self.nnz)
        su = np.zeros(n)
        sp = np.zeros(n)
        phi = np.array(phi, dtype=np.float64, copy=True)
        opt = lambda v: None if v is None else _p(np.ascontiguousarray(v, np.float64))  # noqa: E731
        keep = [None if v is None else np.ascontiguousarray(v, np.float64)
                for v in (den, phi_o, phi_oo)]
        smap = None if srd_map is None else np.ascontiguousarray(srd_map, np.int32)
        rc = lib().orc_transport_assemble(
            C.byref(self.cm), C.byref(self.cg), C.byref(self.cc), _p(self.slot_pn),
            _p(self.slot_np), _p(np.ascontiguousarray(flmass, np.float64)),
            _p(np.ascontiguousarray(dcoef, np.float64)), _p(phi),
            _p(np.ascontiguousarray(grad, np.float64).reshape(-1)), C.c_double(gam), int(scheme),
            _p(np.ascontiguousarray(bc_kind, np.uint8)), None if smap is None else _p(smap),
            C.c_double(urf), int(time_scheme), opt(keep[0]), C.c_double(dt), opt(keep[1]),
            opt(keep[2]), _p(a), _p(su), _p(sp))
        assert rc == 0, rc
        return a, su, sp, phi

    def temperature(self, T, To, Too, den, vis, flmass, wall_frac, viscos, pranl, sigt, urf,
                    gds, dt, tscheme, scheme, solver, itr_max, tol_abs, tol_rel,
                    grad_method="gauss"):
        """calculate_temperature (temperature.f90:35-385) restated: gradient,
        face coefficients, assembly, csrsolve from the current field,
        updateBoundary, and the clip of negative values to `small`.  Returns
        the new field and (a, su, sp, grad, resor: the value csrsolve returns
        in its res0 argument, i.e. the solver's resor)."""
        m = self.mesh
        n = m.n_cells
        T = np.array(T, dtype=np.float64, copy=True)
        if grad_method == "gauss":
            gr = self.grad_gauss(T)
        else:
            gr = self.grad_lsq({"lsq": 0, "lsq_dm": 1, "lsq_qr": 2}[grad_method], T, compat=1)
        kinds, smap = transport_bc(m, wall_frac)
        dcoef = self.transport_dcoef(vis, kinds, viscos, pranl, sigt)
        ts = TIME_SCHEMES[tscheme]
        code = {"cds": 0, "linearUpwind": 1}[scheme]
        a, su, sp, T = self.transport_assemble(flmass, dcoef, T, gr, gds, code, kinds, smap, urf,
                                               ts, den, dt, To, Too)
        out = self.solve(solver, a, su, fi0=T[:n], itr_max=itr_max, tol_abs=tol_abs,
                         tol_rel=tol_rel)
        T[:n] = out["fi"]
        T = self.update_boundary(T)
        if T[:n].min() < 0.0:
            T[:n] = np.maximum(T[:n], orc_small())
        return T, dict(a=a, su=su, sp=sp, grad=gr, resor=out["resor"])

    # -- linear algebra ----------------------------------------------------
    def spmv(self, a, x):
        y = np.zeros(self.mesh.n_cells)
        lib().orc_spmv(C.byref(self.cc), _p(np.ascontiguousarray(a, np.float64)),
                       _p(np.ascontiguousarray(x, np.float64)), _p(y))
        return y

    def levels(self, upper: bool = False):
        lev = np.zeros(self.mesh.n_cells, np.int32)
        nlev = lib().orc_levels(C.byref(self.cc), int(upper), _p(lev))
        return lev, nlev

    def rcm(self, pp_iters: int = 4):
        """Reverse Cuthill-McKee permutation (new id of every cell)."""
        out = np.zeros(self.mesh.n_cells, np.int32)
        rc = lib().orc_rcm(C.byref(self.cc), int(pp_iters), _p(out))
        assert rc == 0, rc
        return out

    def dic_diag(self, a):
        d = np.zeros(self.mesh.n_cells)
        lib().orc_dic_diag(C.byref(self.cc), _p(np.ascontiguousarray(a, np.float64)), _p(d))
        return d

    def dilu_diag(self, a):
        d = np.zeros(self.mesh.n_cells)
        lib().orc_dilu_diag(C.byref(self.cc), _p(np.ascontiguousarray(a, np.float64)), _p(d))
        return d

    def dic_apply(self, a, d, r):
        z = np.zeros(self.mesh.n_cells)
        lib().orc_dic_apply(C.byref(self.cc), _p(np.ascontiguousarray(a, np.float64)),
                            _p(np.ascontiguousarray(d, np.float64)),
                            _p(np.ascontiguousarray(r, np.float64)), _p(z))
        return z

    def solve(self, solver: str, a, rhs, fi0=None, itr_max=1000, tol_abs=0.0,
              tol_rel=1e-10, history=False, mr=4):
        return solve_csr(self.cc, solver, a, rhs, fi0, itr_max, tol_abs, tol_rel, history,
                         self.mesh.n_cells, mr)


TIME_SCHEMES = {"steady": 0, "bdf": 1, "bdf2": 2, "cn": 3}


def orc_small():
    lib().orc_small.restype = C.c_double
    return lib().orc_small()


def transport_bc(mesh, wall_frac):
    """Boundary treatment of calculate_temperature (temperature.f90:168-320)
    per boundary face: kind 2 inlet/outlet/pressure, 0 isothermal wall
    (bcDFraction(7,ib) = 1), 1 adiabatic wall (0), 4 prescribed heat flux
    (wallQFlux), 5 periodic face (its twin patch, usually 'empty', gets no
    term), 3 no term (symmetry, empty, walls with another fraction);
    and, for kinds 0 and 4, the boundary face whose |S|/dn the reference
    reads: srdw(iWall), iWall counting the kind-0/4 faces in patch order
    while srdw lists the faces of the 'wall' patches."""
    nb, nin = mesh.n_bnd, mesh.n_inner
    kinds = np.full(nb, 3, np.uint8)
    smap = np.arange(nb, dtype=np.int32)
    rng = [(int(mesh.start_face[i]) - nin, int(mesh.start_face[i]) - nin + int(mesh.nfaces[i]))
           for i in range(mesh.n_patches)]
    walls = np.concatenate([np.arange(*rng[i]) for i, t in enumerate(mesh.types) if t == "wall"]
                           + [np.zeros(0, np.int64)]).astype(np.int32)
    iwall = 0
    for i, t in enumerate(mesh.types):
        s, e = rng[i]
        if t in ("inlet", "outlet", "pressure"):
            kinds[s:e] = 2
        elif (t == "wall" and wall_frac[i] == 1.0) or t == "wallQFlux":
            kinds[s:e] = 0 if t == "wall" else 4
            if iwall + (e - s) > walls.size:
                raise ValueError("the reference's wall counter runs past its wall faces")
            smap[s:e] = walls[iwall:iwall + (e - s)]
            iwall += e - s
        elif t == "wall" and wall_frac[i] == 0.0:
            kinds[s:e] = 1
        elif t == "periodic":
            kinds[s:e] = 5
    return kinds, smap


_SOLVERS = {"iccg": "orc_iccg", "dpcg": "orc_dpcg", "bicgstab": "orc_bicgstab",
            "gauss-seidel": "orc_gauss_seidel", "iccg-l2": "orc_iccg_l2"}


def csr_struct(ia, ja, diag):
    keep = (np.ascontiguousarray(ia, np.int32), np.ascontiguousarray(ja, np.int32),
            np.ascontiguousarray(diag, np.int32))
    cc = _Csr(len(keep[0]) - 1, len(keep[1]), _p(keep[0]), _p(keep[1]), _p(keep[2]))
    return cc, keep


def solve_csr(cc, solver, a, rhs, fi0, itr_max, tol_abs, tol_rel, history, n, mr=4):
    fi = np.zeros(n) if fi0 is None else np.array(fi0, dtype=np.float64, copy=True)
    info = _Info()
    a = np.ascontiguousarray(a, np.float64)
    rhs = np.ascontiguousarray(rhs, np.float64)
    if solver == "pmgmres":
        # csrsolve hard-codes the restart length mr = 4 (linear_solvers.f90:85)
        hist = np.zeros(max(itr_max * mr, 1))
        rc = lib().orc_pmgmres_ilu(C.byref(cc), _p(a), _p(fi), _p(rhs), int(itr_max), int(mr),
                                   C.c_double(tol_abs), C.c_double(tol_rel), _p(hist),
                                   C.byref(info))
        assert rc >= 0, f"ILU(0) failed at row {-rc - 1}"
    else:
        hist = np.zeros(max(itr_max, 1))
        fn = getattr(lib(), _SOLVERS[solver])
        fn(C.byref(cc), _p(a), _p(fi), _p(rhs), int(itr_max), C.c_double(tol_abs),
           C.c_double(tol_rel), _p(hist), C.byref(info))
    out = dict(fi=fi, iters=info.iters, resor=info.resor, res0=info.res0, resl=info.resl,
               factor=info.factor)
    if history:
        out["hist"] = hist[: info.iters].copy()
    return out


# -- explicit-CSR helpers (0-based ia/ja/diag) -----------------------------
def ilu0(ia, ja, diag, a):
    """ILU(0) factors of ilu_cr (linear_solvers.f90:1267-1385)."""
    cc, keep = csr_struct(ia, ja, diag)
    l = np.zeros(len(keep[1]))
    rc = lib().orc_ilu0(C.byref(cc), _p(np.ascontiguousarray(a, np.float64)), _p(l))
    assert rc == 0, f"ILU(0) failed at row {-rc - 1}"
    return l


def ilu_apply(ia, ja, diag, l, r):
    """z = (LU)^{-1} r of lus_cr (linear_solvers.f90:1386-1477)."""
    cc, keep = csr_struct(ia, ja, diag)
    z = np.zeros(len(keep[0]) - 1)
    lib().orc_ilu_apply(C.byref(cc), _p(np.ascontiguousarray(l, np.float64)),
                        _p(np.ascontiguousarray(r, np.float64)), _p(z))
    return z


def atx(ia, ja, diag, a, x):
    """y = A^T x of atx_cr (linear_solvers.f90:1079-1171)."""
    cc, keep = csr_struct(ia, ja, diag)
    y = np.zeros(len(keep[0]) - 1)
    lib().orc_atx(C.byref(cc), _p(np.ascontiguousarray(a, np.float64)),
                  _p(np.ascontiguousarray(x, np.float64)), _p(y))
    return y


def solve_explicit(solver, ia, ja, diag, a, rhs, fi0=None, itr_max=1000, tol_abs=0.0,
                   tol_rel=1e-10, history=False, mr=4):
    cc, keep = csr_struct(ia, ja, diag)
    return solve_csr(cc, solver, a, rhs, fi0, itr_max, tol_abs, tol_rel, history,
                     len(keep[0]) - 1, mr)
