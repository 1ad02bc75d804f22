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
                ("start_face", C.c_void_p), ("bctype", C.c_void_p)]


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
            bctype=np.ascontiguousarray(mesh.bctype_codes, dtype=np.int32))
        k = self._keep
        self.cm = _Mesh(mesh.n_nodes, mesh.n_cells, mesh.n_inner, mesh.n_faces, mesh.n_patches,
                        _p(k["points"]), _p(k["nnodes"]), _p(k["node"]), _p(k["owner"]),
                        _p(k["neighbour"]), _p(k["nfaces"]), _p(k["start_face"]), _p(k["bctype"]))
        n, F, Fi = mesh.n_cells, mesh.n_faces, mesh.n_inner
        sizes = dict(xc=n, yc=n, zc=n, vol=n, arx=F, ary=F, arz=F, xf=F, yf=F, zf=F,
                     facint=Fi, Df=Fi, dnb=F - Fi, srdb=F - Fi)
        self.geom = {key: np.zeros(sizes[key]) for key in self.GEOM_KEYS}
        self.cg = _Geom(*[_p(self.geom[key]) for key in self.GEOM_KEYS])
        rc = L.orc_geometry(C.byref(self.cm), C.byref(self.cg))
        assert rc == 0, rc
        nnz = L.orc_csr_nnz(C.byref(self.cm))
        assert nnz > 0, nnz
        self.ia = np.zeros(n + 1, np.int32)
        self.ja = np.zeros(nnz, np.int32)
        self.diag = np.zeros(n, np.int32)
        self.slot_pn = np.zeros(Fi, np.int32)
        self.slot_np = np.zeros(Fi, np.int32)
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
