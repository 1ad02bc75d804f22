"""Python mirror of the reference's Fortran module API on top of libfcp.

Names and argument meaning follow the reference: ``read_mesh`` hands a mesh
over (geometry.f90:90), ``create_CSR_matrix`` (sparse_matrix.f90:86),
``laplacian(mu, phi)`` (laplacian.f90:1), ``grad_gauss(u)``
(gradients.f90:1607), ``csrsolve(solver, fi, rhs, ...)`` (linear_solvers.f90:40)
and the explicit-array ``iccg``/``dpcg``/``bicgstab`` (:364, :206, :548).
Arrays are numpy (staged through device memory) or torch CUDA tensors
(used in place).  Indices in this module are 0-based; the C-ABI is 1-based
like the Fortran API and the conversion happens here.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .mesh import Mesh

FCP_PREC_FP64 = 0
FCP_PREC_MIXED = 1
FCP_PREC_FP64_TILE = 2  # FP64 iterations, the mixed path's two-level tile preconditioner
FCP_CRIT_REF_L1 = 0
FCP_CRIT_TRUE_L2 = 1
FCP_GRAD_GAUSS = 0
FCP_GRAD_LSQ = 1
FCP_GRAD_LSQ_DM = 2
FCP_GRAD_LSQ_QR = 3
FCP_GRAD_GAUSS_FVX = 4
FCP_GRAD_GAUSS_CORRECTED = 5
FCP_TUNE_SWEEP_WAVES = 1
FCP_TUNE_SWEEP_SLEEP = 2
FCP_TUNE_SWEEP_KIND = 3
FCP_TUNE_SWEEP_ROWS = 4
FCP_TUNE_SWEEP_TILE = 5
FCP_TUNE_INNER_PRECOND = 6
FCP_TUNE_INNER_TILE = 7
FCP_TUNE_GRAPH = 8
FCP_TUNE_INNER_COARSE = 9
FCP_TUNE_SPMV = 10
FCP_TUNE_TILE_PACK = 11
FCP_TUNE_COARSE_FUSE = 12
FCP_TUNE_ASM_BLOCKS = 13


class FcpError(RuntimeError):
    pass


def _ptr(a):
    """(void*, keepalive) for a numpy array or a torch tensor (host or CUDA)."""
    if a is None:
        return None, None
    if isinstance(a, np.ndarray):
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError("array must be C-contiguous")
        return C.c_void_p(a.ctypes.data), a
    if hasattr(a, "data_ptr"):
        if not a.is_contiguous():
            raise ValueError("tensor must be contiguous")
        return C.c_void_p(a.data_ptr()), a
    raise TypeError(f"unsupported array type {type(a)}")


def _f64(a):
    if isinstance(a, np.ndarray):
        return np.ascontiguousarray(a, dtype=np.float64)
    return a


class Context:
    """One libfcp context (device-resident mesh, geometry, CSR and matrix)."""

    def __init__(self, device: int = 0):
        self.lib = _lib.load()
        h = C.c_void_p()
        rc = self.lib.fcp_init(device, C.byref(h))
        if rc != 0:
            raise FcpError(f"fcp_init({device}) failed with {rc}")
        self.h = h
        self.mesh = None
        self.nnz = 0

    def close(self):
        if self.h:
            self.lib.fcp_finalize(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc < 0:
            msg = self.lib.fcp_last_error(self.h)
            raise FcpError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")
        return rc

    @property
    def stream(self):
        return self.lib.fcp_stream(self.h)

    # -- mesh --------------------------------------------------------------
    def read_mesh(self, mesh: Mesh):
        """Upload a mesh (geometry.f90:90 read_mesh hand-over)."""
        self.mesh = mesh
        node1 = np.where(mesh.node >= 0, mesh.node + 1, 0).astype(np.int32)
        arrs = [np.ascontiguousarray(mesh.points, dtype=np.float64).reshape(-1),
                np.ascontiguousarray(mesh.nnodes, dtype=np.int32),
                np.ascontiguousarray(node1.reshape(-1)),
                np.ascontiguousarray(mesh.owner + 1, dtype=np.int32),
                np.ascontiguousarray(mesh.neighbour + 1, dtype=np.int32),
                np.ascontiguousarray(mesh.nfaces, dtype=np.int32),
                np.ascontiguousarray(mesh.start_face, dtype=np.int32),
                np.ascontiguousarray(mesh.bctype_codes, dtype=np.int32)]
        ptrs = [_ptr(a)[0] for a in arrs]
        rc = self.lib.fcp_mesh_upload(self.h, mesh.n_nodes, mesh.n_cells, mesh.n_inner, mesh.n_bnd,
                                      mesh.n_patches, *ptrs)
        self._check(rc, "fcp_mesh_upload")
        if mesh.twin_start:  # periodic patches (geometry.f90:251-258)
            tw = np.ascontiguousarray(mesh.twin_start, dtype=np.int32)
            self._check(self.lib.fcp_mesh_periodic(self.h, len(tw), _ptr(tw)[0]),
                        "fcp_mesh_periodic")

    def read_mesh_file(self, case_dir: str, fmt: str = "native"):
        """read_mesh from <case_dir>/polyMesh in the native or OpenFOAM format
        (parsed by the library).  Returns (numNodes, numCells, numInnerFaces,
        numBoundaryFaces, numBoundaries) and the patches (nfaces, startFace,
        bctype codes)."""
        sizes = np.zeros(5, np.int32)
        b = case_dir.encode()
        rc = self.lib.fcp_read_polymesh(self.h, b, len(b), {"native": 0, "foam": 1}[fmt],
                                    _ptr(sizes)[0])
        self._check(rc, "fcp_read_polymesh")
        nb = int(sizes[4])
        nf, sf, bt = (np.zeros(max(nb, 1), np.int32) for _ in range(3))
        self._check(self.lib.fcp_mesh_patches(self.h, _ptr(nf)[0], _ptr(sf)[0], _ptr(bt)[0]),
                    "fcp_mesh_patches")
        self.mesh = None
        return tuple(int(v) for v in sizes), (nf[:nb], sf[:nb], bt[:nb])

    def geometry(self):
        m = self.mesh
        n, F, Fi, Fb = m.n_cells, m.n_faces, m.n_inner, m.n_bnd
        sizes = dict(xc=n, yc=n, zc=n, vol=n, arx=F, ary=F, arz=F, xf=F, yf=F, zf=F, facint=Fi,
                     Df=Fi, dnb=Fb, srdb=Fb)
        out = {k: np.zeros(v) for k, v in sizes.items()}
        rc = self.lib.fcp_geometry(self.h, *[_ptr(out[k])[0] for k in sizes])
        self._check(rc, "fcp_geometry")
        return out

    def compute_geometry(self):
        """Geometry on the device without copying anything back."""
        rc = self.lib.fcp_geometry(self.h, *([None] * 14))
        self._check(rc, "fcp_geometry")

    def create_CSR_matrix(self, want_arrays: bool = True):
        m = self.mesh
        nnz = C.c_int()
        if not want_arrays:
            self._check(self.lib.fcp_csr_build(self.h, C.byref(nnz), None, None, None, None, None),
                        "fcp_csr_build")
            self.nnz = nnz.value
            return None
        npr = int(sum(int(m.nfaces[ib]) for ib in m.periodic_patches))
        nn = 2 * (m.n_inner + npr) + m.n_cells
        ia = np.zeros(m.n_cells + 1, np.int32)
        ja = np.zeros(nn, np.int32)
        dg = np.zeros(m.n_cells, np.int32)
        spn = np.zeros(max(m.n_inner + npr, 1), np.int32)
        snp = np.zeros(max(m.n_inner + npr, 1), np.int32)
        rc = self.lib.fcp_csr_build(self.h, C.byref(nnz), *[_ptr(a)[0] for a in (ia, ja, dg, spn, snp)])
        self._check(rc, "fcp_csr_build")
        self.nnz = nnz.value
        return dict(ia=ia - 1, ja=ja - 1, diag=dg - 1, slot_pn=spn[: m.n_inner + npr] - 1,
                    slot_np=snp[: m.n_inner + npr] - 1)

    def levels(self, upper: bool = False):
        lev = np.zeros(self.mesh.n_cells, np.int32)
        nlev = C.c_int()
        self._check(self.lib.fcp_levels(self.h, int(upper), _ptr(lev)[0], C.byref(nlev)), "fcp_levels")
        return lev, nlev.value

    # -- assembly ----------------------------------------------------------
    def laplacian(self, mu, phi, su, a_out=None):
        """laplacian(mu,phi): resident matrix; su updated in place."""
        mu, phi = _f64(mu), _f64(phi)
        rc = self.lib.fcp_laplacian(self.h, _ptr(mu)[0], _ptr(phi)[0], _ptr(a_out)[0], _ptr(su)[0])
        self._check(rc, "fcp_laplacian")
        return a_out

    def laplacian_face(self, mu_f, bc_kind, phi, su, a_out=None):
        mu_f, phi = _f64(mu_f), _f64(phi)
        if isinstance(bc_kind, np.ndarray):
            bc_kind = np.ascontiguousarray(bc_kind, dtype=np.uint8)
        rc = self.lib.fcp_laplacian_face(self.h, _ptr(mu_f)[0], _ptr(bc_kind)[0], _ptr(phi)[0],
                                         _ptr(a_out)[0], _ptr(su)[0])
        self._check(rc, "fcp_laplacian_face")
        return a_out

    def nonorth_correction(self, gamma_f, grad, su):
        rc = self.lib.fcp_nonorth_correction(self.h, _ptr(_f64(gamma_f))[0], _ptr(_f64(grad))[0],
                                             _ptr(su)[0])
        self._check(rc, "fcp_nonorth_correction")

    def convdiff(self, flmass, dcoef, phi, grad, gam, scheme, bc_kind, su, sp, urf=1.0,
                 a_out=None):
        """facefluxsc assembly (scheme 0 'cds', 1 'linearUpwind'); su/sp
        accumulate, phi boundary values of zero-gradient faces are updated."""
        if isinstance(bc_kind, np.ndarray):
            bc_kind = np.ascontiguousarray(bc_kind, dtype=np.uint8)
        rc = self.lib.fcp_convdiff(self.h, _ptr(_f64(flmass))[0], _ptr(_f64(dcoef))[0],
                                   _ptr(phi)[0], _ptr(_f64(grad))[0], float(gam), int(scheme),
                                   _ptr(bc_kind)[0], float(urf), _ptr(a_out)[0], _ptr(su)[0],
                                   _ptr(sp)[0])
        self._check(rc, "fcp_convdiff")
        return a_out

    def convdiff_cell(self, flmass, diffcoef, phi, grad, gam, scheme, bc_kind, su, sp, urf=1.0,
                      a_out=None):
        """convdiff with facefluxscalar (scalar_fluxes.f90:344-464): the
        diffusivity per cell, interpolated to each inner face; a boundary face
        takes its owner's."""
        if isinstance(bc_kind, np.ndarray):
            bc_kind = np.ascontiguousarray(bc_kind, dtype=np.uint8)
        rc = self.lib.fcp_convdiff_cell(self.h, _ptr(_f64(flmass))[0], _ptr(_f64(diffcoef))[0],
                                        _ptr(phi)[0], _ptr(_f64(grad))[0], float(gam), int(scheme),
                                        _ptr(bc_kind)[0], float(urf), _ptr(a_out)[0],
                                        _ptr(su)[0], _ptr(sp)[0])
        self._check(rc, "fcp_convdiff_cell")
        return a_out

    def momentum_assemble(self, u, v, w, p, den, vis, flmass, viscos, urf, uo=None, vo=None,
                          wo=None, uoo=None, voo=None, woo=None, visw=None, dt=1.0,
                          time_scheme="steady", conv_scheme="linearUpwind", gds=1.0, dpdxi=None,
                          grad=None, rhs=None, diag=None, ap=None, spv=None, src=None, a_out=None):
        """calcuvw's assembly (velocity.f90:50-750) before its solves.  u, v, w,
        p updated in place (boundary values); returns (rhs, diag), each (3, n),
        for csrsolve_multi(..., diag=diag)."""
        n = self.mesh.n_cells
        if rhs is None:
            rhs = np.zeros((3, n))
        if diag is None:
            diag = np.zeros((3, n))
        ts = self.TIME_SCHEMES[time_scheme] if isinstance(time_scheme, str) else int(time_scheme)
        cs = {"cds": 0, "linearUpwind": 1}[conv_scheme] if isinstance(conv_scheme, str) \
            else int(conv_scheme)
        urf = np.ascontiguousarray(urf, np.float64)
        args = [u, v, w] + [_f64(x) if x is not None else None for x in (uo, vo, wo, uoo, voo, woo)]
        rc = self.lib.fcp_momentum_assemble(
            self.h, *[_ptr(x)[0] for x in args], _ptr(p)[0], _ptr(_f64(den))[0],
            _ptr(_f64(vis))[0], _ptr(_f64(visw))[0], _ptr(_f64(flmass))[0], float(viscos),
            float(dt), ts, cs, float(gds), _ptr(urf)[0], _ptr(dpdxi)[0], _ptr(grad)[0],
            _ptr(rhs)[0], _ptr(diag)[0], _ptr(ap)[0], _ptr(spv)[0], _ptr(src)[0], _ptr(a_out)[0])
        self._check(rc, "fcp_momentum_assemble")
        return rhs, diag

    def pressure_assemble(self, u, v, w, p, dpdxi, den, apu, apv, apw, flmass, su,
                          const_mflux=True, flomas=0.0, a_out=None):
        """calcp_simple's pressure-correction assembly (calcp_simple.f90:
        69-228); flmass (and, without const_mflux, the outlet values of u, v,
        w) updated in place, su overwritten; the matrix stays resident."""
        rc = self.lib.fcp_pressure_assemble(
            self.h, _ptr(u)[0], _ptr(v)[0], _ptr(w)[0], _ptr(_f64(p))[0], _ptr(_f64(dpdxi))[0],
            _ptr(_f64(den))[0], _ptr(_f64(apu))[0], _ptr(_f64(apv))[0], _ptr(_f64(apw))[0],
            _ptr(flmass)[0], 1 if const_mflux else 0,
            float(flomas), _ptr(a_out)[0], _ptr(su)[0])
        self._check(rc, "fcp_pressure_assemble")
        return a_out

    def pressure_correct(self, pp, flmass, u, v, w, p, dpdxi, apu, apv, apw, urfp, pref_cell,
                         su=None, sv=None, sw=None):
        """calcp_simple's correction from the solved pp (calcp_simple.f90:
        333-419); pref_cell 1-based.  pp, flmass, u, v, w, p, dpdxi in/out."""
        rc = self.lib.fcp_pressure_correct(
            self.h, _ptr(pp)[0], _ptr(flmass)[0], _ptr(u)[0], _ptr(v)[0], _ptr(w)[0], _ptr(p)[0],
            _ptr(dpdxi)[0], _ptr(_f64(apu))[0], _ptr(_f64(apv))[0], _ptr(_f64(apw))[0],
            float(urfp), int(pref_cell), _ptr(su)[0], _ptr(sv)[0], _ptr(sw)[0])
        self._check(rc, "fcp_pressure_correct")

    def ddt(self, scheme, den, dt, phi_o, su, sp, phi_oo=None):
        rc = self.lib.fcp_ddt(self.h, int(scheme), _ptr(_f64(den))[0], float(dt),
                              _ptr(_f64(phi_o))[0], _ptr(_f64(phi_oo))[0], _ptr(su)[0], _ptr(sp)[0])
        self._check(rc, "fcp_ddt")

    TIME_SCHEMES = {"steady": 0, "bdf": 1, "bdf2": 2, "cn": 3}

    def transport(self, phi, vis, flmass, wall_frac, viscos, pranl, sigt, urf, gds, solver,
                  itr_max, tol_abs, tol_rel, time_scheme="steady", den=None, dt=1.0, phi_o=None,
                  phi_oo=None, conv_scheme="linearUpwind", grad_method=FCP_GRAD_GAUSS, clip=True,
                  grad_out=None, a_out=None, su_out=None, sp_out=None, precision=FCP_PREC_FP64,
                  criterion=FCP_CRIT_REF_L1):
        """calculate_temperature for one scalar field: phi[numTotal] updated in
        place; grad_out, a_out, su_out, sp_out (optional) receive the gradient
        and the assembled system.  Returns (status, info) with the solver's
        resor (csrsolve's res0), the cell range before the clip and the solve
        statistics."""
        opts = _lib.SolveOpts()
        self.lib.fcp_solve_opts_default(C.byref(opts))
        opts.precision = precision
        opts.criterion = criterion
        resor, lo, hi = C.c_double(), C.c_double(), C.c_double()
        sb = solver.encode()
        wf = np.ascontiguousarray(wall_frac, dtype=np.float64)
        rc = self.lib.fcp_transport(
            self.h, _ptr(phi)[0], _ptr(_f64(phi_o))[0], _ptr(_f64(phi_oo))[0], _ptr(_f64(den))[0],
            _ptr(_f64(vis))[0], _ptr(_f64(flmass))[0], _ptr(wf)[0], float(viscos), float(pranl),
            float(sigt), float(urf), float(gds), float(dt), self.TIME_SCHEMES[time_scheme],
            {"cds": 0, "linearUpwind": 1}[conv_scheme], int(grad_method), sb, len(sb),
            int(itr_max), float(tol_abs), float(tol_rel), int(bool(clip)), _ptr(grad_out)[0],
            _ptr(a_out)[0], _ptr(su_out)[0], _ptr(sp_out)[0], C.byref(resor), C.byref(lo),
            C.byref(hi), C.byref(opts))
        self._check(rc, "fcp_transport")
        info = {k: getattr(opts, k) for k, _ in _lib.SolveOpts._fields_}
        info.update(status=rc, resor=resor.value, fimin=lo.value, fimax=hi.value)
        return rc, info

    def grad_gauss(self, u, out=None):
        if out is None:
            out = np.zeros((self.mesh.n_total, 3))
        rc = self.lib.fcp_grad_gauss(self.h, _ptr(_f64(u))[0], _ptr(out)[0])
        self._check(rc, "fcp_grad_gauss")
        return out

    # -- discrete adjoint building blocks (outputs accumulate) --------------
    def matrix_transpose(self):
        self._check(self.lib.fcp_matrix_transpose(self.h), "fcp_matrix_transpose")

    def convdiff_backward(self, flmass, dcoef, phi_old, grad_old, gam, scheme, bc_kind, lam,
                          phi_new, dcoef_bar, phi_old_bar, grad_old_bar):
        if isinstance(bc_kind, np.ndarray):
            bc_kind = np.ascontiguousarray(bc_kind, dtype=np.uint8)
        rc = self.lib.fcp_convdiff_backward(
            self.h, _ptr(_f64(flmass))[0], _ptr(_f64(dcoef))[0], _ptr(_f64(phi_old))[0],
            _ptr(_f64(grad_old))[0], float(gam), int(scheme), _ptr(bc_kind)[0], _ptr(_f64(lam))[0],
            _ptr(_f64(phi_new))[0], _ptr(dcoef_bar)[0], _ptr(phi_old_bar)[0], _ptr(grad_old_bar)[0])
        self._check(rc, "fcp_convdiff_backward")

    def laplacian_face_backward(self, mu_f, bc_kind, phi_bc, gamma_f, grad, lam, phi, mu_f_bar,
                                phi_bc_bar, gamma_bar=None, grad_bar=None):
        if isinstance(bc_kind, np.ndarray):
            bc_kind = np.ascontiguousarray(bc_kind, dtype=np.uint8)
        rc = self.lib.fcp_laplacian_face_backward(
            self.h, _ptr(_f64(mu_f))[0], _ptr(bc_kind)[0], _ptr(_f64(phi_bc))[0],
            _ptr(_f64(gamma_f))[0], _ptr(_f64(grad))[0], _ptr(_f64(lam))[0], _ptr(_f64(phi))[0],
            _ptr(mu_f_bar)[0], _ptr(phi_bc_bar)[0], _ptr(gamma_bar)[0], _ptr(grad_bar)[0])
        self._check(rc, "fcp_laplacian_face_backward")

    def grad_gauss_backward(self, grad_bar, phi_bar):
        rc = self.lib.fcp_grad_gauss_backward(self.h, _ptr(_f64(grad_bar))[0], _ptr(phi_bar)[0])
        self._check(rc, "fcp_grad_gauss_backward")

    def rcm(self, pp_iters: int = 4):
        """GPU reverse Cuthill-McKee: new 0-based number of every cell."""
        perm = np.zeros(self.mesh.n_cells, np.int32)
        self._check(self.lib.fcp_rcm(self.h, int(pp_iters), _ptr(perm)[0]), "fcp_rcm")
        return perm - 1

    def create_lsq_grad_matrix(self, method: int, compat: bool = True, want: bool = False):
        """create_lsq_grad_matrix for method 1 lsq, 2 lsq_dm, 3 lsq_qr."""
        out = None
        if want:
            out = np.zeros((self.mesh.n_cells, 18 if method == 3 else 9))
        rc = self.lib.fcp_create_lsq_grad_matrix(self.h, int(method), int(bool(compat)),
                                                 _ptr(out)[0])
        self._check(rc, "fcp_create_lsq_grad_matrix")
        return out

    def grad(self, method: int, phi, out=None, gold=None):
        """grad(phi, dPhidxi): method 0 Gauss, 1-3 least squares, 4 the
        two-sweep Gauss of fvxGradient.f90 (:1549-1645), 5 its
        grad_gauss_corrected (:1664-1758) started from gold (n, 3) -- or, with
        gold None, from the cell rows of out on entry."""
        if out is None:
            out = np.zeros((self.mesh.n_total, 3))
        if gold is not None:
            n = self.mesh.n_cells
            out[:n] = np.asarray(gold, np.float64).reshape(-1, 3) if isinstance(out, np.ndarray) \
                else gold.reshape(-1, 3)
        rc = self.lib.fcp_grad(self.h, int(method), _ptr(_f64(phi))[0], _ptr(out)[0])
        self._check(rc, "fcp_grad")
        return out

    # -- LIS interface (LIS_linear_solvers.f90:38-200) ----------------------
    def lis_create_solver(self, options: str):
        b = options.encode()
        self._check(self.lib.fcp_lis_create_solver(self.h, b, len(b)), "fcp_lis_create_solver")

    def lis_spsolve(self, fi, rhs, chvar: str = "x"):
        """lis_spsolve on the resident matrix: fi (current field in, solution
        out).  Returns (status, dict(res0, resid, iter))."""
        res0, resid, it = C.c_double(), C.c_double(), C.c_int()
        b = chvar.encode()
        rc = self.lib.fcp_lis_spsolve(self.h, _ptr(fi)[0], _ptr(_f64(rhs))[0], b, len(b),
                                      C.byref(res0), C.byref(resid), C.byref(it))
        self._check(rc, "fcp_lis_spsolve")
        return rc, dict(res0=res0.value, resid=resid.value, iter=it.value)

    def lis_destroy_solver(self):
        self._check(self.lib.fcp_lis_destroy_solver(self.h), "fcp_lis_destroy_solver")

    def update_boundary(self, phi):
        self._check(self.lib.fcp_update_boundary(self.h, _ptr(phi)[0]), "fcp_update_boundary")
        return phi

    # -- linear algebra ----------------------------------------------------
    def matrix_set(self, a):
        self._check(self.lib.fcp_matrix_set(self.h, _ptr(_f64(a))[0]), "fcp_matrix_set")

    def spmv(self, x, y=None):
        if y is None:
            y = np.zeros(self.mesh.n_cells)
        self._check(self.lib.fcp_spmv(self.h, _ptr(_f64(x))[0], _ptr(y)[0]), "fcp_spmv")
        return y

    def spmv_transpose(self, x, y=None):
        """y = A^T x (atx_cr)."""
        if y is None:
            y = np.zeros(self.mesh.n_cells)
        self._check(self.lib.fcp_spmv_transpose(self.h, _ptr(_f64(x))[0], _ptr(y)[0]),
                    "fcp_spmv_transpose")
        return y

    def autotune(self, solver: str = "iccg"):
        """fcp_autotune: picks the tile size of the FP32 inner solves'
        preconditioner for this mesh; returns (tile_rows, cached)."""
        t, c = C.c_int(0), C.c_int(0)
        sb = solver.encode()
        self._check(self.lib.fcp_autotune(self.h, sb, len(sb), C.byref(t), C.byref(c)),
                    "fcp_autotune")
        return t.value, bool(c.value)

    def sweep_tiles(self):
        """Tile of every cell in the tile-local preconditioner (0-based)."""
        t = np.zeros(self.mesh.n_cells, np.int32)
        self._check(self.lib.fcp_sweep_tiles(self.h, _ptr(t)[0]), "fcp_sweep_tiles")
        return t

    def precond_apply(self, r, kind: int = 0):
        """z = M^{-1} r: kind 0 DIC, 1 DILU (returns (z, inverted pivots)),
        2 ILU(0) of pmgmres_ilu (returns (z, None)), 3 / 4 tile-local DIC /
        DILU of the FP32 inner solves in FP64, 5 / 6 the same in FP32 by the
        inner solves' kernel, 7 kind 5 plus the coarse correction, 8 / 9
        kinds 5 / 7 by the same kernel in FP64 (FCP_PREC_FP64_TILE's
        preconditioner) (returns (z, pivots))."""
        n = self.mesh.n_cells
        z = np.zeros(n)
        d = np.zeros(n) if kind != 2 else None
        rc = self.lib.fcp_precond_apply(self.h, kind, _ptr(_f64(r))[0], _ptr(z)[0],
                                        _ptr(d)[0] if d is not None else None)
        self._check(rc, "fcp_precond_apply")
        return z, d

    def ilu_factor(self, nnz):
        """ILU(0) factors of the resident matrix (ilu_cr layout, CSR order)."""
        l = np.zeros(nnz)
        self._check(self.lib.fcp_ilu_factor(self.h, _ptr(l)[0]), "fcp_ilu_factor")
        return l

    def csrsolve(self, solver: str, fi, rhs, itr_max=1000, tol_abs=0.0, tol_rel=1e-10,
                 chvar="phi", verbose=0, precision=FCP_PREC_FP64, criterion=FCP_CRIT_REF_L1,
                 inner_tol=1e-4, max_outer=30, check_every=8):
        """csrsolve on the resident matrix; fi updated in place.  Returns
        (status, info dict)."""
        opts = _lib.SolveOpts()
        self.lib.fcp_solve_opts_default(C.byref(opts))
        opts.precision = precision
        opts.criterion = criterion
        opts.inner_tol = inner_tol
        opts.max_outer = max_outer
        opts.check_every = check_every
        res0 = C.c_double()
        sb = solver.encode()
        cb = chvar.encode()
        rc = self.lib.fcp_csrsolve(self.h, sb, len(sb), _ptr(fi)[0], _ptr(_f64(rhs))[0],
                                   C.byref(res0), int(itr_max), float(tol_abs), float(tol_rel), cb,
                                   len(cb), int(verbose), C.byref(opts))
        self._check(rc, "fcp_csrsolve")
        info = {k: getattr(opts, k) for k, _ in _lib.SolveOpts._fields_}
        info["status"] = rc
        return rc, info

    def csrsolve_multi(self, solver: str, fi, rhs, itr_max=1000, tol_abs=0.0, tol_rel=1e-10,
                       verbose=0, precision=FCP_PREC_FP64, criterion=FCP_CRIT_REF_L1,
                       inner_tol=1e-4, max_outer=30, check_every=8, diag=None):
        """Batched csrsolve of nrhs columns sharing the resident matrix.  fi and
        rhs are (nrhs, n) row-major arrays (numpy or torch), i.e. Fortran
        fi(n, nrhs); fi is updated in place.  diag (nrhs, n), optional: each
        column's own diagonal (fcp_csrsolve_multi_diag, the momentum
        components' solves).  Returns (status, info, res0[nrhs])."""
        nrhs, n = rhs.shape
        opts = _lib.SolveOpts()
        self.lib.fcp_solve_opts_default(C.byref(opts))
        opts.precision = precision
        opts.criterion = criterion
        opts.inner_tol = inner_tol
        opts.max_outer = max_outer
        opts.check_every = check_every
        res0 = (C.c_double * nrhs)()
        sb = solver.encode()
        if diag is None:
            rc = self.lib.fcp_csrsolve_multi(self.h, sb, len(sb), int(nrhs), _ptr(fi)[0],
                                             int(fi.shape[1]), _ptr(_f64(rhs))[0], res0,
                                             int(itr_max), float(tol_abs), float(tol_rel),
                                             int(verbose), C.byref(opts))
            self._check(rc, "fcp_csrsolve_multi")
        else:
            rc = self.lib.fcp_csrsolve_multi_diag(self.h, sb, len(sb), int(nrhs), _ptr(fi)[0],
                                                  int(fi.shape[1]), _ptr(_f64(rhs))[0],
                                                  _ptr(_f64(diag))[0], res0, int(itr_max),
                                                  float(tol_abs), float(tol_rel), int(verbose),
                                                  C.byref(opts))
            self._check(rc, "fcp_csrsolve_multi_diag")
        info = {k: getattr(opts, k) for k, _ in _lib.SolveOpts._fields_}
        info["status"] = rc
        return rc, info, np.array(res0[:])

    def tune(self, param: int, value: int):
        self._check(self.lib.fcp_tune_set(self.h, int(param), int(value)), "fcp_tune_set")

    def profile(self, enable: bool):
        self._check(self.lib.fcp_profile(self.h, int(enable)), "fcp_profile")

    def profile_report(self):
        buf = C.create_string_buffer(1 << 16)
        self.lib.fcp_profile_report(self.h, buf, len(buf))
        out = {}
        for line in buf.value.decode().splitlines():
            name, cnt, ms = line.split()
            out[name] = (int(cnt), float(ms))
        return out


def explicit_solve(solver: str, ia, ja, a, diag, fi, rhs, itr_max, tol_abs, tol_rel, verbose=0,
                   mr=4):
    """iccg/dpcg/bicgstab/GaussSeidel/pmgmres_ilu(n,nnz,ia,ja,a,diag,fi,rhs,
    resor,...) with 0-based ia/ja/diag (converted to the 1-based C-ABI).  fi
    updated in place; mr is the restart length of pmgmres."""
    lib = _lib.load()
    n = len(diag)
    ia1 = np.ascontiguousarray(np.asarray(ia) + 1, np.int32)
    ja1 = np.ascontiguousarray(np.asarray(ja) + 1, np.int32)
    dg1 = np.ascontiguousarray(np.asarray(diag) + 1, np.int32)
    a = np.ascontiguousarray(a, np.float64)
    rhs = np.ascontiguousarray(rhs, np.float64)
    resor = C.c_double()
    if solver == "pmgmres":
        rc = lib.fcp_pmgmres(n, len(ja1), _ptr(ia1)[0], _ptr(ja1)[0], _ptr(a)[0], _ptr(dg1)[0],
                             _ptr(fi)[0], _ptr(rhs)[0], C.byref(resor), int(itr_max), int(mr),
                             float(tol_abs), float(tol_rel), int(verbose))
    else:
        fn = {"iccg": lib.fcp_iccg, "dpcg": lib.fcp_dpcg, "bicgstab": lib.fcp_bicgstab,
              "gauss-seidel": lib.fcp_gauss_seidel, "gauss_seidel": lib.fcp_gauss_seidel}[solver]
        rc = fn(n, len(ja1), _ptr(ia1)[0], _ptr(ja1)[0], _ptr(a)[0], _ptr(dg1)[0], _ptr(fi)[0],
                _ptr(rhs)[0], C.byref(resor), int(itr_max), float(tol_abs), float(tol_rel),
                int(verbose))
    if rc < 0:
        raise FcpError(f"fcp_{solver} failed with {rc}")
    return rc, resor.value
