"""fcp_dpcg_noalloc, the allocation-free variant of the explicit-array dpcg
entry point, against fcp_dpcg itself: two calls in a row (the second on a
smaller system, so stale scratch contents of the first would show), fixed
iteration counts (the same Krylov trajectory up to reduction order) and
converged solves; and its rejections (too many rows, host arrays, index
arrays that are not a CSR matrix), which happen before any solve kernel."""
import ctypes as C

import numpy as np
import pytest
import torch

from fcp.mesh import hex_mesh
from oracle import Oracle

pytestmark = pytest.mark.gpu


def _system(N, seed):
    mesh = hex_mesh(N, jitter=0.2, seed=seed)
    o = Oracle(mesh)
    n = mesh.n_cells
    mu = -(10.0 ** np.sin(np.arange(n) * 0.3))
    a, _ = o.laplacian(mu, np.zeros(mesh.n_total))
    b = np.cos(np.arange(n) * 0.1) * mesh.n_cells ** -0.5
    return o.ia.copy(), o.ja.copy(), o.diag.copy(), a, b


def _lib():
    from fcp import _lib as L
    return L.load()


def _noalloc(sysm, itr_max, tol_abs, tol_rel, ja_override=None):
    ia, ja, diag, a, b = sysm
    n, nnz = len(diag), len(ja)
    dev = torch.device("cuda:0")
    T = lambda v, dt: torch.from_numpy(np.ascontiguousarray(v)).to(dev, dt)  # noqa: E731
    ia1 = T(ia + 1, torch.int32)
    ja1 = T((ja if ja_override is None else ja_override) + 1, torch.int32)
    dg1 = T(diag + 1, torch.int32)
    ad, bd = T(a, torch.float64), T(b, torch.float64)
    x = torch.zeros(n, dtype=torch.float64, device=dev)
    resor = C.c_double(-1.0)
    torch.cuda.synchronize()
    rc = _lib().fcp_dpcg_noalloc(n, nnz, ia1.data_ptr(), ja1.data_ptr(), ad.data_ptr(),
                                 dg1.data_ptr(), x.data_ptr(), bd.data_ptr(), C.byref(resor),
                                 int(itr_max), float(tol_abs), float(tol_rel),
                                 torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return rc, x.cpu().numpy(), resor.value


def _existing(sysm, itr_max, tol_abs, tol_rel):
    from fcp import api
    ia, ja, diag, a, b = sysm
    x = np.zeros(len(diag))
    rc, resor = api.explicit_solve("dpcg", ia, ja, a, diag, x, b, itr_max, tol_abs, tol_rel)
    return rc, x, resor


@pytest.mark.parametrize("N,seed", [(24, 3), (16, 5)])
def test_noalloc_dpcg_matches_dpcg(N, seed):
    print(f"[case] no-alloc dpcg {N}^3 seed {seed}", flush=True)
    sysm = _system(N, seed)
    for itr_max, tol_rel in ((1, 0.0), (7, 0.0), (40, 0.0), (5000, 1e-10)):
        rc0, x0, r0 = _existing(sysm, itr_max, 0.0, tol_rel)
        rc1, x1, r1 = _noalloc(sysm, itr_max, 0.0, tol_rel)
        assert rc1 == rc0, (itr_max, rc0, rc1)
        scale = np.abs(x0).max()
        tol = 1e-11 if tol_rel == 0.0 else 1e-9
        assert np.abs(x1 - x0).max() <= tol * scale, (itr_max, np.abs(x1 - x0).max(), scale)
        assert abs(r1 - r0) <= 1e-12 * abs(r0), (r0, r1)


def test_noalloc_two_calls_in_a_row_second_smaller():
    print("[case] no-alloc dpcg 24^3 then 16^3", flush=True)
    big, small = _system(24, 3), _system(16, 5)
    for sysm in (big, small):
        rc0, x0, r0 = _existing(sysm, 5000, 0.0, 1e-10)
        rc1, x1, r1 = _noalloc(sysm, 5000, 0.0, 1e-10)
        assert rc0 == 0 and rc1 == 0
        assert np.abs(x1 - x0).max() <= 1e-9 * np.abs(x0).max()
        assert abs(r1 - r0) <= 1e-12 * abs(r0)


def test_noalloc_rejections():
    print("[case] no-alloc dpcg rejections", flush=True)
    lib = _lib()
    sysm = _system(6, 1)
    ia, ja, diag, a, b = sysm
    n, nnz = len(diag), len(ja)
    maxrows = 1 << 18
    d = torch.zeros(8, dtype=torch.float64, device="cuda:0")
    i = torch.ones(8, dtype=torch.int32, device="cuda:0")
    resor = C.c_double()
    rc = lib.fcp_dpcg_noalloc(maxrows + 1, 8, i.data_ptr(), i.data_ptr(), d.data_ptr(),
                              i.data_ptr(), d.data_ptr(), d.data_ptr(), C.byref(resor), 10, 0.0,
                              1e-8, None)
    assert rc == -1 and b"FCP_NOALLOC_MAX_ROWS" in lib.fcp_last_error(None)
    # host arrays
    h = [np.ascontiguousarray(v) for v in (ia + 1, ja + 1, a, diag + 1, np.zeros(n), b)]
    h[0], h[1], h[3] = (v.astype(np.int32) for v in (h[0], h[1], h[3]))
    rc = lib.fcp_dpcg_noalloc(n, nnz, *[v.ctypes.data for v in h], C.byref(resor), 10, 0.0,
                              1e-8, None)
    assert rc == -1 and b"device arrays" in lib.fcp_last_error(None)
    # a column index out of range, and a diagonal pointer to an off-diagonal
    bad = ja.copy()
    bad[3] = n + 4
    rc, _, _ = _noalloc(sysm, 10, 0.0, 1e-8, ja_override=bad)
    assert rc == -1 and b"CSR" in lib.fcp_last_error(None)
    bad = ja.copy()
    bad[diag[2]] = 0
    rc, _, _ = _noalloc(sysm, 10, 0.0, 1e-8, ja_override=bad)
    assert rc == -1
    # and the next valid call works
    rc0, x0, _ = _existing(sysm, 5000, 0.0, 1e-10)
    rc1, x1, _ = _noalloc(sysm, 5000, 0.0, 1e-10)
    assert rc0 == rc1 == 0 and np.abs(x1 - x0).max() <= 1e-9 * np.abs(x0).max()
