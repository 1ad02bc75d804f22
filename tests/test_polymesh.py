"""Polyhedral test meshes from merged hexahedra (CPU): the merged cells are
closed, conserve volume, and give CSR rows longer than the eight entries the
SELL products issue together."""
import numpy as np

from fcp.mesh import hex_mesh, merge_cells, staggered_pairs
from oracle import Oracle


def test_merged_cells_closed_and_volume_conserving():
    mesh = hex_mesh(7, jitter=0.2)
    group = staggered_pairs(7)
    pm = merge_cells(mesh, group)
    o, o0 = Oracle(pm), Oracle(mesh)
    g = o.geom
    ar = np.stack([g["arx"], g["ary"], g["arz"]], 1)
    s = np.zeros((pm.n_cells, 3))
    np.add.at(s, pm.owner, ar)
    np.add.at(s, pm.neighbour, -ar[:pm.n_inner])
    assert np.abs(s).max() < 1e-15  # every merged cell is closed
    vol = np.bincount(group, weights=o0.geom["vol"])
    assert np.allclose(g["vol"], vol, rtol=1e-13, atol=0)
    # owner < neighbour, inner faces sorted, one face per pair
    assert np.all(pm.owner[:pm.n_inner] < pm.neighbour)
    key = pm.owner[:pm.n_inner].astype(np.int64) * pm.n_cells + pm.neighbour
    assert np.all(np.diff(key) > 0)
    rows = np.diff(o.ia)
    assert rows.max() == 11 and (rows > 8).sum() > 0
    # the merged pair's centroid is the volume-weighted mean of the two hexes'
    for k in ("xc", "yc", "zc"):
        c = np.bincount(group, weights=o0.geom[k] * o0.geom["vol"]) / vol
        assert np.allclose(g[k], c, rtol=0, atol=1e-14)


def test_merged_poisson_oracle_converges():
    pm = merge_cells(hex_mesh(6, jitter=0.2), staggered_pairs(6))
    o = Oracle(pm)
    from fcp import workloads
    g = o.geom
    inp = workloads.poisson_jittered_inputs(pm, g["vol"], g["xc"], g["yc"], g["zc"])
    a, su = o.laplacian_face(inp["mu_f"], inp["bc_kind"], inp["phi_bc"], inp["source"] * g["vol"])
    r = o.solve("iccg", a, su, itr_max=2000, tol_abs=0.0, tol_rel=1e-12)
    assert np.linalg.norm(su - o.spmv(a, r["fi"])) <= 1e-11 * np.linalg.norm(su)
