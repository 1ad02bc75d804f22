"""Face-based polyhedral meshes: synthetic generators and native polyMesh I/O.

The in-memory layout mirrors the reference's ``geometry`` module globals
(``src/mesh/geometry.f90:19-83``) but is 0-based:

* ``points``     (nNodes, 3) float64
* ``nnodes``     (nFaces,)   int32   number of vertices of each face
* ``node``       (nFaces, NOMAX) int32, 0-based vertex ids, padded with -1
* ``owner``      (nFaces,)   int32   0-based owner cell of every face
* ``neighbour``  (nInner,)   int32   0-based neighbour cell of every inner face
* patches: ``names``, ``types``, ``nfaces``, ``start_face`` where
  ``start_face`` is the 0-based id of the patch's first face, which is the
  same number the reference stores as ``startFace`` (faces of patch ``ib``
  are ``startFace(ib)+1 .. startFace(ib)+nfaces(ib)`` in 1-based numbering,
  ``geometry.f90:78``/``laplacian.f90:78``).

Faces are in the OpenFOAM upper-triangular order the reference's meshes use:
inner faces first, sorted by owner then neighbour with owner < neighbour,
then boundary faces patch by patch.

Synthetic meshes: ``hex_mesh`` builds the unit cube split into N^3 hexahedra,
cell id ``i + N*j + N*N*k``, node id ``i + (N+1)*j + (N+1)^2*k``, patches in
the order x-, x+, y-, y+, z-, z+ with N^2 faces each.  Interior nodes can be
jittered by a seeded uniform offset of +-``jitter*h`` per coordinate;
boundary nodes stay on the cube faces.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np

from .rng import uniform01

NOMAX = 6  # max vertices per face, geometry.f90:40

# Boundary-type codes used across the C-ABI (text names as in polyMesh/boundary,
# geometry.f90:234-261 and updateBoundary.f90:31-44).
BCTYPE_CODES = {
    "wall": 1,
    "symmetry": 2,
    "inlet": 3,
    "outlet": 4,
    "pressure": 5,
    "empty": 6,
    "periodic": 7,
    "wallQFlux": 8,
}


@dataclass
class Mesh:
    points: np.ndarray
    nnodes: np.ndarray
    node: np.ndarray
    owner: np.ndarray
    neighbour: np.ndarray
    n_cells: int
    names: list = field(default_factory=list)
    types: list = field(default_factory=list)
    nfaces: np.ndarray = None
    start_face: np.ndarray = None
    # start face of the twin of every 'periodic' patch, in patch order (the
    # fifth column of a periodic patch's line in polyMesh/boundary,
    # geometry.f90:251-258; 0-based like start_face)
    twin_start: list = field(default_factory=list)

    @property
    def n_nodes(self) -> int:
        return int(self.points.shape[0])

    @property
    def n_faces(self) -> int:
        return int(self.owner.shape[0])

    @property
    def n_inner(self) -> int:
        return int(self.neighbour.shape[0])

    @property
    def n_bnd(self) -> int:
        return self.n_faces - self.n_inner

    @property
    def n_total(self) -> int:
        return self.n_cells + self.n_bnd

    @property
    def n_patches(self) -> int:
        return len(self.names)

    @property
    def bctype_codes(self) -> np.ndarray:
        return np.array([BCTYPE_CODES[t] for t in self.types], dtype=np.int32)

    @property
    def periodic_patches(self) -> list:
        return [ib for ib, t in enumerate(self.types) if t == "periodic"]

    def patch_of_bface(self) -> np.ndarray:
        """Patch index of every boundary face (length nBnd)."""
        out = np.empty(self.n_bnd, dtype=np.int32)
        for ib in range(self.n_patches):
            s = int(self.start_face[ib]) - self.n_inner
            out[s:s + int(self.nfaces[ib])] = ib
        return out

    def check(self) -> None:
        """Structural sanity checks mirroring geometry.f90:276-279."""
        assert self.node.shape == (self.n_faces, NOMAX)
        assert self.nnodes.min() >= 3 and self.nnodes.max() <= NOMAX
        assert int(np.sum(self.nfaces)) == self.n_bnd
        assert self.owner.min() >= 0 and self.owner.max() < self.n_cells
        if self.n_inner:
            assert self.neighbour.min() >= 0 and self.neighbour.max() < self.n_cells
            assert np.all(self.neighbour != self.owner[: self.n_inner])
        per = self.periodic_patches
        assert len(self.twin_start) == len(per)
        for ib, t in zip(per, self.twin_start):  # the twin is a whole patch of equal size
            tw = [jb for jb in range(self.n_patches) if int(self.start_face[jb]) == int(t)]
            assert len(tw) == 1 and int(self.nfaces[tw[0]]) == int(self.nfaces[ib]), (ib, t)


def hex_mesh(N: int, jitter: float = 0.0, seed: int = 2,
             types=("wall",) * 6) -> Mesh:
    """Unit cube, N^3 hexahedra, optional interior-node jitter (fraction of h).
    A 'periodic' patch takes the other patch of its axis as its twin (the
    reference's convention: the twin is listed with its own type, usually
    'empty')."""
    N = int(N)
    M = N + 1
    h = 1.0 / N
    ii, jj, kk = np.meshgrid(np.arange(M), np.arange(M), np.arange(M), indexing="ij")
    # node id = i + M*j + M*M*k  -> flatten in (k, j, i) order
    ii = ii.transpose(2, 1, 0).reshape(-1)
    jj = jj.transpose(2, 1, 0).reshape(-1)
    kk = kk.transpose(2, 1, 0).reshape(-1)
    # Lattice coordinates i/N (correctly rounded; planes 0 and 1 exact).
    pts = np.stack([ii / N, jj / N, kk / N], axis=1).astype(np.float64)
    if jitter:
        nid = np.arange(M ** 3, dtype=np.uint64)
        interior = (ii > 0) & (ii < N) & (jj > 0) & (jj < N) & (kk > 0) & (kk < N)
        for c in range(3):
            u = uniform01(seed, 100 + c, nid)
            off = (2.0 * u - 1.0) * (jitter * h)
            pts[:, c] = np.where(interior, pts[:, c] + off, pts[:, c])

    def nid3(i, j, k):
        return (i + M * j + M * M * k).astype(np.int32)

    ci, cj, ck = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    ci = ci.transpose(2, 1, 0).reshape(-1)
    cj = cj.transpose(2, 1, 0).reshape(-1)
    ck = ck.transpose(2, 1, 0).reshape(-1)
    cell = (ci + N * cj + N * N * ck).astype(np.int64)
    ncell = N ** 3

    # Inner faces in upper-triangular order: for every owner cell, its
    # higher-numbered neighbours ascending: +x (c+1), +y (c+N), +z (c+N^2).
    fx_own = cell[ci < N - 1]
    fy_own = cell[cj < N - 1]
    fz_own = cell[ck < N - 1]
    # per-cell face slots: 3 possible, build sort key owner*3 + dir
    own = np.concatenate([fx_own, fy_own, fz_own])
    dirn = np.concatenate([np.zeros(fx_own.size, np.int64), np.ones(fy_own.size, np.int64),
                           np.full(fz_own.size, 2, np.int64)])
    order = np.argsort(own * 3 + dirn, kind="stable")
    own = own[order]
    dirn = dirn[order]
    oi = own % N
    oj = (own // N) % N
    ok = own // (N * N)
    nb = own + np.where(dirn == 0, 1, np.where(dirn == 1, N, N * N))

    def quad_x(i, j, k):  # face at x=i, normal +x
        return np.stack([nid3(i, j, k), nid3(i, j + 1, k), nid3(i, j + 1, k + 1), nid3(i, j, k + 1)], 1)

    def quad_y(i, j, k):  # face at y=j, normal +y
        return np.stack([nid3(i, j, k), nid3(i, j, k + 1), nid3(i + 1, j, k + 1), nid3(i + 1, j, k)], 1)

    def quad_z(i, j, k):  # face at z=k, normal +z
        return np.stack([nid3(i, j, k), nid3(i + 1, j, k), nid3(i + 1, j + 1, k), nid3(i, j + 1, k)], 1)

    inner_nodes = np.empty((own.size, 4), np.int32)
    mx = dirn == 0
    my = dirn == 1
    mz = dirn == 2
    inner_nodes[mx] = quad_x(oi[mx] + 1, oj[mx], ok[mx])
    inner_nodes[my] = quad_y(oi[my], oj[my] + 1, ok[my])
    inner_nodes[mz] = quad_z(oi[mz], oj[mz], ok[mz] + 1)

    # Boundary patches x-, x+, y-, y+, z-, z+ ; faces ordered by owner cell id.
    bnodes = []
    bown = []
    names = ["xmin", "xmax", "ymin", "ymax", "zmin", "zmax"]
    for p in range(6):
        axis = p // 2
        hi = p % 2
        sel = [ci, cj, ck][axis] == (N - 1 if hi else 0)
        c = cell[sel]
        i, j, k = ci[sel], cj[sel], ck[sel]
        if axis == 0:
            q = quad_x(i + hi, j, k)
        elif axis == 1:
            q = quad_y(i, j + hi, k)
        else:
            q = quad_z(i, j, k + hi)
        if not hi:  # outward normal is -axis: reverse vertex order
            q = q[:, ::-1]
        bnodes.append(q)
        bown.append(c)
    nb_faces = np.full(6, N * N, dtype=np.int32)
    n_inner = own.size
    start = n_inner + np.concatenate([[0], np.cumsum(nb_faces)[:-1]]).astype(np.int32)

    allq = np.concatenate([inner_nodes] + bnodes, axis=0)
    nf = allq.shape[0]
    node = np.full((nf, NOMAX), -1, dtype=np.int32)
    node[:, :4] = allq
    nnodes = np.full(nf, 4, dtype=np.int32)
    owner = np.concatenate([own] + bown).astype(np.int32)
    twins = [int(start[p ^ 1]) for p in range(6) if types[p] == "periodic"]
    m = Mesh(points=pts, nnodes=nnodes, node=node, owner=owner,
             neighbour=nb.astype(np.int32), n_cells=ncell, names=names,
             types=list(types), nfaces=nb_faces, start_face=start.astype(np.int32),
             twin_start=twins)
    m.check()
    return m


def cell_centres_lattice(N: int) -> np.ndarray:
    """Centres of the undistorted lattice cells (for analytic test fields)."""
    c = (np.arange(N) + 0.5) / N
    x = np.tile(c, N * N)
    y = np.tile(np.repeat(c, N), N)
    z = np.repeat(c, N * N)
    return np.stack([x, y, z], 1)


# --------------------------------------------------------------------------
# Native polyMesh format (geometry.f90:118-406): 1-based text files
# points / faces / owner / neighbour / boundary / size.
# --------------------------------------------------------------------------

def write_native(mesh: Mesh, case_dir: str) -> None:
    d = os.path.join(case_dir, "polyMesh")
    os.makedirs(d, exist_ok=True)
    np.savetxt(os.path.join(d, "points"), mesh.points, fmt="%.17g")
    with open(os.path.join(d, "faces"), "w") as f:
        for k in range(mesh.n_faces):
            nn = int(mesh.nnodes[k])
            f.write(str(nn) + " " + " ".join(str(int(v) + 1) for v in mesh.node[k, :nn]) + "\n")
    np.savetxt(os.path.join(d, "owner"), mesh.owner + 1, fmt="%d")
    np.savetxt(os.path.join(d, "neighbour"), mesh.neighbour + 1, fmt="%d")
    with open(os.path.join(d, "boundary"), "w") as f:
        f.write("# bcname bctype nFaces startFace [startFaceTwin]\n")
        twin = dict(zip(mesh.periodic_patches, mesh.twin_start))
        for ib in range(mesh.n_patches):
            extra = f" {int(twin[ib])}" if ib in twin else ""
            f.write(f"{mesh.names[ib]} {mesh.types[ib]} {int(mesh.nfaces[ib])} "
                    f"{int(mesh.start_face[ib])}{extra}\n")
    with open(os.path.join(d, "size"), "w") as f:
        for v in (mesh.n_nodes, mesh.n_cells, mesh.n_inner, mesh.n_bnd, mesh.n_faces):
            f.write(f"{v}\n")


def read_native(case_dir: str) -> Mesh:
    d = os.path.join(case_dir, "polyMesh")
    with open(os.path.join(d, "size")) as f:
        n_nodes, n_cells, n_inner, n_bnd, n_faces = [int(f.readline()) for _ in range(5)]
    pts = np.loadtxt(os.path.join(d, "points"), dtype=np.float64).reshape(n_nodes, 3)
    node = np.full((n_faces, NOMAX), -1, dtype=np.int32)
    nnodes = np.zeros(n_faces, dtype=np.int32)
    with open(os.path.join(d, "faces")) as f:
        for k in range(n_faces):
            v = [int(t) for t in f.readline().split()]
            nnodes[k] = v[0]
            node[k, : v[0]] = np.array(v[1: 1 + v[0]]) - 1
    owner = np.loadtxt(os.path.join(d, "owner"), dtype=np.int64).reshape(-1).astype(np.int32) - 1
    nb = np.loadtxt(os.path.join(d, "neighbour"), dtype=np.int64).reshape(-1).astype(np.int32) - 1
    names, types, nf, sf, tw = _read_boundary(os.path.join(d, "boundary"))
    m = Mesh(points=pts, nnodes=nnodes, node=node, owner=owner, neighbour=nb,
             n_cells=n_cells, names=names, types=types,
             nfaces=np.array(nf, np.int32), start_face=np.array(sf, np.int32), twin_start=tw)
    m.check()
    return m


def _read_boundary(path):
    """polyMesh/boundary table: name type nFaces startFace, plus startFaceTwin
    on 'periodic' lines (geometry.f90:234-258); '#' lines are comments."""
    names, types, nf, sf, tw = [], [], [], [], []
    with open(path) as f:
        for line in f:
            s = line.strip()
            if not s or s.startswith("#"):
                continue
            t = s.split()
            names.append(t[0]); types.append(t[1]); nf.append(int(t[2])); sf.append(int(t[3]))
            if t[1] == "periodic":
                tw.append(int(t[4]))
    return names, types, nf, sf, tw


# --------------------------------------------------------------------------
# OpenFOAM polyMesh (0-based lists with FoamFile headers), the format
# read_mesh_openfoam reads (geometry.f90:846-1300).  polyMesh/boundary uses
# the reference's simplified table "name type nFaces startFace".
# --------------------------------------------------------------------------

def _foam_list(path):
    """Return the entries of the first top-level '(...)' list after the count."""
    with open(path) as f:
        txt = f.read()
    # drop the FoamFile dictionary and comments
    import re
    txt = re.sub(r"/\*.*?\*/", " ", txt, flags=re.S)
    txt = re.sub(r"//[^\n]*", " ", txt)
    i = txt.find("FoamFile")
    if i >= 0:
        j = txt.find("}", i)
        txt = txt[j + 1:]
    m = re.search(r"(\d+)\s*\(", txt)
    count = int(m.group(1))
    body = txt[m.end():]
    return count, body


def read_foam(case_dir: str) -> Mesh:
    import re
    d = os.path.join(case_dir, "polyMesh")
    npts, body = _foam_list(os.path.join(d, "points"))
    vals = re.findall(r"\(([^()]*)\)", body)[:npts]
    pts = np.array([[float(t) for t in v.split()] for v in vals], dtype=np.float64)
    nf, body = _foam_list(os.path.join(d, "faces"))
    faces = re.findall(r"(\d+)\s*\(([^()]*)\)", body)[:nf]
    node = np.full((nf, NOMAX), -1, dtype=np.int32)
    nnodes = np.zeros(nf, dtype=np.int32)
    for k, (cnt, ids) in enumerate(faces):
        v = [int(t) for t in ids.split()]
        nnodes[k] = int(cnt)
        node[k, : len(v)] = v
    no, body = _foam_list(os.path.join(d, "owner"))
    owner = np.array(body.split(")")[0].split(), dtype=np.int64)[:no].astype(np.int32)
    nn, body = _foam_list(os.path.join(d, "neighbour"))
    nb = np.array(body.split(")")[0].split(), dtype=np.int64)[:nn].astype(np.int32)
    names, types, nfs, sfs, tw = _read_boundary(os.path.join(d, "boundary"))
    n_cells = int(max(owner.max(), nb.max() if nb.size else 0)) + 1
    m = Mesh(points=pts, nnodes=nnodes, node=node, owner=owner, neighbour=nb,
             n_cells=n_cells, names=names, types=types,
             nfaces=np.array(nfs, np.int32), start_face=np.array(sfs, np.int32), twin_start=tw)
    m.check()
    return m


# Kuhn split of a cube into 6 tetrahedra along its (0,0,0)-(1,1,1) diagonal:
# tet p walks from corner 000 to 111 adding the unit steps of axis order
# _KUHN[p]; every cube uses the same diagonal, so the split is conforming.
_KUHN = ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0))


def permute_cells(mesh: Mesh, new_of_old: np.ndarray) -> Mesh:
    """Renumber cells (cell c becomes new_of_old[c]) and restore the
    upper-triangular face order: inner faces get owner < neighbour (face
    reversed when the pair swaps) sorted by (owner, neighbour); boundary
    faces stay in their patch, sorted by owner.  Points are unchanged."""
    new = np.asarray(new_of_old, dtype=np.int64)
    Fi = mesh.n_inner
    o = new[mesh.owner[:Fi]]
    nb = new[mesh.neighbour]
    swap = o > nb
    lo = np.where(swap, nb, o)
    hi = np.where(swap, o, nb)
    nodes = mesh.node[:Fi].copy()
    if swap.any():
        nn = mesh.nnodes[:Fi]
        for k in np.unique(nn[swap]):
            sel = swap & (nn == k)
            nodes[sel, :k] = nodes[sel, :k][:, ::-1]
    order = np.lexsort((hi, lo))
    inner_nodes = nodes[order]
    inner_nn = mesh.nnodes[:Fi][order]
    own = [lo[order]]
    bnodes, bnn = [inner_nodes], [inner_nn]
    for ib in range(mesh.n_patches):
        s = int(mesh.start_face[ib])
        e = s + int(mesh.nfaces[ib])
        bo = new[mesh.owner[s:e]]
        ordb = np.argsort(bo, kind="stable")
        own.append(bo[ordb])
        bnodes.append(mesh.node[s:e][ordb])
        bnn.append(mesh.nnodes[s:e][ordb])
    m = Mesh(points=mesh.points, nnodes=np.concatenate(bnn).astype(np.int32),
             node=np.concatenate(bnodes).astype(np.int32),
             owner=np.concatenate(own).astype(np.int32), neighbour=hi[order].astype(np.int32),
             n_cells=mesh.n_cells, names=list(mesh.names), types=list(mesh.types),
             nfaces=mesh.nfaces.copy(), start_face=mesh.start_face.copy(),
             twin_start=list(mesh.twin_start))
    m.check()
    return m


def merge_cells(mesh: Mesh, group) -> Mesh:
    """Polyhedral mesh whose cell c is the union of the cells i of ``mesh``
    with ``group[i] == c`` (no reference counterpart: a generator of test
    meshes with more faces per cell than the synthetic hexes and tets have).

    Faces between cells of one group are removed; every other face is kept
    with its vertices.  Inner faces are re-oriented so that the owner is the
    lower new cell (vertex order reversed where the pair swaps) and re-sorted
    by owner, then neighbour, as the reference's meshes are; boundary faces
    keep their patches and order.  Two cells must share at most one face
    (the reference's CSR keeps one entry per face, ``sparse_matrix.f90:108-
    175``), which is checked.  Periodic patches are not supported.
    """
    assert not mesh.periodic_patches
    group = np.asarray(group, dtype=np.int64)
    assert group.shape == (mesh.n_cells,)
    nc = int(group.max()) + 1
    assert np.array_equal(np.unique(group), np.arange(nc))
    ni = mesh.n_inner
    o = group[mesh.owner[:ni]]
    nb = group[mesh.neighbour]
    keep = np.nonzero(o != nb)[0]
    fo, fn = np.minimum(o, nb)[keep], np.maximum(o, nb)[keep]
    order = np.lexsort((fn, fo))
    pair = fo[order] * nc + fn[order]
    assert np.all(np.diff(pair) > 0), "two cells share more than one face"
    inner = keep[order]
    node_in = mesh.node[inner].copy()
    nn_in = mesh.nnodes[inner]
    for r in np.nonzero(o[inner] > nb[inner])[0]:
        m = int(nn_in[r])
        node_in[r, :m] = node_in[r, :m][::-1]
    nf_new = len(inner) + mesh.n_bnd
    owner = np.empty(nf_new, dtype=np.int32)
    owner[:len(inner)] = fo[order]
    owner[len(inner):] = group[mesh.owner[ni:]]
    out = Mesh(points=mesh.points.copy(),
               nnodes=np.concatenate([nn_in, mesh.nnodes[ni:]]).astype(np.int32),
               node=np.concatenate([node_in, mesh.node[ni:]]).astype(np.int32),
               owner=owner, neighbour=fn[order].astype(np.int32), n_cells=nc,
               names=list(mesh.names), types=list(mesh.types), nfaces=mesh.nfaces.copy(),
               start_face=(np.asarray(mesh.start_face) - (ni - len(inner))).astype(np.int32),
               twin_start=[])
    out.check()
    return out


def staggered_pairs(N: int) -> np.ndarray:
    """Cell groups for ``merge_cells`` on ``hex_mesh(N)``: along x, cells
    i, i+1 merge into one polyhedron for i = s, s+2, ... with s = (j+k) mod 2,
    so a merged cell's y and z neighbours are two different cells on each
    side (ten faces, ten neighbours, rows of eleven CSR entries); cells left
    over at the ends of a row stay single."""
    i, j, k = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    i, j, k = (a.transpose(2, 1, 0).ravel() for a in (i, j, k))  # cell i + N j + N^2 k
    s = (j + k) % 2
    lead = (i - s) % 2 == 0
    paired = (i >= s) & np.where(lead, i + 1 < N, i - 1 >= s)
    first = np.where(paired & ~lead, i - 1, i)  # the x-lower cell of the group
    key = first + N * j + N * N * k
    _, group = np.unique(key, return_inverse=True)
    return group.astype(np.int64)


def fisher_yates(n: int, seed: int, stream: int = 200) -> np.ndarray:
    """Permutation of range(n) by Fisher-Yates with SplitMix64 draws:
    for i = n-1 .. 1, j = floor(u_i (i+1)), swap(a[i], a[j])."""
    u = uniform01(seed, stream, np.arange(n, dtype=np.uint64))
    j = np.minimum(np.floor(u * (np.arange(n) + 1)).astype(np.int64), np.arange(n))
    a = list(range(n))
    jl = j.tolist()
    for i in range(n - 1, 0, -1):
        k = jl[i]
        a[i], a[k] = a[k], a[i]
    return np.array(a, dtype=np.int64)


def tet_mesh(N: int, jitter: float = 0.0, seed: int = 3, perm_seed=None,
             types=("wall",) * 6) -> Mesh:
    """Unit cube, N^3 cubes each split into 6 Kuhn tetrahedra (cell id
    6*cube + p before any permutation, cube = i + N*j + N*N*k), interior
    nodes jittered by +-jitter*h (seed), then -- with perm_seed -- the cell
    order permuted by ``fisher_yates(n, perm_seed)`` (cell c becomes perm[c]).
    Patches x-, x+, y-, y+, z-, z+ with 2 N^2 triangles each."""
    N = int(N)
    base = hex_mesh(1, types=types)  # for patch names only
    M = N + 1
    h = 1.0 / N
    ii, jj, kk = np.meshgrid(np.arange(M), np.arange(M), np.arange(M), indexing="ij")
    ii = ii.transpose(2, 1, 0).reshape(-1)
    jj = jj.transpose(2, 1, 0).reshape(-1)
    kk = kk.transpose(2, 1, 0).reshape(-1)
    lat = np.stack([ii, jj, kk], axis=1).astype(np.int64)
    pts = (lat / N).astype(np.float64)
    if jitter:
        nid = np.arange(M ** 3, dtype=np.uint64)
        interior = (ii > 0) & (ii < N) & (jj > 0) & (jj < N) & (kk > 0) & (kk < N)
        for c in range(3):
            u = uniform01(seed, 100 + c, nid)
            off = (2.0 * u - 1.0) * (jitter * h)
            pts[:, c] = np.where(interior, pts[:, c] + off, pts[:, c])
    ci, cj, ck = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    ci = ci.transpose(2, 1, 0).reshape(-1).astype(np.int64)
    cj = cj.transpose(2, 1, 0).reshape(-1).astype(np.int64)
    ck = ck.transpose(2, 1, 0).reshape(-1).astype(np.int64)
    nc = N ** 3
    tets = np.empty((nc, 6, 4), np.int64)
    for p, ax in enumerate(_KUHN):
        off = np.zeros(3, np.int64)
        verts = []
        for s in range(4):
            verts.append((ci + off[0]) + M * (cj + off[1]) + M * M * (ck + off[2]))
            if s < 3:
                off[ax[s]] += 1
        tets[:, p, :] = np.stack(verts, axis=1)
    tets = tets.reshape(-1, 4)
    ncell = 6 * nc
    # the four faces of every tet, oriented away from the opposite vertex
    fa = np.empty((ncell, 4, 3), np.int64)
    opp = np.empty((ncell, 4), np.int64)
    for m in range(4):
        others = [q for q in range(4) if q != m]
        fa[:, m, :] = tets[:, others]
        opp[:, m] = tets[:, m]
    fa = fa.reshape(-1, 3)
    opp = opp.reshape(-1)
    fcell = np.repeat(np.arange(ncell, dtype=np.int64), 4)
    a, b, c = lat[fa[:, 0]], lat[fa[:, 1]], lat[fa[:, 2]]
    nrm = np.cross(b - a, c - a)
    flip = np.einsum("ij,ij->i", nrm, lat[opp] - a) > 0
    fa[flip, 1], fa[flip, 2] = fa[flip, 2], fa[flip, 1].copy()
    del a, b, c, nrm, flip, opp
    key = np.sort(fa, axis=1)
    order = np.lexsort((key[:, 2], key[:, 1], key[:, 0]))
    ks = key[order]
    same = np.all(ks[1:] == ks[:-1], axis=1)
    first = np.nonzero(same)[0]          # pair (order[first], order[first+1])
    paired = np.zeros(order.size, bool)
    paired[first] = True
    paired[first + 1] = True
    r1, r2 = order[first], order[first + 1]
    c1, c2 = fcell[r1], fcell[r2]
    own_rec = np.where(c1 < c2, r1, r2)
    own = np.minimum(c1, c2)
    nbr = np.maximum(c1, c2)
    io = np.lexsort((nbr, own))
    inner_nodes = fa[own_rec[io]]
    own, nbr = own[io], nbr[io]
    brec = order[~paired]
    bl = lat[fa[brec]]
    patch = np.full(brec.size, -1, np.int64)
    for p in range(6):
        axis, hi = p // 2, p % 2
        on = np.all(bl[:, :, axis] == (N if hi else 0), axis=1)
        patch[on] = p
    assert (patch >= 0).all()
    bo = fcell[brec]
    ob = np.lexsort((bo, patch))
    bnodes = fa[brec[ob]]
    bown = bo[ob]
    nb_faces = np.bincount(patch, minlength=6).astype(np.int32)
    start = own.size + np.concatenate([[0], np.cumsum(nb_faces)[:-1]]).astype(np.int32)
    allq = np.concatenate([inner_nodes, bnodes], axis=0)
    node = np.full((allq.shape[0], NOMAX), -1, dtype=np.int32)
    node[:, :3] = allq
    m = Mesh(points=pts, nnodes=np.full(allq.shape[0], 3, np.int32), node=node,
             owner=np.concatenate([own, bown]).astype(np.int32), neighbour=nbr.astype(np.int32),
             n_cells=ncell, names=list(base.names), types=list(types), nfaces=nb_faces,
             start_face=start.astype(np.int32))
    m.check()
    if perm_seed is not None:
        m = permute_cells(m, fisher_yates(ncell, perm_seed))
    return m
