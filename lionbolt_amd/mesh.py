"""Tetrahedral meshes in the reference's conventions.

Conventions mirrored from the reference (they decide normals, neighbours and the
interface-mass permutation, so they are kept exactly):

* local faces of a 4-node tetrahedron are nodes ``[1,2,3], [1,2,4], [1,3,4],
  [2,3,4]`` (1-based), ``src/geometry/gmshprocessing.py:36-42`` and
  ``src/geometry/FiniteElementAnalysis.f90:129-132``;
* two elements are neighbours across a local face when the sorted global node
  triples agree (``gmshprocessing.py:144-194,258-288``); a boundary face has no
  neighbour (the reference stores 0, this module stores -1);
* element-node (spatial dof) numbering is ``sdof = offset(e) + k`` with
  ``offset(e) = 4 (e-1)`` for an all-tetrahedron mesh (``ReadGMSH.f90:259-266``).

The synthetic workload is the jittered cube: an ``n^3`` hex grid on ``[0,a]^3``
with nodes numbered x-fastest, each hex split into the 6 Kuhn tetrahedra that
share the diagonal ``v000-v111`` (element id ``6*hex + t``), interior nodes
jittered per coordinate by ``U[-f h, f h]`` from SplitMix64 stream ``seed``,
boundary nodes fixed, every tetrahedron positively oriented (the reference's
mass-eliminated matrices scale by ``1/detJ``, ``FiniteElementAnalysis.f90:245``).
"""
import itertools

import numpy as np

from . import rng

# Local face -> local nodes (0-based form of gmshprocessing.py:36-42).
FACE_NODES = np.array([[0, 1, 2], [0, 1, 3], [0, 2, 3], [1, 2, 3]], dtype=np.int64)
# Local node opposite each local face.
FACE_OPPOSITE = np.array([3, 2, 1, 0], dtype=np.int64)

_KUHN_PERMS = list(itertools.permutations((0, 1, 2)))


def _perm_parity(p):
    inv = sum(1 for i in range(3) for j in range(i + 1, 3) if p[i] > p[j])
    return inv % 2


class TetMesh:
    """A conforming linear-tetrahedron mesh (all arrays 0-based).

    ``rg``  (NK, 3) float64 node coordinates
    ``node`` (NE, 4) int64 element -> global nodes
    ``nbr``  (NE, 4) int64 neighbour across each local face, -1 on the boundary
    ``mat``  (NE,)  int64 1-based material index
    """

    def __init__(self, rg, node, mat=None, nbr=None):
        self.rg = np.ascontiguousarray(rg, dtype=np.float64)
        self.node = np.ascontiguousarray(node, dtype=np.int64)
        self.NE = int(self.node.shape[0])
        self.NK = int(self.rg.shape[0])
        self.mat = (np.ones(self.NE, dtype=np.int64) if mat is None
                    else np.ascontiguousarray(mat, dtype=np.int64))
        self.nbr = build_neighbors(self.node) if nbr is None else np.ascontiguousarray(nbr, dtype=np.int64)

    @property
    def NENK(self):
        return 4 * self.NE

    @property
    def Nm(self):
        return int(self.mat.max())

    @property
    def offset(self):
        return 4 * np.arange(self.NE + 1, dtype=np.int64)

    @property
    def connectivity(self):
        """Global node of each spatial dof, ``sdof = 4 e + k``."""
        return self.node.reshape(-1)

    def centroids(self):
        return self.rg[self.node].mean(axis=1)

    def volumes(self):
        x = self.rg[self.node]
        j = np.stack([x[:, 1] - x[:, 0], x[:, 2] - x[:, 0], x[:, 3] - x[:, 0]], axis=2)
        return np.linalg.det(j) / 6.0

    def dof_coords(self):
        """Coordinates of every spatial dof, shape (NENK, 3)."""
        return self.rg[self.connectivity]


def build_neighbors(node):
    """Neighbour table from sorted face triples (``gmshprocessing.py:144-288``)."""
    node = np.asarray(node, dtype=np.int64)
    ne = node.shape[0]
    faces = np.sort(node[:, FACE_NODES], axis=2).reshape(-1, 3)   # (NE*4, 3)
    nk = int(node.max()) + 1
    key = (faces[:, 0] * nk + faces[:, 1]) * nk + faces[:, 2]
    order = np.argsort(key, kind="stable")
    ks = key[order]
    nbr = np.full(ne * 4, -1, dtype=np.int64)
    same = np.nonzero(ks[1:] == ks[:-1])[0]
    if same.size and np.any(same[1:] == same[:-1] + 1):
        raise ValueError("non-manifold mesh: a face is shared by more than two elements")
    a = order[same]
    b = order[same + 1]
    nbr[a] = b // 4
    nbr[b] = a // 4
    return nbr.reshape(ne, 4)


def jittered_cube(n, a=10.0, jitter=0.2, seed=1001, slab=None):
    """The jittered Kuhn-split cube described in the module docstring.

    ``slab=(z0, z1)`` makes elements whose centroid z lies in ``[z0, z1]`` material 2.
    """
    h = a / n
    m = n + 1
    ix, iy, iz = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    # x-fastest node numbering: id = ix + m*(iy + m*iz)
    ix = ix.transpose(2, 1, 0).reshape(-1)
    iy = iy.transpose(2, 1, 0).reshape(-1)
    iz = iz.transpose(2, 1, 0).reshape(-1)
    rg = np.stack([ix, iy, iz], axis=1).astype(np.float64) * h
    interior = (ix > 0) & (ix < n) & (iy > 0) & (iy < n) & (iz > 0) & (iz < n)
    if jitter:
        u = rng.uniform(seed, 3 * m ** 3).reshape(-1, 3)
        rg[interior] += (2.0 * u[interior] - 1.0) * (jitter * h)

    hx, hy, hz = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    hx = hx.transpose(2, 1, 0).reshape(-1)
    hy = hy.transpose(2, 1, 0).reshape(-1)
    hz = hz.transpose(2, 1, 0).reshape(-1)

    def vid(dx, dy, dz):
        return (hx + dx) + m * ((hy + dy) + m * (hz + dz))

    tets = []
    for p in _KUHN_PERMS:
        step = [0, 0, 0]
        path = [vid(0, 0, 0)]
        for axis in p[:2]:
            step[axis] = 1
            path.append(vid(*step))
        path.append(vid(1, 1, 1))
        if _perm_parity(p):
            path[1], path[2] = path[2], path[1]
        tets.append(np.stack(path, axis=1))
    node = np.stack(tets, axis=1).reshape(-1, 4)   # element id = 6*hex + t
    mesh = TetMesh(rg, node)
    if slab is not None:
        zc = mesh.centroids()[:, 2]
        mesh.mat = np.where((zc >= slab[0]) & (zc <= slab[1]), 2, 1).astype(np.int64)
    vol = mesh.volumes()
    if not np.all(vol > 0):
        raise ValueError("jittered cube produced a non-positive tetrahedron")
    return mesh


def read_gmsh41(path):
    """Read the 4-node tetrahedra of an ASCII GMSH 4.1 file as the reference does.

    ``ReadGMSH.f90:75-176``: the material of the b-th 3D element block is the
    physical tag of the b-th volume entity in ``$Entities`` (used as given, no
    renumbering; a volume with no physical tag means material 1, which then must
    hold for every volume; more than one tag is an error).  Elements are taken in
    file order, block by block.  Nodes are numbered in ``$Nodes`` order (the
    reference writes them 1..NK in that order, ``gmshprocessing.py:65-84``, and
    indexes them by GMSH node tag; a dense renumbering of the tags is the same
    map when the tags run 1..NK in order).  Neighbours come from the sorted
    global node triples of the faces (``gmshprocessing.py:144-288``).
    """
    with open(path) as fh:
        lines = fh.read().split("\n")
    pos = {}
    for i, line in enumerate(lines):
        if line.startswith("$") and not line.startswith("$End"):
            pos[line.strip()] = i
    volinfo = []
    if "$Entities" in pos:
        i = pos["$Entities"] + 1
        npnt, ncur, nsur, nvol = (int(t) for t in lines[i].split()[:4])
        i += 1 + npnt + ncur + nsur
        for _ in range(nvol):
            t = lines[i].split()
            nphys = int(t[7])
            if nphys > 1:
                raise ValueError("ReadGMSH: a volume has more than one physical tag")
            volinfo.append(int(t[8]) if nphys == 1 else 1)
            i += 1
        if any(int(lines[pos["$Entities"] + 2 + npnt + ncur + nsur + b].split()[7]) == 0 for b in range(nvol)) \
                and any(v != 1 for v in volinfo):
            raise ValueError("ReadGMSH: the mesh does not specify a material for every volume")
    # Nodes
    i = pos["$Nodes"] + 1
    nblocks, nnodes = (int(t) for t in lines[i].split()[:2])
    i += 1
    tags = np.empty(nnodes, dtype=np.int64)
    xyz = np.empty((nnodes, 3))
    c = 0
    for _ in range(nblocks):
        t = lines[i].split()
        cnt = int(t[3])
        i += 1
        tags[c:c + cnt] = [int(lines[i + j]) for j in range(cnt)]
        i += cnt
        xyz[c:c + cnt] = np.array([lines[i + j].split()[:3] for j in range(cnt)], dtype=np.float64)
        i += cnt
        c += cnt
    remap = np.full(int(tags.max()) + 1, -1, dtype=np.int64)
    remap[tags] = np.arange(nnodes)
    # Elements: 3D blocks in file order, block b takes volinfo[b]
    i = pos["$Elements"] + 1
    nblocks = int(lines[i].split()[0])
    i += 1
    conn, mats = [], []
    b = 0
    for _ in range(nblocks):
        dim, etag, etype, cnt = (int(t) for t in lines[i].split()[:4])
        i += 1
        if dim == 3:
            if etype != 4:
                raise ValueError("only 4-node tetrahedra are supported (ReadGMSH.f90:160-166)")
            blk = np.array([lines[i + j].split()[1:5] for j in range(cnt)], dtype=np.int64)
            conn.append(remap[blk])
            mats.append(np.full(cnt, volinfo[b] if b < len(volinfo) else 1, dtype=np.int64))
            b += 1
        i += cnt
    return TetMesh(xyz, np.concatenate(conn), np.concatenate(mats))


def union(meshes):
    """Disjoint union of meshes for a batch of independent problems.

    Nodes and elements are concatenated in order; neighbours are shifted by the
    element offset; element materials become ``p + 1`` for problem ``p`` (so each
    problem has its own cross sections).  Returns the mesh and the element
    offsets ``elem_off`` (length P + 1)."""
    rg, node, nbr, mat, eoff = [], [], [], [], [0]
    noff = 0
    for p, m in enumerate(meshes):
        rg.append(m.rg)
        node.append(m.node + noff)
        nbr.append(np.where(m.nbr >= 0, m.nbr + eoff[-1], -1))
        mat.append(np.full(m.NE, p + 1, dtype=np.int64))
        noff += m.NK
        eoff.append(eoff[-1] + m.NE)
    return (TetMesh(np.concatenate(rg), np.concatenate(node), np.concatenate(mat), np.concatenate(nbr)),
            np.array(eoff, dtype=np.int64))
