"""Mesh container and structured mesh generators (host side, NumPy).

Mirrors the parts of folax's `Mesh` (fol/mesh_input_output/mesh.py:39-181) the hot path uses:
node coordinates (n_nodes, 3), connectivity per element type (int32), node sets and the orientation
fix.  File readers (meshio `.med`, Kratos `.mdpa`) are not part of this package.
"""
from __future__ import annotations

import numpy as np

from . import _lib


class Mesh:
    """Container with the reference getters (mesh.py:144-172)."""

    def __init__(self, io_name: str = "mesh", file_name: str = "mesh.none", case_dir: str = ".", scale_factor=1.0):
        self.name = io_name
        self.file_name = file_name
        self.node_ids = np.zeros(0, dtype=np.int64)
        self.nodes_coordinates = np.zeros((0, 3))
        self.elements_nodes = {}
        self.node_sets = {}
        self.is_initialized = False

    def Initialize(self):
        if self.is_initialized:
            return
        self.CheckAndOrientElements()
        self.is_initialized = True

    def CheckAndOrientElements(self):
        """Swap local nodes 0 and 1 of elements whose detJ at the one-point rule is negative
        (mesh.py:123-142); done by the library's host routine."""
        n_total = 0
        for etype, conn in list(self.elements_nodes.items()):
            if etype in _lib.ELEMENT_TYPES:
                new, n = _lib.orient_elements(_lib.ELEMENT_TYPES[etype], self.nodes_coordinates, conn)
                self.elements_nodes[etype] = new
                n_total += n
        return n_total

    def GetNodesIds(self):
        return self.node_ids

    def GetNumberOfNodes(self):
        return len(self.node_ids)

    def GetNodesCoordinates(self):
        return self.nodes_coordinates

    def GetNodesX(self):
        return self.nodes_coordinates[:, 0]

    def GetNodesY(self):
        return self.nodes_coordinates[:, 1]

    def GetNodesZ(self):
        return self.nodes_coordinates[:, 2]

    def GetElementsIds(self, element_type):
        return np.arange(len(self.elements_nodes[element_type]))

    def GetNumberOfElements(self, element_type):
        return len(self.elements_nodes[element_type])

    def GetElementsNodes(self, element_type):
        return self.elements_nodes[element_type]

    def GetNodeSet(self, set_name):
        return self.node_sets[set_name]


def _box_nodes(nx, ny, nz, Lx, Ly, Lz):
    xs, ys, zs = np.linspace(0, Lx, nx + 1), np.linspace(0, Ly, ny + 1), np.linspace(0, Lz, nz + 1)
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)


def _face_sets(coords, Lx, Ly, Lz):
    nid = np.arange(coords.shape[0])
    return {"left": nid[coords[:, 0] == 0.0], "right": nid[coords[:, 0] == Lx],
            "front": nid[coords[:, 1] == 0.0], "back": nid[coords[:, 1] == Ly],
            "bottom": nid[coords[:, 2] == 0.0], "top": nid[coords[:, 2] == Lz]}


def create_2D_square_mesh(L, N):
    """create_2D_square_mesh (fol/tools/usefull_functions.py:213-258): N nodes per side on [0,L]^2,
    node id = j + i*N (x fastest), element e = i*(N-1)+j -> [id, id+1, id+N+1, id+N] (CCW),
    node sets left = arange(0, N^2, N), right = arange(N-1, N^2, N)."""
    Ne = N - 1
    x = np.linspace(0.0, L, N)
    X, Y = np.meshgrid(x, x)
    m = Mesh("square_io", "square.")
    m.node_ids = np.arange(N * N)
    m.nodes_coordinates = np.stack([X.ravel(), Y.ravel(), np.zeros(N * N)], axis=1)
    i, j = np.meshgrid(np.arange(Ne), np.arange(Ne), indexing="ij")
    base = (i * N + j).ravel()
    m.elements_nodes = {"quad": np.stack([base, base + 1, base + N + 1, base + N], axis=1).astype(np.int32)}
    m.node_sets = {"left": np.arange(0, N * N, N), "right": np.arange(N - 1, N * N, N),
                   "bottom": np.arange(0, N), "top": np.arange(N * (N - 1), N * N)}
    m.is_initialized = True
    return m


def create_3D_box_mesh(Nx, Ny, Nz, Lx=1.0, Ly=1.0, Lz=1.0, case_dir=None):
    """Structured HEX8 box (the reference builds it with gmsh, usefull_functions.py:146-211; that
    numbering is not reproducible offline).  Nodes lexicographic x fastest; element local order
    (---),(+--),(++-),(-+-),(--+),(+-+),(+++),(-++) as hexahedra_3d_8.py:79-88."""
    m = Mesh("box_io", "box.")
    m.nodes_coordinates = _box_nodes(Nx, Ny, Nz, Lx, Ly, Lz)
    m.node_ids = np.arange(m.nodes_coordinates.shape[0])
    k, j, i = np.meshgrid(np.arange(Nz), np.arange(Ny), np.arange(Nx), indexing="ij")
    sx, sy, sz = 1, Nx + 1, (Nx + 1) * (Ny + 1)
    b = (i * sx + j * sy + k * sz).ravel()
    m.elements_nodes = {"hexahedron": np.stack([b, b + sx, b + sx + sy, b + sy, b + sz, b + sx + sz,
                                                b + sx + sy + sz, b + sy + sz], axis=1).astype(np.int32)}
    m.node_sets = _face_sets(m.nodes_coordinates, Lx, Ly, Lz)
    m.is_initialized = True
    return m


_KUHN = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]


def create_3D_tet_box_mesh(n, L=1.0, jitter=0.0, seed=None):
    """Cube of n^3 hexes, each split into 6 Kuhn tetrahedra (vertex path 000 -> 111 along each
    axis permutation); interior nodes jittered by U(-jitter*h, jitter*h)^3 (boundary nodes stay on
    the faces), re-drawn until every detJ > 0; orientation fixed as mesh.py:123-142."""
    coords = _box_nodes(n, n, n, L, L, L)
    h = L / n
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    off = [1, n + 1, (n + 1) ** 2]
    b = (i * off[0] + j * off[1] + k * off[2]).ravel().astype(np.int64)
    tets = []
    for p in _KUHN:
        p1 = b + off[p[0]]
        p2 = p1 + off[p[1]]
        tets.append(np.stack([b, p1, p2, p2 + off[p[2]]], axis=1))
    conn = np.stack(tets, axis=1).reshape(-1, 4).astype(np.int32)
    if jitter > 0.0:
        rng = np.random.default_rng(seed)
        interior = np.all((coords > 1e-12 * L) & (coords < L * (1 - 1e-12)), axis=1)
        base = coords
        while True:
            c2 = base.copy()
            c2[interior] += rng.uniform(-jitter * h, jitter * h, size=(int(interior.sum()), 3))
            conn2, _ = _lib.orient_elements(_lib.TET4, c2, conn)
            X = c2[conn2]
            det = np.linalg.det(np.stack([X[:, 1] - X[:, 0], X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]], axis=2))
            if np.all(det > 0.0):
                coords, conn = c2, conn2
                break
    else:
        conn, _ = _lib.orient_elements(_lib.TET4, coords, conn)
    m = Mesh("tet_box_io", "tet_box.")
    m.nodes_coordinates = coords
    m.node_ids = np.arange(coords.shape[0])
    m.elements_nodes = {"tetra": conn}
    m.node_sets = _face_sets(coords, L, L, L)
    m.is_initialized = True
    return m


def concatenate_meshes(meshes, element_type):
    """Disjoint union of meshes of one element type: node ids are offset per mesh; returns
    (coords, conn, node_offsets, element_offsets).  Used to run many uneven meshes in one launch."""
    coords, conns, noff, eoff = [], [], [0], [0]
    for m in meshes:
        c = m.GetElementsNodes(element_type).astype(np.int64) + noff[-1]
        coords.append(m.GetNodesCoordinates())
        conns.append(c)
        noff.append(noff[-1] + m.GetNumberOfNodes())
        eoff.append(eoff[-1] + c.shape[0])
    return (np.concatenate(coords), np.concatenate(conns).astype(np.int32), np.asarray(noff, dtype=np.int64),
            np.asarray(eoff, dtype=np.int64))
