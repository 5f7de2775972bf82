"""Mesh container, Kratos `.mdpa` reader, legacy VTK writer and structured mesh generators (host
side, NumPy).

Mirrors folax's `Mesh` (fol/mesh_input_output/mesh.py:39-181): node coordinates (n_nodes, 3),
connectivity per element type (int32), node sets, the orientation fix, `.mdpa` input
(mesh.py:96-111, 188-242) and `Finalize` export.  The meshio formats (`.med` and the rest) need
meshio/h5py, which this environment does not have: they raise NotImplementedError.
"""
from __future__ import annotations

import numpy as np

from . import _lib


# Kratos element names -> element type names (the same table as mesh.py:15-37: it is the file
# format's vocabulary).
_MDPA_TYPES = {
    "Line2D2": "line", "Line3D2": "line", "Triangle2D3": "triangle", "Triangle3D3": "triangle",
    "Quadrilateral2D4": "quad", "Quadrilateral3D4": "quad", "Tetrahedra3D4": "tetra", "Hexahedra3D8": "hexahedron",
    "Prism3D6": "wedge", "Line2D3": "line3", "Triangle2D6": "triangle6", "Triangle3D6": "triangle6",
    "Quadrilateral2D9": "quad9", "Quadrilateral3D9": "quad9", "Tetrahedra3D10": "tetra10",
    "Hexahedra3D27": "hexahedron27", "Point2D": "vertex", "Point3D": "vertex", "Quadrilateral2D8": "quad8",
    "Quadrilateral3D8": "quad8", "Hexahedra3D20": "hexahedron20",
}

# legacy VTK cell type ids
_VTK_CELL = {"vertex": 1, "line": 3, "triangle": 5, "quad": 9, "tetra": 10, "hexahedron": 12, "wedge": 13,
             "line3": 21, "triangle6": 22, "quad8": 23, "tetra10": 24, "hexahedron20": 25, "quad9": 28,
             "hexahedron27": 29}


def read_mdpa(path, scale_factor=1.0):
    """Kratos `.mdpa` text: `Begin Nodes` rows `id x y z`, `Begin Elements <Name>` rows
    `id property n1 .. nk` (1-based), `Begin SubModelPart <name>` with a `SubModelPartNodes` block.
    Returns (coords (n,3) float64, {type: int32 conn (0-based)}, {set name: int64 node ids}).
    Node ids are taken in file order (mesh.py:215-217 uses positions, not the ids)."""
    coords, elems, sets = None, {}, {}
    with open(path, "r") as f:
        lines = f.read().splitlines()
    i, n = 0, len(lines)
    while i < n:
        env = lines[i].strip()
        i += 1
        if env.startswith("Begin Nodes"):
            j = i
            while j < n and "End Nodes" not in lines[j]:
                j += 1
            data = np.array(" ".join(lines[i:j]).split(), dtype=np.float64)
            coords = data.reshape(-1, 4)[:, 1:] * scale_factor
            i = j + 1
        elif env.startswith("Begin Elements"):
            name = env[len("Begin Elements"):].strip()
            etype = next((v for k, v in _MDPA_TYPES.items() if k in name), None)
            rows = []
            while i < n and not lines[i].startswith("End Elements"):
                vals = [int(v) for v in lines[i].split()]
                if vals:
                    rows.append(vals[2:])
                i += 1
            i += 1
            conn = np.asarray(rows, dtype=np.int64).reshape(len(rows), -1) - 1
            conn = conn.astype(np.int32)
            elems[etype] = conn if etype not in elems else np.vstack([elems[etype], conn])
        elif env.startswith("Begin SubModelPart ") and not env.startswith("Begin SubModelPartNodes"):
            name = env[len("Begin SubModelPart "):].strip()
            if i < n and lines[i].strip().startswith("Begin SubModelPartNodes"):
                i += 1
                ids = []
                while i < n and not lines[i].strip().startswith("End SubModelPartNodes"):
                    if lines[i].strip():
                        ids.append(int(lines[i].strip()) - 1)
                    i += 1
                sets[name] = np.asarray(ids, dtype=np.int64)
    if coords is None:
        raise ValueError(f"{path}: no `Begin Nodes` block")
    return coords, elems, sets


def write_vtk(path, coords, elements_nodes, point_data=None):
    """Legacy ASCII VTK unstructured grid with optional scalar / vector point data."""
    cells = [(t, np.asarray(c)) for t, c in elements_nodes.items() if t in _VTK_CELL and len(c)]
    ncell = sum(len(c) for _, c in cells)
    size = sum(c.size + len(c) for _, c in cells)
    with open(path, "w") as f:
        f.write("# vtk DataFile Version 4.2\nfolax_amd mesh\nASCII\nDATASET UNSTRUCTURED_GRID\n")
        f.write(f"POINTS {coords.shape[0]} double\n")
        np.savetxt(f, np.asarray(coords, dtype=np.float64).reshape(-1, 3))
        f.write(f"CELLS {ncell} {size}\n")
        for _, c in cells:
            np.savetxt(f, np.hstack([np.full((len(c), 1), c.shape[1]), c]), fmt="%d")
        f.write(f"CELL_TYPES {ncell}\n")
        for t, c in cells:
            np.savetxt(f, np.full(len(c), _VTK_CELL[t]), fmt="%d")
        if point_data:
            f.write(f"POINT_DATA {coords.shape[0]}\n")
            for name, v in point_data.items():
                v = np.asarray(v, dtype=np.float64).reshape(coords.shape[0], -1)
                if v.shape[1] == 1:
                    f.write(f"SCALARS {name} double 1\nLOOKUP_TABLE default\n")
                    np.savetxt(f, v)
                else:
                    f.write(f"VECTORS {name} double\n")
                    np.savetxt(f, np.hstack([v, np.zeros((v.shape[0], 3 - v.shape[1]))]) if v.shape[1] < 3 else v)


class Mesh:
    """Container with the reference getters (mesh.py:144-172).  Reads `.mdpa` files on
    Initialize (mesh.py:59-120), then applies the orientation fix."""

    def __init__(self, io_name: str = "mesh", file_name: str = "mesh.none", case_dir: str = ".", scale_factor=1.0):
        self.name = io_name
        self.file_name = file_name
        self.mesh_format = file_name.split(".")[1] if "." in file_name else ""
        self.case_dir = case_dir
        self.scale_factor = scale_factor
        self.node_ids = np.zeros(0, dtype=np.int64)
        self.nodes_coordinates = np.zeros((0, 3))
        self.elements_nodes = {}
        self.node_sets = {}
        self.point_data = {}
        self.is_initialized = False

    def GetName(self):
        return self.name

    def Initialize(self):
        if self.is_initialized:
            return
        import os
        path = os.path.join(self.case_dir, self.file_name)
        if self.mesh_format == "mdpa":
            self.nodes_coordinates, self.elements_nodes, self.node_sets = read_mdpa(path, self.scale_factor)
            self.node_ids = np.arange(self.nodes_coordinates.shape[0])
        elif os.path.exists(path):
            raise NotImplementedError(f"reading .{self.mesh_format} needs meshio/h5py, which are not installed")
        self.CheckAndOrientElements()
        self.is_initialized = True

    def __setitem__(self, key, value):
        self.point_data[key] = np.asarray(value)

    def __getitem__(self, key):
        return self.point_data[key]

    def HasPointData(self, data_name):
        """mesh.py:174-175."""
        return data_name in self.point_data

    def Finalize(self, export_dir: str = ".", export_format: str = "vtk"):
        """Writes the mesh and the assigned point data (mesh.py:177-180); legacy VTK only."""
        import os
        if export_format != "vtk":
            raise NotImplementedError("only the legacy VTK format is written (meshio is not installed)")
        write_vtk(os.path.join(export_dir, self.file_name.split(".")[0] + ".vtk"), self.nodes_coordinates,
                  self.elements_nodes, self.point_data)

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
