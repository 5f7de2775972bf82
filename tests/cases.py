"""Seeded test cases shared by the oracle tests and the GPU parity tests.

A case is a mesh (coords, int32 conn, node sets), a physics setting, a nodal (or per-element)
field, a nodal solution and Dirichlet data, generated with numpy.random.default_rng(seed).
"""
import numpy as np

from oracle import fe

PHYS_CODE = {"thermal": 0, "mechanical": 1, "neohooke": 2}


def mesh_for(etype, n, seed=0, jitter=0.15):
    """Small distorted structured meshes: QUAD4 n x n, HEX8 n^3, TET4 n^3 x 6 (interior jitter)."""
    rng = np.random.default_rng(seed)
    if etype == "quad":
        coords, conn, sets = fe.square_quad_mesh(n + 1, 1.0)
        h = 1.0 / n
        inner = np.all((coords[:, :2] > 1e-12) & (coords[:, :2] < 1 - 1e-12), axis=1)
        coords = coords.copy()
        coords[inner, :2] += rng.uniform(-jitter * h, jitter * h, size=(int(inner.sum()), 2))
    elif etype == "hexahedron":
        coords, conn, sets = fe.box_hex_mesh(n, n, n)
        h = 1.0 / n
        inner = np.all((coords > 1e-12) & (coords < 1 - 1e-12), axis=1)
        coords = coords.copy()
        coords[inner] += rng.uniform(-jitter * h, jitter * h, size=(int(inner.sum()), 3))
    else:
        coords, conn, sets = fe.box_tet_mesh(n, 1.0, jitter=jitter, seed=seed)
    return coords, np.asarray(conn, dtype=np.int32), sets


def make_case(kind, etype, n=4, seed=0, beta=0.0, c=1.0, elemental=False, u_scale=None, body=None, num_gp=None):
    rng = np.random.default_rng(1000 + seed)
    coords, conn, sets = mesh_for(etype, n, seed)
    ph = fe.Physics(kind, etype, num_gp=num_gp, beta=beta, c=c, E=1.7, nu=0.3, body=body)
    Nn = coords.shape[0]
    ndof = ph.ndpn * Nn
    field = rng.uniform(0.5, 2.0, size=conn.shape[0] if elemental else Nn)
    if u_scale is None:
        u_scale = 0.05 if kind == "neohooke" else 1.0
    u = rng.uniform(-1.0, 1.0, size=ndof) * u_scale
    if kind == "thermal":
        u = rng.uniform(0.0, 1.0, size=ndof)
    w = rng.standard_normal(ndof)
    if kind == "thermal":
        dofs_list, bc = ["T"], {"T": {"left": 1.0, "right": 0.0}}
    elif ph.dim == 2:
        dofs_list, bc = ["Ux", "Uy"], {"Ux": {"left": 0.0, "right": 0.05}, "Uy": {"left": 0.0, "right": 0.05}}
    else:
        dofs_list, bc = ["Ux", "Uy", "Uz"], {"Ux": {"bottom": 0.0}, "Uy": {"bottom": 0.0},
                                             "Uz": {"bottom": 0.0, "top": 0.01}}
    didx, dval = fe.dirichlet_dofs(sets, dofs_list, bc)
    return dict(kind=kind, etype=etype, ph=ph, coords=coords, conn=conn, sets=sets, field=field, u=u, w=w,
                didx=didx, dval=dval, elemental=elemental, beta=beta, c=c, body=body)


def femhot_context(case):
    """Device context for a case (imported lazily: needs torch + the built library)."""
    from folax_amd.context import FemhotContext
    ph = case["ph"]
    body = case["body"]
    bf = (0.0, 0.0, 0.0)
    if body is not None:
        b = np.atleast_1d(np.asarray(body, dtype=np.float64)).reshape(-1)
        bf = tuple(list(b) + [0.0] * (3 - b.size))
    return FemhotContext(case["etype"], PHYS_CODE[case["kind"]], case["coords"], case["conn"], num_gp=ph.num_gp,
                         young_modulus=ph.E, poisson_ratio=ph.nu, beta=ph.beta, c=ph.c, body_force=bf,
                         dirichlet_idx=case["didx"], dirichlet_val=case["dval"])
