"""CPU tests of the chunk planner (host only, no GPU): femhot_chunk_plan_check builds the plan the
chunk kernels index LDS and the partial slots with, and checks every invariant they rely on (element
permutation, segment runs, chunk-local connectivity against the mesh, incidence runs and per-corner
ranks, coordinate records, zero / shared nodes, partial-slot layout). Here it runs on every element
type, on element orders the planner keeps or replaces (Morton chunks), on uneven segments, unused
nodes, collapsed elements and chunk sizes from 1 element up; for plans in the caller's order the
figures it reports are recounted with numpy. The same checks run under AddressSanitizer and
UndefinedBehaviorSanitizer from tests/sanitize/."""
import numpy as np
import pytest

from folax_amd import _lib as L
from folax_amd.mesh import concatenate_meshes, create_2D_square_mesh, create_3D_box_mesh, create_3D_tet_box_mesh


def _recount(conn, Nn, epc, seg_off):
    """Figures of a plan whose chunks are runs of the caller's order: runs of at most epc elements
    that never cross a segment."""
    chunks = []
    for s in range(len(seg_off) - 1):
        for i in range(seg_off[s], seg_off[s + 1], epc):
            chunks.append(np.unique(conn[i:min(i + epc, seg_off[s + 1])]))
    touch = np.zeros(Nn, dtype=np.int64)
    for c in chunks:
        touch[c] += 1
    return dict(n_chunks=len(chunks), max_nodes=max(len(c) for c in chunks), chunk_nodes=sum(len(c) for c in chunks),
                shared_nodes=int((touch > 1).sum()), slots=int(touch[touch > 1].sum()), zero_nodes=int((touch == 0).sum()))


def _check(et, coords, conn, epc, seg_off=None):
    st = L.chunk_plan_check(et, coords, conn, epc, seg_off)
    so = [0, conn.shape[0]] if seg_off is None else list(seg_off)
    assert st["max_elems"] == min(epc, max(b - a for a, b in zip(so[:-1], so[1:])))
    if st["caller_order"]:
        ref = _recount(conn, coords.shape[0], epc, so)
        for k, v in ref.items():
            assert st[k] == v, (k, st[k], v)
    return st


def _meshes():
    q = create_2D_square_mesh(1.0, 9)
    h = create_3D_box_mesh(5, 4, 3)
    t = create_3D_tet_box_mesh(4, 1.0, jitter=0.2, seed=3)
    return [(L.QUAD4, q.GetNodesCoordinates(), q.GetElementsNodes("quad")),
            (L.HEX8, h.GetNodesCoordinates(), h.GetElementsNodes("hexahedron")),
            (L.TET4, t.GetNodesCoordinates(), t.GetElementsNodes("tetra"))]


@pytest.mark.parametrize("epc", [1, 3, 7, 32, 128, 1024])
def test_plan_invariants_every_element_type(epc):
    for et, coords, conn in _meshes():
        st = _check(et, coords, conn, epc)
        assert st["zero_nodes"] == 0
        if epc >= conn.shape[0]:
            assert st["n_chunks"] == 1 and st["shared_nodes"] == 0 and st["slots"] == 0


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_plan_invariants_shuffled_order(seed):
    """A random element order is not local, so the planner cuts Morton chunks instead; the
    invariants hold either way."""
    rng = np.random.default_rng(seed)
    orders = []
    for et, coords, conn in _meshes():
        conn = conn[rng.permutation(conn.shape[0])]
        for epc in (16, 100):
            orders.append(_check(et, coords, conn, epc)["caller_order"])
    assert 0 in orders


def test_plan_invariants_uneven_segments():
    """Meshes of very different sizes in one plan (one segment each, as the batched uneven workload
    runs them), with an empty segment in the middle: no chunk crosses a segment boundary."""
    ms = [create_3D_tet_box_mesh(n, 1.0, jitter=0.1, seed=n) for n in (1, 2, 5, 3)]
    coords, conn, noff, eoff = concatenate_meshes(ms, "tetra")
    seg = np.concatenate([eoff[:2], eoff[1:2], eoff[2:]])  # segment 1 is empty
    for epc in (5, 48, 200, 4096):
        st = _check(L.TET4, coords, conn, epc, seg)
        assert st["n_chunks"] >= 4


def test_plan_invariants_unused_nodes_and_collapsed_elements():
    """Nodes in no element are listed for clearing; an element naming a node twice (a collapsed
    hexahedron) gives that chunk node two incidences from one element."""
    h = create_3D_box_mesh(3, 3, 3)
    coords = np.concatenate([h.GetNodesCoordinates(), np.full((5, 3), 7.0)])
    conn = h.GetElementsNodes("hexahedron").copy()
    conn[4, 5] = conn[4, 4]
    for epc in (2, 9, 27):
        st = _check(L.HEX8, coords, conn, epc)
        assert st["zero_nodes"] == 5 + (1 if not np.isin(h.GetElementsNodes("hexahedron")[4, 5], conn) else 0)


def test_plan_limits_and_arguments():
    t = create_3D_tet_box_mesh(16, 1.0, seed=1)  # 24,576 tets: 98,304 corners in one chunk
    with pytest.raises(L.FemhotError) as e:
        L.chunk_plan_check(L.TET4, t.GetNodesCoordinates(), t.GetElementsNodes("tetra"), 1 << 20)
    assert e.value.code == L.ERR_UNSUPPORTED
    with pytest.raises(L.FemhotError) as e:
        L.chunk_plan_check(L.TET4, t.GetNodesCoordinates(), t.GetElementsNodes("tetra"), 0)
    assert e.value.code == L.ERR_ARG
    bad = t.GetElementsNodes("tetra").copy()
    bad[3, 2] = t.GetNumberOfNodes()
    with pytest.raises(L.FemhotError) as e:
        L.chunk_plan_check(L.TET4, t.GetNodesCoordinates(), bad, 64)
    assert e.value.code == L.ERR_INDEX
