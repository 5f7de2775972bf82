"""The allocation-free potential-gradient entry point (femhot_energy_grad_noalloc) through the C ABI.

It runs the same chunk kernels with the same configuration as femhot_energy_grad, only with its
scratch in the library's fixed per-GPU array, so its outputs equal femhot_energy_grad's bit for bit.
Two calls in a row on one GPU, the second on a smaller mesh (the shared array keeps the first call's
contents), then the refusals: a plan that is not built, and a call whose scratch exceeds the array.
"""
import numpy as np
import pytest

from oracle import fe

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

CHUNK, EFIELD = 0x40, 0x2


def rel(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _mesh(n, seed):
    from folax_amd import _lib as L
    from folax_amd.context import FemhotContext
    coords, conn, _ = fe.box_tet_mesh(n, 1.0, jitter=0.2, seed=seed)
    conn = np.asarray(conn, dtype=np.int32)
    rng = np.random.default_rng(seed + 1)
    E = np.exp(rng.uniform(0.0, np.log(100.0), size=conn.shape[0]))
    u = rng.uniform(-0.01, 0.01, size=3 * coords.shape[0])
    ctx = FemhotContext("tetra", L.MECHANICAL, coords, conn, young_modulus=1.0, poisson_ratio=0.3)
    return ctx, coords, conn, E, u


def _call(ctx, name, E, u, batch, dtype=torch.float32):
    """One call of `name` (femhot_energy_grad or femhot_energy_grad_noalloc) through ctypes."""
    from folax_amd import _lib as L
    from folax_amd.context import _ptr, _stream
    Nn, Ne = ctx.n_nodes, ctx.n_elems
    en = torch.empty(batch, dtype=torch.float64, device="cuda")
    gu = torch.empty(batch, 3 * Nn, dtype=dtype, device="cuda")
    gd = torch.empty(batch, Ne, dtype=dtype, device="cuda")
    rc = getattr(ctx._lib, name)(ctx._h, L.F32 if dtype == torch.float32 else L.F64, batch, _ptr(E), Ne, _ptr(u),
                                 3 * Nn, _ptr(en), _ptr(gu), 3 * Nn, _ptr(gd), Ne, CHUNK | EFIELD, _stream())
    L.check(rc)
    return en, gu, gd


def test_noalloc_equals_energy_grad_on_two_calls_second_smaller():
    from folax_amd import _lib as L
    for n, seed, prepared_by_config in ((14, 71, True), (6, 73, False)):
        print(f"[no-alloc] tet mesh n={n}, plan {'set by config' if prepared_by_config else 'built by a call'}",
              flush=True)
        ctx, coords, conn, E, u = _mesh(n, seed)
        Et = torch.as_tensor(E, dtype=torch.float32, device="cuda").view(1, -1)
        ut = torch.as_tensor(u, dtype=torch.float32, device="cuda").view(1, -1)
        if prepared_by_config:
            ctx.set_chunk_config(L.OP_ENERGY_GRAD, torch.float32, 1, 512, 256, 4, flags=EFIELD)
        ref = _call(ctx, "femhot_energy_grad", Et, ut, 1)  # builds the default plan when not configured
        got = _call(ctx, "femhot_energy_grad_noalloc", Et, ut, 1)
        torch.cuda.synchronize()
        for a, b in zip(got, ref):
            assert torch.equal(a, b)
        ph = fe.Physics("mechanical", "tetra", E=1.0, nu=0.3)
        Pi, r, gde = fe.energy_grad(ph, coords[conn], E, u[fe.element_dofs(conn, 3)])
        assert rel(got[1].cpu().numpy(), fe.scatter_nodal(conn, r, coords.shape[0], 3)) <= 1e-5
        assert rel(got[2].cpu().numpy(), gde) <= 1e-5
        assert abs(float(got[0].sum()) - Pi.sum()) <= 1e-5 * abs(Pi.sum())
        ctx.close()


def test_noalloc_refuses_an_unbuilt_plan_and_oversized_scratch():
    from folax_amd import _lib as L
    ctx, coords, conn, E, u = _mesh(8, 75)
    E64 = torch.as_tensor(E, dtype=torch.float64, device="cuda").view(1, -1)
    u64 = torch.as_tensor(u, dtype=torch.float64, device="cuda").view(1, -1)
    with pytest.raises(L.FemhotError) as ei:  # no FP64 call has built a plan yet
        _call(ctx, "femhot_energy_grad_noalloc", E64, u64, 1, dtype=torch.float64)
    assert ei.value.code == L.ERR_ARG and "not built" in str(ei.value)
    # 12 chunks of 256 tets share well over 700 node slots, so 30000 samples need far more than
    # 30000 x 700 x 12 bytes = 252 MB of partials: beyond the fixed 64 MiB, refused before any launch
    B = 30000
    ctx.set_chunk_config(L.OP_ENERGY_GRAD, torch.float32, B, 256, 256, 0, flags=EFIELD)
    Eb = torch.as_tensor(E, dtype=torch.float32, device="cuda").view(1, -1).expand(B, -1).contiguous()
    ub = torch.as_tensor(u, dtype=torch.float32, device="cuda").view(1, -1).expand(B, -1).contiguous()
    with pytest.raises(L.FemhotError) as ei:
        _call(ctx, "femhot_energy_grad_noalloc", Eb, ub, B)
    assert ei.value.code == L.ERR_SHAPE and "scratch" in str(ei.value)
    ctx.close()
