"""Full-size workloads on the GPU, checked through properties that hold at any size.

The oracle finishes these sizes too slowly to compare every output, so each workload is pinned by
what the mathematics fixes: closed-form CSR structure, null spaces of the elastic tangent,
symmetry, consistency of the residual with the assembled tangent, homogeneity of the energies
(Euler's theorem), FP32 against FP64, finite differences, and the FP64 oracle on a few samples
or, for the tetrahedral batch, the C restatement of the oracle on every element.

Workloads (all inputs seeded):
  hex64   HEX8 64^3 linear elasticity, two-phase nodal field (1 / 10), 823,875 dofs: CSR pattern,
          tangent, Jacobi-PCG in FP64 and FP32+FP64-refinement.
  quad128 QUAD4 128^2 x 64 samples, nonlinear thermal k = k0 (1 + 0.5 T^2): residual-norm loss and
          its exact gradient, FP32 and FP64.
  hex48   HEX8 48^3 x 16 samples, Neo-Hookean, displacements from a 4-layer tanh MLP: energy loss
          and its gradient through the network, FP32 and FP64.
  tet4    32 jittered TET4 cubes (10,368 ... 384,000 tets, 4,400,448 in all), per-element Young's
          modulus, one disjoint mesh: potential and gradient (g_u, g_E) on the chunk path.
"""
import numpy as np
import pytest

from oracle import fe

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def trel(a, b):
    a, b = a.double().reshape(-1), b.double().reshape(-1)
    return float(torch.linalg.norm(a - b) / torch.linalg.norm(b).clamp_min(1e-300))


def csr_matvec(rp, ci, vals, x):
    """y = A x by a segmented sum in torch (independent of the library's SpMV)."""
    rows = torch.repeat_interleave(torch.arange(rp.numel() - 1, device=x.device), rp[1:] - rp[:-1])
    return torch.zeros_like(x).index_add_(0, rows, vals * x[ci.long()])


# ----------------------------------------------------------------------------------------------
# hex64: pattern, tangent, PCG
# ----------------------------------------------------------------------------------------------
@pytest.fixture(scope="module")
def hex64():
    from folax_amd.context import FemhotContext
    coords, conn, sets = fe.box_hex_mesh(64, 64, 64)
    E = fe.grf_two_phase((65, 65, 65), 8.0, 2001)
    bc = {"Ux": {"bottom": 0.0}, "Uy": {"bottom": 0.0}, "Uz": {"bottom": 0.0, "top": 0.01}}
    didx, dval = fe.dirichlet_dofs(sets, ["Ux", "Uy", "Uz"], bc)
    ctx = FemhotContext("hexahedron", 1, coords, conn, num_gp=2, young_modulus=1.0, poisson_ratio=0.3,
                        dirichlet_idx=didx, dirichlet_val=dval)
    u = np.random.default_rng(2002).uniform(0.0, 1.0, size=3 * coords.shape[0])
    u = fe.apply_dirichlet(u, didx, dval)
    dev = torch.device("cuda")
    Et = torch.as_tensor(E, dtype=torch.float64, device=dev)
    ut = torch.as_tensor(u, dtype=torch.float64, device=dev)
    rp, ci = ctx.pattern(dev)
    vals, R = ctx.tangent_csr(Et, ut)
    torch.cuda.synchronize()
    yield dict(ctx=ctx, coords=coords, didx=didx, E=Et, u=ut, rp=rp, ci=ci, vals=vals, R=R)
    ctx.close()


def test_hex64_pattern_closed_form(hex64):
    """nnz = (3N-2)^3 * 9; row lengths 3 * prod(stencil widths); columns strictly increasing; the
    exact neighbour set on sampled rows."""
    ctx, rp, ci = hex64["ctx"], hex64["rp"], hex64["ci"]
    assert ctx.nnz() == (3 * 65 - 2) ** 3 * 9 == 64_701_513
    n1 = 65
    w = np.where((np.arange(n1) == 0) | (np.arange(n1) == n1 - 1), 2, 3)
    node_len = 3 * (w[None, None, :] * w[None, :, None] * w[:, None, None]).ravel()  # (k, j, i), x fastest
    assert np.array_equal((rp[1:] - rp[:-1]).cpu().numpy(), np.repeat(node_len, 3))
    d = (ci[1:].long() - ci[:-1].long())
    d[rp[1:-1] - 1] = 1  # row boundaries
    assert bool((d > 0).all())
    rng = np.random.default_rng(7)
    rpn = rp.cpu().numpy()
    for row in rng.integers(0, 3 * n1 ** 3, size=300):
        node = row // 3
        i, j, k = node % n1, (node // n1) % n1, node // n1 ** 2
        nb = [(a + n1 * b + n1 * n1 * c) for c in range(max(k - 1, 0), min(k + 2, n1))
              for b in range(max(j - 1, 0), min(j + 2, n1)) for a in range(max(i - 1, 0), min(i + 2, n1))]
        want = np.sort((3 * np.asarray(nb)[:, None] + np.arange(3)[None]).ravel())
        got = ci[rpn[row]:rpn[row + 1]].cpu().numpy()
        assert np.array_equal(got, want), row


def test_hex64_tangent_null_space_symmetry_and_residual(hex64):
    """Rigid translations and infinitesimal rotations are in the null space of every free row (columns
    are kept, Dirichlet rows hold only the diagonal); the free block is symmetric; R = K u on free rows
    (linear elasticity, no body force); the library SpMV equals an independent segmented sum."""
    rp, ci, vals, R, u = hex64["rp"], hex64["ci"], hex64["vals"], hex64["R"], hex64["u"]
    coords = torch.as_tensor(hex64["coords"], dtype=torch.float64, device="cuda")
    n = u.numel()
    dmask = torch.zeros(n, dtype=torch.bool, device="cuda")
    dmask[torch.as_tensor(hex64["didx"], device="cuda")] = True
    rows = torch.repeat_interleave(torch.arange(n, device="cuda"), rp[1:] - rp[:-1])
    diag = torch.zeros(n, dtype=torch.float64, device="cuda").index_add_(0, rows, vals * (ci.long() == rows))
    absrow = torch.zeros(n, dtype=torch.float64, device="cuda").index_add_(0, rows, vals.abs())
    omega = torch.tensor([0.3, -0.2, 0.5], dtype=torch.float64, device="cuda")
    modes = []
    for c in range(3):
        t = torch.zeros(n, dtype=torch.float64, device="cuda")
        t[c::3] = 1.0
        modes.append(t)
    modes.append(torch.linalg.cross(omega.expand_as(coords), coords).reshape(-1))
    for t in modes:
        y = csr_matvec(rp, ci, vals, t)
        scale = absrow * t.abs().max()
        assert float((y[~dmask].abs() / scale[~dmask]).max()) <= 1e-12
        assert float((y[dmask] - diag[dmask] * t[dmask]).abs().max()) <= 1e-12 * float(scale.max())
    g = torch.Generator(device="cuda").manual_seed(3)
    x = torch.randn(n, dtype=torch.float64, device="cuda", generator=g).masked_fill(dmask, 0.0)
    z = torch.randn(n, dtype=torch.float64, device="cuda", generator=g).masked_fill(dmask, 0.0)
    Kz, Kx = csr_matvec(rp, ci, vals, z), csr_matvec(rp, ci, vals, x)
    a, b = float(x @ Kz), float(z @ Kx)
    assert abs(a - b) <= 1e-12 * float(torch.linalg.norm(x) * torch.linalg.norm(Kz))
    Ku = csr_matvec(rp, ci, vals, u)
    assert trel(R[~dmask], Ku[~dmask]) <= 1e-12
    assert float(R[dmask].abs().max()) == 0.0
    assert trel(hex64["ctx"].spmv(rp, ci, vals, u), Ku) <= 1e-14


def test_hex64_row_owner_assembly(hex64):
    """Row-owner tangent assembly at full size: every value written once from a fixed summation order,
    so two calls give the same bits; equal to the atomic assembly within FP64 rounding; with
    ACCUMULATE it adds onto what the values hold."""
    from folax_amd import _lib as L
    ctx, E, u = hex64["ctx"], hex64["E"], hex64["u"]
    v1, R1 = ctx.tangent_csr(E, u, flags=L.SCATTER_ROWS)
    v2, _ = ctx.tangent_csr(E, u, want_R=False, flags=L.SCATTER_ROWS)
    torch.cuda.synchronize()
    assert torch.equal(v1, v2)
    assert trel(v1, hex64["vals"]) <= 1e-13
    assert trel(R1, hex64["R"]) <= 1e-13
    acc = v1.clone()
    ctx.tangent_csr(E, u, want_R=False, flags=L.SCATTER_ROWS | L.ACCUMULATE, out_vals=acc)
    torch.cuda.synchronize()
    assert torch.equal(acc, 2.0 * v1)


@pytest.mark.parametrize("cp", [False, True])
@pytest.mark.parametrize("mode", ["f64", "mixed"])
def test_hex64_pcg(hex64, mode, cp):
    """K du = -R solved to 1e-10: the true residual (independent SpMV) meets the tolerance, Dirichlet
    increments stay zero, and the FP64 and mixed-precision solutions agree; cp: the node-blocked SpMV
    (FEMHOT_PCG_CONTEXT_PATTERN) against the row-wise one."""
    rp, ci, vals, R = hex64["rp"], hex64["ci"], hex64["vals"], hex64["R"]
    b = -R
    x, it, rr, ok = hex64["ctx"].pcg(rp, ci, vals, b, tol=1e-10, maxiter=20000, mode=mode, context_pattern=cp)
    assert ok and rr <= 1e-10
    true = float(torch.linalg.norm(b - csr_matvec(rp, ci, vals, x)) / torch.linalg.norm(b))
    assert true <= 2e-10, (mode, it, true)
    assert float(x[torch.as_tensor(hex64["didx"], device="cuda")].abs().max()) == 0.0
    if mode == "f64" and not cp:
        hex64["x64"] = x
    else:
        assert "x64" in hex64
        assert trel(x, hex64["x64"]) <= 1e-4   # both within cond(K) * 1e-10 of the exact solution


# ----------------------------------------------------------------------------------------------
# quad128: residual-norm loss and its exact gradient
# ----------------------------------------------------------------------------------------------
def _quad128_inputs(B=64):
    from folax_amd.mesh import create_2D_square_mesh
    m = create_2D_square_mesh(1.0, 129)
    coords = m.GetNodesCoordinates()
    f = np.arange(1, 5)
    K0 = np.stack([fe.fourier_field(coords, np.random.default_rng(3000 + b).standard_normal(16), f, f)
                   for b in range(B)])
    T = np.stack([np.random.default_rng(3100 + b).uniform(0.0, 1.0, size=coords.shape[0]) for b in range(B)])
    return m, K0, T


@pytest.fixture(scope="module")
def quad128():
    from folax_amd import loss_functions as lf
    m, K0, T = _quad128_inputs()
    bc = {"T": {"left": 1.0, "right": 0.0}}
    loss = lf.ThermalLoss2DQuad("thermal_128", {"dirichlet_bc_dict": bc, "beta": 0.5, "c": 2}, m)
    loss.Initialize()
    out = dict(m=m, K0=K0, T=T, loss=loss, bc=bc)
    for dt in (torch.float64, torch.float32):
        P = torch.as_tensor(K0, dtype=dt, device="cuda").requires_grad_(True)
        Tt = torch.as_tensor(T, dtype=dt, device="cuda").requires_grad_(True)
        L = loss.ComputeBatchResidualNormLoss(P, Tt)
        L.backward()
        out[dt] = (float(L), Tt.grad.detach().double(), P.grad.detach().double())
    return out


def test_quad128_fp32_matches_fp64(quad128):
    L64, gT64, gk64 = quad128[torch.float64]
    L32, gT32, gk32 = quad128[torch.float32]
    assert abs(L32 - L64) <= 1e-5 * abs(L64)
    assert trel(gT32, gT64) <= 1e-5
    assert trel(gk32, gk64) <= 1e-5


def test_quad128_fp64_matches_oracle_on_samples(quad128):
    """The FP64 gradient of samples 0 and 63 against the oracle at the full 128^2 size."""
    m, K0, T = quad128["m"], quad128["K0"], quad128["T"]
    _, gT64, gk64 = quad128[torch.float64]
    coords, conn = m.GetNodesCoordinates(), m.GetElementsNodes("quad")
    ph = fe.Physics("thermal", "quad", num_gp=2, beta=0.5, c=2.0)
    idx, val = fe.dirichlet_dofs(m.node_sets, ["T"], quad128["bc"])
    for b in (0, 63):
        Tb = fe.apply_dirichlet(T[b], idx, val)
        _, gu, gd = fe.residual_norm_loss_grad(ph, coords, conn, K0[b], Tb, idx)
        assert rel(gT64[b].cpu().numpy(), gu) <= 1e-12
        assert rel(gk64[b].cpu().numpy(), gd) <= 1e-12


def test_quad128_gradient_finite_differences(quad128):
    """100 central differences (step 1e-6, seed 3200) of the FP64 loss along random directions in
    (T, k0) against the kernel gradient."""
    loss, K0, T = quad128["loss"], quad128["K0"], quad128["T"]
    _, gT, gk = quad128[torch.float64]
    P = torch.as_tensor(K0, dtype=torch.float64, device="cuda")
    Tt = torch.as_tensor(T, dtype=torch.float64, device="cuda")
    g = torch.Generator(device="cuda").manual_seed(3200)
    h = 1e-6
    worst = 0.0
    with torch.no_grad():
        for _ in range(100):
            vT = torch.randn(Tt.shape, dtype=torch.float64, device="cuda", generator=g)
            vk = torch.randn(P.shape, dtype=torch.float64, device="cuda", generator=g)
            lp = float(loss.ComputeBatchResidualNormLoss(P + h * vk, Tt + h * vT))
            lm = float(loss.ComputeBatchResidualNormLoss(P - h * vk, Tt - h * vT))
            fd = (lp - lm) / (2 * h)
            an = float((gT * vT).sum() + (gk * vk).sum())
            worst = max(worst, abs(fd - an) / abs(an))
    assert worst <= 1e-6, worst


# ----------------------------------------------------------------------------------------------
# hex48: Neo-Hookean energy through an MLP
# ----------------------------------------------------------------------------------------------
NU_48 = 5.0 / 12.0       # lambda = 5, mu = 1 for the E = 1 phase
E0_48 = 17.0 / 6.0


@pytest.fixture(scope="module")
def hex48():
    import copy
    from folax_amd import loss_functions as lf
    from folax_amd.mesh import create_3D_box_mesh
    B = 16
    m = create_3D_box_mesh(48, 48, 48)
    Nn = m.GetNumberOfNodes()
    E = np.stack([E0_48 * fe.grf_two_phase((49, 49, 49), 8.0, 4000 + b) for b in range(B)])
    bc = {"Ux": {"bottom": 0.0}, "Uy": {"bottom": 0.0}, "Uz": {"bottom": 0.0, "top": 0.01}}
    loss = lf.NeoHookeMechanicalLoss3DHexa("nh_48", {"dirichlet_bc_dict": bc, "num_gp": 2,
                                                      "material_dict": {"young_modulus": 1.0, "poisson_ratio": NU_48}}, m)
    loss.Initialize()
    torch.manual_seed(4100)
    mlp = torch.nn.Sequential(torch.nn.Linear(Nn, 256), torch.nn.Tanh(), torch.nn.Linear(256, 256), torch.nn.Tanh(),
                              torch.nn.Linear(256, 256), torch.nn.Tanh(), torch.nn.Linear(256, 3 * Nn)).cuda()
    out = dict(m=m, E=E, loss=loss)
    for dt, net in ((torch.float32, mlp), (torch.float64, copy.deepcopy(mlp).double())):
        Et = torch.as_tensor(E, dtype=dt, device="cuda").requires_grad_(True)
        U = net(Et.detach() / (10 * E0_48)) * 1e-3
        U.retain_grad()
        Lm, (lo, hi, mean) = loss.ComputeBatchLoss(Et, U)
        Lm.backward()
        pg = torch.cat([p.grad.reshape(-1).double() for p in net.parameters()])
        out[dt] = dict(L=float(Lm), U=U.detach(), gU=U.grad.detach().double(), gE=Et.grad.detach().double(), pg=pg,
                       lohi=(float(lo), float(hi)))
        net.zero_grad(set_to_none=True)
    return out


def test_hex48_fp32_matches_fp64(hex48):
    a, b = hex48[torch.float32], hex48[torch.float64]
    assert abs(a["L"] - b["L"]) <= 1e-5 * abs(b["L"])
    assert b["lohi"][0] <= b["L"] <= b["lohi"][1]
    assert trel(a["gE"], b["gE"]) <= 1e-5
    assert trel(a["gU"], b["gU"]) <= 1e-4   # the FP32 network's own rounding moves U itself
    assert trel(a["pg"], b["pg"]) <= 1e-3


def test_hex48_fp32_kernel_error_at_the_same_displacements(hex48):
    """The FP32 kernels' own error at full size: the FP32 network's displacements U32 (rounded to
    FP32 once) go through both the FP32 and the FP64 energy and energy-gradient kernels, so the
    difference is the element arithmetic alone (the network's rounding does not enter).  Loss,
    dL/dU and dL/dE within 1e-5 of FP64."""
    loss = hex48["loss"]
    a = hex48[torch.float32]
    U32 = a["U"]
    out = {}
    for dt in (torch.float32, torch.float64):
        Et = torch.as_tensor(hex48["E"], dtype=torch.float32, device="cuda").to(dt).requires_grad_(True)
        Ut = U32.to(dt).detach().requires_grad_(True)
        Lm, _ = loss.ComputeBatchLoss(Et, Ut)
        Lm.backward()
        out[dt] = (float(Lm), Ut.grad.double(), Et.grad.double())
    L32, gU32, gE32 = out[torch.float32]
    L64, gU64, gE64 = out[torch.float64]
    assert abs(L32 - a["L"]) <= 1e-7 * abs(a["L"])  # the fixture's own FP32 forward pass
    assert abs(L32 - L64) <= 1e-5 * abs(L64)
    assert trel(gU32, gU64) <= 1e-5
    assert trel(gE32, gE64) <= 1e-5


def test_hex48_energy_is_homogeneous_in_the_field(hex48):
    """psi scales with the interpolated field (kappa, mu proportional to N.d): E . dL/dE = L."""
    for dt, tol in ((torch.float64, 1e-12), (torch.float32, 1e-5)):
        r = hex48[dt]
        Et = torch.as_tensor(hex48["E"], dtype=torch.float64, device="cuda")
        assert abs(float((Et * r["gE"]).sum()) - r["L"]) <= tol * abs(r["L"])


def test_hex48_displacement_gradient_finite_differences(hex48):
    """FP64 central differences of the batch energy along random displacement directions (free dofs)
    against dL/dU from the energy-gradient kernel."""
    loss = hex48["loss"]
    r = hex48[torch.float64]
    U = r["U"]
    Et = torch.as_tensor(hex48["E"], dtype=torch.float64, device="cuda")
    didx = torch.as_tensor(loss.dirichlet_indices, device="cuda")
    g = torch.Generator(device="cuda").manual_seed(4200)
    h = 1e-3
    with torch.no_grad():
        for _ in range(6):
            v = torch.randn(U.shape, dtype=torch.float64, device="cuda", generator=g) * 1e-3
            v[:, didx] = 0.0
            lp, _ = loss.ComputeBatchLoss(Et, U + h * v)
            lm, _ = loss.ComputeBatchLoss(Et, U - h * v)
            fd = (float(lp) - float(lm)) / (2 * h)
            an = float((r["gU"] * v).sum())
            assert abs(fd - an) <= 1e-5 * abs(an), (fd, an)


# ----------------------------------------------------------------------------------------------
# tet4: the uneven batch of tetrahedral meshes on the chunk path
# ----------------------------------------------------------------------------------------------
@pytest.fixture(scope="module")
def tet4():
    import bench
    from folax_amd import _lib as L
    from folax_amd.context import FemhotContext
    W = bench.build_workload()
    ctx = FemhotContext("tetra", L.MECHANICAL, W["coords"], W["conn"], young_modulus=1.0, poisson_ratio=0.3,
                        segment_offsets=W["eoff"])
    out = dict(W=W, ctx=ctx)
    for dt in (torch.float64, torch.float32):
        E = torch.as_tensor(W["E"], dtype=dt, device="cuda").view(1, -1)
        u = torch.as_tensor(W["u"], dtype=dt, device="cuda").view(1, -1)
        out[dt] = ctx.energy_grad(E, u, flags=L.ELEMENT_FIELD | L.SCATTER_CHUNK)
        out[(dt, "plain")] = ctx.energy_grad(E, u, flags=L.ELEMENT_FIELD)
    torch.cuda.synchronize()
    yield out
    ctx.close()


def test_tet4_matches_c_restatement(tet4):
    """Both FP64 paths against the C/OpenMP restatement of the oracle on all 4.4 M tets."""
    from oracle import cpuref
    W = tet4["W"]
    nm = len(W["ns"])
    seg = np.repeat(np.arange(nm), np.diff(W["eoff"])).astype(np.int64)
    ref = (np.zeros(W["u"].shape[0]), np.zeros(W["conn"].shape[0]), np.zeros(nm))
    cpuref.tet4_elastic_energy_grad(W["coords"], W["conn"], W["E"], W["u"], seg_of_elem=seg, n_seg=nm, out=ref)
    for key in (torch.float64, (torch.float64, "plain")):
        en, gu, gd = tet4[key]
        assert rel(gu.cpu().numpy(), ref[0]) <= 1e-12
        assert rel(gd.cpu().numpy(), ref[1]) <= 1e-12
        assert rel(en.cpu().numpy(), ref[2]) <= 1e-12
    en, gu, gd = tet4[torch.float32]
    assert rel(gu.cpu().numpy(), ref[0]) <= 1e-5
    assert rel(gd.cpu().numpy(), ref[1]) <= 1e-5
    assert rel(en.cpu().numpy(), ref[2]) <= 1e-5


def test_tet4_euler_identities(tet4):
    """Pi = 1/2 u^T K(E) u per mesh is homogeneous of degree 2 in u and 1 in E: per segment,
    u . g_u = 2 Pi and E . g_E = Pi."""
    W = tet4["W"]
    nm = len(W["ns"])
    noff, eoff = W["noff"], W["eoff"]
    for dt, tol in ((torch.float64, 1e-12), (torch.float32, 1e-5)):
        en, gu, gd = tet4[dt]
        en, gu, gd = en.reshape(-1).cpu().numpy(), gu.reshape(-1).double().cpu().numpy(), gd.reshape(-1).double().cpu().numpy()
        for s in range(nm):
            us = W["u"][3 * noff[s]:3 * noff[s + 1]]
            Es = W["E"][eoff[s]:eoff[s + 1]]
            assert abs(us @ gu[3 * noff[s]:3 * noff[s + 1]] - 2 * en[s]) <= tol * abs(en[s]), (dt, s)
            assert abs(Es @ gd[eoff[s]:eoff[s + 1]] - en[s]) <= tol * abs(en[s]), (dt, s)


def test_tet4_chunk_is_deterministic(tet4):
    """The chunk path has no atomics: a second call returns the same bits."""
    from folax_amd import _lib as L
    W, ctx = tet4["W"], tet4["ctx"]
    E = torch.as_tensor(W["E"], dtype=torch.float32, device="cuda").view(1, -1)
    u = torch.as_tensor(W["u"], dtype=torch.float32, device="cuda").view(1, -1)
    en, gu, gd = ctx.energy_grad(E, u, flags=L.ELEMENT_FIELD | L.SCATTER_CHUNK)
    en0, gu0, gd0 = tet4[torch.float32]
    assert torch.equal(gu, gu0) and torch.equal(gd, gd0) and torch.equal(en, en0)
