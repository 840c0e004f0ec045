"""The mixed-precision thermal run on the GPU (F in BF16 / FP16 / FP32, FP32
exchange accumulation, FP64 Adams update) against an all-FP64 CPU run from
the same F, and the per-step energy balance."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

U = np.array([[1.0, 0, 0, 0, 0], [3 / 2, -1 / 2, 0, 0, 0], [23 / 12, -4 / 3, 5 / 12, 0, 0],
              [55 / 24, -59 / 24, 37 / 24, -3 / 8, 0],
              [1901 / 720, -1387 / 360, 109 / 30, -637 / 360, 251 / 720]])


def fp64_run(F, prob, row, cols, G, T0, n_steps, dt, record=()):
    """All-FP64 CPU restatement: SB_FF (thermal_equation/form_factors.cpp:109-128),
    conduction (CSR), heat sources, Adams (thermal_solution/cpu_adams.cpp:147-182)."""
    from oracle import oracle as O

    n = len(T0)
    F = np.ascontiguousarray(F, np.float64)
    src = np.zeros(n)
    np.add.at(src, prob["src_mesh"], prob["src_power"])
    E = np.zeros((5, n))
    T = np.array(T0, np.float64)
    rows = np.repeat(np.arange(n), np.diff(row))
    out = {}
    for step in range(n_steps):
        power = 5.670400e-8 * T ** 4 * prob["area"]
        em = prob["eps_front"] * power + prob["eps_rear"] * power
        ab = F.T @ em
        flow = np.zeros(n)
        np.add.at(flow, rows, G * (T[cols] - T[rows]))
        net = ab - em + flow + src
        E[step % 5] = net / prob["C"]
        T = O.adams_combine_f64(step, dt, T, E)
        if step + 1 in record:
            out[step + 1] = T.copy()
    return T, out


def setup(gpu, cells, n_rays_per_tri, seed=3):
    import torch

    from form_factors_amd import scenes

    prob = scenes.thermal_plates(cells=cells)
    sc = gpu.Scene(prob["faces"], prob["meshes"])
    sc.set_emission(n_rays=n_rays_per_tri * len(prob["faces"]), frame=gpu.FRAME_ORTHONORMAL)
    sc.build_grid()
    n = len(prob["meshes"])
    F = sc.ff_fused(seed)
    torch.cuda.synchronize()
    row, cols, G = gpu.conduction_graph(prob["faces"], prob["meshes"], prob["conductivity"], prob["thickness"])
    prob["rays_per_mesh"] = sc.rays_per_mesh().astype(np.float64)
    return prob, F, (row, cols, G), n


def make_run(gpu, prob, F, csr, dtype):
    th = gpu.Thermal(prob["C"], prob["area"], prob["eps_front"], prob["eps_rear"])
    S, scale = gpu.pack_exchange(F, dtype, row_mul=prob["rays_per_mesh"])
    th.set_exchange(S, dtype, scale)
    th.set_conduction(*csr)
    th.set_sources(prob["src_mesh"], prob["src_power"])
    return th


def test_thermal_reduced_10k_steps_vs_fp64(gpu):
    prob, F, csr, n = setup(gpu, cells=12, n_rays_per_tri=2000)
    T0 = np.full(n, 300.0)
    n_steps, dt = 10_000, 0.1
    Fh = F.cpu().numpy()
    T_ref, _ = fp64_run(Fh, prob, *csr, T0, n_steps, dt)
    # radiation to space dominates the 1 kW source: the plates cool by ~100 K,
    # and the heated patch stays warmer than the rim
    assert T_ref.max() < 250 and T_ref.max() - T_ref.min() > 5
    errs = {}
    for dtype in (gpu.F32, gpu.BF16, gpu.F16):
        th = make_run(gpu, prob, F, csr, dtype)
        T, bal = th.run(n_steps, dt, T0=T0, balance=True)
        errs[dtype] = np.abs(T - T_ref).max() / np.abs(T_ref).max()
        # energy balance: stored power = the Adams combination of net powers
        net, stored = bal[:, 0], bal[:, 1]
        for s in range(n_steps):
            k = min(s, 4)
            expect = sum(U[k, l] * net[s - l] for l in range(k + 1))
            assert abs(stored[s] - expect) <= 1e-9 * (np.abs(net).max() + 1e-300), (dtype, s)
        th.close()
    print("max relative T error vs FP64:", errs)
    # count-valued rows are exact in the packed BF16/FP16 storage
    assert errs[gpu.F32] < 1e-6
    assert errs[gpu.BF16] <= 1e-6 and errs[gpu.F16] <= 1e-6


def test_pack_exchange_rows(gpu):
    """Count rows become exact counts with scale 1/n; other rows keep mul 1
    (BF16) or a power of two putting the row maximum in [16384, 32768) (FP16);
    the scaled product equals F^T E."""
    import torch

    rng = np.random.default_rng(5)
    n = 96
    n_rays = rng.integers(50, 5000, n).astype(np.float32)
    counts = rng.poisson(1.5, (n, n)).astype(np.float32)
    counts[np.arange(n), rng.integers(0, n, n)] = 1.0  # every row has a count-1 cell
    counts[:, 0] = np.minimum(counts[:, 0], 200)
    F = (counts * (np.float32(1.0) / n_rays[:, None])).astype(np.float32)  # the reference's normalisation
    F[7] = rng.random(n).astype(np.float32) * 1e-3  # a non-count row
    F[9] = 0.0  # an empty row
    Fd = torch.from_numpy(F).cuda()
    S, scale = gpu.pack_exchange(Fd, gpu.BF16)
    S = S.float().cpu().numpy().T  # stored transposed: S[j, m]
    scale = scale.cpu().numpy()
    rows = [m for m in range(n) if m not in (7, 9)]
    assert np.array_equal(S[rows], counts[rows])
    assert np.allclose(scale[rows], 1.0 / n_rays[rows], rtol=1e-6)
    assert scale[7] == 1.0 and scale[9] == 1.0
    S16, scale16 = gpu.pack_exchange(Fd, gpu.F16)
    S16 = S16.float().cpu().numpy().T
    scale16 = scale16.cpu().numpy()
    assert np.array_equal(S16[rows], counts[rows])
    mul7 = 1.0 / scale16[7]
    assert mul7 == 2.0 ** np.round(np.log2(mul7)) and 16384 <= F[7].max() * mul7 < 32768 + 1e-3
    # with explicit multipliers
    S2, sc2 = gpu.pack_exchange(Fd, gpu.BF16, row_mul=n_rays)
    assert np.array_equal(S2.float().cpu().numpy().T[rows], counts[rows])
    nz = [m for m in range(n) if m != 9]  # an all-zero row keeps scale 1
    assert np.allclose(sc2.cpu().numpy()[nz], 1.0 / n_rays[nz]) and sc2.cpu().numpy()[9] == 1.0
    # the product through ffk_sbff
    T = torch.from_numpy(rng.uniform(250, 350, n)).cuda()
    A = torch.from_numpy(rng.uniform(0.5, 2.0, n)).cuda()
    ef = torch.full((n,), 0.8, dtype=torch.float64, device="cuda")
    er = torch.full((n,), 0.3, dtype=torch.float64, device="cuda")
    Sd, scale_d = gpu.pack_exchange(Fd, gpu.BF16)
    em = torch.zeros(n, dtype=torch.float64, device="cuda")
    ab = torch.zeros(n, dtype=torch.float64, device="cuda")
    p = gpu._ptr
    work = torch.empty(n, dtype=torch.float32, device="cuda")
    gpu.check(gpu.lib().ffk_sbff(n, p(Sd), gpu.BF16, p(scale_d), p(T), p(A), p(ef), p(er), p(em), p(ab), p(work),
                                 gpu._stream()), "ffk_sbff")
    torch.cuda.synchronize()
    E = (0.8 + 0.3) * 5.670400e-8 * T.cpu().numpy() ** 4 * A.cpu().numpy()
    ref = F.astype(np.float64).T @ E
    F_rt = F.copy()
    F_rt[7] = S[7].astype(np.float64)  # BF16-rounded values of the non-count row
    ref_rt = F_rt.astype(np.float64).T @ E
    got = ab.cpu().numpy()
    assert np.allclose(em.cpu().numpy(), E, rtol=1e-12)
    assert np.abs(got - ref_rt).max() <= 2e-6 * np.abs(ref).max()


@pytest.mark.parametrize("name", ["plates", "sphere"])
def test_sbff_matches_reference_equation(gpu, name):
    """ffk_sbff on the reference's own F against the emission and absorption
    of its form_factors_equation (thermal_equation/form_factors.cpp:109-128).
    With the rays per mesh as row multipliers the storage holds the hit counts:
    exact in FP32 and FP16 (counts <= 2048) and in BF16 up to 256. Without
    them, rows whose counts are not all multiples of the smallest one are
    stored as rounded F: 2^-11 (FP16) and 2^-8 (BF16) relative."""
    import torch

    from tests.conftest import golden

    g = golden("sbff.npz")
    F = np.ascontiguousarray(g[f"{name}_F"])
    n = len(F)
    em_ref = g[f"{name}_emission"].astype(np.float64)
    ab_ref = g[f"{name}_absorption"].astype(np.float64)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).cuda()  # noqa: E731
    T, A = dev(g[f"{name}_T"]), dev(g[f"{name}_area"])
    ef, er = dev(g[f"{name}_eps"][:, 0]), dev(g[f"{name}_eps"][:, 1])
    Fd = torch.from_numpy(F).cuda()
    p = gpu._ptr
    faces, meshes = g[f"{name}_faces"], g[f"{name}_meshes"]
    w, total = gpu.area_weights(faces)
    fr, rr = gpu.emitted_counts(w, int(g[f"{name}_n_rays"]), total)
    per_face = fr.astype(np.int64) + rr
    rays = np.array([per_face[a:a + c].sum() for a, c, _ in meshes], np.float64)
    counts = F.astype(np.float64) * rays[:, None]
    assert np.abs(counts - np.round(counts)).max() < 1e-3  # F rows are counts / rays of the mesh
    bf16_exact = counts.max() <= 256
    cases = [(gpu.F32, rays, 2e-6), (gpu.F16, rays, 2e-6), (gpu.BF16, rays, 2e-6 if bf16_exact else 2.0 ** -8),
             (gpu.F32, None, 2e-6), (gpu.F16, None, 2.0 ** -11), (gpu.BF16, None, 2.0 ** -8)]
    for dtype, row_mul, tol in cases:
        print(f"sbff {name} dtype {dtype} row_mul {row_mul is not None}", flush=True)
        S, scale = gpu.pack_exchange(Fd, dtype, row_mul=row_mul)
        em = torch.zeros(n, dtype=torch.float64, device="cuda")
        ab = torch.zeros(n, dtype=torch.float64, device="cuda")
        work = torch.empty(n, dtype=torch.float32, device="cuda")
        gpu.check(gpu.lib().ffk_sbff(n, p(S), dtype, p(scale), p(T), p(A), p(ef), p(er), p(em), p(ab), p(work),
                                     gpu._stream()), "ffk_sbff")
        torch.cuda.synchronize()
        em, ab = em.cpu().numpy(), ab.cpu().numpy()
        assert np.all(np.abs(em - em_ref) <= 1e-6 * em_ref), dtype
        assert np.all(np.abs(ab - ab_ref) <= tol * ab_ref), (dtype, row_mul is not None, np.abs(ab / ab_ref - 1).max())


def test_plain_bf16_rounding_is_coherent_and_packing_removes_it(gpu):
    """Rounding count-valued F entries to BF16 gives every cell of equal count
    the same relative error; the packed storage has none."""
    import torch

    n, n_rays = 64, 1000
    rng = np.random.default_rng(2)
    counts = rng.poisson(2.0, (n, n)).astype(np.float32)
    counts[:, 3] = 1.0
    F = (counts * (np.float32(1.0) / np.float32(n_rays))).astype(np.float32)
    Fd = torch.from_numpy(F).cuda()
    plain = gpu.convert_f(Fd, gpu.BF16).float().cpu().numpy().astype(np.float64)
    rel = (plain - F) / np.where(F > 0, F, 1)
    ones = counts == 1
    assert np.ptp(rel[ones]) == 0 and abs(rel[ones][0]) > 1e-4  # one common error for all count-1 cells
    S, scale = gpu.pack_exchange(Fd, gpu.BF16)
    back = S.float().cpu().numpy().T.astype(np.float64) * scale.cpu().numpy()[:, None]
    assert np.abs(back - F).max() <= 3e-7 * F.max()


def test_thermal_graph_replay_equals_eager(gpu):
    prob, F, csr, n = setup(gpu, cells=6, n_rays_per_tri=500)
    T0 = np.linspace(280, 320, n)
    a = make_run(gpu, prob, F, csr, gpu.BF16)
    b = make_run(gpu, prob, F, csr, gpu.BF16)
    Ta = a.run(55, 0.1, T0=T0, use_graph=True)
    Tb = b.run(55, 0.1, T0=T0, use_graph=False)
    assert np.array_equal(Ta, Tb)
    # continuing a run equals one longer run
    Ta2 = a.run(45, 0.1, use_graph=True)
    Tc = b.run(45, 0.1, use_graph=False)
    assert np.array_equal(Ta2, Tc)
    a.close()
    b.close()


@pytest.mark.parametrize("unroll,max_blocks", [(2, 0), (8, 0), (4, 3)])
def test_thermal_launch_shapes_agree(gpu, unroll, max_blocks):
    """Every launch shape computes each row in the same order: identical T."""
    prob, F, csr, n = setup(gpu, cells=8, n_rays_per_tri=500)
    T0 = np.linspace(280, 320, n)
    a = make_run(gpu, prob, F, csr, gpu.F16)
    b = make_run(gpu, prob, F, csr, gpu.F16)
    b.set_tuning(unroll, max_blocks)
    assert np.array_equal(a.run(30, 0.1, T0=T0), b.run(30, 0.1, T0=T0))
    a.close()
    b.close()


def test_thermal_full_size_prefix_vs_fp64(gpu):
    """Config C5 size: 2 x 100 x 100 cells (20,000 meshes, 40,000 triangles)."""
    prob, F, csr, n = setup(gpu, cells=100, n_rays_per_tri=200)
    assert n == 20_000
    T0 = np.full(n, 300.0)
    Fh = F.cpu().numpy()
    _, rec = fp64_run(Fh, prob, *csr, T0, 20, 0.1, record=(20,))
    for dtype in (gpu.BF16, gpu.F16):
        th = make_run(gpu, prob, F, csr, dtype)
        T = th.run(20, 0.1, T0=T0)
        err = np.abs(T - rec[20]).max() / rec[20].max()
        assert err <= 1e-4, (dtype, err)
        th.close()


def test_device_fp64_reference_matches_cpu_fp64(gpu):
    """The torch-FP64 device restatement used at full size equals the CPU
    FP64 run (oracle Adams combination) to rounding."""
    from tests.thermal_ref import fp64_run_device

    prob, F, csr, n = setup(gpu, cells=12, n_rays_per_tri=2000)
    T0 = np.linspace(290, 310, n)
    T_cpu, _ = fp64_run(F.cpu().numpy(), prob, *csr, T0, 300, 0.1)
    T_dev = fp64_run_device(F, prob, *csr, T0, 300, 0.1)
    assert np.abs(T_dev - T_cpu).max() <= 1e-12 * T_cpu.max()


def test_thermal_full_size_10k_steps_vs_fp64(gpu):
    """The mixed-precision workload at its stated size and length: 20,000
    meshes (two 100 x 100-cell plates), 10,000 Adams steps of 0.1 s, F stored
    in BF16 and FP16 (FP32 accumulation, FP64 update) against the all-FP64
    device run from the same F: max relative T error <= 1e-4."""
    from tests.thermal_ref import fp64_run_device

    prob, F, csr, n = setup(gpu, cells=100, n_rays_per_tri=500)
    assert n == 20_000
    T0 = np.full(n, 300.0)
    print("C5 full size: FP64 device reference, 10,000 steps", flush=True)
    T_ref = fp64_run_device(F, prob, *csr, T0, 10_000, 0.1)
    assert T_ref.max() - T_ref.min() > 1.0  # the heated patch stands out
    errs = {}
    for dtype in (gpu.BF16, gpu.F16):
        print(f"C5 full size: dtype {dtype}, 10,000 steps", flush=True)
        th = make_run(gpu, prob, F, csr, dtype)
        T = th.run(10_000, 0.1, T0=T0)
        errs[dtype] = float(np.abs(T - T_ref).max() / np.abs(T_ref).max())
        th.close()
    print("max relative T error vs FP64 at 20,000 meshes x 10,000 steps:", errs)
    assert all(e <= 1e-4 for e in errs.values()), errs


def test_thermal_counts_above_256_vs_fp64(gpu):
    """Counts above 256 per cell (200 meshes, 1e6 rays per mesh): BF16 keeps
    8 significant bits and FP16 11, so the stored counts round; the error of
    a 10,000-step run against FP64 is bounded and reported."""
    from tests.thermal_ref import fp64_run_device

    prob, F, csr, n = setup(gpu, cells=10, n_rays_per_tri=500_000)
    counts = (F.double().cpu().numpy() * prob["rays_per_mesh"][:, None])
    print(f"counts per cell up to {counts.max():.0f}", flush=True)
    assert counts.max() > 2048
    T0 = np.full(n, 300.0)
    T_ref = fp64_run_device(F, prob, *csr, T0, 10_000, 0.1)
    errs = {}
    for dtype in (gpu.F32, gpu.BF16, gpu.F16):
        th = make_run(gpu, prob, F, csr, dtype)
        T = th.run(10_000, 0.1, T0=T0)
        errs[dtype] = float(np.abs(T - T_ref).max() / np.abs(T_ref).max())
        th.close()
    print("counts > 256: max relative T error vs FP64:", errs, flush=True)
    assert errs[gpu.F32] <= 1e-7
    assert errs[gpu.F16] <= 1e-4
    assert errs[gpu.BF16] <= 1e-3
