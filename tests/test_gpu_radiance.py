"""radiance_equation::calculate on the GPU (ffk_radiance) against the oracle's
restatement of thermal_equation/radiance_cpu.cpp on the identical Philox rays:
emission bit for bit, absorption equal to the double-precision sum of the same
float ray powers (the GPU sums them exactly in 64-bit fixed point)."""
import numpy as np
import pytest

from tests.conftest import bits, golden

pytestmark = pytest.mark.gpu


def _partial_meshes():
    from form_factors_amd import scenes

    faces = scenes.icosphere(1, inward=True)
    # faces 40..79 belong to no mesh: they emit nothing and absorb nothing
    meshes = np.array([[i, 2, 0] for i in range(0, 40, 2)], np.int32)[::-1].copy()
    return faces, meshes


def _cases():
    from form_factors_amd import scenes

    rng = np.random.default_rng(11)
    g = golden("radiance_multi.npz")
    out = [("multi_out_of_order", g["faces"], g["meshes"], g["eps"], g["T"], 60_000)]
    faces, _ = scenes.ff_box()
    m = scenes.per_face_meshes(12)
    out.append(("box", faces, m, rng.uniform(0.1, 1, (12, 2)).astype(np.float32),
                rng.uniform(250, 400, 12).astype(np.float32), 100_000))
    faces, meshes = _partial_meshes()
    out.append(("partial_meshes", faces, meshes, rng.uniform(0.1, 1, (len(meshes), 2)).astype(np.float32),
                rng.uniform(250, 400, len(meshes)).astype(np.float32), 50_000))
    faces, meshes = scenes.confetti_scene(300, seed=5)
    out.append(("confetti", faces, meshes, rng.uniform(0.1, 1, (len(meshes), 2)).astype(np.float32),
                rng.uniform(250, 400, len(meshes)).astype(np.float32), 200_000))
    # about 86 rays per face with the mesh list reversed: a block's rays span
    # several faces, and their walk entries (mesh order) run against face order
    meshes = np.ascontiguousarray(meshes[::-1])
    out.append(("confetti_reversed", faces, meshes, rng.uniform(0.1, 1, (len(meshes), 2)).astype(np.float32),
                rng.uniform(250, 400, len(meshes)).astype(np.float32), 200_000))
    # 20 K meshes next to 2000 K ones: ray powers (proportional to T^3, since
    # the ray counts follow T eps A) span six decades in one fixed-point scale
    faces, meshes = scenes.confetti_scene(300, seed=7)
    T = np.where(rng.random(len(meshes)) < 0.5, 20.0, 2000.0).astype(np.float32)
    out.append(("wide_T_20K_2000K", faces, meshes, rng.uniform(0.1, 1, (len(meshes), 2)).astype(np.float32), T,
                200_000))
    # meshes that share faces (the reference walks a shared face once per
    # mesh, each time with that mesh's temperature and emissivities; a hit on
    # it counts for the last mesh): more walk entries than faces
    faces, meshes = scenes.confetti_scene(200, seed=9)
    extra = np.array([[f0, n, 0] for f0, n in ((0, 40), (30, 60), (100, 7), (5, 200), (150, 90))], np.int32)
    meshes = np.concatenate([meshes, extra]).astype(np.int32)
    out.append(("overlapping_meshes", faces, meshes, rng.uniform(0.1, 1, (len(meshes), 2)).astype(np.float32),
                rng.uniform(250, 400, len(meshes)).astype(np.float32), 150_000))
    return out


# Per-mesh precision of the fixed-point sums: ray powers are grouped in bands
# of 2^8 below the largest one, each band with its own scale, so every ray is
# rounded by at most max(rays * 2^-55, 2^-52) of its power and so is each
# mesh's sum, however widely the powers spread (200,000 rays: 6e-12).
def rel_tol(n_rays):
    return n_rays * 2.0 ** -55 + 2.0 ** -50


@pytest.mark.parametrize("frame", [0, 1])
@pytest.mark.parametrize("case", range(7))
def test_radiance_matches_oracle(gpu, oracle, case, frame):
    name, faces, meshes, eps, T, n_rays = _cases()[case]
    print(f"radiance case {name} frame {frame}", flush=True)
    sc = gpu.Scene(faces, meshes)
    sc.build_grid()
    seed = 0xA11CE + case
    em, ab = sc.radiance_step(eps, T, n_rays, seed, frame=frame)
    em_o, ab_o, ab64 = oracle.radiance(faces, meshes, eps, T, n_rays, rng=oracle.RNG_PHILOX, seed=seed,
                                       frame=frame, caster=0, n_threads=8)
    assert np.array_equal(bits(em), bits(em_o))
    assert ab64.sum() > 0
    assert np.abs(ab - ab64).max() <= 1e-12 * ab64.max()
    # per mesh, relative to that mesh's own absorption (zero stays zero)
    rel = np.abs(ab - ab64) / np.where(ab64 > 0, ab64, 1)
    assert np.all(np.abs(ab - ab64) <= rel_tol(n_rays) * ab64), rel.max()
    # the reference's float accumulation in ray order differs only by rounding
    assert np.allclose(ab, ab_o, rtol=2e-4, atol=1e-6 * ab64.max())


def test_radiance_homogeneous_and_repeatable(gpu):
    from form_factors_amd import scenes

    faces = scenes.icosphere(2, inward=True)
    meshes = scenes.per_face_meshes(len(faces))
    eps = np.full((len(meshes), 2), 0.6, np.float32)
    T = np.full(len(meshes), 300.0, np.float32)
    sc = gpu.Scene(faces, meshes)
    sc.build_grid()
    em1, ab1 = sc.radiance_step(eps, T, 300_000, 7)
    em1b, ab1b = sc.radiance_step(eps, T, 300_000, 7)
    em2, ab2 = sc.radiance_step(eps, 2 * T, 300_000, 7)
    assert np.array_equal(ab1, ab1b)  # deterministic
    assert np.array_equal(ab2, 16 * ab1) and np.array_equal(em2, 16 * em1)
    # closed sphere with inward normals: every front ray is absorbed; the rear
    # rays leave through the outside, so half of the emission comes back
    assert abs(ab1.sum() / em1.astype(np.float64).sum() - 0.5) < 1e-3


@pytest.mark.parametrize("case", [0, 3, 4, 6])
def test_radiance_schedules_agree(gpu, case):
    """Per-ray casting, lane refill and pooled tests deposit the same fixed
    point sums: identical absorption bits in every setting."""
    name, faces, meshes, eps, T, n_rays = _cases()[case]
    sc = gpu.Scene(faces, meshes)
    sc.build_grid()
    out = []
    for refill, pool, walk in [(0, 0, 0), (16, 0, 0), (16, 4, 0), (1, 1, 0), (64, 64, 0), (16, 4, 1), (16, 4, 2),
                              (32, 2, 3)]:
        print(f"radiance {name} refill={refill} pool={pool} walk={walk}", flush=True)
        sc.set_refill(refill, pool)
        sc.set_walk(walk)
        em, ab = sc.radiance_step(eps, T, n_rays, 99)
        out.append(ab)
    for ab in out[1:]:
        assert np.array_equal(bits(ab), bits(out[0]))
    sc.close()


def test_radiance_errors(gpu):
    from form_factors_amd import ffk, scenes

    faces, meshes = scenes.ff_parallel_planes(1.0)
    sc = gpu.Scene(faces, meshes)
    with pytest.raises(ffk.FFKError):  # no emission layout / grid yet
        sc.radiance(1, np.zeros(2 * len(faces), np.float32))
    sc.build_grid()
    sc.set_emission(n_rays=1000)
    bad = np.full(2 * len(faces), -1.0, np.float32)
    with pytest.raises(ffk.FFKError):
        sc.radiance(1, bad)
    zero = sc.radiance(1, np.zeros(2 * len(faces), np.float32))
    assert (zero == 0).all()


def test_radiance_step_equals_explicit_calls(gpu):
    """ffk_radiance_step (cached areas, layout set again only when the counts
    change) equals the explicit weights -> counts -> layout -> powers ->
    device pass sequence bit for bit over several steps, including a repeated
    temperature (layout kept) and a frame change; a fused form-factor launch
    afterwards (its tile list rebuilt for the last layout) equals one on a
    fresh scene with that layout."""
    import torch

    from form_factors_amd import ffk, scenes

    faces, meshes = scenes.confetti_scene(300, seed=7)
    rng = np.random.default_rng(3)
    eps = rng.uniform(0.1, 1, (len(meshes), 2)).astype(np.float32)
    Ts = [rng.uniform(250, 400, len(meshes)).astype(np.float32) for _ in range(3)]
    steps = [(Ts[0], 0), (Ts[1], 0), (Ts[1], 0), (Ts[2], 1), (Ts[0], 1)]
    a = gpu.Scene(faces, meshes)
    a.build_grid()
    b = gpu.Scene(faces, meshes)
    b.build_grid()
    for k, (T, frame) in enumerate(steps):
        print(f"radiance step {k} frame {frame}", flush=True)
        em_a, ab_a = a.radiance_step(eps, T, 120_000, 100 + k, frame=frame)
        w, total = ffk.radiance_weights(faces, meshes, eps, T)
        front, rear = ffk.emitted_counts(w, 120_000, total)
        b.set_emission(front=front, rear=rear, frame=frame)
        p, em_b = ffk.radiance_powers(faces, meshes, eps, T, front, rear)
        ab_b = b.radiance(100 + k, p)
        assert np.array_equal(bits(em_a), bits(em_b)), k
        assert np.array_equal(ab_a.view(np.uint64), ab_b.view(np.uint64)), k
        assert np.array_equal(a.front, front) and np.array_equal(a.rear, rear) and a.frame == frame
        assert a.n_rays == int(front.sum() + rear.sum())
    N = len(meshes)
    Fa = torch.empty((N, N), dtype=torch.float32, device="cuda")
    Fb = torch.empty((N, N), dtype=torch.float32, device="cuda")
    a.ff_fused(9, F=Fa)
    c = gpu.Scene(faces, meshes)
    c.set_emission(front=a.front, rear=a.rear, frame=a.frame)
    c.build_grid()
    c.ff_fused(9, F=Fb)
    torch.cuda.synchronize()
    assert torch.equal(Fa.view(torch.int32), Fb.view(torch.int32))
    for s in (a, b, c):
        s.close()


def test_radiance_large_walk_table(gpu, oracle):
    """More than 65,536 walk entries (67,600 faces in 33,800 meshes): the walk
    positions come from the multi-block scan instead of the one-block scan;
    emission bit for bit and absorption against the double-precision sum, as
    in test_radiance_matches_oracle."""
    from form_factors_amd import scenes

    faces, meshes = scenes.two_plates_thermal(cells=130, gap=5.0, size=10.0)
    assert int(meshes[:, 1].sum()) > 65536
    rng = np.random.default_rng(21)
    eps = rng.uniform(0.1, 1, (len(meshes), 2)).astype(np.float32)
    T = rng.uniform(250, 400, len(meshes)).astype(np.float32)
    n_rays = 60_000
    sc = gpu.Scene(faces, meshes)
    sc.build_grid()
    em, ab = sc.radiance_step(eps, T, n_rays, 0xB16, frame=1)
    em_o, ab_o, ab64 = oracle.radiance(faces, meshes, eps, T, n_rays, rng=oracle.RNG_PHILOX, seed=0xB16, frame=1,
                                       caster=1, n_threads=8)
    assert np.array_equal(bits(em), bits(em_o))
    assert ab64.sum() > 0
    assert np.all(np.abs(ab - ab64) <= rel_tol(n_rays) * ab64)
    sc.close()
