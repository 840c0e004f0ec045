"""GPU parity of the hot path, through the C ABI: casting, emission, tally and
the fused form-factor kernel against the CPU oracle and the golden vectors.
Integer and index results are compared exactly, floats bit for bit."""
import numpy as np
import pytest

from conftest import bits, golden

pytestmark = pytest.mark.gpu


def t(a, dtype=None):
    import torch

    return torch.as_tensor(np.ascontiguousarray(a), device="cuda")


def make_scene(ffk, faces, meshes=None, n_rays=None, frame=0, dims=(0, 0, 0), front=None, rear=None):
    sc = ffk.Scene(faces, meshes)
    if n_rays is not None or front is not None:
        sc.set_emission(n_rays=n_rays, front=front, rear=rear, frame=frame)
    sc.build_grid(dims)
    return sc


# ------------------------------------------------------------- casting --

@pytest.mark.parametrize("name", ["plates", "sphere3", "confetti"])
@pytest.mark.parametrize("method", [0, 1], ids=["naive", "grid"])
def test_cast_matches_reference_golden(gpu, name, method):
    g = golden(f"cast_{name}.npz")
    print(f"cast {name} method={method}", flush=True)
    sc = make_scene(gpu, g["faces"])
    idx, pts = sc.cast(t(g["rays"]), method=method)
    idx, pts = idx.cpu().numpy(), pts.cpu().numpy()
    assert np.array_equal(idx, g["naive_idx"])
    hit = idx >= 0
    assert np.array_equal(bits(pts[hit]), bits(g["naive_pts"][hit]))


@pytest.mark.parametrize("dims", [(1, 1, 1), (3, 5, 2), (16, 16, 16), (64, 64, 64), (0, 0, 0)])
def test_grid_cast_independent_of_resolution(gpu, oracle, dims):
    from form_factors_amd import scenes

    rng = np.random.default_rng(3)
    faces = np.concatenate([scenes.confetti(600, rng), scenes.cube_shell(5)]).astype(np.float32)
    o = rng.uniform(-0.3, 1.3, (20000, 3))
    rays = np.concatenate([o, o + rng.normal(0, 1, (20000, 3))], 1).astype(np.float32)
    print(f"grid dims {dims}", flush=True)
    sc = make_scene(gpu, faces, dims=dims)
    idx, pts = sc.cast(t(rays), method=1)
    ref_i, ref_p = oracle.naive_cast(faces, rays, n_threads=8)
    idx, pts = idx.cpu().numpy(), pts.cpu().numpy()
    assert np.array_equal(idx, ref_i)
    assert np.array_equal(bits(pts[idx >= 0]), bits(ref_p[ref_i >= 0]))


@pytest.mark.parametrize("n", [1, 100, 5000, 300_000])
def test_grid_cast_pooled_equals_per_ray(gpu, oracle, n):
    """The pooled-test cast (lane refill, tests spread over the wavefront)
    returns the per-ray cast's indices and points bit for bit, for every
    setting and ray count (one partial block up to many full ones); rays that
    miss the grid box, zero directions and rays starting on faces included."""
    from form_factors_amd import scenes

    rng = np.random.default_rng(n)
    faces = np.concatenate([scenes.confetti(400, rng), scenes.cube_shell(4)]).astype(np.float32)
    o = rng.uniform(-0.5, 1.5, (n, 3))
    e = o + rng.normal(0, 1, (n, 3))
    m = min(n, 10)
    e[:m // 2] = o[:m // 2]                                   # zero direction
    o[m // 2:m] = faces[:m - m // 2, :3] * 0.5 + faces[:m - m // 2, 3:6] * 0.5  # on a face edge
    rays = np.concatenate([o, e], 1).astype(np.float32)
    sc = make_scene(gpu, faces, dims=(24, 24, 24))
    sc.set_refill(0, 0)
    ref_i, ref_p = sc.cast(t(rays), method=1)
    ref_i, ref_p = ref_i.cpu().numpy(), ref_p.cpu().numpy()
    if n <= 5000:
        want_i, _ = oracle.naive_cast(faces, rays, n_threads=8)
        assert np.array_equal(ref_i, want_i)
    for refill, pool, walk in [(16, 4, 0), (1, 1, 0), (64, 64, 0), (8, 2, 0), (32, 16, 0), (16, 4, 1), (1, 1, 1),
                              (16, 4, 2), (1, 1, 3), (64, 64, 2), (32, 2, 3)]:
        print(f"pooled cast n={n} refill={refill} pool={pool} walk={walk}", flush=True)
        sc.set_refill(refill, pool)
        sc.set_walk(walk)
        idx, pts = sc.cast(t(rays), method=1)
        idx, pts = idx.cpu().numpy(), pts.cpu().numpy()
        assert np.array_equal(idx, ref_i), (refill, pool)
        hit = idx >= 0
        assert np.array_equal(bits(pts[hit]), bits(ref_p[hit])), (refill, pool)
    sc.close()


def test_cast_reference_kats(gpu):
    from test_oracle import reference_cast_kats

    for name, faces, rays, expect in reference_cast_kats():
        for method in (0, 1):
            print(f"kat {name} method={method}", flush=True)
            sc = make_scene(gpu, faces)
            idx, pts = sc.cast(t(rays), method=method)
            idx, pts = idx.cpu().numpy(), pts.cpu().numpy()
            assert (idx >= 0).all(), name
            d = pts - expect
            assert (np.einsum("ij,ij->i", d, d) < 1e-8).all(), name


def test_cast_edge_cases(gpu, oracle):
    # degenerate faces (H2 threshold), duplicate faces (ties to the lower
    # index), rays parallel to faces, zero-length and backwards rays
    s = np.float32(0.01)
    faces = np.array([[0, 0, 0, s, 0, 0, 0, s, 0],            # degenerate by the 1e-8 test
                      [0, 0, 1, 1, 0, 1, 0, 1, 1],
                      [0, 0, 1, 1, 0, 1, 0, 1, 1],            # duplicate of face 1
                      [0, 0, 2, 1, 0, 2, 1, 1, 2]], np.float32)
    rays = np.array([[0.002, 0.002, 0.5, 0.002, 0.002, -1],   # through the degenerate face only
                     [0.2, 0.2, 0, 0.2, 0.2, 0.5],            # hits the duplicate pair
                     [0.2, 0.2, 1, 0.7, 0.2, 1],              # in the plane of face 1
                     [0.2, 0.2, 3, 0.2, 0.2, 3],              # zero direction
                     [0.2, 0.2, 0.5, 0.2, 0.2, 0.4],          # away from everything
                     [5, 5, 5, 6, 6, 6]], np.float32)
    ref_i, ref_p = oracle.naive_cast(faces, rays)
    for method in (0, 1):
        sc = make_scene(gpu, faces)
        idx, pts = sc.cast(t(rays), method=method)
        idx = idx.cpu().numpy()
        assert np.array_equal(idx, ref_i), (method, idx, ref_i)
    assert ref_i[1] == 1 and ref_i[0] == -1


def test_empty_scene_is_an_error(gpu):
    with pytest.raises(gpu.FFKError):
        gpu.Scene(np.zeros((0, 9), np.float32))


# ------------------------------------------------------------- emission --

@pytest.mark.parametrize("frame", [0, 1], ids=["reference_frame", "orthonormal"])
@pytest.mark.parametrize("name", ["plates", "sphere3", "confetti"])
def test_malley_emission_bit_exact(gpu, oracle, name, frame):
    g = golden(f"cast_{name}.npz")
    faces = g["faces"]
    w, total = oracle.area_weights(faces)
    sc = make_scene(gpu, faces, n_rays=int(g["n_rays"]), frame=frame)
    rays = sc.emit_malley(seed=0x1234_5678_9ABC).cpu().numpy()
    want = oracle.malley_emit(faces, w, total, int(g["n_rays"]), rng=oracle.RNG_PHILOX, seed=0x1234_5678_9ABC,
                              frame=frame)
    assert rays.shape == want.shape
    assert np.array_equal(bits(rays), bits(want))


def test_parallel_emission_bit_exact(gpu, oracle):
    g = golden("parallel.npz")
    rays = gpu.emit_parallel(g["direction"], g["center"], g["distance"], g["radius"], 20000, seed=99).cpu().numpy()
    want = oracle.parallel_emit(g["direction"], g["center"], g["distance"], g["radius"], 20000,
                                rng=oracle.RNG_PHILOX, seed=99)
    assert np.array_equal(bits(rays), bits(want))


# ---------------------------------------------------- form factors ------

def ff_cases():
    from form_factors_amd import scenes

    return [("parallel_c1", *scenes.ff_parallel_planes(1.0), 40_000),
            ("perpendicular", *scenes.ff_perpendicular_planes(), 40_000),
            ("box", *scenes.ff_box(), 40_000),
            ("plates5", *scenes.parallel_plates(5), 100_000),
            ("sphere2", scenes.icosphere(2), scenes.per_face_meshes(320), 160_000),
            ("confetti", *scenes.confetti_scene(150, cells=2, seed=4), 50_000)]


@pytest.mark.parametrize("case", ff_cases(), ids=lambda c: c[0])
@pytest.mark.parametrize("frame", [0, 1], ids=["reference_frame", "orthonormal"])
def test_fused_ff_bit_exact_vs_oracle(gpu, oracle, case, frame):
    name, faces, meshes, n_rays = case
    print(f"ff {name} frame={frame}", flush=True)
    sc = make_scene(gpu, faces, meshes, n_rays=n_rays, frame=frame)
    F = sc.ff_fused(seed=42).cpu().numpy()
    r, want = oracle.ff_calculate(faces, meshes, n_rays, rng=oracle.RNG_PHILOX, seed=42, frame=frame, n_threads=8)
    assert r == 0
    assert np.array_equal(bits(F), bits(want))


@pytest.mark.parametrize("hist", [0, 1, 2])
@pytest.mark.parametrize("rays_per_tile", [0, 1000, 4096])
def test_fused_ff_tiling_and_histograms(gpu, oracle, hist, rays_per_tile):
    from form_factors_amd import scenes

    faces, meshes = scenes.parallel_plates(4)
    meshes = scenes.plane_grid_meshes(4, 4)
    meshes = np.concatenate([meshes, meshes + np.array([32, 0, 0], np.int32)])
    sc = make_scene(gpu, faces, meshes, n_rays=200_000)
    sc.set_tuning(block=256, rays_per_tile=rays_per_tile, hist=hist)
    F = sc.ff_fused(seed=5).cpu().numpy()
    r, want = oracle.ff_calculate(faces, meshes, 200_000, seed=5)
    assert np.array_equal(bits(F), bits(want))


@pytest.mark.parametrize("refill,pool,walk", [(16, 0, 0), (16, 4, 0), (32, 2, 1), (32, 2, 3)])
@pytest.mark.parametrize("rays_per_tile", [1000, 4096])
@pytest.mark.parametrize("order", ["face_order", "reversed"])
def test_fused_ff_split_multi_face_meshes(gpu, oracle, refill, pool, walk, rays_per_tile, order):
    """Meshes of several faces split over several tiles: a tile starts inside
    a mesh, its lanes' face hints miss as refill hands them rays a block
    apart, and the lookup searches only the tile's own faces. With the mesh
    list reversed the tally walk attributes rays to meshes in list order
    (cpu_system.cpp:165-191), so tiles cover faces of other meshes. F is the
    oracle's bit for bit."""
    from form_factors_amd import scenes

    faces, _ = scenes.parallel_plates(6)  # 2 x 72 faces
    first = np.arange(0, len(faces), 4, dtype=np.int32)
    meshes = np.stack([first, np.full_like(first, 4), np.zeros_like(first)], 1)  # 36 meshes of 4 faces
    if order == "reversed":
        meshes = np.ascontiguousarray(meshes[::-1])
    print(f"split meshes {order} rpt={rays_per_tile} sched={refill}/{pool}/{walk}", flush=True)
    sc = make_scene(gpu, faces, meshes, n_rays=150_000)
    sc.set_tuning(block=256, rays_per_tile=rays_per_tile, hist=2)
    sc.set_refill(refill, pool)
    sc.set_walk(walk)
    F = sc.ff_fused(seed=23).cpu().numpy()
    r, want = oracle.ff_calculate(faces, meshes, 150_000, seed=23)
    assert r == 0
    assert np.array_equal(bits(F), bits(want))


@pytest.mark.parametrize("refill,pool,walk", [(1, 0, 0), (16, 0, 0), (48, 0, 0), (64, 0, 0), (1, 1, 0), (16, 4, 0),
                                              (16, 1, 0), (64, 64, 0), (8, 2, 0), (16, 4, 1), (1, 1, 1), (64, 8, 1),
                                              (16, 4, 2), (1, 1, 3), (32, 2, 3)])
@pytest.mark.parametrize("hist,rays_per_tile", [(0, 0), (2, 1000), (1, 0)])
def test_fused_ff_refill_bit_exact(gpu, oracle, refill, pool, walk, hist, rays_per_tile):
    """Lanes that take new rays as soon as enough of them are idle, with the
    triangle tests of waiting lanes pooled over the wavefront or not, cast the
    same rays: F is the oracle's bit for bit (global tally, split tiles and
    direct rows)."""
    from form_factors_amd import scenes

    faces, meshes = scenes.confetti_scene(150, cells=2, seed=13)
    meshes = np.ascontiguousarray(meshes)
    sc = make_scene(gpu, faces, meshes, n_rays=120_000, frame=1)
    sc.set_tuning(block=256, rays_per_tile=rays_per_tile, hist=hist)
    sc.set_refill(refill, pool)
    sc.set_walk(walk)
    F = sc.ff_fused(seed=17).cpu().numpy()
    r, want = oracle.ff_calculate(faces, meshes, 120_000, seed=17, frame=1, n_threads=8)
    assert np.array_equal(bits(F), bits(want))


def test_unfused_pipeline_equals_fused(gpu, oracle):
    from form_factors_amd import scenes

    faces, meshes = scenes.confetti_scene(200, cells=2, seed=9)
    sc = make_scene(gpu, faces, meshes, n_rays=80_000, frame=1)
    rays = sc.emit_malley(seed=77)
    idx, _ = sc.cast(rays, method=1, points=False)
    tally = sc.tally(idx)
    F1 = sc.normalise(tally).cpu().numpy()
    F2 = sc.ff_fused(seed=77).cpu().numpy()
    assert np.array_equal(bits(F1), bits(F2))
    # the same hits from the naive kernel
    idx2, _ = sc.cast(rays, method=0, points=False)
    assert np.array_equal(idx.cpu().numpy(), idx2.cpu().numpy())


def test_meshes_out_of_face_order_follow_the_reference_walk(gpu, oracle):
    # The reference attributes ray n to the mesh its mesh-order walk is in
    # (cpu_system.cpp:165-191), which differs from the emitting face's mesh
    # when meshes are not in face order; a face outside every mesh emits but
    # is never counted as hit.
    from form_factors_amd import scenes

    faces = scenes.plane_grid_faces(1, 1, 3, 3)
    faces = np.concatenate([faces, scenes.translate(faces, (0, 0, 0.7))]).astype(np.float32)
    meshes = np.array([[20, 4, 0], [0, 6, 0], [6, 10, 0], [16, 2, 0], [24, 11, 0]], np.int32)
    meshes[4, 1] = 12  # faces 24..35
    sc = make_scene(gpu, faces, meshes, n_rays=60_000)
    F = sc.ff_fused(seed=8).cpu().numpy()
    r, want = oracle.ff_calculate(faces, meshes, 60_000, seed=8)
    assert np.array_equal(bits(F), bits(want))


def test_ff_analytic_parallel_plates_c1(gpu, oracle):
    # Config C1 at full size: 400 triangles (one mesh each), 10k rays per
    # triangle (4e6 rays), grid 16^3; F bit for bit against the oracle fed
    # the identical rays, and the analytic plate-to-plate value.
    from form_factors_amd import scenes

    faces, meshes = scenes.parallel_plates(10)
    sc = make_scene(gpu, faces, meshes, n_rays=4_000_000, dims=(16, 16, 16))
    assert sc.n_rays == 4_000_000
    F32 = sc.ff_fused(seed=1).cpu().numpy()
    print("C1 full size: oracle on 4e6 rays", flush=True)
    r, want = oracle.ff_calculate(faces, meshes, 4_000_000, rng=oracle.RNG_PHILOX, seed=1,
                                  frame=oracle.FRAME_REFERENCE, caster=0, n_threads=16)
    assert r == 0 and np.array_equal(bits(F32), bits(want))
    F = F32.astype(np.float64)
    plate = 2 * F[:200, 200:].sum() / 200
    # analytic plate-to-plate F for unit squares at distance 1 (the
    # reference halves it for two-sided emission, test/form_factors.cpp:192-196)
    assert abs(plate - 0.19982) < 3e-3
    # symmetry of the two plates
    assert abs(2 * F[200:, :200].sum() / 200 - plate) < 3e-3


def test_tally_above_2_24_is_exact_not_saturated(gpu, oracle):
    """A cell count above 2^24: the reference's float `F += 1.f` stops at 2^24
    (the oracle's or_ff_tally restates it), while the device tallies integers
    and rounds the exact count once -- a deliberate deviation."""
    import torch

    from form_factors_amd import scenes

    faces, _ = scenes.parallel_plates(cells=1, gap=1.0)  # 4 faces
    meshes = np.array([[0, 2, 0], [2, 2, 0]], np.int32)
    n0 = (1 << 24) + 5
    front = np.array([n0, 0, 1, 0], np.int32)
    rear = np.zeros(4, np.int32)
    sc = gpu.Scene(faces, meshes)
    n = sc.set_emission(front=front, rear=rear)
    assert n == n0 + 1
    hit = torch.full((n,), 2, dtype=torch.int32, device="cuda")  # every ray of mesh 0 hits face 2 (mesh 1)
    hit[n0:] = -1
    print("tally of 2^24 + 5 hits in one cell", flush=True)
    F = sc.normalise(sc.tally(hit)).cpu().numpy()
    ratio = np.float32(1) / np.float32(n0)
    assert F[0, 1] == np.float32(np.float32(n0) * ratio)  # exact count, rounded once
    F_ref = oracle.ff_tally(meshes, 4, front, rear, hit.cpu().numpy())
    assert F_ref[0, 1] == np.float32(np.float32(1 << 24) * ratio)  # the reference's saturated sum
    assert F[0, 1] != F_ref[0, 1] and F[1].sum() == 0 and F_ref[1].sum() == 0
    sc.close()
