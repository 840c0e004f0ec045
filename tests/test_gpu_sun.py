"""Distant-source absorption on the GPU (ffk_sun) against the oracle's
restatement of thermal_equation/parallel_rays_cpu.cpp on the identical Philox
rays: the absorbed power per mesh equals the double-precision sum of the
reference's per-ray float values."""
import numpy as np
import pytest

from tests.conftest import golden

pytestmark = pytest.mark.gpu


def _cases():
    from form_factors_amd import scenes

    g = golden("sun.npz")
    out = [("confetti", g["faces"], g["meshes"], g["absorb"], float(g["power"]), g["direction"], 40_000)]
    faces = scenes.icosphere(3, inward=False)
    meshes = scenes.per_face_meshes(len(faces))
    rng = np.random.default_rng(2)
    d = np.array([0.3, -0.2, -0.93], np.float32)
    d /= np.linalg.norm(d)
    out.append(("sphere", faces, meshes, rng.uniform(0.2, 1, (len(meshes), 2)).astype(np.float32), 1361.0, d,
                100_000))
    return out


@pytest.mark.parametrize("aperture", [0, 1])
@pytest.mark.parametrize("case", range(2))
def test_sun_matches_oracle(gpu, oracle, case, aperture):
    name, faces, meshes, absorb, power, d, n = _cases()[case]
    print(f"sun case {name} aperture {aperture}", flush=True)
    sc = gpu.Scene(faces, meshes)
    sc.build_grid()
    seed = 0x5A + case
    got = sc.sun(absorb, power, d, n, seed, aperture=aperture)
    ab, ab64 = oracle.sun(faces, meshes, absorb, power, d, n, rng=oracle.RNG_PHILOX, seed=seed, aperture=aperture,
                          caster=0, n_threads=8)
    assert ab64.sum() > 0
    assert np.abs(got - ab64).max() <= 1e-12 * ab64.max()
    assert np.allclose(got, ab, rtol=1e-4, atol=1e-6 * ab64.max())


def test_sun_sphere_intercepts_its_cross_section(gpu):
    """Absorbance 1 on a closed sphere: the absorbed power is the power times
    the sphere's cross-section (to the polyhedron's area and the MC error)."""
    from form_factors_amd import scenes

    faces = scenes.icosphere(4, inward=False)
    meshes = scenes.per_face_meshes(len(faces))
    sc = gpu.Scene(faces, meshes)
    sc.build_grid()
    absorb = np.ones((len(meshes), 2), np.float32)
    d = np.array([0, 0, -1], np.float32)
    got = sc.sun(absorb, 1000.0, d, 2_000_000, 9)
    c, r0 = gpu.bsphere(faces)
    r = np.float32(r0 * np.float32(1.1))
    counts = sc.sun_counts(d, c, np.float32(np.float32(1.1) + r), r, 2_000_000, 9)
    # outward normals: hits are on front sides, except a ray that slips
    # through a shared edge (the triangle test's float boundaries) and meets
    # the inside of the far side
    assert counts[:, 1].sum() <= 1e-5 * counts.sum()
    assert abs(got.sum() / (1000.0 * np.pi) - 1) < 0.01


def test_sun_square_aperture_full_c3_scene(gpu, oracle):
    """The full confetti scene (2000 confetti + 2028-triangle cube shell of
    side 1.1) under 1e7 parallel rays through the 2 x 2 square aperture: the
    shell's projection s^2 (|dx|+|dy|+|dz|) over the aperture area 4 gives the
    hit fraction 0.4961 (binomial 3-sigma); on a 2e5-ray prefix the hits per
    mesh equal the oracle's naive cast of the identical rays."""
    from form_factors_amd import scenes

    faces, meshes = scenes.confetti_scene(2000, cells=13, seed=1)
    assert len(faces) == 4028
    sc = gpu.Scene(faces, meshes)
    sc.build_grid()
    c, _ = gpu.bsphere(faces)
    # the shell's projection (radius s sqrt(3)/2 = 0.953 around the cube centre) lies in the square
    assert np.abs(c.astype(np.float64) - 0.5).max() < 0.04
    d = np.array([0.36, -0.48, -0.8], np.float32)
    d /= np.linalg.norm(d)
    n = 10_000_000
    print("sun C3 full size, square aperture, 1e7 rays", flush=True)
    counts = sc.sun_counts(d, c, 2.1, 1.0, n, 5, aperture=1)
    p = 1.1 * 1.1 * float(np.abs(d.astype(np.float64)).sum()) / 4.0
    got = float(counts.sum()) / n
    assert abs(got - p) < 3 * np.sqrt(p * (1 - p) / n), (got, p)
    m = 200_000
    print("sun C3 prefix vs oracle naive cast", flush=True)
    pre = sc.sun_counts(d, c, 2.1, 1.0, m, 5, aperture=1).sum(1)
    rays = oracle.parallel_emit2(d, c, 2.1, 1.0, 1, m, seed=5)
    idx, _ = oracle.naive_cast(faces, rays, n_threads=16)
    assert np.array_equal(meshes[:, 0], np.arange(len(faces)))  # one mesh per face
    want = np.bincount(idx[idx >= 0], minlength=len(meshes))
    assert np.array_equal(pre.astype(np.int64), want)
