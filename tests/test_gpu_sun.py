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
