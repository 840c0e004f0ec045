"""Host arithmetic of the radiance equation (weights, per-ray powers,
emission) against the oracle's restatement of thermal_equation/radiance_cpu.cpp,
bit for bit."""
import numpy as np
import pytest

from tests.conftest import bits, golden


@pytest.mark.parametrize("name", ["radiance.npz", "radiance_multi.npz"])
def test_radiance_weights_powers_emission(oracle, name):
    from form_factors_amd import ffk

    g = golden(name)
    w, total = ffk.radiance_weights(g["faces"], g["meshes"], g["eps"], g["T"])
    w_o, total_o = oracle.radiance_weights(g["faces"], g["meshes"], g["eps"], g["T"])
    assert np.array_equal(bits(w), bits(w_o)) and bits(np.float32(total)) == bits(np.float32(total_o))
    front, rear = ffk.emitted_counts(w, int(g["n_rays"]), total)
    p, em = ffk.radiance_powers(g["faces"], g["meshes"], g["eps"], g["T"], front, rear)
    # the reference's emission array for this scene
    assert np.array_equal(bits(em), bits(g["emission"]))
    # powers: front_emission / front_rays in float32, the reference's
    # expression order (radiance_cpu.cpp:156-178)
    sigma = np.float32(5.670400e-8)
    fa = np.ascontiguousarray(g["faces"], np.float32).reshape(-1, 9)
    for m, (f0, n, _) in enumerate(g["meshes"]):
        t = np.float32(g["T"][m])
        pd = sigma * (t * t * t * t)
        fd, rd = pd * np.float32(g["eps"][m, 0]), pd * np.float32(g["eps"][m, 1])
        for f in range(f0, f0 + n):
            a_f = np.float32(oracle.lib().or_triangle_area(fa[f].ctypes.data_as(oracle._f)))
            exp_f = (fd * a_f) / np.float32(front[f]) if front[f] else np.float32(0)
            exp_r = (rd * a_f) / np.float32(rear[f]) if rear[f] else np.float32(0)
            assert bits(p[2 * f]) == bits(exp_f) and bits(p[2 * f + 1]) == bits(exp_r), f


def test_radiance_powers_zero_for_silent_sides():
    from form_factors_amd import ffk, scenes

    faces, meshes = scenes.ff_parallel_planes(1.0)
    eps = np.array([[0.5, 0.0], [0.0, 0.0]], np.float32)
    T = np.array([300, 300], np.float32)
    w, total = ffk.radiance_weights(faces, meshes, eps, T)
    front, rear = ffk.emitted_counts(w, 1000, total)
    assert rear.sum() == 0 and front[2:].sum() == 0
    p, em = ffk.radiance_powers(faces, meshes, eps, T, front, rear)
    assert (p[1::2] == 0).all() and (p[4:] == 0).all() and em[1] == 0


def test_bsphere_host_matches_reference():
    """ffk_bsphere (host) is math::triangles_bsphere bit for bit."""
    from form_factors_amd import ffk

    g = golden("sun.npz")
    for k in range(3):
        c, r = ffk.bsphere(g[f"bs{k}_faces"])
        assert np.array_equal(bits(c), bits(g[f"bs{k}_center"])) and bits(r) == bits(g[f"bs{k}_radius"][0])
