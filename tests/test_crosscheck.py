"""The perfTest-style cross-check helpers (perfTest/perfTest.cpp:42-153)."""
import numpy as np

from form_factors_amd import crosscheck


def test_error_stats_follows_calculate_error():
    idx_ref = np.array([3, 3, -1, 5, 7, -1], np.int32)
    idx = np.array([3, 4, -1, 5, 7, 2], np.int32)
    p_ref = np.zeros((6, 3), np.float32)
    p = np.zeros((6, 3), np.float32)
    p[0] = (3, 4, 0)       # same face: distance 5
    p[1] = (100, 0, 0)     # different face: a mismatch, no distance
    p[3] = (0, 0, 0.5)     # same face: distance 0.5
    p_ref[4] = (1, 1, 1)   # same face: distance sqrt(3)
    p[4] = (2, 2, 2)
    s = crosscheck.error_stats(idx_ref, p_ref, idx, p)
    assert s["mismatch"] == 2  # rays 1 and 5 (a miss against a hit is a mismatch too)
    want = np.float32(np.float32(np.float32(5) + np.float32(0.5)) + np.sqrt(np.float32(3)))
    assert s["total_distance"] == float(want)
    assert s["max_distance"] == 5.0
    assert s["average_distance"] == float(np.float32(want) / np.float32(6))  # over all rays


def test_perftest_scene_geometry():
    faces, rays = crosscheck.perftest_scene(2000, 20.0, seed=3)
    assert faces.shape == (2000, 9) and rays.shape == (200_000, 6)
    v = faces.reshape(-1, 3, 3)
    # each triangle: three unit-sphere points around a common centre inside the ball of radius 20
    c = v.mean(1)
    assert np.linalg.norm(c, axis=1).max() < 21
    end, origin = rays[:, 3:].astype(np.float64), rays[:, :3].astype(np.float64)
    assert np.allclose(np.linalg.norm(end, axis=1), 22, rtol=1e-5)
    assert np.allclose(origin, 1.1 * end, rtol=1e-6)
    # directions uniform on the sphere: z and the azimuth uniform
    z = end[:, 2] / 22
    assert abs(z.mean()) < 0.01 and abs((z * z).mean() - 1 / 3) < 0.01
